"""ctypes wrapper of the CPU restatement of PAR_Word2Vec_cpu.cpp.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the CPU baseline.  The
product package never imports it."""
from __future__ import annotations

import ctypes as C
import subprocess
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "lib"

i32, i64, u64, f32, f64, vp, cp = C.c_int32, C.c_int64, C.c_uint64, C.c_float, C.c_double, C.c_void_p, C.c_char_p


class Params(C.Structure):
    _fields_ = [("dim", i32), ("window", i32), ("negative", i32), ("batch_size", i32), ("iter", i32),
                ("threads", i32), ("alpha", f32), ("sample", f32), ("seed", u64), ("bf16", i32), ("strong_rng", i32),
                ("alpha_per_sentence", i32), ("reserved", i32)]


class Stats(C.Structure):
    _fields_ = [("raw_words", u64), ("trained_tokens", u64), ("sentences", u64), ("pairs", u64),
                ("batches", u64), ("negatives", u64), ("rep_pairs", u64), ("loss_attr", f64), ("loss_rep", f64),
                ("draws", u64), ("final_rng", u64), ("final_alpha", f32)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class Trace(C.Structure):
    _fields_ = [("cap_tokens", i64), ("cap_batches", i64), ("cap_negs", i64), ("cap_sent", i64),
                ("tokens", vp), ("b", vp), ("sent_len", vp), ("sent_alpha", vp), ("rep_b", vp), ("negs", vp),
                ("n_tokens", i64), ("n_batches", i64), ("n_negs", i64), ("n_sent", i64), ("overflow", i32)]


class Curve(C.Structure):
    _fields_ = [("bucket_words", i64), ("n_buckets", i64), ("loss", vp), ("pairs", vp), ("loss_attr", vp),
                ("pairs_attr", vp)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


_libs: dict[str, C.CDLL] = {}


def lib(flavour: str = "parity") -> C.CDLL:
    name = "liboracle.so" if flavour == "parity" else "liboracle_fast.so"
    if name in _libs:
        return _libs[name]
    path = LIB / name
    if not path.exists():
        raise RuntimeError(f"{path} missing; build the oracle (make -C oracle)")
    L = C.CDLL(str(path))
    L.orc_lcg.restype = u64
    L.orc_lcg.argtypes = [u64]
    L.orc_set_mutation.argtypes = [i32]
    L.orc_mutation.restype = i32
    L.orc_exp_resolution.restype = f32
    L.orc_exp_table.argtypes = [vp]
    L.orc_init_net.argtypes = [i64, i32, vp, vp]
    L.orc_unigram_table_n.argtypes = [vp, i32, vp, i64]
    L.orc_unigram_table.argtypes = [vp, i32, vp]
    L.orc_vocab_learn.restype = vp
    L.orc_vocab_learn.argtypes = [cp, i32]
    L.orc_vocab_read.restype = vp
    L.orc_vocab_read.argtypes = [cp, cp, i32]
    L.orc_vocab_save.argtypes = [vp, cp]
    L.orc_vocab_size.restype = i32
    L.orc_vocab_size.argtypes = [vp]
    L.orc_vocab_train_words.restype = u64
    L.orc_vocab_train_words.argtypes = [vp]
    L.orc_vocab_count.restype = C.c_uint32
    L.orc_vocab_count.argtypes = [vp, i32]
    L.orc_vocab_word.restype = cp
    L.orc_vocab_word.argtypes = [vp, i32]
    L.orc_vocab_counts.argtypes = [vp, vp]
    L.orc_vocab_search.restype = i32
    L.orc_vocab_search.argtypes = [vp, cp]
    L.orc_vocab_free.argtypes = [vp]
    L.orc_load_stream.restype = i64
    L.orc_load_stream.argtypes = [vp, cp, u64, u64, vp]
    L.orc_train.argtypes = [vp, vp, u64, vp, i32, u64, vp, vp, vp, vp, vp, vp, vp]
    L.orc_train_hogwild.restype = f64
    L.orc_train_hogwild.argtypes = [vp, vp, u64, vp, i32, u64, vp, vp, vp, vp, vp]
    L.orc_train_hogwild_file.restype = f64
    L.orc_train_hogwild_file.argtypes = [vp, vp, cp, u64, u64, vp, vp, vp, vp, vp, vp]
    L.orc_vocab_file_size.restype = u64
    L.orc_vocab_file_size.argtypes = [vp]
    for g in ("orc_gemm_nt", "orc_gemm_nn", "orc_gemm_tn"):
        getattr(L, g).argtypes = [C.c_int, C.c_int, C.c_int, vp, C.c_int, vp, C.c_int, vp, C.c_int]
    L.orc_save_model.argtypes = [cp, vp, vp, i32, i32]
    L.orc_window_loss_f64.restype = f64
    L.orc_window_loss_f64.argtypes = [i64, i32, i32, i32, vp, vp, vp, vp, vp]
    L.orc_window_grad_f64.argtypes = [i64, i32, i32, i32, vp, vp, vp, vp, vp, vp, vp]
    _libs[name] = L
    return L


# Planted misreadings (ORC_MUT_* in par_sgns_oracle.h), for the tests that
# show each known-answer pin fails under the misreading it guards against.
MUTATIONS = {"alpha_after_load": 1, "alpha_ge": 2, "no_rejection": 3, "repb_2w": 4, "window_incl": 5,
             "pair_order": 6, "unigram_linear": 7, "subsample_flip": 8, "table_index": 9, "wca_local": 10,
             "rep_label": 11, "batch_end": 12}


class mutation:
    """Context manager: run the parity oracle with one planted misreading."""

    def __init__(self, name: str):
        self.m = MUTATIONS[name]

    def __enter__(self):
        lib().orc_set_mutation(self.m)
        return self

    def __exit__(self, *exc):
        lib().orc_set_mutation(0)


def P(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(vp)


def exp_table() -> np.ndarray:
    t = np.empty(1001, np.float32)
    lib().orc_exp_table(P(t))
    return t


def init_net(V: int, d: int) -> tuple[np.ndarray, np.ndarray]:
    a = np.empty((V, d), np.float32)
    b = np.empty((V, d), np.float32)
    lib().orc_init_net(V, d, P(a), P(b))
    return a, b


def unigram_table(counts: np.ndarray, table_size: int = 100_000_000) -> np.ndarray:
    c = np.ascontiguousarray(counts, np.uint32)
    t = np.empty(table_size, np.int32)
    lib().orc_unigram_table_n(P(c), c.size, P(t), table_size)
    return t


class OracleVocab:
    def __init__(self, h):
        if not h:
            raise RuntimeError("oracle vocabulary failed")
        self.h = h

    @classmethod
    def learn(cls, path: str, min_count: int = 5):
        return cls(lib().orc_vocab_learn(path.encode(), min_count))

    @classmethod
    def read(cls, vocab_path: str, train_path: str, min_count: int = 5, flavour: str = "parity"):
        """ReadVocab (cpu:301-335) from a "word count" file; the training
        file only gives the byte size the thread slices are cut from."""
        v = cls(lib(flavour).orc_vocab_read(vocab_path.encode(), train_path.encode(), min_count))
        v.flavour = flavour
        return v

    @property
    def file_size(self) -> int:
        return int(lib().orc_vocab_file_size(self.h))

    def __len__(self):
        return lib().orc_vocab_size(self.h)

    @property
    def train_words(self) -> int:
        return int(lib().orc_vocab_train_words(self.h))

    def counts(self) -> np.ndarray:
        out = np.empty(len(self), np.uint32)
        lib().orc_vocab_counts(self.h, P(out))
        return out

    def word(self, i: int) -> str:
        return lib().orc_vocab_word(self.h, i).decode(errors="surrogateescape")

    def stream(self, path: str) -> np.ndarray:
        cap = self.train_words
        out = np.empty(cap + 1, np.int32)
        n = lib().orc_load_stream(self.h, path.encode(), 0, cap, P(out))
        return out[:n]

    def save_model(self, path: str, w_in: np.ndarray, binary: bool) -> None:
        lib().orc_save_model(path.encode(), self.h, P(np.ascontiguousarray(w_in, np.float32)), w_in.shape[1],
                             int(binary))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_vocab_free(self.h)
            self.h = None


@dataclass
class TrainResult:
    w_in: np.ndarray
    w_out: np.ndarray
    stats: dict
    trace: dict | None = None
    curve: tuple[np.ndarray, np.ndarray] | None = None
    curve_attr: tuple[np.ndarray, np.ndarray] | None = None  # attraction pairs only


def train(stream: np.ndarray, counts: np.ndarray, train_words: int, dim: int, window: int, negative: int,
          batch_size: int, iters: int, alpha: float, sample: float, w_in: np.ndarray | None = None,
          w_out: np.ndarray | None = None, table: np.ndarray | None = None, want_trace: bool = False,
          bucket_words: int = 0, n_buckets: int = 0, seed: int = 0, bf16: bool = False,
          strong_rng: bool = False, alpha_per_sentence: bool = False) -> TrainResult:
    """Single-thread deterministic run (reference with -threads 1).  seed and
    bf16 are test-only options: another LCG seed than thread 0's, and bf16
    tables with stochastic rounding at the GPU trainer's write-back points;
    strong_rng replaces the reference's LCG by a well-mixed generator;
    alpha_per_sentence sets each sentence's rate from the words before it."""
    L = lib()
    counts = np.ascontiguousarray(counts, np.uint32)
    V = counts.size
    s = np.empty(stream.size + 1, np.int32)
    s[:-1] = stream
    s[-1] = 0
    if w_in is None:
        w_in, w_out = init_net(V, dim)
    w_in = np.ascontiguousarray(w_in, np.float32).copy()
    w_out = np.ascontiguousarray(w_out, np.float32).copy()
    if table is None:
        table = unigram_table(counts)
    et = exp_table()
    p = Params(dim, window, negative, batch_size, iters, 1, alpha, sample, seed, int(bf16), int(strong_rng),
               int(alpha_per_sentence), 0)
    if bf16:  # tables start as the bf16 values the GPU's upload rounds them to (nearest even)
        for a in (w_in, w_out):
            b = a.view(np.uint32)
            b[...] = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    st = Stats()
    tr = None
    keep = []
    if want_trace:
        cap_t = stream.size * iters + 1000
        cap_b = cap_t // max(batch_size, 1) + stream.size // 1 + 1000
        cap_n = cap_b * (2 * window) * max(negative, 1)
        arrs = dict(tokens=np.empty(cap_t, np.int32), b=np.empty(cap_t, np.uint8),
                    sent_len=np.empty(cap_t, np.int32), sent_alpha=np.empty(2 * cap_t, np.float32),
                    rep_b=np.empty(cap_b, np.int32), negs=np.empty(max(cap_n, 1), np.int32))
        keep.append(arrs)
        tr = Trace(cap_t, cap_b, cap_n, cap_t, *(P(arrs[k]) for k in
                                                  ("tokens", "b", "sent_len", "sent_alpha", "rep_b", "negs")),
                   0, 0, 0, 0, 0)
    cv = None
    if n_buckets:
        cl, cpairs = np.zeros(n_buckets), np.zeros(n_buckets)
        al, apairs = np.zeros(n_buckets), np.zeros(n_buckets)
        cv = Curve(bucket_words, n_buckets, P(cl), P(cpairs), P(al), P(apairs))
    rc = L.orc_train(C.byref(p), P(s), stream.size, P(counts), V, train_words, P(table), P(et), P(w_in),
                     P(w_out), C.byref(st), C.byref(tr) if tr is not None else None,
                     C.byref(cv) if cv is not None else None)
    if rc != 0:
        raise RuntimeError(f"orc_train failed: {rc}")
    trace = None
    if tr is not None:
        if tr.overflow:
            raise RuntimeError("trace capacity exceeded")
        a = keep[0]
        trace = dict(tokens=a["tokens"][: tr.n_tokens].copy(), b=a["b"][: tr.n_tokens].copy(),
                     sent_len=a["sent_len"][: tr.n_sent].copy(), sent_alpha=a["sent_alpha"][: 2 * tr.n_sent].copy(),
                     rep_b=a["rep_b"][: tr.n_batches].copy(), negs=a["negs"][: tr.n_negs].copy())
    curve = (cl, cpairs) if n_buckets else None
    curve_attr = (al, apairs) if n_buckets else None
    return TrainResult(w_in, w_out, st.as_dict(), trace, curve, curve_attr)


def train_hogwild(stream: np.ndarray, counts: np.ndarray, train_words: int, dim: int, window: int, negative: int,
                  batch_size: int, iters: int, alpha: float, sample: float, threads: int,
                  table: np.ndarray | None = None, flavour: str = "fast",
                  w_out_out: list | None = None) -> tuple[float, dict, np.ndarray]:
    """OpenMP Hogwild run over an in-memory stream cut into `threads` equal
    word slices; returns (training seconds, stats, W_in).  The timing flavour
    by default; the parity flavour with one thread is orc_train's schedule."""
    L = lib(flavour)
    counts = np.ascontiguousarray(counts, np.uint32)
    V = counts.size
    s = np.ascontiguousarray(stream, np.int32)
    w_in, w_out = init_net(V, dim)
    if table is None:
        table = unigram_table(counts)
    et = exp_table()
    p = Params(dim, window, negative, batch_size, iters, threads, alpha, sample, 0, 0, 0, 0, 0)
    st = Stats()
    secs = L.orc_train_hogwild(C.byref(p), P(s), s.size, P(counts), V, train_words, P(table), P(et), P(w_in),
                               P(w_out), C.byref(st))
    if w_out_out is not None:
        w_out_out.append(w_out)
    return secs, st.as_dict(), w_in


def train_hogwild_file(vocab: OracleVocab, path: str, file_words: int, train_words: int, dim: int, window: int,
                       negative: int, batch_size: int, iters: int, alpha: float, sample: float, threads: int,
                       table: np.ndarray | None = None, flavour: str = "fast") -> dict:
    """The Hogwild trainer as the reference's harness runs it (cpu:397-416):
    thread `id` seeks to file_size * id / threads in the text file `path`,
    loads up to ceil(file_words / threads) words and trains them; the time
    covers loading and training (cpu:395, 702-706).  Returns seconds,
    load_max (the longest per-thread load), stats, W_in and W_out."""
    L = lib(flavour)
    counts = vocab.counts()
    V = counts.size
    w_in, w_out = init_net(V, dim)
    if table is None:
        table = unigram_table(counts)
    et = exp_table()
    p = Params(dim, window, negative, batch_size, iters, threads, alpha, sample, 0, 0, 0, 0, 0)
    st = Stats()
    load = C.c_double(0.0)
    secs = L.orc_train_hogwild_file(C.byref(p), vocab.h, path.encode(), file_words, train_words, P(table), P(et),
                                    P(w_in), P(w_out), C.byref(st), C.byref(load))
    if secs < 0:
        raise RuntimeError(f"orc_train_hogwild_file: cannot read {path}")
    return dict(seconds=secs, load_max=load.value, stats=st.as_dict(), w_in=w_in, w_out=w_out)


def write_text_corpus(path: str, stream: np.ndarray, words: list[str]) -> int:
    """Writes vocabulary-id `stream` as text: id 0 ("</s>") ends a line,
    other ids are `words[id]` followed by a space.  Returns the byte size."""
    enc = [w.encode() for w in words]
    tok = np.array([b"\n"] + [e + b" " for e in enc[1:]], dtype=object)
    n = 0
    with open(path, "wb") as f:
        for lo in range(0, stream.size, 1 << 20):
            b = b"".join(tok[stream[lo:lo + (1 << 20)]].tolist())
            f.write(b)
            n += len(b)
    return n


def write_vocab(path: str, words: list[str], counts: np.ndarray) -> None:
    """The "word count" lines ReadVocab (cpu:301-335) reads."""
    with open(path, "w") as f:
        for w, c in zip(words, counts.tolist()):
            f.write(f"{w} {c}\n")


def window_loss_f64(centre, ctx, neg, w_in: np.ndarray, w_out: np.ndarray) -> float:
    c = np.ascontiguousarray(centre, np.int32)
    x = np.ascontiguousarray(ctx, np.int32)
    n = np.ascontiguousarray(neg, np.int32)
    a = np.ascontiguousarray(w_in, np.float64)
    b = np.ascontiguousarray(w_out, np.float64)
    nw = c.size
    return lib().orc_window_loss_f64(nw, x.size // nw, n.size // nw, a.shape[1], P(c), P(x), P(n), P(a), P(b))


def window_grad_f64(centre, ctx, neg, w_in: np.ndarray, w_out: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    c = np.ascontiguousarray(centre, np.int32)
    x = np.ascontiguousarray(ctx, np.int32)
    n = np.ascontiguousarray(neg, np.int32)
    a = np.ascontiguousarray(w_in, np.float64)
    b = np.ascontiguousarray(w_out, np.float64)
    gi = np.zeros_like(a)
    go = np.zeros_like(b)
    nw = c.size
    lib().orc_window_grad_f64(nw, x.size // nw, n.size // nw, a.shape[1], P(c), P(x), P(n), P(a), P(b), P(gi), P(go))
    return gi, go
