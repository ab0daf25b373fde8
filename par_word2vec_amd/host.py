"""Host side of the trainer: vocabulary, id streams, subsampling into the
sentence CSR, the exact random schedule of the deterministic mode, and
synthetic corpora.  Every function is a thin wrapper over libpar_host.so
(include/par_host.h); the arithmetic lives in C++."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _native as N


class HostError(RuntimeError):
    pass


def _chk(rc: int, what: str) -> int:
    if rc < 0:
        raise HostError(f"{what}: {N.host_error()} (code {rc})")
    return rc


class Vocab:
    """Vocabulary of LearnVocabFromTrainFile / SortVocab
    (PAR_Word2Vec_cpu.cpp:253-292, 204-226): id 0 is "</s>", the rest sorted
    by count descending with first-appearance tie order."""

    def __init__(self, handle, raw_vocab: int | None = None):
        if not handle:
            raise HostError(N.host_error())
        self._h = handle
        self.raw_vocab = raw_vocab

    @classmethod
    def learn_text(cls, path: str, min_count: int = 5) -> "Vocab":
        return cls(N.host().ph_vocab_learn_text(path.encode(), min_count))

    @classmethod
    def read(cls, vocab_path: str, min_count: int = 5) -> "Vocab":
        return cls(N.host().ph_vocab_read(vocab_path.encode(), min_count))

    @classmethod
    def from_raw(cls, raw: np.ndarray, raw_vocab: int, min_count: int = 5) -> "Vocab":
        raw = np.ascontiguousarray(raw, dtype=np.int32)
        return cls(N.host().ph_vocab_from_raw(N.ptr(raw), raw.size, raw_vocab, min_count), raw_vocab)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            N.host().ph_vocab_free(h)
            self._h = None

    def __len__(self) -> int:
        return N.host().ph_vocab_size(self._h)

    @property
    def train_words(self) -> int:
        return int(N.host().ph_vocab_train_words(self._h))

    def counts(self) -> np.ndarray:
        out = np.empty(len(self), dtype=np.uint32)
        _chk(N.host().ph_vocab_counts(self._h, N.ptr(out)), "counts")
        return out

    def word(self, i: int) -> str:
        w = N.host().ph_vocab_word(self._h, i)
        if w is None:
            raise IndexError(i)
        return w.decode(errors="surrogateescape")

    def lookup(self, word: str) -> int:
        return int(N.host().ph_vocab_lookup(self._h, word.encode()))

    def raw_map(self) -> np.ndarray:
        out = np.empty(self.raw_vocab + 1, dtype=np.int32)
        _chk(N.host().ph_vocab_raw_map(self._h, N.ptr(out), self.raw_vocab), "raw_map")
        return out

    def save(self, path: str) -> None:
        _chk(N.host().ph_vocab_save(self._h, path.encode()), "save vocab")

    def save_embeddings(self, path: str, w_in: np.ndarray, binary: bool = False) -> None:
        w = np.ascontiguousarray(w_in, dtype=np.float32)
        _chk(N.host().ph_save_embeddings(path.encode(), self._h, N.ptr(w), w.shape[1], int(binary)), "save")


def load_text_stream(vocab: Vocab, path: str) -> np.ndarray:
    """loadStream (cpu:359-370) over the whole file."""
    n = _chk(N.host().ph_load_text_stream(vocab._h, path.encode(), None, 0), "stream size")
    out = np.empty(n, dtype=np.int32)
    _chk(N.host().ph_load_text_stream(vocab._h, path.encode(), N.ptr(out), n), "stream")
    return out


def map_raw_stream(vocab: Vocab, raw: np.ndarray) -> np.ndarray:
    raw = np.ascontiguousarray(raw, dtype=np.int32)
    n = _chk(N.host().ph_map_raw_stream(vocab._h, N.ptr(raw), raw.size, None, 0), "map size")
    out = np.empty(n, dtype=np.int32)
    _chk(N.host().ph_map_raw_stream(vocab._h, N.ptr(raw), raw.size, N.ptr(out), n), "map")
    return out


@dataclass
class SentenceCSR:
    tokens: np.ndarray    # int32 kept tokens
    sent_off: np.ndarray  # int64 [S+1]
    sent_raw: np.ndarray  # int64 [S]: raw offset of each sentence start

    @property
    def n_sent(self) -> int:
        return self.sent_off.size - 1


def subsample(stream: np.ndarray, counts: np.ndarray, train_words: int, sample: float, seed: int,
              max_len: int = 1000) -> SentenceCSR:
    stream = np.ascontiguousarray(stream, dtype=np.int32)
    counts = np.ascontiguousarray(counts, dtype=np.uint32)
    nt = np.zeros(1, dtype=np.int64)
    lib = N.host()
    S = _chk(lib.ph_subsample(N.ptr(stream), stream.size, N.ptr(counts), counts.size, train_words, sample,
                              seed, max_len, None, 0, None, None, 0, N.ptr(nt)), "subsample count")
    tok = np.empty(int(nt[0]), dtype=np.int32)
    off = np.empty(S + 1, dtype=np.int64)
    raw = np.empty(max(S, 1), dtype=np.int64)
    S2 = _chk(lib.ph_subsample(N.ptr(stream), stream.size, N.ptr(counts), counts.size, train_words, sample,
                               seed, max_len, N.ptr(tok), tok.size, N.ptr(off), N.ptr(raw), S, N.ptr(nt)),
              "subsample")
    assert S2 == S
    return SentenceCSR(tok, off, raw[:S])


def unigram_entries(counts: np.ndarray, table_size: int = 100_000_000) -> np.ndarray:
    counts = np.ascontiguousarray(counts, dtype=np.uint32)
    out = np.empty(counts.size, dtype=np.int64)
    _chk(N.host().ph_unigram_entries(N.ptr(counts), counts.size, table_size, N.ptr(out)), "unigram")
    return out


@dataclass
class Plan:
    tokens: np.ndarray
    b: np.ndarray
    sent_off: np.ndarray
    sent_alpha: np.ndarray
    batch_off: np.ndarray
    neg_off: np.ndarray
    negs: np.ndarray
    draws: int
    final_rng: int
    final_alpha: float
    raw_words: int

    @property
    def n_sent(self) -> int:
        return self.sent_off.size - 1


def build_plan(stream: np.ndarray, counts: np.ndarray, train_words: int, window: int, negative: int,
               batch_size: int, iters: int, alpha: float, sample: float) -> Plan:
    """Exact single-thread schedule of the reference (cpu:397-627)."""
    stream = np.ascontiguousarray(stream, dtype=np.int32)
    counts = np.ascontiguousarray(counts, dtype=np.uint32)
    lib = N.host()
    p = lib.ph_plan_build(N.ptr(stream), stream.size, N.ptr(counts), counts.size, train_words, window,
                          negative, batch_size, iters, alpha, sample)
    if not p:
        raise HostError(N.host_error())
    try:
        info = N.PhPlanInfo()
        _chk(lib.ph_plan_info_get(p, N.C.byref(info)), "plan info")
        tok = np.empty(info.n_tokens, np.int32)
        b = np.empty(info.n_tokens, np.uint8)
        so = np.empty(info.n_sent + 1, np.int64)
        sa = np.empty(2 * info.n_sent, np.float32)
        bo = np.empty(info.n_sent + 1, np.int64)
        no = np.empty(info.n_batches + 1, np.int64)
        ng = np.empty(max(info.n_negs, 1), np.int32)
        _chk(lib.ph_plan_copy(p, N.ptr(tok), N.ptr(b), N.ptr(so), N.ptr(sa), N.ptr(bo), N.ptr(no), N.ptr(ng)),
             "plan copy")
        return Plan(tok, b, so, sa, bo, no, ng[: info.n_negs], int(info.draws), int(info.final_rng),
                    float(info.final_alpha), int(info.raw_words))
    finally:
        lib.ph_plan_free(p)


def exp_table() -> np.ndarray:
    out = np.empty(1001, dtype=np.float32)
    N.host().ph_exp_table(N.ptr(out))
    return out


def init_net(vocab_size: int, dim: int) -> tuple[np.ndarray, np.ndarray]:
    """InitNet (cpu:337-357)."""
    w_in = np.empty((vocab_size, dim), dtype=np.float32)
    w_out = np.empty((vocab_size, dim), dtype=np.float32)
    N.host().ph_init_net(vocab_size, dim, N.ptr(w_in), N.ptr(w_out))
    return w_in, w_out


def synth_zipf(n_tokens: int, raw_vocab: int, zipf_s: float, line_len: int = 1000,
               lognorm: tuple[float, float] | None = None, len_clip: tuple[int, int] = (5, 2000),
               n_clusters: int = 0, cluster_size: int = 0, span: int = 3, p_replace: float = 0.0,
               band: tuple[int, int] | None = None, seed: int = 1) -> tuple[np.ndarray, np.ndarray]:
    """Zipf corpus with planted clusters; returns (raw stream, cluster raw ids
    [n_clusters, cluster_size])."""
    n_cw = n_clusters * cluster_size
    if band is None:
        lo = max(100, raw_vocab // 100)
        band = (lo, min(raw_vocab + 1, lo + 4 * max(n_cw, 1)))
    mu, sigma = lognorm if lognorm else (0.0, 0.0)
    min_lines = n_tokens // max(1, (len_clip[0] if lognorm else line_len)) + 2
    cap = n_tokens + min_lines
    out = np.empty(cap, dtype=np.int32)
    cw = np.empty(max(n_cw, 1), dtype=np.int32)
    n = _chk(N.host().ph_synth_zipf(n_tokens, raw_vocab, zipf_s, line_len, mu, sigma, len_clip[0], len_clip[1],
                                    n_clusters, cluster_size, band[0], band[1], span, p_replace, seed,
                                    N.ptr(out), cap, N.ptr(cw)), "synth zipf")
    return out[:n], cw[:n_cw].reshape(n_clusters, cluster_size)


def synth_sbm_walks(n_nodes: int = 10_000, n_comm: int = 20, exponent: float = 2.5, mean_degree: float = 40.0,
                    p_intra: float = 0.85, walks_per_node: int = 10, walk_len: int = 80,
                    seed: int = 3) -> tuple[np.ndarray, np.ndarray]:
    cap = n_nodes * walks_per_node * (walk_len + 1)
    out = np.empty(cap, dtype=np.int32)
    comm = np.empty(n_nodes, dtype=np.int32)
    n = _chk(N.host().ph_synth_sbm_walks(n_nodes, n_comm, exponent, mean_degree, p_intra, walks_per_node,
                                         walk_len, seed, N.ptr(out), cap, N.ptr(comm)), "synth sbm")
    return out[:n], comm


def synth_sbm_graph(n_nodes: int = 10_000, n_comm: int = 20, exponent: float = 2.5, mean_degree: float = 40.0,
                    p_intra: float = 0.85, seed: int = 3) -> tuple[np.ndarray, int, int]:
    """(degree per node, edges, edges inside a community) of the graph that
    synth_sbm_walks walks with the same arguments and seed."""
    import ctypes as C
    deg = np.empty(n_nodes, dtype=np.int64)
    intra = C.c_int64(0)
    m = _chk(N.host().ph_synth_sbm_graph(n_nodes, n_comm, exponent, mean_degree, p_intra, seed, N.ptr(deg),
                                         C.byref(intra)), "synth sbm graph")
    return deg, int(m), intra.value


def save_corpus(path: str, vocab: Vocab, stream: np.ndarray) -> None:
    """Binary corpus: vocabulary + id stream in one file (ph_corpus_save)."""
    s = np.ascontiguousarray(stream, dtype=np.int32)
    _chk(N.host().ph_corpus_save(path.encode(), vocab._h, N.ptr(s), s.size), "save corpus")


def load_corpus(path: str) -> tuple[Vocab, np.ndarray]:
    """(vocabulary, id stream) of a binary corpus written by save_corpus."""
    import ctypes as C
    n = C.c_int64(0)
    h = N.host().ph_corpus_vocab(path.encode(), C.byref(n))
    vocab = Vocab(h)
    out = np.empty(n.value, dtype=np.int32)
    _chk(N.host().ph_corpus_ids(path.encode(), N.ptr(out), out.size), "corpus ids")
    return vocab, out


def write_raw_text(path: str, raw: np.ndarray) -> None:
    raw = np.ascontiguousarray(raw, dtype=np.int32)
    _chk(N.host().ph_write_raw_text(path.encode(), N.ptr(raw), raw.size), "write text")
