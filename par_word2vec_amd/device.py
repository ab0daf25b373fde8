"""Python face of the device C-ABI (include/par_sgns.h).  `SgnsDevice` owns a
handle: device tables, sampler, corpus or plan, and one HIP stream.  There is
no fallback: a missing library or a failed call raises."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N

FP32, BF16 = 0, 1
DETERMINISTIC, PLAIN, PERSISTENT = 0, 1, 2
MODES = {"det": DETERMINISTIC, "deterministic": DETERMINISTIC, "plain": PLAIN, "persistent": PERSISTENT}
PRECISIONS = {"fp32": FP32, "bf16": BF16}
WIN_MFMA, WIN_FP64 = 1, 2
ROUNDINGS = {"stochastic": 0, "nearest": 1, "nearest_out": 2, "nearest_in": 3}
# repulsion batch evaluation: matrix cores on bf16 tables unless "vector" is asked for
REPULSIONS = {"auto": 0, "vector": 1, "matrix": 2}
# bf16 matrix-core write-back (par_sgns_write_back): atomics, plain stores of W_out rows, of W_in
# rows, of both tables, or of rows at or past hot_rows
WRITE_BACKS = {"atomic": 0, "store_out": 1, "store_in": 2, "store": 3, "hot": 4}


class DeviceError(RuntimeError):
    pass


class SgnsDevice:
    def __init__(self, vocab_size: int, dim: int, window: int, negative: int, batch_size: int,
                 precision: str = "fp32", device: int = 0, rounding: str = "stochastic"):
        self.lib = N.sgns()
        self.cfg = N.ParSgnsConfig(vocab_size, dim, window, negative, batch_size, PRECISIONS[precision],
                                   device, ROUNDINGS[rounding])
        self.vocab_size, self.dim = vocab_size, dim
        self.precision = precision
        h = C.c_void_p()
        rc = self.lib.par_sgns_create(C.byref(self.cfg), C.byref(h))
        self._h = h
        if rc != 0:
            msg = self.lib.par_sgns_last_error(h).decode() if h else "invalid configuration"
            self.close()
            raise DeviceError(f"par_sgns_create: {msg} (code {rc})")

    # ------------------------------------------------------------ plumbing
    def _chk(self, rc: int, what: str) -> None:
        if rc != 0:
            raise DeviceError(f"{what}: {self.lib.par_sgns_last_error(self._h).decode()} (code {rc})")

    def close(self) -> None:
        if getattr(self, "_h", None):
            self.lib.par_sgns_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def row_stride(self) -> int:
        return int(self.lib.par_sgns_row_stride(self._h))

    # -------------------------------------------------------------- uploads
    def set_exp_table(self, table: np.ndarray) -> None:
        t = np.ascontiguousarray(table, dtype=np.float32)
        self._chk(self.lib.par_sgns_set_exp_table(self._h, N.ptr(t), t.size), "set_exp_table")

    def set_sampler(self, entries: np.ndarray) -> None:
        e = np.ascontiguousarray(entries, dtype=np.int64)
        self._chk(self.lib.par_sgns_set_sampler(self._h, N.ptr(e), e.size), "set_sampler")

    def sample_negatives(self, n: int, seed: int = 1, sent: int = 0) -> np.ndarray:
        out = np.empty(n, dtype=np.int32)
        self._chk(self.lib.par_sgns_sample_negatives(self._h, seed, sent, n, N.ptr(out)), "sample_negatives")
        return out

    def set_weights(self, w_in: np.ndarray | None, w_out: np.ndarray | None) -> None:
        a = None if w_in is None else np.ascontiguousarray(w_in, dtype=np.float32)
        b = None if w_out is None else np.ascontiguousarray(w_out, dtype=np.float32)
        for x in (a, b):
            if x is not None and x.shape != (self.vocab_size, self.dim):
                raise ValueError(f"weights must be [{self.vocab_size}, {self.dim}]")
        self._chk(self.lib.par_sgns_set_weights(self._h, N.ptr(a), N.ptr(b)), "set_weights")

    def get_weights(self) -> tuple[np.ndarray, np.ndarray]:
        a = np.empty((self.vocab_size, self.dim), dtype=np.float32)
        b = np.empty_like(a)
        self._chk(self.lib.par_sgns_get_weights(self._h, N.ptr(a), N.ptr(b)), "get_weights")
        return a, b

    def set_corpus(self, tokens: np.ndarray, sent_off: np.ndarray, sent_raw: np.ndarray) -> None:
        t = np.ascontiguousarray(tokens, dtype=np.int32)
        o = np.ascontiguousarray(sent_off, dtype=np.int64)
        r = np.ascontiguousarray(sent_raw, dtype=np.int64)
        self._chk(self.lib.par_sgns_set_corpus(self._h, N.ptr(t), t.size, N.ptr(o), N.ptr(r), o.size - 1),
                  "set_corpus")
        self.n_sent = o.size - 1

    def set_plan(self, plan) -> None:
        p = plan
        negs = p.negs if p.negs.size else np.zeros(1, np.int32)
        self._chk(self.lib.par_sgns_set_plan(self._h, N.ptr(p.tokens), N.ptr(p.b), p.tokens.size, N.ptr(p.sent_off),
                                             N.ptr(p.sent_alpha), p.sent_off.size - 1, N.ptr(p.batch_off),
                                             N.ptr(p.neg_off), p.neg_off.size - 1, N.ptr(negs), p.negs.size),
                  "set_plan")

    # ------------------------------------------------------------- training
    def train(self, mode: str | int, sent_begin: int, sent_end: int, alpha: float, train_words: int,
              iters: int = 1, epoch: int = 0, seed: int = 1, compute_loss: bool = False,
              waves_per_block: int = 0, blocks: int = 0, use_graph: bool = False, planned: bool = False,
              prepare_only: bool = False, repulsion: str = "auto", sentences_per_dequeue: int = 0,
              min_sentences_per_wave: int = 0, write_back: str = "atomic", hot_rows: int = 0) -> dict:
        m = MODES[mode] if isinstance(mode, str) else int(mode)
        args = N.ParSgnsTrainArgs(m, iters, epoch, int(compute_loss), alpha, waves_per_block, blocks,
                                  int(use_graph), int(planned), REPULSIONS[repulsion], train_words, seed,
                                  sent_begin, sent_end, sentences_per_dequeue, min_sentences_per_wave,
                                  WRITE_BACKS[write_back], hot_rows)
        if prepare_only:
            self._chk(self.lib.par_sgns_prepare(self._h, C.byref(args)), "prepare")
            return {}
        st = N.ParSgnsStats()
        self._chk(self.lib.par_sgns_train(self._h, C.byref(args), C.byref(st)), "train")
        return st.as_dict()

    def autotune(self, sent_begin: int, sent_end: int, alpha: float, train_words: int, iters: int = 1,
                 epoch: int = 0, seed: int = 1, reps: int = 3, repulsion: str = "auto",
                 min_sentences_per_wave: int = 0, write_back: str = "atomic", hot_rows: int = 0) -> dict:
        """Times the persistent kernel's launch shapes on the sentence range
        (par_sgns_autotune; the tables are left as they were) and makes the
        fastest the handle's default.  Returns {"best": point, "points": [...]}."""
        args = N.ParSgnsTrainArgs(MODES["persistent"], iters, epoch, 0, alpha, 0, 0, 0, 0, REPULSIONS[repulsion],
                                  train_words, seed, sent_begin, sent_end, 0, min_sentences_per_wave,
                                  WRITE_BACKS[write_back], hot_rows)
        cap = 64
        pts = (N.ParSgnsTunePoint * cap)()
        n = C.c_int32(0)
        best = N.ParSgnsTunePoint()
        self._chk(self.lib.par_sgns_autotune(self._h, C.byref(args), reps, pts, cap, C.byref(n), C.byref(best)),
                  "autotune")
        return {"best": best.as_dict(), "points": [pts[i].as_dict() for i in range(min(n.value, cap))]}

    def set_launch_shape(self, waves_per_block: int = 0, waves_per_cu: int = 0, sentences_per_dequeue: int = 0) -> None:
        """Default persistent launch shape from a saved tuning (all 0: built-in default)."""
        self._chk(self.lib.par_sgns_set_launch_shape(self._h, waves_per_block, waves_per_cu, sentences_per_dequeue),
                  "set_launch_shape")

    # ------------------------------------------------- loss and its adjoint
    def _windows(self, fn, centre, ctx, neg, flags) -> float:
        c = np.ascontiguousarray(centre, dtype=np.int32)
        x = np.ascontiguousarray(ctx, dtype=np.int32)
        n = np.ascontiguousarray(neg, dtype=np.int32)
        nw = c.size
        nctx = x.shape[1] if x.ndim == 2 else x.size // max(nw, 1)
        nneg = n.shape[1] if n.ndim == 2 else n.size // max(nw, 1)
        out = C.c_double(0.0)
        self._chk(fn(self._h, N.ptr(c), N.ptr(x), N.ptr(n), nw, nctx, nneg, flags, C.byref(out)), fn.__name__)
        return out.value

    def loss_forward(self, centre, ctx, neg, flags: int = 0) -> float:
        return self._windows(self.lib.par_sgns_loss_forward, centre, ctx, neg, flags)

    def loss_forward_noalloc(self, centre, ctx, neg, flags: int = 0) -> float:
        """loss_forward staged in the library's fixed per-device scratch (no
        device allocation; at most 2,097,088 window indices per call)."""
        return self._windows(self.lib.par_sgns_loss_forward_noalloc, centre, ctx, neg, flags)

    def loss_adjoint(self, centre, ctx, neg, flags: int = 0) -> float:
        return self._windows(self.lib.par_sgns_loss_adjoint, centre, ctx, neg, flags)

    def get_grads(self) -> tuple[np.ndarray, np.ndarray]:
        a = np.empty((self.vocab_size, self.dim), dtype=np.float32)
        b = np.empty_like(a)
        self._chk(self.lib.par_sgns_get_grads(self._h, N.ptr(a), N.ptr(b)), "get_grads")
        return a, b

    def last_kernel_ms(self, kind: int = 0) -> tuple[float, int]:
        ms, n = C.c_double(), C.c_int32()
        self._chk(self.lib.par_sgns_last_kernel_time(self._h, kind, C.byref(ms), C.byref(n)), "kernel time")
        return ms.value, n.value
