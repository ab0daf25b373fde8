"""ctypes bindings of the two native libraries (include/par_host.h and
include/par_sgns.h).  Loading fails loudly when a library is missing: there
is no Python or CPU fallback for any entry point."""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_DIR = Path(__file__).resolve().parent / "lib"

i32, i64, u64, u32, f32, f64 = C.c_int32, C.c_int64, C.c_uint64, C.c_uint32, C.c_float, C.c_double
vp = C.c_void_p
cp = C.c_char_p


class NativeLibraryMissing(RuntimeError):
    pass


def _load(name: str) -> C.CDLL:
    path = LIB_DIR / name
    if not path.exists():
        raise NativeLibraryMissing(
            f"{path} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` first")
    return C.CDLL(str(path), mode=C.RTLD_GLOBAL)


# ----------------------------------------------------------------- structs
class ParSgnsConfig(C.Structure):
    _fields_ = [("vocab_size", i32), ("dim", i32), ("window", i32), ("negative", i32),
                ("batch_size", i32), ("precision", i32), ("device", i32), ("rounding", i32)]


class ParSgnsTrainArgs(C.Structure):
    _fields_ = [("mode", i32), ("iter", i32), ("epoch", i32), ("compute_loss", i32),
                ("alpha", f32), ("waves_per_block", i32), ("blocks", i32), ("use_graph", i32),
                ("schedule", i32), ("repulsion", i32),
                ("train_words", u64), ("seed", u64), ("sent_begin", i64), ("sent_end", i64),
                ("sentences_per_dequeue", i32), ("min_sentences_per_wave", i32), ("write_back", i32),
                ("hot_rows", i32)]


class ParSgnsTunePoint(C.Structure):
    _fields_ = [("waves_per_block", i32), ("waves_per_cu", i32), ("sentences_per_dequeue", i32),
                ("reserved", i32), ("kernel_ms", f64)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class ParSgnsStats(C.Structure):
    _fields_ = [("loss_attr", f64), ("loss_rep", f64), ("tokens", u64), ("sentences", u64),
                ("pairs", u64), ("batches", u64), ("negatives", u64), ("rep_pairs", u64),
                ("kernel_ms", f64), ("alg_bytes", f64), ("alg_flops", f64), ("blocks", i32),
                ("waves_per_block", i32), ("sentences_per_dequeue", i32), ("reserved", i32)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class PhPlanInfo(C.Structure):
    _fields_ = [("n_sent", i64), ("n_tokens", i64), ("n_batches", i64), ("n_negs", i64),
                ("draws", u64), ("final_rng", u64), ("raw_words", u64), ("final_alpha", f32)]


# symbol -> (restype, argtypes); every symbol the headers declare
HOST_SYMBOLS = {
    "ph_last_error": (cp, []),
    "ph_vocab_learn_text": (vp, [cp, i32]),
    "ph_vocab_read": (vp, [cp, i32]),
    "ph_vocab_save": (C.c_int, [vp, cp]),
    "ph_vocab_from_raw": (vp, [vp, i64, i32, i32]),
    "ph_vocab_free": (None, [vp]),
    "ph_vocab_size": (i32, [vp]),
    "ph_vocab_train_words": (u64, [vp]),
    "ph_vocab_counts": (C.c_int, [vp, vp]),
    "ph_vocab_word": (cp, [vp, i32]),
    "ph_vocab_lookup": (i32, [vp, cp]),
    "ph_vocab_raw_map": (C.c_int, [vp, vp, i32]),
    "ph_load_text_stream": (i64, [vp, cp, vp, i64]),
    "ph_map_raw_stream": (i64, [vp, vp, i64, vp, i64]),
    "ph_subsample": (i64, [vp, i64, vp, i32, u64, f32, u64, i32, vp, i64, vp, vp, i64, vp]),
    "ph_unigram_entries": (C.c_int, [vp, i32, i64, vp]),
    "ph_plan_build": (vp, [vp, i64, vp, i32, u64, i32, i32, i32, i32, f32, f32]),
    "ph_plan_info_get": (C.c_int, [vp, vp]),
    "ph_plan_copy": (C.c_int, [vp, vp, vp, vp, vp, vp, vp, vp]),
    "ph_plan_free": (None, [vp]),
    "ph_exp_table": (None, [vp]),
    "ph_init_net": (None, [i64, i32, vp, vp]),
    "ph_synth_zipf": (i64, [i64, i32, f64, i32, f64, f64, i32, i32, i32, i32, i32, i32, i32, f64, u64,
                           vp, i64, vp]),
    "ph_synth_sbm_walks": (i64, [i32, i32, f64, f64, f64, i32, i32, u64, vp, i64, vp]),
    "ph_synth_sbm_graph": (i64, [i32, i32, f64, f64, f64, u64, vp, vp]),
    "ph_write_raw_text": (C.c_int, [cp, vp, i64]),
    "ph_corpus_save": (C.c_int, [cp, vp, vp, i64]),
    "ph_corpus_vocab": (vp, [cp, vp]),
    "ph_corpus_ids": (i64, [cp, vp, i64]),
    "ph_save_embeddings": (C.c_int, [cp, vp, vp, i32, i32]),
}

SGNS_SYMBOLS = {
    "par_sgns_create": (C.c_int, [vp, vp]),
    "par_sgns_destroy": (C.c_int, [vp]),
    "par_sgns_last_error": (cp, [vp]),
    "par_sgns_row_stride": (i32, [vp]),
    "par_sgns_set_exp_table": (C.c_int, [vp, vp, i32]),
    "par_sgns_set_sampler": (C.c_int, [vp, vp, i32]),
    "par_sgns_sample_negatives": (C.c_int, [vp, u64, i64, i64, vp]),
    "par_sgns_set_weights": (C.c_int, [vp, vp, vp]),
    "par_sgns_get_weights": (C.c_int, [vp, vp, vp]),
    "par_sgns_set_corpus": (C.c_int, [vp, vp, i64, vp, vp, i64]),
    "par_sgns_set_plan": (C.c_int, [vp, vp, vp, i64, vp, vp, i64, vp, vp, i64, vp, i64]),
    "par_sgns_train": (C.c_int, [vp, vp, vp]),
    "par_sgns_prepare": (C.c_int, [vp, vp]),
    "par_sgns_autotune": (C.c_int, [vp, vp, i32, vp, i32, vp, vp]),
    "par_sgns_set_launch_shape": (C.c_int, [vp, i32, i32, i32]),
    "par_sgns_loss_forward": (C.c_int, [vp, vp, vp, vp, i64, i32, i32, i32, vp]),
    "par_sgns_loss_adjoint": (C.c_int, [vp, vp, vp, vp, i64, i32, i32, i32, vp]),
    "par_sgns_loss_forward_noalloc": (C.c_int, [vp, vp, vp, vp, i64, i32, i32, i32, vp]),
    "par_sgns_get_grads": (C.c_int, [vp, vp, vp]),
    "par_sgns_last_kernel_time": (C.c_int, [vp, i32, vp, vp]),
}

_host = None
_sgns = None


def _bind(lib: C.CDLL, table: dict) -> C.CDLL:
    for name, (res, args) in table.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def host() -> C.CDLL:
    global _host
    if _host is None:
        _host = _bind(_load("libpar_host.so"), HOST_SYMBOLS)
    return _host


def sgns() -> C.CDLL:
    global _sgns
    if _sgns is None:
        _sgns = _bind(_load("libpar_sgns.so"), SGNS_SYMBOLS)
    return _sgns


def ptr(a: np.ndarray | None):
    """Pointer of a C-contiguous numpy array (None -> NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(vp)


def host_error() -> str:
    return (host().ph_last_error() or b"").decode()
