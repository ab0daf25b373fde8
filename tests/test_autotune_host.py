"""Saved tunings of the launch-shape autotuner: a tuning applies only to the
kernel shape it was measured on (host logic; no GPU)."""
import json

from par_word2vec_amd import autotune as A
from par_word2vec_amd.corpora import Workload


def _w(**kw):
    base = dict(name="t", kind="zipf", gen={}, dim=300, window=5, negative=10, sample=1e-5, batch_size=24)
    base.update(kw)
    return Workload(**base)


def test_load_matches_kernel_shape(tmp_path, monkeypatch):
    monkeypatch.setattr(A, "TUNED", tmp_path)
    w = _w()
    best = {"waves_per_block": 2, "waves_per_cu": 8, "sentences_per_dequeue": 2, "kernel_ms": 1.0}
    A.path_for("t", "bf16").write_text(json.dumps({"key": A.shape_key(w, "bf16"), "best": best, "points": [best]}))
    assert A.load("t", w, "bf16")["best"] == best
    assert A.load("t", _w(window=8), "bf16") is None       # another ring depth
    assert A.load("t", _w(dim=128), "bf16") is None        # another kernel instance
    assert A.load("t", w, "fp32") is None                  # no file for it
    assert A.load("other", w, "bf16") is None
    assert A.load("t", w, "bf16", write_back="store_out") is None   # another write-back instance
    A.path_for("t", "bf16", write_back="store_out").write_text(
        json.dumps({"key": A.shape_key(w, "bf16", write_back="store_out"), "best": best, "points": [best]}))
    assert A.load("t", w, "bf16", write_back="store_out")["best"] == best
    assert A.path_for("t", "bf16", write_back="store_out") != A.path_for("t", "bf16")


def test_apply_passes_the_best_shape():
    calls = []

    class Dev:
        def set_launch_shape(self, *a):
            calls.append(a)

    b = A.apply(Dev(), {"best": {"waves_per_block": 4, "waves_per_cu": 8, "sentences_per_dequeue": 1, "kernel_ms": 2.0}})
    assert calls == [(4, 8, 1)] and b["waves_per_cu"] == 8


def test_saved_tunings_match_their_workloads():
    """The committed tunings apply to the workloads they were measured on
    (bench.py then launches with them instead of the built-in default)."""
    from par_word2vec_amd.corpora import WORKLOADS
    for name in ("1b", "text8", "sbm"):
        for wb in ("atomic", "store_out"):
            t = A.load(name, WORKLOADS[name], "bf16", write_back=wb)
            assert t is not None, (name, wb)
            b = t["best"]
            assert 1 <= b["waves_per_block"] <= 8 and b["waves_per_cu"] % b["waves_per_block"] == 0
            assert b in t["points"]
