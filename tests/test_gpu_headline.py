"""The headline workload under test: the 1B-word-shaped corpus at full size
(800M Zipf(1.0) tokens over 555,000 words, log-normal line lengths, d 300,
window 5, 10 negatives, sample 1e-5), one epoch in ~10M-raw-word chunks, with
the benchmark's two variants: bf16 tables on the persistent kernel (matrix-core
repulsion, stochastic rounding, one hipGraph per chunk, the saved launch shape)
and fp32 tables on the plain kernel, both with the benchmark's write-back (W_in
rows by atomics, W_out rows by plain stores).

Bars (the metric's "at equal planted-cluster score", with scores that do not
saturate):
  every kept token trained exactly once, every non-empty sentence once
  FP64 training loss per pair of bf16-persistent within 1 % of fp32-plain
  FP64 SGNS loss per pair of 65,536 held-out windows within 1 %
  planted-cluster recall@10 within 0.01
  planted-cluster cosine margin of bf16 at least 0.92 of fp32's: measured
    0.940-0.942 (0.2458-0.2461 against 0.2609-0.2615 over seeds and fp32
    schedulers).  The gap is a property of bf16 tables with stochastic
    rounding, not of the persistent scheduler (bf16 plain: 0.2486): every
    rounding adds noise of variance ~ulp * |update|, which lowers the mates'
    cosine by ~1 % and raises the nearest non-mates' by ~1 %, and the margin,
    a difference of the two, moves 6 %.  The vector repulsion, which rounds a
    row once per block of negatives instead of once per batch, loses 11 %.
"""
import numpy as np
import pytest

from par_word2vec_amd import autotune, corpora, evaluate as E, host as H
from par_word2vec_amd.device import SgnsDevice
from par_word2vec_amd.trainer import chunk_bounds

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def onebillion():
    c = corpora.make("1b")
    w = c.workload
    print(f"[headline] corpus: {c.stream.size} raw words, V={len(c.vocab)}", flush=True)
    csr = H.subsample(c.stream, c.counts, c.train_words, w.sample, seed=1000003)
    held = corpora.heldout_stream(c, 3_000_000, seed=99)
    windows = E.heldout_windows(held, c.counts, c.train_words, w.sample, w.window, 5, 65_536, seed=8)
    return c, csr, windows


WRITE_BACK = "store_out"  # bench.py's default


def _epoch(c, csr, mode, precision):
    w = c.workload
    with SgnsDevice(len(c.vocab), w.dim, w.window, w.negative, w.batch_size, precision) as dev:
        dev.set_exp_table(H.exp_table())
        dev.set_weights(*H.init_net(len(c.vocab), w.dim))
        dev.set_sampler(H.unigram_entries(c.counts))
        dev.set_corpus(csr.tokens, csr.sent_off, csr.sent_raw)
        if mode == "persistent":
            saved = autotune.load("1b", w, precision, "auto", WRITE_BACK)
            if saved is not None:
                autotune.apply(dev, saved)
        tot = dict(tokens=0, sentences=0, pairs=0, rep_pairs=0, loss_attr=0.0, loss_rep=0.0, kernel_ms=0.0)
        for b, e in chunk_bounds(csr.sent_raw, 10_000_000):
            st = dev.train(mode, b, e, w.alpha, c.train_words, seed=1, compute_loss=True,
                           use_graph=mode == "persistent", write_back=WRITE_BACK)
            for k in tot:
                tot[k] += st[k]
        w_in, w_out = dev.get_weights()
    return tot, w_in, w_out


def test_headline_bf16_persistent_matches_fp32_plain(onebillion):
    c, csr, windows = onebillion
    n_nonempty = int(np.count_nonzero(np.diff(csr.sent_off) > 0))
    res = {}
    for mode, prec in (("plain", "fp32"), ("persistent", "bf16")):
        print(f"[headline] {prec} {mode}: training one epoch", flush=True)
        st, w_in, w_out = _epoch(c, csr, mode, prec)
        assert st["tokens"] == csr.tokens.size and st["sentences"] == n_nonempty
        res[prec] = dict(
            loss=(st["loss_attr"] + st["loss_rep"]) / (st["pairs"] + st["rep_pairs"]),
            heldout=E.heldout_loss(w_in, w_out, *windows),
            margin=E.cluster_margin(w_in, c.groups),
            recall=E.planted_cluster_recall(w_in, c.groups),
            ms=st["kernel_ms"])
        print(f"[headline] {prec} {mode}: {res[prec]}", flush=True)
    a, b = res["fp32"], res["bf16"]
    assert abs(b["loss"] - a["loss"]) <= 0.01 * a["loss"]
    assert abs(b["heldout"] - a["heldout"]) <= 0.01 * a["heldout"]
    assert a["margin"] > 0.2 and b["margin"] >= 0.92 * a["margin"]
    assert abs(b["recall"] - a["recall"]) <= 0.01
    # the training moved the held-out objective well below its value at the start
    w_in0, _ = H.init_net(len(c.vocab), c.workload.dim)
    start = E.heldout_loss(w_in0, np.zeros_like(w_in0), *windows)
    assert a["heldout"] < 0.9 * start and b["heldout"] < 0.9 * start
