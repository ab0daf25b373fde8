"""Per-batch gradient error on the DeepWalk SBM workload (10,000 nodes, 20
communities, 10 walks of 80 per node; d 128, window 10, 5 negatives), BF16
tables with FP32 accumulation against the FP64 oracle gradient.

After one epoch of training (fp32 tables, and bf16 tables with stochastic
rounding, persistent kernel), 8 batches of 8,192 training windows are drawn
from the walks (a centre, its 2 x 10 walk neighbours, 5 negatives from
unigram^0.75), and each batch's SGNS gradient dL/dW_in, dL/dW_out is taken by
the window adjoint kernel and by the oracle in FP64:

  bf16-trained tables    the bf16 kernel (vector and matrix-core adjoint)
                         against the FP64 gradient of the same tables: the
                         error of FP32 accumulation (and, on the matrix cores,
                         of the logit gradient as a bf16 hi + lo pair)
  fp32-trained tables    the bf16 kernel on those tables rounded to bf16
                         against the FP64 gradient of the unrounded tables:
                         the storage rounding (2^-9 relative per entry) too
  fp32-trained tables    the fp32 kernel, the same FP64 reference

Bars: relative L2 error per batch below 1e-5 (fp32 storage and arithmetic),
1e-5 for bf16 storage with exact tables (FP32 accumulation only), 1e-4 on the
matrix cores, and 1e-2 where the storage rounding enters.  Community
recovery (kNN majority accuracy) of both trained tables at least 0.99.
"""
import numpy as np
import pytest

import oracle as O
from par_word2vec_amd import autotune, corpora, evaluate as E, host as H
from par_word2vec_amd.device import SgnsDevice, WIN_MFMA
from par_word2vec_amd.trainer import chunk_bounds

pytestmark = pytest.mark.gpu

N_BATCH, PER_BATCH = 8, 8192


@pytest.fixture(scope="module")
def trained():
    c = corpora.make("sbm")
    w = c.workload
    csr = H.subsample(c.stream, c.counts, c.train_words, w.sample, seed=1000003)
    out = {}
    for prec in ("fp32", "bf16"):
        with SgnsDevice(len(c.vocab), w.dim, w.window, w.negative, w.batch_size, prec) as dev:
            dev.set_exp_table(H.exp_table())
            dev.set_weights(*H.init_net(len(c.vocab), w.dim))
            dev.set_sampler(H.unigram_entries(c.counts))
            dev.set_corpus(csr.tokens, csr.sent_off, csr.sent_raw)
            saved = autotune.load("sbm", w, prec, "auto")
            if saved is not None:
                autotune.apply(dev, saved)
            for b, e in chunk_bounds(csr.sent_raw, 10_000_000):
                dev.train("persistent", b, e, w.alpha, c.train_words, seed=1, use_graph=True)
            out[prec] = dev.get_weights()
    centre, ctx, neg = E.heldout_windows(c.stream, c.counts, c.train_words, w.sample, w.window, w.negative,
                                         N_BATCH * PER_BATCH, seed=17)
    return c, out, (centre, ctx, neg)


def _adjoint(c, prec, tables, batch, flags=0):
    w = c.workload
    with SgnsDevice(len(c.vocab), w.dim, w.window, w.negative, w.batch_size, prec) as dev:
        dev.set_weights(*tables)
        dev.loss_adjoint(*batch, flags=flags)
        gi, go = dev.get_grads()
    return np.concatenate([gi.ravel(), go.ravel()]).astype(np.float64)


def _rel(g, ref):
    return float(np.linalg.norm(g - ref) / np.linalg.norm(ref))


def test_sbm_gradient_error_per_batch(trained):
    c, tabs, (centre, ctx, neg) = trained
    acc = {k: float(E.community_knn_accuracy(tabs[k][0], c.labels)) for k in tabs}
    print(f"[sbm grad] community kNN accuracy {acc}", flush=True)
    assert acc["fp32"] >= 0.99 and acc["bf16"] >= 0.99
    errs = {k: [] for k in ("bf16 tables: bf16 kernel", "bf16 tables: bf16 matrix kernel",
                            "fp32 tables: fp32 kernel", "fp32 tables: bf16 kernel (rounded tables)",
                            "fp32 tables: bf16 matrix kernel (rounded tables)")}
    for k in range(N_BATCH):
        s = slice(k * PER_BATCH, (k + 1) * PER_BATCH)
        batch = (centre[s], ctx[s], neg[s])
        flat = (centre[s], ctx[s].reshape(-1), neg[s].reshape(-1))
        for prec_tab in ("bf16", "fp32"):
            w_in, w_out = tabs[prec_tab]
            gi, go = O.window_grad_f64(*flat, w_in.astype(np.float64), w_out.astype(np.float64))
            ref = np.concatenate([gi.ravel(), go.ravel()])
            if prec_tab == "bf16":
                errs["bf16 tables: bf16 kernel"].append(_rel(_adjoint(c, "bf16", (w_in, w_out), batch), ref))
                errs["bf16 tables: bf16 matrix kernel"].append(
                    _rel(_adjoint(c, "bf16", (w_in, w_out), batch, WIN_MFMA), ref))
            else:
                errs["fp32 tables: fp32 kernel"].append(_rel(_adjoint(c, "fp32", (w_in, w_out), batch), ref))
                errs["fp32 tables: bf16 kernel (rounded tables)"].append(
                    _rel(_adjoint(c, "bf16", (w_in, w_out), batch), ref))
                errs["fp32 tables: bf16 matrix kernel (rounded tables)"].append(
                    _rel(_adjoint(c, "bf16", (w_in, w_out), batch, WIN_MFMA), ref))
    for k, v in errs.items():
        print(f"[sbm grad] {k}: rel L2 per batch max {max(v):.2e}, mean {np.mean(v):.2e}", flush=True)
    assert max(errs["fp32 tables: fp32 kernel"]) < 1e-5
    assert max(errs["bf16 tables: bf16 kernel"]) < 1e-5
    assert max(errs["bf16 tables: bf16 matrix kernel"]) < 1e-4
    assert max(errs["fp32 tables: bf16 kernel (rounded tables)"]) < 1e-2
    assert max(errs["fp32 tables: bf16 matrix kernel (rounded tables)"]) < 1e-2
