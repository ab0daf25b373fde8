"""GPU parity of the training kernels against the oracle (CPU restatement of
PAR_Word2Vec_cpu.cpp).  All calls go through the C-ABI of libpar_sgns.so.

Tolerances:
  deterministic mode  bit-identical W_in/W_out to the oracle; FP64 loss
                      within 1e-9 relative (only exp/log1p libraries differ)
  Hogwild modes       mean loss per pair within 1% of the oracle's and
                      planted-cluster recall@10 within 0.02 absolute of it
                      (different random streams; the quality gate)."""
import numpy as np
import pytest

import oracle as O
from par_word2vec_amd import corpora, evaluate as E, host as H
from par_word2vec_amd.device import SgnsDevice

pytestmark = pytest.mark.gpu


def _setup(c, precision="fp32", seed_weights=None):
    w = c.workload
    V = len(c.vocab)
    dev = SgnsDevice(V, w.dim, w.window, w.negative, w.batch_size, precision)
    dev.set_exp_table(H.exp_table())
    w_in, w_out = H.init_net(V, w.dim) if seed_weights is None else seed_weights
    dev.set_weights(w_in, w_out)
    return dev, w_in, w_out


def _small_corpus(seed=7, n=60_000, V=1500, line=200):
    raw, cw = H.synth_zipf(n, V, 1.0, line_len=line, n_clusters=5, cluster_size=10, p_replace=0.3, seed=seed)
    vocab = H.Vocab.from_raw(raw, V, 5)
    stream = H.map_raw_stream(vocab, raw)
    return vocab, stream


@pytest.mark.parametrize("dim,window,negative,batch,iters,sample", [
    (100, 5, 5, 24, 1, 1e-3),
    (64, 3, 2, 7, 2, 1e-4),
    (130, 8, 5, 16, 1, 0.0),
    (300, 5, 10, 24, 1, 1e-4),
])
def test_det_bitwise_vs_oracle(dim, window, negative, batch, iters, sample):
    print(f"[det] dim={dim} W={window} neg={negative} B={batch} iter={iters} sample={sample}", flush=True)
    vocab, stream = _small_corpus()
    counts, T = vocab.counts(), vocab.train_words
    plan = H.build_plan(stream, counts, T, window, negative, batch, iters, 0.025, sample)
    ref = O.train(stream, counts, T, dim, window, negative, batch, iters, 0.025, sample)
    V = len(vocab)
    with SgnsDevice(V, dim, window, negative, batch, "fp32") as dev:
        dev.set_exp_table(H.exp_table())
        w_in, w_out = H.init_net(V, dim)
        dev.set_weights(w_in, w_out)
        dev.set_plan(plan)
        st = dev.train("det", 0, plan.n_sent, 0.025, T, iters=iters, compute_loss=True)
        g_in, g_out = dev.get_weights()
    assert st["pairs"] == ref.stats["pairs"]
    assert st["negatives"] == ref.stats["negatives"]
    assert st["tokens"] == ref.stats["trained_tokens"]
    assert np.array_equal(g_in, ref.w_in), f"W_in differs in {np.count_nonzero(g_in != ref.w_in)} entries"
    assert np.array_equal(g_out, ref.w_out), f"W_out differs in {np.count_nonzero(g_out != ref.w_out)} entries"
    np.testing.assert_allclose(st["loss_attr"], ref.stats["loss_attr"], rtol=1e-9)
    np.testing.assert_allclose(st["loss_rep"], ref.stats["loss_rep"], rtol=1e-9)


def test_det_zipf1m_full():
    """The 1M-token workload end to end, deterministic mode vs oracle."""
    print("[det] zipf1m full", flush=True)
    c = corpora.make("zipf1m")
    w = c.workload
    plan = H.build_plan(c.stream, c.counts, c.train_words, w.window, w.negative, w.batch_size, 1, w.alpha, w.sample)
    ref = O.train(c.stream, c.counts, c.train_words, w.dim, w.window, w.negative, w.batch_size, 1, w.alpha, w.sample)
    dev, _, _ = _setup(c)
    dev.set_plan(plan)
    st = dev.train("det", 0, plan.n_sent, w.alpha, c.train_words, compute_loss=True)
    g_in, g_out = dev.get_weights()
    dev.close()
    rel = np.abs(g_in - ref.w_in).max() / np.abs(ref.w_in).max()
    assert rel <= 1e-5
    assert np.array_equal(g_in, ref.w_in) and np.array_equal(g_out, ref.w_out)
    assert st["pairs"] == ref.stats["pairs"]


@pytest.mark.parametrize("mode,precision", [("plain", "fp32"), ("persistent", "fp32"), ("persistent", "bf16")])
def test_hogwild_quality_zipf1m(mode, precision):
    print(f"[hogwild] zipf1m {mode} {precision}", flush=True)
    c = corpora.make("zipf1m")
    w = c.workload
    ref = O.train(c.stream, c.counts, c.train_words, w.dim, w.window, w.negative, w.batch_size, 1, w.alpha, w.sample)
    ref_recall = E.planted_cluster_recall(ref.w_in, c.groups)
    ref_loss = (ref.stats["loss_attr"] + ref.stats["loss_rep"]) / (ref.stats["pairs"] + ref.stats["rep_pairs"])
    dev, _, _ = _setup(c, precision)
    dev.set_sampler(H.unigram_entries(c.counts))
    csr = H.subsample(c.stream, c.counts, c.train_words, w.sample, seed=11)
    dev.set_corpus(csr.tokens, csr.sent_off, csr.sent_raw)
    st = dev.train(mode, 0, csr.n_sent, w.alpha, c.train_words, compute_loss=True, seed=5)
    g_in, _ = dev.get_weights()
    dev.close()
    assert np.isfinite(g_in).all()
    loss = (st["loss_attr"] + st["loss_rep"]) / (st["pairs"] + st["rep_pairs"])
    recall = E.planted_cluster_recall(g_in, c.groups)
    print(f"loss/pair {loss:.5f} vs oracle {ref_loss:.5f}; recall {recall:.4f} vs {ref_recall:.4f}; "
          f"tokens {st['tokens']} vs {ref.stats['trained_tokens']}", flush=True)
    assert abs(st["tokens"] - ref.stats["trained_tokens"]) / ref.stats["trained_tokens"] < 0.01
    assert abs(loss - ref_loss) / ref_loss < 0.01
    assert recall >= ref_recall - 0.02
