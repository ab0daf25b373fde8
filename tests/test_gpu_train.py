"""GPU parity of the training kernels against the oracle (CPU restatement of
PAR_Word2Vec_cpu.cpp).  Every call goes through the C-ABI of libpar_sgns.so.

Bars:
  deterministic mode           W_in/W_out bit-identical to the oracle; FP64 loss
                               within 1e-9 relative (only exp/log1p differ)
  Hogwild kernel, exact plan   one wavefront executing the oracle's own random
                               schedule: tables within 1e-4 of the update scale,
                               losses within 1e-6 relative (only the order of
                               float additions differs)
  Hogwild kernel, Philox       its own random streams: loss per pair within 1%
                               of the oracle's over the epoch and per tenth of
                               the corpus within 3%, planted-cluster recall@10
                               within 0.015 of the oracle's
"""
import numpy as np
import pytest

import oracle as O
from par_word2vec_amd import corpora, evaluate as E, host as H
from par_word2vec_amd.device import DeviceError, SgnsDevice
from par_word2vec_amd.trainer import chunk_bounds

pytestmark = pytest.mark.gpu


def _small_corpus(seed=7, n=60_000, V=1500, line=200):
    raw, _ = H.synth_zipf(n, V, 1.0, line_len=line, n_clusters=5, cluster_size=10, p_replace=0.3, seed=seed)
    vocab = H.Vocab.from_raw(raw, V, 5)
    return vocab, H.map_raw_stream(vocab, raw)


@pytest.fixture(scope="module")
def zipf1m():
    c = corpora.make("zipf1m")
    w = c.workload
    bucket = 100_100
    ref = O.train(c.stream, c.counts, c.train_words, w.dim, w.window, w.negative, w.batch_size, 1, w.alpha,
                  w.sample, bucket_words=bucket, n_buckets=10)
    return c, ref, bucket


@pytest.mark.parametrize("dim,window,negative,batch,iters,sample", [
    (100, 5, 5, 24, 1, 1e-3),
    (64, 3, 2, 7, 2, 1e-4),
    (130, 8, 5, 16, 1, 0.0),
    (300, 5, 10, 24, 1, 1e-4),
    (512, 3, 5, 24, 1, 1e-3),   # the largest dimension
    (64, 31, 3, 64, 1, 1e-3),   # the largest window and batch
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


def test_det_zipf1m_full(zipf1m):
    """The 1M-token workload end to end, deterministic mode vs oracle."""
    c, ref, _ = zipf1m
    w = c.workload
    plan = H.build_plan(c.stream, c.counts, c.train_words, w.window, w.negative, w.batch_size, 1, w.alpha, w.sample)
    with SgnsDevice(len(c.vocab), w.dim, w.window, w.negative, w.batch_size, "fp32") as dev:
        dev.set_exp_table(H.exp_table())
        dev.set_weights(*H.init_net(len(c.vocab), w.dim))
        dev.set_plan(plan)
        st = dev.train("det", 0, plan.n_sent, w.alpha, c.train_words, compute_loss=True)
        g_in, g_out = dev.get_weights()
    rel = np.abs(g_in - ref.w_in).max() / np.abs(ref.w_in).max()
    assert rel <= 1e-5
    assert np.array_equal(g_in, ref.w_in) and np.array_equal(g_out, ref.w_out)
    assert st["pairs"] == ref.stats["pairs"]


@pytest.mark.parametrize("precision_dim,window,negative,batch,sample", [
    ((100, "fp32"), 5, 5, 24, 1e-3),
    ((300, "fp32"), 5, 10, 24, 1e-3),
    ((128, "fp32"), 8, 5, 16, 1e-4),
    ((100, "fp32"), 5, 0, 24, 1e-3),
    ((75, "fp32"), 4, 6, 20, 1e-3),    # odd dimension: a pair straddles the row's end
    ((200, "fp32"), 5, 5, 24, 1e-3),   # two pairs per lane, lanes 36..63 past the row
    ((512, "fp32"), 3, 5, 24, 1e-3),   # the largest dimension: four pairs per lane
    ((64, "fp32"), 31, 3, 64, 1e-3),   # the largest window (63 ring slots) and batch
    ((48, "fp32"), 6, 64, 8, 1e-3),    # the most negatives: up to 704 per batch
])
def test_hogwild_kernel_exact_schedule(precision_dim, window, negative, batch, sample):
    """The Hogwild kernel, one wavefront, driven by the oracle's schedule:
    window ring, repeated-centre carry and snapshot repulsion must reproduce
    the reference arithmetic up to the order of float additions."""
    dim, precision = precision_dim
    print(f"[planned] dim={dim} W={window} neg={negative} B={batch}", flush=True)
    vocab, stream = _small_corpus()
    counts, T = vocab.counts(), vocab.train_words
    V = len(vocab)
    plan = H.build_plan(stream, counts, T, window, negative, batch, 1, 0.025, sample)
    w_in0, _ = H.init_net(V, dim)
    w_out0 = (np.random.default_rng(1).random((V, dim)).astype(np.float32) - 0.5) / dim
    res = {}
    for mode in ("det", "persistent"):
        with SgnsDevice(V, dim, window, negative, batch, precision) as dev:
            dev.set_exp_table(H.exp_table())
            dev.set_weights(w_in0, w_out0)
            dev.set_plan(plan)
            st = dev.train(mode, 0, plan.n_sent, 0.025, T, compute_loss=True, blocks=1, waves_per_block=1,
                           planned=mode != "det")
            res[mode] = dev.get_weights() + (st,)
    (di, do, ds), (fi, fo, fs) = res["det"], res["persistent"]
    assert np.abs(fi - di).max() <= 1e-4 * np.abs(di - w_in0).max()
    assert np.abs(fo - do).max() <= 1e-4 * np.abs(do - w_out0).max()
    for k in ("pairs", "negatives", "rep_pairs", "tokens"):
        assert fs[k] == ds[k]
    # the float-reordering error of a repulsion sum grows like the square root
    # of its length, K <= (2W-1)*negative terms per batch row: 1e-6 up to 90
    # terms (every workload of the reference's flags), scaled beyond
    rtol = 1e-6 * max(1.0, ((2 * window - 1) * negative / 90) ** 0.5)
    np.testing.assert_allclose(fs["loss_attr"], ds["loss_attr"], rtol=rtol)
    np.testing.assert_allclose(fs["loss_rep"], ds["loss_rep"], rtol=rtol)


def _planned_runs(vocab, stream, dim, window, negative, batch, sample, runs):
    """Train one epoch of the oracle's exact schedule in one wavefront for each
    (mode, precision, repulsion) in `runs`, from the same initial tables."""
    counts, T = vocab.counts(), vocab.train_words
    V = len(vocab)
    plan = H.build_plan(stream, counts, T, window, negative, batch, 1, 0.025, sample)
    # initial tables exactly representable in bf16, so every run starts equal
    w_in0, _ = H.init_net(V, dim)
    w_out0 = (np.random.default_rng(1).random((V, dim)).astype(np.float32) - 0.5) / dim
    w_in0, w_out0 = ((x.view(np.uint32) & np.uint32(0xFFFF0000)).view(np.float32) for x in (w_in0, w_out0))
    out = {}
    for mode, precision, rep in runs:
        with SgnsDevice(V, dim, window, negative, batch, precision, rounding="nearest") as dev:
            dev.set_exp_table(H.exp_table())
            dev.set_weights(w_in0, w_out0)
            dev.set_plan(plan)
            st = dev.train(mode, 0, plan.n_sent, 0.025, T, compute_loss=True, blocks=1, waves_per_block=1,
                           planned=mode != "det", repulsion=rep)
            out[(mode, precision, rep)] = dev.get_weights() + (st,)
    return out, w_in0, w_out0


def _rms(x):
    return float(np.sqrt(np.mean(np.square(x, dtype=np.float64))))


@pytest.mark.parametrize("dim,window,negative,batch,single", [
    (300, 5, 10, 24, False),
    (100, 5, 5, 24, False),
    (128, 8, 5, 32, False),   # the largest batch of the matrix path
    (64, 8, 25, 24, False),   # up to 375 negatives: several negative groups
    (300, 5, 10, 24, True),   # one-word sentences: repulsion only
    (75, 4, 6, 20, False),    # odd dimension: the last written pair straddles d
    (512, 3, 5, 32, False),   # the largest dimension on the matrix path
])
def test_matrix_repulsion_exact_schedule(dim, window, negative, batch, single):
    """bf16 tables, round-to-nearest write-backs, the oracle's exact schedule
    in one wavefront.  The matrix-core repulsion (G on bf16 MFMA with fp32
    sums, the gradient rounded to bf16 for dC and dN, one write-back per row
    and batch) against the fp32 deterministic run: its error must stay at the
    bf16 storage scale (RMS within 8% of the RMS update: round-to-nearest
    drops updates below half an ulp, the dominant error of bf16 tables) and be
    no larger than that of the vector repulsion on the same bf16 tables, which
    does the same arithmetic in fp32 but rounds a row once per block of
    negatives.  Losses are evaluated on the bf16 tables, which differ from the
    fp32 run at that storage scale.  With one-word sentences every update of a
    frequent negative is a write-back of its own, so W_out's storage error is
    that of round-to-nearest alone and equal for both paths (so are the
    losses, about 2.5% above fp32): only the relative bars apply to it."""
    print(f"[matrix] dim={dim} W={window} neg={negative} B={batch} single={single}", flush=True)
    if single:
        raw, _ = H.synth_zipf(20_000, 1500, 1.0, line_len=1, seed=9)
        vocab = H.Vocab.from_raw(raw, 1500, 5)
        stream = H.map_raw_stream(vocab, raw)
    else:
        vocab, stream = _small_corpus()
    runs = [("det", "fp32", "auto"), ("persistent", "bf16", "matrix"), ("persistent", "bf16", "vector")]
    res, w_in0, w_out0 = _planned_runs(vocab, stream, dim, window, negative, batch, 1e-3, runs)
    di, do, ds = res[runs[0]]
    mi, mo, ms = res[runs[1]]
    vi, vo, vs = res[runs[2]]
    for k in ("pairs", "negatives", "rep_pairs", "tokens"):
        assert ms[k] == ds[k] == vs[k]
    for name, det, mat, vec, w0 in (("W_in", di, mi, vi, w_in0), ("W_out", do, mo, vo, w_out0)):
        upd = _rms(det - w0)
        em, ev = _rms(mat - det), _rms(vec - det)
        print(f"  {name}: rms update {upd:.3e}, matrix err {em:.3e}, vector err {ev:.3e}", flush=True)
        if not (single and name == "W_out"):
            assert em <= 0.08 * upd
        assert em <= ev + 1e-3 * upd
    for k in ("loss_attr", "loss_rep"):
        print(f"  {k}: det {ds[k]:.6e} matrix {ms[k]:.6e} vector {vs[k]:.6e}", flush=True)
        if single:  # storage-dominated: as close to fp32 as the vector path, within 5%
            assert abs(ms[k] - ds[k]) <= abs(vs[k] - ds[k]) + 1e-3 * abs(ds[k])
            np.testing.assert_allclose(ms[k], ds[k], rtol=5e-2)
        else:
            np.testing.assert_allclose(ms[k], ds[k], rtol=5e-3)


def test_matrix_repulsion_batch_limit():
    """Batches above 32 positions keep one snapshot only on the vector path:
    the matrix path is refused and AUTO runs the vector path."""
    vocab, stream = _small_corpus()
    runs = [("persistent", "bf16", "auto"), ("persistent", "bf16", "vector")]
    res, _, _ = _planned_runs(vocab, stream, 64, 5, 5, 40, 1e-3, runs)
    (ai, ao, _), (vi, vo, _) = res[runs[0]], res[runs[1]]
    assert np.array_equal(ai, vi) and np.array_equal(ao, vo)
    with SgnsDevice(len(vocab), 64, 5, 5, 40, "bf16") as dev:
        dev.set_exp_table(H.exp_table())
        dev.set_weights(*H.init_net(len(vocab), 64))
        dev.set_plan(H.build_plan(stream, vocab.counts(), vocab.train_words, 5, 5, 40, 1, 0.025, 1e-3))
        with pytest.raises(DeviceError, match="batch size"):
            dev.train("persistent", 0, 10, 0.025, vocab.train_words, planned=True, repulsion="matrix")


def test_matrix_repulsion_needs_bf16():
    vocab, stream = _small_corpus()
    plan = H.build_plan(stream, vocab.counts(), vocab.train_words, 5, 5, 24, 1, 0.025, 1e-3)
    with SgnsDevice(len(vocab), 64, 5, 5, 24, "fp32") as dev:
        dev.set_exp_table(H.exp_table())
        dev.set_weights(*H.init_net(len(vocab), 64))
        dev.set_plan(plan)
        with pytest.raises(DeviceError, match="bf16"):
            dev.train("persistent", 0, plan.n_sent, 0.025, vocab.train_words, planned=True, repulsion="matrix")


def test_hogwild_philox_loss_curve_and_recall(zipf1m):
    """fp32 persistent kernel with its own Philox streams and on-device alias
    sampling, trained chunk by chunk (a tenth of the corpus each, the default
    grid): loss per pair of every tenth and of the epoch against the oracle's
    loss curve, and planted-cluster recall."""
    c, ref, bucket = zipf1m
    w = c.workload
    ref_recall = E.planted_cluster_recall(ref.w_in, c.groups)
    ref_total = (ref.stats["loss_attr"] + ref.stats["loss_rep"]) / (ref.stats["pairs"] + ref.stats["rep_pairs"])
    curve_loss, curve_pairs = ref.curve
    csr = H.subsample(c.stream, c.counts, c.train_words, w.sample, seed=11)
    chunks = chunk_bounds(csr.sent_raw, bucket)
    assert len(chunks) == 10
    with SgnsDevice(len(c.vocab), w.dim, w.window, w.negative, w.batch_size, "fp32") as dev:
        dev.set_exp_table(H.exp_table())
        dev.set_weights(*H.init_net(len(c.vocab), w.dim))
        dev.set_sampler(H.unigram_entries(c.counts))
        dev.set_corpus(csr.tokens, csr.sent_off, csr.sent_raw)
        per = [dev.train("persistent", b, e, w.alpha, c.train_words, compute_loss=True, seed=5) for b, e in chunks]
        g_in, _ = dev.get_weights()
    gl = np.array([(s["loss_attr"] + s["loss_rep"]) / (s["pairs"] + s["rep_pairs"]) for s in per])
    rl = curve_loss / curve_pairs
    total = sum(s["loss_attr"] + s["loss_rep"] for s in per) / sum(s["pairs"] + s["rep_pairs"] for s in per)
    recall = E.planted_cluster_recall(g_in, c.groups)
    print(f"loss/pair {total:.5f} vs {ref_total:.5f}; curve rel err {np.round(gl / rl - 1, 4)}; "
          f"recall {recall:.4f} vs {ref_recall:.4f}", flush=True)
    assert abs(total / ref_total - 1) < 0.01
    assert np.all(np.abs(gl / rl - 1) < 0.03)
    assert abs(recall - ref_recall) < 0.015
    assert sum(s["tokens"] for s in per) == csr.tokens.size


def test_plain_mode_counts(zipf1m):
    """One workgroup per sentence: every sentence and token is trained once;
    counters are consistent with the CSR."""
    c, _, _ = zipf1m
    w = c.workload
    csr = H.subsample(c.stream, c.counts, c.train_words, w.sample, seed=11)
    with SgnsDevice(len(c.vocab), w.dim, w.window, w.negative, w.batch_size, "fp32") as dev:
        dev.set_exp_table(H.exp_table())
        dev.set_weights(*H.init_net(len(c.vocab), w.dim))
        dev.set_sampler(H.unigram_entries(c.counts))
        dev.set_corpus(csr.tokens, csr.sent_off, csr.sent_raw)
        st = dev.train("plain", 0, csr.n_sent, w.alpha, c.train_words, compute_loss=True, seed=5)
        g_in, g_out = dev.get_weights()
    L = np.diff(csr.sent_off)
    assert st["sentences"] == csr.n_sent and st["tokens"] == csr.tokens.size
    assert st["batches"] == int(np.sum((L + w.batch_size - 1) // w.batch_size))
    assert st["negatives"] % w.negative == 0 and st["negatives"] > 0
    assert np.isfinite(g_in).all() and np.isfinite(g_out).all()


def test_sampler_distribution():
    """On-device alias table + Philox against the reference's unigram^0.75
    table distribution with "</s>" slots spread uniformly (cpu:606-608)."""
    c = corpora.make("zipf1m")
    w = c.workload
    V = len(c.vocab)
    ent = H.unigram_entries(c.counts)
    p = ent / ent.sum()
    q = p.copy()
    q[1:] += p[0] / (V - 1)
    q[0] = 0
    with SgnsDevice(V, w.dim, w.window, w.negative, w.batch_size, "fp32") as dev:
        dev.set_sampler(ent)
        s = dev.sample_negatives(40_000_000, seed=5)
    h = np.bincount(s, minlength=V) / s.size
    tv = 0.5 * np.abs(h - q).sum()
    assert h[0] == 0.0
    assert tv < 0.012, tv  # sampling noise alone is about 0.4 * sqrt(V / n) = 0.006
    assert np.abs(h - q).max() < 2e-4
