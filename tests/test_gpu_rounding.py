"""Stochastic rounding of bf16 tables, through the training kernel.

One wavefront trains the oracle's exact schedule with a learning rate so small
that every update is a small fraction of a bf16 ulp (about 0.3 %), starting
from the same bf16-representable tables on an fp32 handle and on bf16 handles.
Then, per table element, the bf16 change D16 is a sum of rounding events whose
expectation is the fp32 change D32 (unbiased rounding), while round-to-nearest
drops every such update (stagnation).

The statistics use the entries whose starting magnitude lies in [2^-8, 2^-6)
(the bulk of both tables: ulp 2^-15 or 2^-14), where one write-back adds at
most ~0.15 ulp, and that moved by at least half an ulp in fp32 over the pass.
Bars:
  stochastic rounding   over six rounding-noise seeds: the mean of the
                        least-squares slopes of D16 on D32 within max(5 %, 3
                        standard errors of that mean) of 1, and the mean of
                        (D16 - D32) / ulp within 4 standard errors of 0 -- on
                        the matrix-core and vector repulsion paths, with atomic
                        and plain-store write-backs.  One seed's slope spreads
                        by about +-0.05 around the mean (0.90-1.12 over eight
                        seeds on the vector path), so one seed cannot be held
                        to 5 %.
  round-to-nearest      slope below 0.5 (the test sees a biased rounding; it is
                        exactly 0 here, whatever the seed)
"""
import numpy as np
import pytest

from par_word2vec_amd import host as H
from par_word2vec_amd.device import SgnsDevice

pytestmark = pytest.mark.gpu

DIM, W, NEG, B = 64, 5, 5, 24
ALPHA = 5e-5


def _bf16(a):
    b = np.ascontiguousarray(a, np.float32).view(np.uint32)
    return ((b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000).view(np.float32)


def _ulp(x):
    """bf16 ulp at each value (2^(e-7) for |x| in [2^e, 2^(e+1)))."""
    e = np.floor(np.log2(np.maximum(np.abs(x), 1e-30)))
    return np.exp2(e - 7)


@pytest.fixture(scope="module")
def setup():
    raw, _ = H.synth_zipf(200_000, 800, 1.0, line_len=200, seed=41)
    vocab = H.Vocab.from_raw(raw, 800, 5)
    stream = H.map_raw_stream(vocab, raw)
    counts, T = vocab.counts(), vocab.train_words
    plan = H.build_plan(stream, counts, T, W, NEG, B, 1, ALPHA, 1e-3)
    V = len(vocab)
    rng = np.random.default_rng(3)
    w_in = _bf16(H.init_net(V, DIM)[0])
    w_out = _bf16((rng.random((V, DIM), dtype=np.float32) - 0.5) / DIM)
    with SgnsDevice(V, DIM, W, NEG, B, "fp32") as dev:
        dev.set_exp_table(H.exp_table())
        dev.set_weights(w_in, w_out)
        dev.set_plan(plan)
        dev.train("persistent", 0, plan.n_sent, ALPHA, T, planned=True, blocks=1, waves_per_block=1)
        f_in, f_out = dev.get_weights()
    return vocab, plan, T, (w_in, w_out), (f_in, f_out)


def _stats(d16, d32, w0):
    ulp = _ulp(w0)
    m = (np.abs(w0) >= 2.0 ** -8) & (np.abs(w0) < 2.0 ** -6) & (np.abs(d32) >= 0.5 * ulp)
    x, y = d32[m] / ulp[m], d16[m] / ulp[m]
    slope = float((x * y).sum() / (x * x).sum())
    r = y - x
    return slope, float(r.mean()), float(r.std() / np.sqrt(r.size)), int(m.sum())


@pytest.mark.parametrize("rounding,repulsion,write_back", [
    ("stochastic", "matrix", "atomic"),
    ("stochastic", "matrix", "store_out"),
    ("stochastic", "vector", "atomic"),
    ("nearest", "matrix", "atomic"),
])
def test_stochastic_rounding_is_unbiased(setup, rounding, repulsion, write_back):
    vocab, plan, T, (w_in, w_out), (f_in, f_out) = setup
    print(f"[rounding] {rounding} {repulsion} {write_back}", flush=True)
    d32 = np.concatenate([(f_in - w_in).ravel(), (f_out - w_out).ravel()]).astype(np.float64)
    w0 = np.concatenate([w_in.ravel(), w_out.ravel()]).astype(np.float64)
    slopes, means, ses = [], [], []
    for seed in range(1, 7 if rounding == "stochastic" else 2):  # the seed keys the rounding-noise streams
        with SgnsDevice(len(vocab), DIM, W, NEG, B, "bf16", rounding=rounding) as dev:
            dev.set_exp_table(H.exp_table())
            dev.set_weights(w_in, w_out)
            dev.set_plan(plan)
            dev.train("persistent", 0, plan.n_sent, ALPHA, T, planned=True, blocks=1, waves_per_block=1,
                      repulsion=repulsion, write_back=write_back, seed=seed)
            b_in, b_out = dev.get_weights()
        d16 = np.concatenate([(b_in - w_in).ravel(), (b_out - w_out).ravel()]).astype(np.float64)
        slope, mean, se, n = _stats(d16, d32, w0)
        assert n > 2000
        slopes.append(slope)
        means.append(mean)
        ses.append(se)
    k = len(slopes)
    slope, mean = float(np.mean(slopes)), float(np.mean(means))
    se_slope = float(np.std(slopes, ddof=1) / np.sqrt(k)) if k > 1 else 0.0
    se_mean = float(np.sqrt(np.sum(np.square(ses))) / k)
    print(f"[rounding] {rounding} {repulsion} {write_back}: slope {slope:.4f} (se {se_slope:.4f}; seeds "
          f"{np.round(slopes, 4).tolist()}), mean error {mean:+.4f} ulp (se {se_mean:.4f}) over {n} elements",
          flush=True)
    if rounding == "stochastic":
        assert abs(slope - 1) < max(0.05, 3 * se_slope)
        assert abs(mean) < 4 * se_mean
    else:
        assert slope < 0.5


@pytest.mark.parametrize("rounding", ["nearest_out", "nearest_in"])
def test_rounding_per_table(setup, rounding):
    """Round-to-nearest on one table only reaches that table alone: it
    stagnates (slope below 0.5 for every seed; exactly 0 measured, as with
    nearest on both), while the other keeps stochastic rounding (mean slope of
    three rounding-noise seeds above 0.75; one seed gave 0.88 and 1.06).  Unbiasedness itself is the test above: here the fp32 run is only
    approximately the stochastic table's reference, since the other table stays
    at its starting values and the gradients depend on it (an element of W_in
    drifts by up to a few percent of its value over the pass in fp32)."""
    vocab, plan, T, (w_in, w_out), (f_in, f_out) = setup
    s_in, s_out = [], []  # (slope, elements) per rounding-noise seed
    for seed in (1, 2, 3):
        with SgnsDevice(len(vocab), DIM, W, NEG, B, "bf16", rounding=rounding) as dev:
            dev.set_exp_table(H.exp_table())
            dev.set_weights(w_in, w_out)
            dev.set_plan(plan)
            dev.train("persistent", 0, plan.n_sent, ALPHA, T, planned=True, blocks=1, waves_per_block=1,
                      repulsion="matrix", write_back="atomic", seed=seed)
            b_in, b_out = dev.get_weights()
        a = _stats((b_in - w_in).ravel().astype(np.float64), (f_in - w_in).ravel().astype(np.float64),
                   w_in.ravel().astype(np.float64))
        b = _stats((b_out - w_out).ravel().astype(np.float64), (f_out - w_out).ravel().astype(np.float64),
                   w_out.ravel().astype(np.float64))
        s_in.append((a[0], a[3]))
        s_out.append((b[0], b[3]))
    print(f"[rounding] {rounding}: W_in slopes {[round(x, 4) for x, _ in s_in]} ({s_in[0][1]} elements), "
          f"W_out slopes {[round(x, 4) for x, _ in s_out]} ({s_out[0][1]} elements)", flush=True)
    sr, rn = (s_in, s_out) if rounding == "nearest_out" else (s_out, s_in)
    assert min(n for _, n in sr) > 1000 and min(n for _, n in rn) > 1000
    assert np.mean([x for x, _ in sr]) > 0.75
    assert max(x for x, _ in rn) < 0.5
