"""Loss-forward and adjoint kernels on the 65,536-window workload: windows of
one centre, 2x5 contexts and 5 shared negatives over a 100,000 x 300 table,
both tables U(-0.5/300, 0.5/300).

Bars:
  forward, FP32 accumulation   within 1e-6 relative of the FP64 oracle
  forward, FP64 accumulation   within 1e-11 relative of the FP64 oracle
  adjoint (FP32 atomics)       within 1e-5 of max|grad| of the FP64 oracle gradient
  adjoint vs finite differences   FP64 central differences (h = 1e-4) on 1,000
                               random coordinates of referenced rows, 1e-4
                               relative + 1e-7 absolute
  adjoint vs the CPU pair update  -alpha * dL/dW equals the oracle's
                               attraction update (cpu:551-565) up to the
                               sigmoid table's quantisation (1% relative)
  matrix cores (bf16 tables)   forward within 1e-6 relative (exact bf16 products,
                               fp32 sums); adjoint within 1e-4 of max|grad| (the
                               logit gradient enters the products as bf16 hi + lo)
  matrix cores (fp32 tables)   forward within 1e-6 relative, adjoint within 1e-5 of
                               max|grad| (fp32 products and sums throughout)
"""
import numpy as np
import pytest

import oracle as O
from par_word2vec_amd.device import DeviceError, SgnsDevice, WIN_FP64, WIN_MFMA

pytestmark = pytest.mark.gpu

V, D, NW, NCTX, NNEG = 100_000, 300, 65_536, 10, 5


def bf16_round(x):
    b = np.ascontiguousarray(x, np.float32).view(np.uint32)
    b = (b + np.uint32(0x7FFF) + ((b >> np.uint32(16)) & np.uint32(1))) & np.uint32(0xFFFF0000)
    return b.view(np.float32)


@pytest.fixture(scope="module")
def windows():
    rng = np.random.default_rng(2024)
    w_in = rng.uniform(-0.5 / D, 0.5 / D, (V, D)).astype(np.float32)
    w_out = rng.uniform(-0.5 / D, 0.5 / D, (V, D)).astype(np.float32)
    centre = rng.integers(1, V, NW).astype(np.int32)
    ctx = rng.integers(1, V, (NW, NCTX)).astype(np.int32)
    neg = rng.integers(1, V, (NW, NNEG)).astype(np.int32)
    return w_in, w_out, centre, ctx, neg


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_forward_matches_fp64_oracle(windows, precision):
    w_in, w_out, centre, ctx, neg = windows
    if precision == "bf16":
        w_in, w_out = bf16_round(w_in), bf16_round(w_out)
    ref = O.window_loss_f64(centre, ctx, neg, w_in, w_out)
    with SgnsDevice(V, D, 5, NNEG, 24, precision) as dev:
        dev.set_weights(w_in, w_out)
        l32 = dev.loss_forward(centre, ctx, neg)
        l64 = dev.loss_forward(centre, ctx, neg, flags=WIN_FP64)
    print(f"[{precision}] loss fp32 {l32:.9f} fp64 {l64:.12f} oracle {ref:.12f}", flush=True)
    assert l32 == pytest.approx(ref, rel=1e-6)
    assert l64 == pytest.approx(ref, rel=1e-11)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_adjoint_matches_fp64_gradient(windows, precision):
    w_in, w_out, centre, ctx, neg = windows
    if precision == "bf16":
        w_in, w_out = bf16_round(w_in), bf16_round(w_out)
    gi_ref, go_ref = O.window_grad_f64(centre, ctx, neg, w_in, w_out)
    with SgnsDevice(V, D, 5, NNEG, 24, precision) as dev:
        dev.set_weights(w_in, w_out)
        loss = dev.loss_adjoint(centre, ctx, neg)
        gi, go = dev.get_grads()
        ms, _ = dev.last_kernel_ms(2)
    print(f"[{precision}] adjoint {ms:.3f} ms; max|dgi| {np.abs(gi - gi_ref).max():.3e} of {np.abs(gi_ref).max():.3e}",
          flush=True)
    assert loss == pytest.approx(O.window_loss_f64(centre, ctx, neg, w_in, w_out), rel=1e-6)
    assert np.abs(gi - gi_ref).max() <= 1e-5 * np.abs(gi_ref).max()
    assert np.abs(go - go_ref).max() <= 1e-5 * np.abs(go_ref).max()


def test_matrix_forward_and_adjoint(windows):
    """The matrix-core window path against the FP64 oracle on the same bf16
    tables, and its time beside the vector path."""
    w_in, w_out, centre, ctx, neg = windows
    w_in, w_out = bf16_round(w_in), bf16_round(w_out)
    ref = O.window_loss_f64(centre, ctx, neg, w_in, w_out)
    gi_ref, go_ref = O.window_grad_f64(centre, ctx, neg, w_in, w_out)
    with SgnsDevice(V, D, 5, NNEG, 24, "bf16") as dev:
        dev.set_weights(w_in, w_out)
        lf = dev.loss_forward(centre, ctx, neg, flags=WIN_MFMA)
        f_ms, _ = dev.last_kernel_ms(1)
        dev.loss_forward(centre, ctx, neg)
        fv_ms, _ = dev.last_kernel_ms(1)
        la = dev.loss_adjoint(centre, ctx, neg, flags=WIN_MFMA)
        a_ms, _ = dev.last_kernel_ms(2)
        gi, go = dev.get_grads()
        dev.loss_adjoint(centre, ctx, neg)
        av_ms, _ = dev.last_kernel_ms(2)
    # useful flops: 2*nctx*(1+nneg)*d per window forward, twice that for the two adjoint products
    f_flop = 2.0 * NCTX * (1 + NNEG) * D * NW
    print(f"[matrix] forward {f_ms:.3f} ms (vector {fv_ms:.3f}), adjoint {a_ms:.3f} ms (vector {av_ms:.3f}); "
          f"useful {f_flop / f_ms / 1e9:.2f} / {3 * f_flop / a_ms / 1e9:.2f} TFLOP/s of 2500 bf16; "
          f"max|dgi| {np.abs(gi - gi_ref).max():.3e} of {np.abs(gi_ref).max():.3e}", flush=True)
    assert lf == pytest.approx(ref, rel=1e-6)
    assert la == pytest.approx(ref, rel=1e-6)
    assert np.abs(gi - gi_ref).max() <= 1e-4 * np.abs(gi_ref).max()
    assert np.abs(go - go_ref).max() <= 1e-4 * np.abs(go_ref).max()


def test_fp32_matrix_forward_and_adjoint(windows):
    """The one-tile window kernel on fp32 tables (fp32 matrix cores) against
    the FP64 oracle, at the fp32 vector path's bars."""
    w_in, w_out, centre, ctx, neg = windows
    ref = O.window_loss_f64(centre, ctx, neg, w_in, w_out)
    gi_ref, go_ref = O.window_grad_f64(centre, ctx, neg, w_in, w_out)
    with SgnsDevice(V, D, 5, NNEG, 24, "fp32") as dev:
        dev.set_weights(w_in, w_out)
        lf = dev.loss_forward(centre, ctx, neg, flags=WIN_MFMA)
        f_ms, _ = dev.last_kernel_ms(1)
        dev.loss_forward(centre, ctx, neg)
        fv_ms, _ = dev.last_kernel_ms(1)
        la = dev.loss_adjoint(centre, ctx, neg, flags=WIN_MFMA)
        a_ms, _ = dev.last_kernel_ms(2)
        gi, go = dev.get_grads()
    print(f"[fp32 matrix] forward {f_ms:.3f} ms (vector {fv_ms:.3f}), adjoint {a_ms:.3f} ms; "
          f"max|dgi| {np.abs(gi - gi_ref).max():.3e} of {np.abs(gi_ref).max():.3e}", flush=True)
    assert lf == pytest.approx(ref, rel=1e-6)
    assert la == pytest.approx(ref, rel=1e-6)
    assert np.abs(gi - gi_ref).max() <= 1e-5 * np.abs(gi_ref).max()
    assert np.abs(go - go_ref).max() <= 1e-5 * np.abs(go_ref).max()


def test_matrix_window_errors(windows):
    w_in, w_out, centre, ctx, neg = windows
    with SgnsDevice(V, D, 5, NNEG, 24, "fp32") as dev:
        dev.set_weights(w_in, w_out)
        wide = np.concatenate([ctx[:8], ctx[:8]], axis=1)  # 1 + 20 + 5 rows
        with pytest.raises(DeviceError, match="16 rows"):
            dev.loss_forward(centre[:8], wide, neg[:8], flags=WIN_MFMA)
    with SgnsDevice(V, D, 5, NNEG, 24, "bf16") as dev:
        dev.set_weights(w_in, w_out)
        with pytest.raises(DeviceError, match="fp32"):
            dev.loss_forward(centre[:8], ctx[:8], neg[:8], flags=WIN_MFMA | WIN_FP64)


def test_adjoint_finite_differences(windows):
    """FP64 central differences of the oracle's loss restricted to the
    windows that touch the perturbed row (all other terms cancel)."""
    w_in, w_out, centre, ctx, neg = windows
    with SgnsDevice(V, D, 5, NNEG, 24, "fp32") as dev:
        dev.set_weights(w_in, w_out)
        dev.loss_adjoint(centre, ctx, neg)
        gi, go = dev.get_grads()
    rng = np.random.default_rng(7)
    h = 1e-4
    a = w_in.astype(np.float64)
    b = w_out.astype(np.float64)
    bad = []
    for t in range(1000):
        table_in = t % 2 == 0
        if table_in:
            r = int(ctx.reshape(-1)[rng.integers(0, ctx.size)])
            hit = np.flatnonzero((ctx == r).any(1))
        else:
            pool = np.concatenate([centre, neg.reshape(-1)])
            r = int(pool[rng.integers(0, pool.size)])
            hit = np.flatnonzero((centre == r) | (neg == r).any(1))
        j = int(rng.integers(0, D))
        tab = a if table_in else b
        old = tab[r, j]
        tab[r, j] = old + h
        lp = O.window_loss_f64(centre[hit], ctx[hit], neg[hit], a, b)
        tab[r, j] = old - h
        lm = O.window_loss_f64(centre[hit], ctx[hit], neg[hit], a, b)
        tab[r, j] = old
        fd = (lp - lm) / (2 * h)
        g = float(gi[r, j] if table_in else go[r, j])
        if abs(fd - g) > 1e-4 * abs(fd) + 1e-7:
            bad.append((table_in, r, j, fd, g))
    assert not bad, bad[:5]


def test_adjoint_vs_cpu_pair_update():
    """Sentence [a, b], window 1, no negatives: the oracle applies the pair
    (context b, centre a) first; rows W_out[a] and W_in[b] change only there.
    Its update equals -alpha * dL/dW of the window (centre a, context b) with
    the table sigmoid in place of the exact one."""
    d, alpha = 300, 0.025
    rng = np.random.default_rng(3)
    w_in = rng.uniform(-0.5, 0.5, (3, d)).astype(np.float32) / 8
    w_out = rng.uniform(-0.5, 0.5, (3, d)).astype(np.float32) / 8
    r = O.train(np.array([1, 2, 0], np.int32), np.array([1, 1, 1], np.uint32), 3, d, 1, 5, 24, 1, alpha, 0.0,
                w_in=w_in, w_out=w_out, table=np.zeros(100_000_000, np.int32))
    upd_out = r.w_out[1].astype(np.float64) - w_out[1]
    upd_in = r.w_in[2].astype(np.float64) - w_in[2]
    with SgnsDevice(3, d, 1, 0, 24, "fp32") as dev:
        dev.set_weights(w_in, w_out)
        dev.loss_adjoint(np.array([1], np.int32), np.array([[2]], np.int32), np.zeros((1, 0), np.int32))
        gi, go = dev.get_grads()
    np.testing.assert_allclose(-alpha * go[1], upd_out, rtol=1e-2, atol=1e-9)
    np.testing.assert_allclose(-alpha * gi[2], upd_in, rtol=1e-2, atol=1e-9)
    assert np.abs(upd_out).max() > 1e-6


def test_gradient_error_per_batch(windows):
    """Per-batch gradient error of each storage/arithmetic choice against the
    FP64 gradient of the unrounded fp32 tables (one batch of 65,536 windows):
    fp32 tables with fp32 sums, bf16 storage with fp32 sums, and bf16 storage
    on the matrix cores.  Relative L2 error over both gradient tables; the bf16
    rows carry the storage rounding of the tables (2^-9 relative per entry)."""
    w_in, w_out, centre, ctx, neg = windows
    gi_ref, go_ref = O.window_grad_f64(centre, ctx, neg, w_in, w_out)
    ref = np.concatenate([gi_ref.ravel(), go_ref.ravel()])
    errs = {}
    for name, prec, flags in [("fp32", "fp32", 0), ("bf16", "bf16", 0), ("bf16-matrix", "bf16", WIN_MFMA)]:
        with SgnsDevice(V, D, 5, NNEG, 24, prec) as dev:
            dev.set_weights(w_in, w_out)  # bf16 handles round on upload
            dev.loss_adjoint(centre, ctx, neg, flags=flags)
            gi, go = dev.get_grads()
        g = np.concatenate([gi.ravel(), go.ravel()]).astype(np.float64)
        errs[name] = (np.linalg.norm(g - ref) / np.linalg.norm(ref), np.abs(g - ref).max() / np.abs(ref).max())
    print("[grad error per batch] " + "; ".join(f"{k}: rel L2 {a:.2e}, max {b:.2e}" for k, (a, b) in errs.items()),
          flush=True)
    assert errs["fp32"][0] < 1e-5
    assert errs["bf16"][0] < 1e-2 and errs["bf16-matrix"][0] < 1e-2


@pytest.mark.parametrize("nctx,nneg,dim,vocab", [
    (1, 0, 64, 5000),     # one-tile kernel, two rows, 2 steps of 32 dimensions
    (7, 7, 130, 5000),    # one-tile kernel, 15 rows, odd padding
    (15, 0, 500, 5000),   # one-tile kernel, 16 rows, 16 steps
    (10, 5, 300, 40),     # one-tile kernel, rows repeated inside windows (atomics add them all)
    (12, 7, 300, 5000),   # 20 rows: one context tile + the Y tile
    (24, 3, 200, 5000),   # two context tiles
    (32, 7, 512, 5000),   # largest window and row: two context tiles, 16 steps, two waves per block
])
def test_matrix_window_shapes(nctx, nneg, dim, vocab):
    """Both matrix-core window kernels (one 16-row tile when 1 + nctx + nneg
    <= 16, else context tiles + a Y tile) against the FP64 oracle on the same
    bf16 tables, at the matrix path's bars; windows of at most 16 rows also on
    fp32 tables (fp32 matrix cores) at the fp32 bars."""
    rng = np.random.default_rng(nctx * 100 + nneg)
    nw = 4096
    w_in = bf16_round(rng.uniform(-0.5, 0.5, (vocab, dim)).astype(np.float32) / 16)
    w_out = bf16_round(rng.uniform(-0.5, 0.5, (vocab, dim)).astype(np.float32) / 16)
    centre = rng.integers(0, vocab, nw).astype(np.int32)
    ctx = rng.integers(0, vocab, (nw, nctx)).astype(np.int32)
    neg = rng.integers(0, vocab, (nw, nneg)).astype(np.int32)
    ref = O.window_loss_f64(centre, ctx, neg, w_in, w_out)
    gi_ref, go_ref = O.window_grad_f64(centre, ctx, neg, w_in, w_out)
    with SgnsDevice(vocab, dim, 5, max(nneg, 1), 24, "bf16") as dev:
        dev.set_weights(w_in, w_out)
        lf = dev.loss_forward(centre, ctx, neg, flags=WIN_MFMA)
        la = dev.loss_adjoint(centre, ctx, neg, flags=WIN_MFMA)
        gi, go = dev.get_grads()
    print(f"[matrix {nctx}x{nneg} d{dim} V{vocab}] loss {lf:.9f} / {la:.9f} oracle {ref:.9f}; "
          f"max|dgi| {np.abs(gi - gi_ref).max():.3e}, max|dgo| {np.abs(go - go_ref).max():.3e}", flush=True)
    assert lf == pytest.approx(ref, rel=1e-6)
    assert la == pytest.approx(ref, rel=1e-6)
    assert np.abs(gi - gi_ref).max() <= 1e-4 * np.abs(gi_ref).max()
    assert np.abs(go - go_ref).max() <= 1e-4 * np.abs(go_ref).max()
    if 1 + nctx + nneg > 16:
        return
    with SgnsDevice(vocab, dim, 5, max(nneg, 1), 24, "fp32") as dev:
        dev.set_weights(w_in, w_out)
        lf = dev.loss_forward(centre, ctx, neg, flags=WIN_MFMA)
        la = dev.loss_adjoint(centre, ctx, neg, flags=WIN_MFMA)
        gi, go = dev.get_grads()
    print(f"[fp32 matrix {nctx}x{nneg} d{dim} V{vocab}] loss {lf:.9f} / {la:.9f}; "
          f"max|dgi| {np.abs(gi - gi_ref).max():.3e}, max|dgo| {np.abs(go - go_ref).max():.3e}", flush=True)
    assert lf == pytest.approx(ref, rel=1e-6)
    assert la == pytest.approx(ref, rel=1e-6)
    assert np.abs(gi - gi_ref).max() <= 1e-5 * np.abs(gi_ref).max()
    assert np.abs(go - go_ref).max() <= 1e-5 * np.abs(go_ref).max()


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
@pytest.mark.parametrize("flags", [0, WIN_MFMA])
def test_empty_window_batch(flags, precision):
    """No windows: zero loss and zero gradients on every path."""
    with SgnsDevice(1000, 300, 5, 5, 24, precision) as dev:
        rng = np.random.default_rng(5)
        dev.set_weights(rng.uniform(-0.1, 0.1, (1000, 300)).astype(np.float32),
                        rng.uniform(-0.1, 0.1, (1000, 300)).astype(np.float32))
        e = np.zeros(0, np.int32)
        assert dev.loss_forward(e, e.reshape(0, 10), e.reshape(0, 5), flags=flags) == 0.0
        assert dev.loss_adjoint(e, e.reshape(0, 10), e.reshape(0, 5), flags=flags) == 0.0
        gi, go = dev.get_grads()
        assert not gi.any() and not go.any()
