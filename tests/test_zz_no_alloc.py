"""The allocation-free window loss (par_sgns_loss_forward_noalloc) against
par_sgns_loss_forward, through the C-ABI.

Bar: the same FP64 loss up to the order of the per-wavefront FP64 atomic
additions (1e-12 relative), on two calls in a row, the second with a smaller
batch (the shared scratch keeps the first call's contents); no device memory
is allocated by the call; a batch past the scratch's capacity is rejected
with PAR_SGNS_EINVAL and a message naming the limit.
"""
import numpy as np
import pytest

from par_word2vec_amd import host as H
from par_word2vec_amd.device import WIN_MFMA, DeviceError, SgnsDevice

pytestmark = pytest.mark.gpu

V, DIM, NCTX, NNEG = 20_000, 300, 10, 5
LIMIT = (8 << 20) // 4 - 64  # indices the 8 MiB scratch stages after its 256-byte loss area


def _windows(n, seed):
    rng = np.random.default_rng(seed)
    return (rng.integers(0, V, n).astype(np.int32), rng.integers(0, V, n * NCTX).astype(np.int32),
            rng.integers(0, V, n * NNEG).astype(np.int32))


def _tables():
    rng = np.random.default_rng(5)
    return ((rng.random((V, DIM), dtype=np.float32) - 0.5) / DIM,
            (rng.random((V, DIM), dtype=np.float32) - 0.5) / DIM)


@pytest.mark.parametrize("precision,flags", [("fp32", 0), ("bf16", 0), ("bf16", WIN_MFMA)])
def test_noalloc_forward_matches_forward(precision, flags):
    import torch

    print(f"[noalloc] {precision} flags={flags}", flush=True)
    with SgnsDevice(V, DIM, 5, NNEG, 24, precision) as dev:
        dev.set_exp_table(H.exp_table())
        dev.set_weights(*_tables())
        big, small = _windows(65_536, 1), _windows(777, 2)
        warm = dev.loss_forward_noalloc(*small, flags=flags)  # code objects loaded before measuring
        free0 = torch.cuda.mem_get_info(0)[0]
        a1 = dev.loss_forward_noalloc(*big, flags=flags)
        a2 = dev.loss_forward_noalloc(*small, flags=flags)
        free1 = torch.cuda.mem_get_info(0)[0]
        b1 = dev.loss_forward(*big, flags=flags)
        b2 = dev.loss_forward(*small, flags=flags)
    assert free1 == free0, f"device memory changed by {free0 - free1} bytes"
    assert np.isfinite(a1) and a1 > 0 and np.isfinite(a2) and a2 > 0
    np.testing.assert_allclose(a1, b1, rtol=1e-12)
    np.testing.assert_allclose(a2, b2, rtol=1e-12)
    np.testing.assert_allclose(warm, b2, rtol=1e-12)


def test_noalloc_rejects_batches_past_the_scratch():
    print("[noalloc] limit", flush=True)
    n_ok = LIMIT // (1 + NCTX + NNEG)
    with SgnsDevice(V, DIM, 5, NNEG, 24, "fp32") as dev:
        dev.set_exp_table(H.exp_table())
        dev.set_weights(*_tables())
        ok = _windows(n_ok, 3)
        got = dev.loss_forward_noalloc(*ok)
        np.testing.assert_allclose(got, dev.loss_forward(*ok), rtol=1e-12)
        with pytest.raises(DeviceError, match="allocation-free entry point stages at most"):
            dev.loss_forward_noalloc(*_windows(n_ok + 1, 4))
        # the handle stays usable after the rejection
        small = _windows(100, 6)
        np.testing.assert_allclose(dev.loss_forward_noalloc(*small), dev.loss_forward(*small), rtol=1e-12)


def test_noalloc_two_handles_one_device():
    """Handles on one device share the scratch: serialised calls from two
    handles each get their own loss."""
    print("[noalloc] two handles", flush=True)
    w1 = _tables()
    w2 = (w1[0] * DIM, w1[1] * DIM)  # U(-0.5, 0.5): logits of order 1, a clearly different loss
    with SgnsDevice(V, DIM, 5, NNEG, 24, "fp32") as d1, SgnsDevice(V, DIM, 5, NNEG, 24, "fp32") as d2:
        for d, w in ((d1, w1), (d2, w2)):
            d.set_exp_table(H.exp_table())
            d.set_weights(*w)
        x = _windows(4096, 7)
        l1 = d1.loss_forward_noalloc(*x)
        l2 = d2.loss_forward_noalloc(*x)
        np.testing.assert_allclose(l1, d1.loss_forward(*x), rtol=1e-12)
        np.testing.assert_allclose(l2, d2.loss_forward(*x), rtol=1e-12)
        assert abs(l1 - l2) > 1e-6 * abs(l1)
