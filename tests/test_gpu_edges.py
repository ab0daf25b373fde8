"""Ragged and empty inputs through the C-ABI (reference rules: a sentence ends
at "</s>" or after 1000 kept words, cpu:447-471; an empty sentence trains
nothing, cpu:539-542, 575).

Bars:
  ragged corpus, deterministic mode   bit-identical tables to the oracle, same
                                      pair / negative / token counts
  ragged corpus, exact schedule       Hogwild kernel within the exact-schedule
                                      bars of test_gpu_train (one wavefront)
  ragged corpus, Philox draws         every kept token trained exactly once in
                                      plain and persistent modes, fp32 and bf16
  empty sentence ranges               every mode returns zero counts and leaves
                                      both tables bit-identical
"""
import numpy as np
import pytest

import oracle as O
from par_word2vec_amd import host as H
from par_word2vec_amd.device import DeviceError, SgnsDevice

pytestmark = pytest.mark.gpu

DIM, W, NEG, B = 64, 5, 5, 24


@pytest.fixture(scope="module")
def ragged():
    # line lengths log-normal(3, 1.6) clipped to [1, 2500]: one-word lines and
    # lines past the 1000-word cut; then empty lines after every 7th line end
    raw, _ = H.synth_zipf(120_000, 1500, 1.0, lognorm=(3.0, 1.6), len_clip=(1, 2500), seed=11)
    vocab = H.Vocab.from_raw(raw, 1500, 5)
    stream = H.map_raw_stream(vocab, raw)
    ends = np.flatnonzero(stream == 0)
    extra = ends[::7] + 1
    stream = np.insert(stream, extra, 0).astype(np.int32)
    lens = np.diff(np.concatenate([[-1], np.flatnonzero(stream == 0)])) - 1
    assert (lens == 0).sum() >= len(extra) and lens.max() > 1000 and (lens == 1).any()
    return vocab, stream


def _dev(vocab, precision, w_in, w_out):
    dev = SgnsDevice(len(vocab), DIM, W, NEG, B, precision)
    dev.set_exp_table(H.exp_table())
    dev.set_weights(w_in, w_out)
    return dev


def test_ragged_det_bitwise_vs_oracle(ragged):
    vocab, stream = ragged
    counts, T = vocab.counts(), vocab.train_words
    print("[edges] ragged det vs oracle", flush=True)
    plan = H.build_plan(stream, counts, T, W, NEG, B, 1, 0.025, 1e-3)
    ref = O.train(stream, counts, T, DIM, W, NEG, B, 1, 0.025, 1e-3)
    with _dev(vocab, "fp32", *H.init_net(len(vocab), DIM)) as dev:
        dev.set_plan(plan)
        st = dev.train("det", 0, plan.n_sent, 0.025, T, compute_loss=True)
        g_in, g_out = dev.get_weights()
    assert st["pairs"] == ref.stats["pairs"]
    assert st["negatives"] == ref.stats["negatives"]
    assert st["tokens"] == ref.stats["trained_tokens"]
    assert np.array_equal(g_in, ref.w_in) and np.array_equal(g_out, ref.w_out)


def test_ragged_exact_schedule(ragged):
    vocab, stream = ragged
    counts, T = vocab.counts(), vocab.train_words
    print("[edges] ragged exact schedule", flush=True)
    plan = H.build_plan(stream, counts, T, W, NEG, B, 1, 0.025, 1e-3)
    w_in0, _ = H.init_net(len(vocab), DIM)
    w_out0 = (np.random.default_rng(2).random((len(vocab), DIM)).astype(np.float32) - 0.5) / DIM
    res = {}
    for mode in ("det", "persistent"):
        with _dev(vocab, "fp32", w_in0, w_out0) as dev:
            dev.set_plan(plan)
            st = dev.train(mode, 0, plan.n_sent, 0.025, T, compute_loss=True, blocks=1, waves_per_block=1,
                           planned=mode != "det")
            res[mode] = dev.get_weights() + (st,)
    (di, do, ds), (fi, fo, fs) = res["det"], res["persistent"]
    assert np.abs(fi - di).max() <= 1e-4 * np.abs(di - w_in0).max()
    assert np.abs(fo - do).max() <= 1e-4 * np.abs(do - w_out0).max()
    for k in ("pairs", "negatives", "rep_pairs", "tokens"):
        assert fs[k] == ds[k]
    np.testing.assert_allclose(fs["loss_attr"], ds["loss_attr"], rtol=1e-6)
    np.testing.assert_allclose(fs["loss_rep"], ds["loss_rep"], rtol=1e-6)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_ragged_philox_counts(ragged, precision):
    vocab, stream = ragged
    counts, T = vocab.counts(), vocab.train_words
    csr = H.subsample(stream, counts, T, 1e-3, seed=5)
    lens = np.diff(csr.sent_off)
    assert lens.max() <= 1000
    n = lens.size
    for mode in ("plain", "persistent"):
        print(f"[edges] ragged philox {precision} {mode}", flush=True)
        with _dev(vocab, precision, *H.init_net(len(vocab), DIM)) as dev:
            dev.set_sampler(H.unigram_entries(counts))
            dev.set_corpus(csr.tokens, csr.sent_off, csr.sent_raw)
            st = dev.train(mode, 0, n, 0.025, T, seed=3, use_graph=mode == "persistent")
        assert st["tokens"] == csr.tokens.size
        assert st["sentences"] == int(np.count_nonzero(lens))


def test_empty_ranges_train_nothing(ragged):
    vocab, stream = ragged
    counts, T = vocab.counts(), vocab.train_words
    plan = H.build_plan(stream, counts, T, W, NEG, B, 1, 0.025, 1e-3)
    csr = H.subsample(stream, counts, T, 1e-3, seed=5)
    w_in0, _ = H.init_net(len(vocab), DIM)
    w_out0 = (np.random.default_rng(3).random((len(vocab), DIM)).astype(np.float32) - 0.5) / DIM
    for precision, mode, planned in [("fp32", "det", False), ("fp32", "plain", False), ("fp32", "persistent", True),
                                     ("bf16", "plain", False), ("bf16", "persistent", False)]:
        print(f"[edges] empty range {precision} {mode} planned={planned}", flush=True)
        with _dev(vocab, precision, w_in0, w_out0) as dev:
            dev.set_sampler(H.unigram_entries(counts))
            dev.set_corpus(csr.tokens, csr.sent_off, csr.sent_raw)
            dev.set_plan(plan)
            before = dev.get_weights()
            for b in (0, 7):
                st = dev.train(mode, b, b, 0.025, T, compute_loss=True, planned=planned,
                               use_graph=mode == "persistent")
                assert st["tokens"] == st["pairs"] == st["negatives"] == st["sentences"] == 0
                assert st["loss_attr"] == 0.0 and st["loss_rep"] == 0.0
            after = dev.get_weights()
        assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])


@pytest.mark.parametrize("floor", [1, 4, 100000])
def test_sentences_per_wave_floor(ragged, floor):
    """min_sentences_per_wave only shapes the default persistent grid: every
    kept token is still trained once, down to one sentence per wave and up to
    a floor that leaves a single wavefront.  The launch shape the call used
    (reported in the stats) follows the floor: at most n / floor waves."""
    vocab, stream = ragged
    counts, T = vocab.counts(), vocab.train_words
    csr = H.subsample(stream, counts, T, 1e-3, seed=5)
    n = csr.sent_off.size - 1
    print(f"[edges] floor {floor}", flush=True)
    with _dev(vocab, "bf16", *H.init_net(len(vocab), DIM)) as dev:
        dev.set_sampler(H.unigram_entries(counts))
        dev.set_corpus(csr.tokens, csr.sent_off, csr.sent_raw)
        base = dev.train("persistent", 0, n, 0.025, T, seed=3, use_graph=True)  # default floor, 24
        st = dev.train("persistent", 0, n, 0.025, T, seed=3, use_graph=True, min_sentences_per_wave=floor)
        assert st["tokens"] == csr.tokens.size
        waves, base_waves = st["blocks"] * st["waves_per_block"], base["blocks"] * base["waves_per_block"]
        assert base_waves <= max(1, n // 24) * base["waves_per_block"]
        assert waves <= max(1, n // floor)
        if floor == 100000:
            assert waves == 1
        if floor < 24:
            assert waves > base_waves
        with pytest.raises(DeviceError):
            dev.train("persistent", 0, n, 0.025, T, min_sentences_per_wave=-1)


@pytest.mark.parametrize("floor", [4, 64])
def test_plain_mode_staleness_bound(ragged, floor):
    """Plain mode with the optional bound keeps at most ceil(n / floor)
    sentences in flight: the range runs as ceil(n / part) consecutive launches
    of part = ceil(n / floor) one-wave blocks (fewer than fit on the device),
    every kept token once.  Without it (0) the range is one launch."""
    vocab, stream = ragged
    counts, T = vocab.counts(), vocab.train_words
    csr = H.subsample(stream, counts, T, 1e-3, seed=5)
    n = csr.sent_off.size - 1
    part = -(-n // floor)
    print(f"[edges] plain floor {floor}: {n} sentences, {part} per launch", flush=True)
    for prec in ("fp32", "bf16"):
        with _dev(vocab, prec, *H.init_net(len(vocab), DIM)) as dev:
            dev.set_sampler(H.unigram_entries(counts))
            dev.set_corpus(csr.tokens, csr.sent_off, csr.sent_raw)
            st = dev.train("plain", 0, n, 0.025, T, seed=3, min_sentences_per_wave=floor)
            assert st["tokens"] == csr.tokens.size and st["sentences"] == n
            assert st["blocks"] == part and st["waves_per_block"] == 1
            assert dev.last_kernel_ms()[1] == -(-n // part)
            st = dev.train("plain", 0, n, 0.025, T, seed=3)
            assert st["tokens"] == csr.tokens.size and st["blocks"] == n and dev.last_kernel_ms()[1] == 1
