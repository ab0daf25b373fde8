"""Launch shape of the persistent kernel: sentences per dequeue and the
autotuner (par_sgns_train_args.sentences_per_dequeue, par_sgns_autotune),
through the C-ABI.

Bars:
  sentences per dequeue   every sentence and token of the range trained exactly
                          once for 1, 2, 3, 4 and 64 sentences per dequeue
                          (counters equal the corpus totals)
  autotune                leaves both tables bit-identical; reports one point
                          per feasible shape, the best being the fastest of
                          them; later default launches train every token once
"""
import numpy as np
import pytest

from par_word2vec_amd import host as H
from par_word2vec_amd.device import DeviceError, SgnsDevice

pytestmark = pytest.mark.gpu


def _corpus(seed=3):
    # log-normal line lengths (5..400): uneven sentences, as the queue sees them
    raw, _ = H.synth_zipf(400_000, 5000, 1.0, lognorm=(3.0, 0.8), len_clip=(5, 400), seed=seed)
    vocab = H.Vocab.from_raw(raw, 5000, 5)
    stream = H.map_raw_stream(vocab, raw)
    counts, T = vocab.counts(), vocab.train_words
    csr = H.subsample(stream, counts, T, 1e-4, seed=17)
    return vocab, counts, T, csr


def _device(vocab, counts, csr, prec, dim=300, window=5, negative=10, batch=24):
    dev = SgnsDevice(len(vocab), dim, window, negative, batch, prec)
    dev.set_exp_table(H.exp_table())
    dev.set_weights(*H.init_net(len(vocab), dim))
    dev.set_sampler(H.unigram_entries(counts))
    dev.set_corpus(csr.tokens, csr.sent_off, csr.sent_raw)
    return dev


@pytest.mark.parametrize("prec", ["bf16", "fp32"])
def test_sentences_per_dequeue_counts(prec):
    vocab, counts, T, csr = _corpus()
    n_sent = csr.sent_off.size - 1
    n_nonempty = int(np.count_nonzero(np.diff(csr.sent_off)))
    with _device(vocab, counts, csr, prec) as dev:
        for g in (1, 2, 3, 4, 64):
            print(f"[dequeue] {prec} sentences_per_dequeue={g}", flush=True)
            st = dev.train("persistent", 0, n_sent, 0.025, T, seed=9, sentences_per_dequeue=g)
            assert st["tokens"] == csr.tokens.size
            assert st["sentences"] == n_nonempty
        with pytest.raises(DeviceError):
            dev.train("persistent", 0, n_sent, 0.025, T, sentences_per_dequeue=65)


@pytest.mark.parametrize("prec", ["bf16", "fp32"])
def test_autotune_restores_tables_and_picks_fastest(prec):
    vocab, counts, T, csr = _corpus(seed=5)
    n_sent = csr.sent_off.size - 1
    with _device(vocab, counts, csr, prec) as dev:
        w0_in, w0_out = dev.get_weights()
        dev.train("persistent", 0, n_sent, 0.025, T, seed=3)  # move the tables off their initial values
        w1_in, w1_out = dev.get_weights()
        assert not np.array_equal(w1_out, w0_out)
        print(f"[autotune] {prec} dim 300", flush=True)
        res = dev.autotune(0, n_sent, 0.025, T, seed=3, reps=1)
        w2_in, w2_out = dev.get_weights()
        assert np.array_equal(w2_in, w1_in) and np.array_equal(w2_out, w1_out)
        pts = res["points"]
        best = res["best"]
        print(f"[autotune] {len(pts)} points, best {best}", flush=True)
        assert len(pts) >= 3
        assert {p["sentences_per_dequeue"] for p in pts} == {1, 2, 4}
        for p in pts:
            assert p["kernel_ms"] > 0
            assert p["waves_per_cu"] % p["waves_per_block"] == 0 and p["waves_per_cu"] <= 8
        assert best in pts
        assert best["kernel_ms"] == min(p["kernel_ms"] for p in pts)
        st = dev.train("persistent", 0, n_sent, 0.025, T, seed=4)  # the tuned shape by default
        assert st["tokens"] == csr.tokens.size
        with pytest.raises(DeviceError):
            dev.autotune(0, n_sent, 0.025, T, reps=0)


def test_set_launch_shape():
    vocab, counts, T, csr = _corpus(seed=6)
    n_sent = csr.sent_off.size - 1
    with _device(vocab, counts, csr, "bf16") as dev:
        for bad in [(3, 4, 1), (2, 9, 1), (1, 4, 0), (0, 4, 1), (2, 4, 65)]:
            with pytest.raises(DeviceError):
                dev.set_launch_shape(*bad)
        for shape in [(2, 4, 2), (8, 8, 4), (1, 6, 1), (0, 0, 0)]:
            print(f"[launch shape] {shape}", flush=True)
            dev.set_launch_shape(*shape)
            st = dev.train("persistent", 0, n_sent, 0.025, T, seed=2, use_graph=True)
            assert st["tokens"] == csr.tokens.size
