"""Repeated uploads and overlapping sentence ranges through the C-ABI.

The handle caches, per sentence range, the longest-first queue order and the
hipGraph captured for a persistent launch.  A later call that re-uploads a
buffer such a graph reads (sampler, plan, exp table, corpus) or orders an
overlapping range must invalidate what it makes stale.  The reference
re-uploads nothing mid-run (text.cu:841-868); the C-ABI promises these calls
may be repeated.

Bars: one wavefront replays sentences in a fixed order with per-sentence
random streams, so a cached-graph run equals an uncached run on a fresh
handle bit for bit; with Philox draws over overlapping ranges every kept
token of the range is trained exactly once per call.
"""
import numpy as np
import pytest

from par_word2vec_amd import host as H
from par_word2vec_amd.device import SgnsDevice

pytestmark = pytest.mark.gpu

DIM, W, NEG, B = 64, 5, 5, 24


@pytest.fixture(scope="module")
def corpus():
    raw, _ = H.synth_zipf(60_000, 1200, 1.0, line_len=120, n_clusters=4, cluster_size=10, p_replace=0.3, seed=31)
    vocab = H.Vocab.from_raw(raw, 1200, 5)
    return vocab, H.map_raw_stream(vocab, raw)


def _dev(vocab, precision, w_in, w_out):
    dev = SgnsDevice(len(vocab), DIM, W, NEG, B, precision)
    dev.set_exp_table(H.exp_table())
    dev.set_weights(w_in, w_out)
    return dev


ONE_WAVE = dict(blocks=1, waves_per_block=1)


def test_graph_dropped_when_plan_reuploaded(corpus):
    """A graph captured over plan 1 must not replay over plan 2's buffers."""
    vocab, stream = corpus
    counts, T = vocab.counts(), vocab.train_words
    V = len(vocab)
    plan1 = H.build_plan(stream, counts, T, W, NEG, B, 1, 0.025, 1e-3)
    half = stream[: int(np.flatnonzero(stream == 0)[len(np.flatnonzero(stream == 0)) // 2]) + 1]
    plan2 = H.build_plan(half, counts, T, W, NEG, B, 1, 0.02, 1e-4)
    n = min(plan1.n_sent, plan2.n_sent)
    assert plan1.tokens.size != plan2.tokens.size
    w0 = H.init_net(V, DIM)
    print("[abi] plan re-upload under a cached graph", flush=True)
    with _dev(vocab, "fp32", *w0) as dev:
        dev.set_plan(plan1)
        dev.train("persistent", 0, n, 0.025, T, planned=True, use_graph=True, **ONE_WAVE)
        mid = dev.get_weights()
        dev.set_plan(plan2)
        st = dev.train("persistent", 0, n, 0.025, T, planned=True, use_graph=True, **ONE_WAVE)
        got = dev.get_weights()
    with _dev(vocab, "fp32", *mid) as ref:
        ref.set_plan(plan2)
        rs = ref.train("persistent", 0, n, 0.025, T, planned=True, use_graph=False, **ONE_WAVE)
        want = ref.get_weights()
    assert st["tokens"] == rs["tokens"] == plan2.sent_off[n]
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])


@pytest.mark.parametrize("what", ["sampler", "exp_table"])
def test_graph_dropped_when_sampler_or_table_reuploaded(corpus, what):
    """A graph captured before set_sampler / set_exp_table must not replay
    with the freed buffers; the cached run equals an uncached one."""
    vocab, stream = corpus
    counts, T = vocab.counts(), vocab.train_words
    V = len(vocab)
    csr = H.subsample(stream, counts, T, 1e-3, seed=5)
    n = csr.sent_off.size - 1
    ent1 = H.unigram_entries(counts)
    ent2 = np.ones_like(ent1)  # a uniform sampler: different draws, same table size
    print(f"[abi] {what} re-upload under a cached graph", flush=True)
    with _dev(vocab, "bf16", *H.init_net(V, DIM)) as dev:
        dev.set_sampler(ent1)
        dev.set_corpus(csr.tokens, csr.sent_off, csr.sent_raw)
        dev.train("persistent", 0, n, 0.025, T, seed=3, use_graph=True, **ONE_WAVE)
        mid = dev.get_weights()
        if what == "sampler":
            dev.set_sampler(ent2)
        else:
            dev.set_exp_table(H.exp_table())
        st = dev.train("persistent", 0, n, 0.025, T, seed=3, use_graph=True, **ONE_WAVE)
        got = dev.get_weights()
    with _dev(vocab, "bf16", *mid) as ref:
        ref.set_sampler(ent2 if what == "sampler" else ent1)
        ref.set_corpus(csr.tokens, csr.sent_off, csr.sent_raw)
        rs = ref.train("persistent", 0, n, 0.025, T, seed=3, use_graph=False, **ONE_WAVE)
        want = ref.get_weights()
    assert st["tokens"] == rs["tokens"] == csr.tokens.size
    assert st["negatives"] == rs["negatives"]
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])


@pytest.mark.parametrize("use_graph", [False, True])
def test_overlapping_ranges_train_each_sentence_once(corpus, use_graph):
    """(0, n), then (0, n/2), then (0, n) again, then (n/4, 3n/4): each call
    trains exactly the kept tokens of its range, and the last full-range call
    equals an uncached full-range call on a fresh handle."""
    vocab, stream = corpus
    counts, T = vocab.counts(), vocab.train_words
    V = len(vocab)
    csr = H.subsample(stream, counts, T, 1e-3, seed=9)
    n = csr.sent_off.size - 1
    off = csr.sent_off
    print(f"[abi] overlapping ranges, use_graph={use_graph}", flush=True)
    with _dev(vocab, "fp32", *H.init_net(V, DIM)) as dev:
        dev.set_sampler(H.unigram_entries(counts))
        dev.set_corpus(csr.tokens, csr.sent_off, csr.sent_raw)
        for b, e in [(0, n), (0, n // 2), (0, n), (n // 4, 3 * n // 4)]:
            if (b, e) == (0, n):  # the last full-range call is compared below
                before = dev.get_weights()
            st = dev.train("persistent", b, e, 0.025, T, seed=4, use_graph=use_graph, **ONE_WAVE)
            assert st["tokens"] == off[e] - off[b], (b, e)
            assert st["sentences"] <= e - b
            if (b, e) == (0, n):
                after = dev.get_weights()
    with _dev(vocab, "fp32", *before) as ref:
        ref.set_sampler(H.unigram_entries(counts))
        ref.set_corpus(csr.tokens, csr.sent_off, csr.sent_raw)
        ref.train("persistent", 0, n, 0.025, T, seed=4, use_graph=False, **ONE_WAVE)
        want = ref.get_weights()
    assert np.array_equal(after[0], want[0]) and np.array_equal(after[1], want[1])
