"""Pins of the host restatement of the device trainer's random schedule
(oracle/philox_schedule.py), which the GPU replay test holds the device's
own draws to: Philox4x32-10 against the Random123 known-answer vectors, the
alias table against the distribution it must realise, and the predicted
schedule against the reference's rules (PAR_Word2Vec_cpu.cpp:534, 588-622)."""
import numpy as np
import pytest

import philox_schedule as PS
from par_word2vec_amd import host as H

# Random123 kat_vectors, philox4x32 with 10 rounds: (counter, key, output)
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_known_answers(ctr, key, out):
    assert tuple(int(v) for v in PS.philox4x32_10(ctr, *key)[:, 0]) == out


def test_philox_vectorised_matches_scalar():
    idx = np.arange(0, 5000, 7)
    many = PS.draw(0x0123456789ABCDEF, 3 << 32 | 5, 2, PS.TAG_NEG, idx)
    for j in (0, 1, 100, idx.size - 1):
        assert (PS.draw(0x0123456789ABCDEF, 3 << 32 | 5, 2, PS.TAG_NEG, int(idx[j]))[:, 0] == many[:, j]).all()


@pytest.mark.parametrize("V,zero_share", [(2, 0.0), (3, 0.0), (17, 0.3), (1000, 0.2), (5000, 0.5)])
def test_alias_table_realises_the_distribution(V, zero_share):
    """Each word's mass -- its own bucket's keep probability plus the rest of
    every bucket aliasing to it, over V -- is the target distribution (word
    0 none, its slots spread over [1, V-1]) up to the float32 keep
    probabilities."""
    rng = np.random.default_rng(V)
    e = rng.integers(1, 100, V).astype(np.int64)
    e[rng.random(V) < zero_share] = 0
    e[0] = 9
    prob, alias, w = PS.alias_table(e)
    assert abs(w.sum() - 1.0) < 1e-12 and w[0] == 0.0
    p64 = prob.astype(np.float64)
    mass = p64.copy()
    np.add.at(mass, alias, 1.0 - p64)
    mass /= V
    n_in = np.bincount(alias, minlength=V) + 1
    assert mass[0] == 0.0
    assert (np.abs(mass - w) <= 6e-8 * n_in / V + 1e-15).all()
    d = PS.draw(99, 4, 0, PS.TAG_NEG, np.arange(200_000))
    h = np.bincount(PS.alias_draw(prob, alias, d[0], d[1]), minlength=V) / 200_000
    assert h[0] == 0.0
    assert 0.5 * np.abs(h - w).sum() < 3.0 * np.sqrt(V / 200_000) + 0.01


def test_alias_table_of_the_unigram_entries():
    tokens, off, raw, T, counts = PS.rejection_corpus()
    prob, alias, w = PS.alias_table(H.unigram_entries(counts))
    assert prob[0] == 0.0 and alias[0] != 0 and (alias >= 0).all() and (alias < counts.size).all()


@pytest.mark.parametrize("window,negative,batch,epoch", [(5, 5, 24, 0), (5, 10, 24, 1), (3, 7, 10, 0)])
def test_predicted_schedule_follows_the_reference_rules(window, negative, batch, epoch):
    tokens, off, raw, T, counts = PS.rejection_corpus()
    V = counts.size
    plan, st = PS.predict_plan(tokens, off, raw, H.unigram_entries(counts), window=window, negative=negative,
                               batch_size=batch, seed=0x1234567890ABCDEF, epoch=epoch, iters=epoch + 1,
                               alpha=0.025, train_words=T)
    n = off.size - 1
    assert plan.b.size == tokens.size and (plan.b < window).all()
    assert plan.batch_off.size == n + 1 and plan.neg_off[-1] == plan.negs.size == st["negatives"]
    full = [s for s in range(n) if off[s + 1] - off[s] == V - 1]
    assert len(full) == 1
    for s in range(n):
        sen = tokens[off[s]:off[s + 1]]
        L = sen.size
        assert plan.batch_off[s + 1] - plan.batch_off[s] == (L + batch - 1) // batch
        for bi in range(plan.batch_off[s + 1] - plan.batch_off[s]):
            bo = plan.batch_off[s] + bi
            negs = plan.negs[plan.neg_off[bo]:plan.neg_off[bo + 1]]
            K = negs.size
            assert K % negative == 0 and K // negative < 2 * window           # rep_b in [0, 2W) (cpu:588-600)
            lo = bi * batch
            if lo <= window - 1 or (L - 1) - lo <= window - 1:
                assert K // negative < window                                  # edge batches: rep_b in [0, W)
            assert ((negs >= 1) & (negs < V)).all()
            # rejected while in the sentence (cpu:604-622); the sentence holding
            # every word keeps each candidate after the last redraw
            assert np.isin(negs, sen).all() if s == full[0] else not np.isin(negs, sen).any()
    a = plan.sent_alpha.reshape(-1, 2)
    assert (a[:, 0] == a[:, 1]).all() and (np.diff(a[:, 1]) <= 0).all() and (a[:, 1] >= np.float32(0.025 * 1e-4)).all()
    assert a[0, 1] == PS.sentence_alpha(0.025, epoch, T, epoch + 1, 0)
