"""CPU restatement of the random schedule of the device's Hogwild trainer
(sgns_fast_kernel drawing its own negatives): given the key, what every
sentence's window shrinks, batch windows and negatives are.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  The product
package never imports it.

The reference draws its whole schedule from one sequential LCG stream per
thread (PAR_Word2Vec_cpu.cpp:534, 588-616).  The device keys every draw by
(key, sentence, epoch, purpose, index) instead, so that no draw depends on
which wave trains which sentence when.  This module restates, in plain
Python/numpy:

* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
  easy as 1, 2, 3", SC'11) and the counter layout of one draw:
  (index, purpose tag, sentence bits 0-31, sentence bits 32-63 ^ epoch << 20),
  key = the 64-bit seed;
* the alias table that par_sgns_set_sampler builds from the reference's
  unigram^0.75 slot counts (Vose's method in double precision), where the
  "</s>" slots' mass is spread uniformly over [1, V-1] -- the reference's
  `if (!sample) sample = next_random % (vocab_size - 1) + 1`
  (PAR_Word2Vec_cpu.cpp:609-610) -- and one draw from it;
* the reference's schedule on those draws: window shrink b of each centre
  (PAR_Word2Vec_cpu.cpp:534), the batch window rep_b (cpu:588-600), K = rep_b
  * negative candidates per batch, each redrawn uniformly over [1, V-1] while
  it occurs in the sentence (cpu:604-622); the device bounds the redraws at
  64 and then keeps the last candidate, which only a sentence holding nearly
  the whole vocabulary reaches;
* the learning rate of a sentence from the training progress at its start,
  in float32 as the device computes it (cpu:443-445).

predict_plan() turns these into an exact schedule (host.Plan) that the
device replays in planned mode: one wave drawing its own negatives and one
wave replaying the predicted plan must leave the same tables, bit for bit.

Pinned by the Random123 known-answer vectors of Philox4x32-10, by the alias
table's exact distribution (each word's mass summed over its own bucket and
the buckets aliasing to it), and on the GPU by the bitwise replay.
"""
from __future__ import annotations

import numpy as np

_MASK = np.uint64(0xFFFFFFFF)
_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = np.uint64(0x9E3779B9), np.uint64(0xBB67AE85)

# purpose tags of the device's draws
TAG_WINDOW, TAG_REP_WINDOW, TAG_NEG, TAG_REDRAW = 1, 2, 3, 4
MAX_REDRAW = 64        # redraws of one candidate before the last one is kept
BATCH_STRIDE = 8192    # index of negative k of batch bi: bi * 8192 + k


def philox4x32_10(ctr, k0: int, k1: int) -> np.ndarray:
    """Philox4x32-10 of counters ctr = (c0, c1, c2, c3) (scalars or equal-length
    arrays) under key (k0, k1); returns a (4, n) uint32 array."""
    x, y, z, w = (np.atleast_1d(np.asarray(c, dtype=np.uint64)) & _MASK for c in ctr)
    ka, kb = np.uint64(k0 & 0xFFFFFFFF), np.uint64(k1 & 0xFFFFFFFF)
    for _ in range(10):
        p0, p1 = _M0 * x, _M1 * z
        x, y, z, w = (p1 >> np.uint64(32)) ^ y ^ ka, p1 & _MASK, (p0 >> np.uint64(32)) ^ w ^ kb, p0 & _MASK
        ka, kb = (ka + _W0) & _MASK, (kb + _W1) & _MASK
    return np.stack([x, y, z, w]).astype(np.uint32)


def draw(seed: int, sent: int, epoch: int, tag: int, idx) -> np.ndarray:
    """The device's draw(seed, sentence, epoch, tag, index): (4, n) uint32."""
    idx = np.atleast_1d(np.asarray(idx, dtype=np.uint64))
    n = idx.size
    hi = ((sent >> 32) ^ (epoch << 20)) & 0xFFFFFFFF
    return philox4x32_10((idx, np.full(n, tag, np.uint64), np.full(n, sent & 0xFFFFFFFF, np.uint64),
                          np.full(n, hi, np.uint64)), seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)


def alias_table(entries) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(keep probability float32[V], alias int32[V], target distribution
    float64[V]) of par_sgns_set_sampler for slot counts `entries`."""
    e = [int(v) for v in entries]
    V = len(e)
    total = 0.0
    for v in e:
        total += float(v)
    uni = float(e[0]) / total / float(V - 1)
    w = [0.0] + [float(e[i]) / total + uni for i in range(1, V)]
    q = [wi * V for wi in w]
    prob = [1.0] * V
    alias = list(range(V))
    small = [i for i in range(V) if q[i] < 1.0]
    large = [i for i in range(V) if not q[i] < 1.0]
    while small and large:
        s, l = small.pop(), large[-1]
        prob[s] = float(np.float32(q[s]))
        alias[s] = l
        q[l] = (q[l] + q[s]) - 1.0
        if q[l] < 1.0:
            large.pop()
            small.append(l)
    for i in large + small:
        prob[i], alias[i] = 1.0, i
    if prob[0] > 0.0 and w[0] == 0.0:
        prob[0] = 0.0
    if alias[0] == 0:
        alias[0] = 1
    return np.array(prob, np.float32), np.array(alias, np.int32), np.array(w)


def alias_draw(prob: np.ndarray, alias: np.ndarray, x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """One alias-table draw per (x, y) word pair of a Philox output: bucket =
    high word of x * V, kept when (y >> 8) / 2^24 < its keep probability."""
    V = prob.size
    bucket = ((x.astype(np.uint64) * np.uint64(V)) >> np.uint64(32)).astype(np.int64)
    u2 = (y >> np.uint32(8)).astype(np.float64) * 2.0 ** -24
    return np.where(u2 < prob[bucket].astype(np.float64), bucket, alias[bucket]).astype(np.int32)


def sentence_alpha(a0: float, epoch: int, train_words: int, iters: int, raw: int) -> np.float32:
    """Learning rate of a sentence starting at progress `raw`, in float32."""
    f = np.float32
    a = f(a0) * (f(1.0) - f(epoch * train_words + raw) / f(iters * train_words + 1))
    fl = f(a0) * f(0.0001)
    return fl if a < fl else a


def predict_plan(tokens: np.ndarray, sent_off: np.ndarray, sent_raw: np.ndarray, entries, *, window: int,
                 negative: int, batch_size: int, seed: int, epoch: int, iters: int, alpha: float,
                 train_words: int):
    """The exact schedule (host.Plan) of the device trainer on this corpus,
    with sentence s of the plan = sentence s of the corpus, trained from
    progress sent_raw[s].  Returns (plan, stats) with stats the counts the
    device reports (tokens, pairs are left to the device)."""
    from par_word2vec_amd.host import Plan

    prob, alias, _ = alias_table(entries)
    V = prob.size
    W, NEG, B = window, negative, batch_size
    n = sent_off.size - 1
    bs, alphas, batch_off, neg_off, negs = [], [], [0], [0], []
    n_neg = 0
    for s in range(n):
        sen = tokens[sent_off[s]:sent_off[s + 1]]
        L = sen.size
        a = sentence_alpha(alpha, epoch, train_words, iters, int(sent_raw[s]))
        alphas += [a, a]
        if L:
            bs.append((draw(seed, s, epoch, TAG_WINDOW, np.arange(L))[0] % np.uint32(W)).astype(np.uint8))
        words = set(int(t) for t in sen)
        nbat = (L + B - 1) // B
        for bi in range(nbat):
            lo = bi * B
            r = int(draw(seed, s, epoch, TAG_REP_WINDOW, bi)[0, 0])
            rep_b = r % W if (lo <= W - 1 or (L - 1) - lo <= W - 1) else r % (2 * W)
            K = rep_b * NEG
            if K:
                idx = bi * BATCH_STRIDE + np.arange(K)
                d = draw(seed, s, epoch, TAG_NEG, idx)
                cand = alias_draw(prob, alias, d[0], d[1])
                for k in range(K):
                    c = int(cand[k])
                    attempt = 1
                    while attempt <= MAX_REDRAW and c in words:
                        r2 = int(draw(seed, s, epoch, TAG_REDRAW + (attempt << 8), int(idx[k]))[0, 0])
                        c = 1 + ((r2 * (V - 1)) >> 32)
                        attempt += 1
                    negs.append(c)
            n_neg += K
            neg_off.append(len(negs))
        batch_off.append(len(neg_off) - 1)
    plan = Plan(np.ascontiguousarray(tokens, np.int32),
                np.concatenate(bs) if bs else np.zeros(0, np.uint8),
                np.ascontiguousarray(sent_off, np.int64), np.array(alphas, np.float32),
                np.array(batch_off, np.int64), np.array(neg_off, np.int64), np.array(negs, np.int32),
                0, 0, 0.0, int(sent_raw[-1]) if n else 0)
    return plan, {"batches": len(neg_off) - 1, "negatives": n_neg}


def rejection_corpus(V: int = 300, seed: int = 11):
    """A corpus for checking the draws: sentences longest first (the device's
    queue order, so queue slot = sentence id), with lengths on both sides of
    a batch (24), of the 64- and 128-token register blocks, and up to the
    1000-token limit; one sentence holding every word of [1, V-1], whose
    candidates all reach the redraw bound; an empty one.  Zipf-like words make
    rejections frequent.  Returns (tokens, sent_off, sent_raw, train_words,
    counts) with a raw span of L + 3 words per sentence."""
    rng = np.random.default_rng(seed)
    lens = [1000, 777, 513, V - 1, 200, 129, 128, 127, 65, 64, 63, 25, 24, 23, 5, 2, 1, 0]
    assert lens == sorted(lens, reverse=True)
    p = 1.0 / np.arange(1, V) ** 1.1
    p /= p.sum()
    sents = [rng.permutation(np.arange(1, V)) if L == V - 1 else rng.choice(V - 1, size=L, p=p) + 1 for L in lens]
    tokens = np.concatenate(sents).astype(np.int32)
    off = np.zeros(len(lens) + 1, np.int64)
    off[1:] = np.cumsum(lens)
    spans = np.array(lens, np.int64) + 3
    raw = np.zeros(len(lens), np.int64)
    raw[1:] = np.cumsum(spans)[:-1]
    counts = np.bincount(tokens, minlength=V).astype(np.int64) + 1
    counts[0] = len(lens)
    return tokens, off, raw, int(spans.sum()), counts
