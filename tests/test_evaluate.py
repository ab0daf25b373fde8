"""Known answers of the embedding-quality scores on crafted embeddings (no GPU).

recall@10 of planted clusters    perfect orthogonal clusters of 20 -> 1; of
                                 6 (5 mates < k) -> 0.5; two words' vectors
                                 swapped between clusters -> (N - 2) / N
community kNN accuracy           orthogonal communities -> 1; two swapped
                                 nodes -> (N - 2) / N
cluster margin                   word = cos(t) u_cluster + sin(t) e_own with
                                 orthonormal directions: mates' cosine is
                                 cos^2 t, non-mates' 0 -> margin cos^2 t
held-out SGNS loss               zero tables -> log 2 per pair; equals the
                                 oracle's FP64 window loss divided by pairs
held-out windows                 contexts are the line neighbours of the
                                 centre; negatives avoid the window
"""
import numpy as np
import pytest

import oracle as O
from par_word2vec_amd import corpora, evaluate as E


def _orthogonal_clusters(n_clusters, m, extra, d=None, jitter=0.0, seed=0):
    """Word 0 unused; clusters of m words along e_c; `extra` words along their
    own directions.  A small private component keeps every cosine distinct."""
    V = 1 + n_clusters * m + extra
    d = d or (1 + n_clusters + extra + V)
    rng = np.random.default_rng(seed)
    w = np.zeros((V, d), np.float32)
    groups = np.arange(1, 1 + n_clusters * m).reshape(n_clusters, m)
    for c in range(n_clusters):
        w[groups[c], c] = 1.0
    for i in range(n_clusters * m + 1, V):
        w[i, n_clusters + (i - n_clusters * m - 1)] = 1.0
    w[0, -1] = 1.0
    if jitter:
        w[np.arange(V), n_clusters + extra + np.arange(V) % (d - n_clusters - extra)] += \
            jitter * rng.random(V).astype(np.float32)
    return w, groups


def test_recall_perfect_and_bounded_by_cluster_size():
    w, g = _orthogonal_clusters(4, 20, 30, jitter=1e-3)
    assert E.planted_cluster_recall(w, g) == 1.0
    w, g = _orthogonal_clusters(6, 6, 30, jitter=1e-3)
    assert E.planted_cluster_recall(w, g) == pytest.approx(0.5, abs=0)  # 5 mates among 10 neighbours


def test_recall_two_swapped_words():
    w, g = _orthogonal_clusters(3, 20, 30, jitter=1e-3)
    a, b = g[0, 0], g[1, 0]
    w[[a, b]] = w[[b, a]]
    # a now sits in cluster 1's direction: none of its 10 neighbours is a mate;
    # every other word still has >= 10 mates at cosine ~1 (the intruder b is
    # pushed below them by a smaller off-axis component)
    w[b] += 0.05 * np.eye(w.shape[1], dtype=np.float32)[-2]
    w[a] += 0.05 * np.eye(w.shape[1], dtype=np.float32)[-2]
    N = g.size
    assert E.planted_cluster_recall(w, g) == pytest.approx((N - 2) / N, abs=1e-12)


def test_community_accuracy_known_answers():
    n_comm, per = 4, 25
    V = 1 + n_comm * per
    d = n_comm + V
    w = np.zeros((V, d), np.float32)
    labels = np.full(V, -1, np.int32)
    for c in range(n_comm):
        ids = 1 + c * per + np.arange(per)
        labels[ids] = c
        w[ids, c] = 1.0
        w[ids, n_comm + ids] = 1e-3 * (1 + np.arange(per))
    w[0, -1] = 1.0
    assert E.community_knn_accuracy(w, labels) == 1.0
    a, b = 1, 1 + per  # one node of community 0 and one of community 1
    w[[a, b]] = w[[b, a]]
    N = n_comm * per
    assert E.community_knn_accuracy(w, labels) == pytest.approx((N - 2) / N, abs=1e-12)


@pytest.mark.parametrize("theta", [0.3, 0.7, 1.2])
def test_cluster_margin_closed_form(theta):
    C, m, extra = 3, 4, 5
    V = 1 + C * m + extra
    d = C + V
    w = np.zeros((V, d), np.float64)
    groups = np.arange(1, 1 + C * m).reshape(C, m)
    for c in range(C):
        w[groups[c], c] = np.cos(theta)
    w[np.arange(V), C + np.arange(V)] += np.where(np.arange(V) <= C * m, np.sin(theta), 1.0)
    w[0] = 0.0
    w[0, C] = 1.0
    # mates: cos^2(theta); every non-mate (other clusters, extra words) is orthogonal
    assert E.cluster_margin(w.astype(np.float32), groups) == pytest.approx(np.cos(theta) ** 2, abs=1e-6)


def test_cluster_margin_drops_when_clusters_blur():
    w, g = _orthogonal_clusters(5, 20, 50, jitter=1e-3)
    m_sharp = E.cluster_margin(w, g)
    rng = np.random.default_rng(1)
    w2 = w + (0.4 / np.sqrt(w.shape[1])) * rng.standard_normal(w.shape).astype(np.float32)  # noise norm ~0.4
    m_blur = E.cluster_margin(w2, g)
    assert m_sharp > 0.99 and m_blur < m_sharp - 0.05
    # recall stays saturated where the margin still moves
    assert E.planted_cluster_recall(w2, g) == 1.0


def test_heldout_loss_zero_tables_and_oracle():
    rng = np.random.default_rng(2)
    V, d, n, W, neg = 50, 8, 300, 3, 4
    centre = rng.integers(1, V, n).astype(np.int32)
    ctx = rng.integers(1, V, (n, 2 * W)).astype(np.int32)
    negs = rng.integers(1, V, (n, neg)).astype(np.int32)
    z = np.zeros((V, d), np.float32)
    assert E.heldout_loss(z, z, centre, ctx, negs) == pytest.approx(np.log(2.0), abs=1e-15)
    w_in = (rng.standard_normal((V, d)) * 0.4).astype(np.float32)
    w_out = (rng.standard_normal((V, d)) * 0.4).astype(np.float32)
    ref = O.window_loss_f64(centre, ctx.reshape(-1), negs.reshape(-1), w_in.astype(np.float64),
                            w_out.astype(np.float64))
    assert E.heldout_loss(w_in, w_out, centre, ctx, negs) == pytest.approx(ref / (n * 2 * W * (1 + neg)), rel=1e-12)


def test_heldout_windows_are_line_neighbourhoods():
    # every token unique (id = position + 1), lines of 30, no subsampling
    lines = [np.arange(1 + 30 * k, 31 + 30 * k) for k in range(20)]
    stream = np.concatenate([np.append(x, 0) for x in lines]).astype(np.int32)
    counts = np.ones(602, np.int64)
    c, x, n = E.heldout_windows(stream, counts, 600, 0.0, 4, 3, 500, seed=3)
    line = (c - 1) // 30
    assert np.all((c - 4 - 1) // 30 == line) and np.all((c + 4 - 1) // 30 == line)
    offs = np.array([-4, -3, -2, -1, 1, 2, 3, 4])
    assert np.array_equal(x, c[:, None] + offs[None, :])
    window = np.concatenate([x, c[:, None]], 1)
    assert not (n[:, :, None] == window[:, None, :]).any()
    assert (n > 0).all()


def test_heldout_stream_follows_the_workload():
    c = corpora.make("zipf1m", scale=0.05)
    h = corpora.heldout_stream(c, 50_000, seed=5)
    assert h.size >= 50_000 * 0.7 and (h == 0).sum() >= 45  # words below min-count are skipped
    assert h.min() >= 0 and h.max() < len(c.vocab)
    # same rank law: the most frequent words of the training stream lead the held-out one too
    top_train = np.argsort(-np.bincount(c.stream, minlength=len(c.vocab))[1:])[:20] + 1
    top_held = np.argsort(-np.bincount(h, minlength=len(c.vocab))[1:])[:20] + 1
    assert len(set(top_train) & set(top_held)) >= 15
    # planted structure: clustered words co-occur with mates far above chance
    g = c.groups
    owner = np.full(len(c.vocab), -1)
    owner[g[g >= 0]] = np.repeat(np.arange(g.shape[0]), g.shape[1])[g.reshape(-1) >= 0]
    a, b = h[:-1], h[1:]
    both = (owner[a] >= 0) & (owner[b] >= 0)
    same_h = (owner[a][both] == owner[b][both]).mean()
    a, b = c.stream[:-1], c.stream[1:]
    both = (owner[a] >= 0) & (owner[b] >= 0)
    same_t = (owner[a][both] == owner[b][both]).mean()
    assert same_h > 0.5 and abs(same_h - same_t) < 0.05
