"""Embedding-quality scores used as the "equal planted-cluster score" gate.

planted_cluster_recall: for every planted-cluster word, the fraction of its
10 nearest neighbours (cosine over the whole vocabulary, itself excluded)
that belong to its cluster; averaged over the clustered words.
cluster_margin: for every planted-cluster word, the mean cosine to its
cluster mates minus the mean cosine to as many nearest non-mates; averaged.
Unlike recall it does not saturate at 1 once every cluster is recovered: it
keeps measuring how far apart the clusters sit.
community_knn_accuracy: for every graph node, the majority community of its
10 nearest nodes; the fraction predicted correctly.
heldout_loss: the SGNS objective (FP64, per pair) of held-out windows --
-log sigmoid(c.x) per context and -log sigmoid(-n.x) per negative, W_in rows
x, W_out rows c and n -- the quantity training lowers.

Scoring is evaluation, not training: it runs as a blocked matrix product in
torch on the GPU when one is present, else in numpy."""
from __future__ import annotations

import numpy as np


def _topk_neighbours(w: np.ndarray, queries: np.ndarray, k: int, candidates: np.ndarray | None = None) -> np.ndarray:
    """Indices (into `candidates`, or into rows of w) of the k most cosine-
    similar rows for each query row index, excluding the query itself."""
    cand = np.arange(w.shape[0]) if candidates is None else candidates
    use_gpu = False
    if cand.size * queries.size > 2e8:  # large problems only: small ones stay in numpy
        try:
            import torch
            use_gpu = torch.cuda.is_available()
        except Exception:  # noqa: BLE001 - torch absent means the numpy path
            use_gpu = False
    norms = np.linalg.norm(w, axis=1, keepdims=True)
    wn = (w / np.maximum(norms, 1e-12)).astype(np.float32)
    out = np.empty((queries.size, k), np.int64)
    if use_gpu:
        import torch
        C = torch.from_numpy(wn[cand]).cuda()
        for lo in range(0, queries.size, 1024):
            q = queries[lo:lo + 1024]
            Q = torch.from_numpy(wn[q]).cuda()
            s = Q @ C.T
            pos = torch.from_numpy(np.searchsorted(cand, q)).cuda()
            hit = torch.from_numpy(np.isin(q, cand)).cuda()
            rows = torch.arange(q.size, device="cuda")
            s[rows[hit], pos[hit]] = -float("inf")
            out[lo:lo + q.size] = torch.topk(s, k, dim=1).indices.cpu().numpy()
        return out
    Cn = wn[cand]
    for lo in range(0, queries.size, 256):
        q = queries[lo:lo + 256]
        s = wn[q] @ Cn.T
        pos = np.searchsorted(cand, q)
        hit = np.isin(q, cand)
        s[np.arange(q.size)[hit], pos[hit]] = -np.inf
        part = np.argpartition(-s, k, axis=1)[:, :k]
        order = np.argsort(-np.take_along_axis(s, part, 1), axis=1)
        out[lo:lo + q.size] = np.take_along_axis(part, order, 1)
    return out


def planted_cluster_recall(w_in: np.ndarray, groups: np.ndarray, k: int = 10) -> float:
    g = np.asarray(groups)
    ids = g.reshape(-1)
    gid = np.repeat(np.arange(g.shape[0]), g.shape[1])
    ok = ids >= 0
    ids, gid = ids[ok], gid[ok]
    owner = np.full(w_in.shape[0], -1, np.int64)
    owner[ids] = gid
    nb = _topk_neighbours(w_in, ids, k)
    return float((owner[nb] == gid[:, None]).mean())


def community_knn_accuracy(w_in: np.ndarray, labels: np.ndarray, k: int = 10) -> float:
    nodes = np.flatnonzero(labels >= 0)
    nb = _topk_neighbours(w_in, nodes, k, candidates=nodes)
    lab = labels[nodes]
    votes = lab[nb]
    n_comm = int(lab.max()) + 1
    counts = np.zeros((nodes.size, n_comm), np.int64)
    np.add.at(counts, (np.repeat(np.arange(nodes.size), k), votes.reshape(-1)), 1)
    return float((counts.argmax(1) == lab).mean())


def _unit(w: np.ndarray) -> np.ndarray:
    return (w / np.maximum(np.linalg.norm(w, axis=1, keepdims=True), 1e-12)).astype(np.float32)


def cluster_margin(w_in: np.ndarray, groups: np.ndarray) -> float:
    """Mean over clustered words of (mean cosine to its m-1 mates) - (mean
    cosine to its m-1 most similar words outside its cluster, itself
    excluded)."""
    g = np.asarray(groups)
    m = g.shape[1]
    ids = g.reshape(-1)
    gid = np.repeat(np.arange(g.shape[0]), m)
    ok = ids >= 0
    wn = _unit(w_in)
    use_gpu = False
    if w_in.shape[0] * ids.size > 2e8:
        try:
            import torch
            use_gpu = torch.cuda.is_available()
        except Exception:  # noqa: BLE001 - torch absent means the numpy path
            use_gpu = False
    margins = np.empty(g.shape[0] * m, np.float64)
    if use_gpu:
        import torch
        Wt = torch.from_numpy(wn).cuda()
        for lo in range(0, ids.size, 1024):
            q = ids[lo:lo + 1024]
            rows = torch.arange(q.size, device="cuda")
            s = torch.from_numpy(wn[np.maximum(q, 0)]).cuda() @ Wt.T
            mates = torch.from_numpy(g[gid[lo:lo + q.size]].astype(np.int64)).cuda()  # [n, m], incl. itself
            mate_ok = mates >= 0
            ms = torch.gather(s, 1, mates.clamp(min=0))
            self_hit = mates == torch.from_numpy(q.astype(np.int64)).cuda()[:, None]
            use = mate_ok & ~self_hit
            mate_mean = (ms * use).sum(1) / use.sum(1).clamp(min=1)
            s.scatter_(1, mates.clamp(min=0), -float("inf"))
            s[rows, torch.from_numpy(np.maximum(q, 0)).cuda()] = -float("inf")
            near = torch.topk(s, m - 1, dim=1).values.mean(1)
            margins[lo:lo + q.size] = (mate_mean - near).double().cpu().numpy()
    else:
        for lo in range(0, ids.size, 256):
            q = ids[lo:lo + 256]
            s = wn[np.maximum(q, 0)] @ wn.T
            mates = g[gid[lo:lo + q.size]]
            r = np.arange(q.size)[:, None]
            ms = s[r, np.maximum(mates, 0)]
            use = (mates >= 0) & (mates != q[:, None])
            mate_mean = (ms * use).sum(1) / np.maximum(use.sum(1), 1)
            s[r, np.maximum(mates, 0)] = -np.inf
            s[np.arange(q.size), np.maximum(q, 0)] = -np.inf
            near = -np.partition(-s, m - 1, axis=1)[:, : m - 1]  # the m-1 largest similarities
            margins[lo:lo + q.size] = mate_mean - near.mean(1)
    return float(margins[ok].mean())


def heldout_windows(stream: np.ndarray, counts: np.ndarray, train_words: int, sample: float, window: int,
                    negative: int, n_windows: int, seed: int = 7) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Windows of a held-out id stream (0 ends a line), shaped like training
    windows: the stream is subsampled with the reference's keep probability
    sqrt(r) + r, r = sample * T / count (PAR_Word2Vec_cpu.cpp:459-463), then
    n_windows centres with all 2*window contexts inside their line are drawn,
    each with `negative` words drawn from unigram^0.75 (cpu:96-117) that are
    not in the window.  Returns (centre [n], ctx [n, 2W], neg [n, negative])."""
    rng = np.random.default_rng(seed)
    cnt = np.asarray(counts, np.float64)
    s = np.asarray(stream)
    if sample > 0:
        r = sample * float(train_words) / np.maximum(cnt[s], 1.0)
        keep = (s == 0) | (rng.random(s.size) < np.sqrt(r) + r)
        s = s[keep]
    ends = np.flatnonzero(s == 0)
    line = np.searchsorted(ends, np.arange(s.size))  # line of each position
    start = np.concatenate([[0], ends[:-1] + 1])[line]
    pos = np.arange(s.size)
    ok = (s != 0) & (pos - window >= start) & (pos + window < ends[np.minimum(line, ends.size - 1)])
    cand = np.flatnonzero(ok)
    if cand.size == 0:
        raise ValueError("no full window in the held-out stream")
    c = rng.choice(cand, size=n_windows, replace=cand.size < n_windows)
    offs = np.concatenate([np.arange(-window, 0), np.arange(1, window + 1)])
    ctx = s[c[:, None] + offs[None, :]]
    p = cnt ** 0.75
    p[0] = 0.0
    p /= p.sum()
    neg = rng.choice(cnt.size, size=(n_windows, negative), p=p)
    clash = (neg[:, :, None] == np.concatenate([ctx, s[c][:, None]], 1)[:, None, :]).any(2)
    while clash.any():  # redraw negatives that hit the window (cf. cpu:609-617)
        neg[clash] = rng.choice(cnt.size, size=int(clash.sum()), p=p)
        clash = (neg[:, :, None] == np.concatenate([ctx, s[c][:, None]], 1)[:, None, :]).any(2)
    return s[c].astype(np.int32), ctx.astype(np.int32), neg.astype(np.int32)


def heldout_loss(w_in: np.ndarray, w_out: np.ndarray, centre: np.ndarray, ctx: np.ndarray, neg: np.ndarray) -> float:
    """FP64 SGNS loss per pair of the windows (see heldout_windows)."""
    total, pairs = 0.0, 0
    for lo in range(0, centre.size, 4096):
        cc, xx, nn = centre[lo:lo + 4096], ctx[lo:lo + 4096], neg[lo:lo + 4096]
        X = w_in[xx].astype(np.float64)                       # [n, 2W, d]
        Y = w_out[np.concatenate([cc[:, None], nn], 1)].astype(np.float64)  # [n, 1+neg, d]
        F = np.einsum("nkd,njd->nkj", X, Y)
        total += np.logaddexp(0.0, -F[:, :, 0]).sum() + np.logaddexp(0.0, F[:, :, 1:]).sum()
        pairs += F.size
    return total / pairs
