"""Embedding-quality scores used as the "equal planted-cluster score" gate.

planted_cluster_recall: for every planted-cluster word, the fraction of its
10 nearest neighbours (cosine over the whole vocabulary, itself excluded)
that belong to its cluster; averaged over the clustered words.
community_knn_accuracy: for every graph node, the majority community of its
10 nearest nodes; the fraction predicted correctly.

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
