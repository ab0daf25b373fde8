"""The five synthetic workloads the trainer is measured and checked on.

Each entry fixes the corpus generator and the trainer flags.  Values not
given by the workload description are fixed here: generator seeds, the
cluster rank band ([max(100, V/100), +4*clustered words)), `w<id>` word
strings, every line ending with a newline, min-count 5 and batch size 24
(PAR_Word2Vec_cpu.cpp:66; SC19_AE_test_cases/Table5.sh:35)."""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import host as H


@dataclass
class Workload:
    name: str
    kind: str                 # "zipf" | "sbm"
    gen: dict
    dim: int
    window: int
    negative: int
    sample: float
    alpha: float = 0.025
    iters: int = 1
    batch_size: int = 24
    min_count: int = 5
    extra: dict = field(default_factory=dict)


WORKLOADS = {
    "zipf1m": Workload("zipf1m: 1M Zipf(1.0) tokens, V 10000, 1000-token lines, 50x20 clusters p=0.3",
                       "zipf", dict(n_tokens=1_000_000, raw_vocab=10_000, zipf_s=1.0, line_len=1000, n_clusters=50,
                                    cluster_size=20, span=3, p_replace=0.3, seed=101),
                       dim=100, window=5, negative=5, sample=1e-4),
    "text8": Workload("text8-shaped: 17M Zipf(1.05) tokens, V 71290, 1000-token lines, 200x25 clusters p=0.25",
                      "zipf", dict(n_tokens=17_000_000, raw_vocab=71_290, zipf_s=1.05, line_len=1000,
                                   n_clusters=200, cluster_size=25, span=3, p_replace=0.25, seed=102),
                      dim=128, window=8, negative=5, sample=1e-4),
    "sbm": Workload("DeepWalk SBM: 10000 nodes, 20 communities, degree exp 2.5 mean 40, 85% intra, 10x80 walks",
                    "sbm", dict(n_nodes=10_000, n_comm=20, exponent=2.5, mean_degree=40.0, p_intra=0.85,
                                walks_per_node=10, walk_len=80, seed=103),
                    dim=128, window=10, negative=5, sample=1e-4),
    "1b": Workload("1B-word-shaped: 800M Zipf(1.0) tokens, V 555000, log-normal(3.0,0.8) lines in [5,2000], "
                   "1000x20 clusters p=0.2",
                   "zipf", dict(n_tokens=800_000_000, raw_vocab=555_000, zipf_s=1.0, lognorm=(3.0, 0.8),
                                len_clip=(5, 2000), n_clusters=1000, cluster_size=20, span=3, p_replace=0.2, seed=104),
                   dim=300, window=5, negative=10, sample=1e-5),
}


@dataclass
class Corpus:
    workload: Workload
    raw: np.ndarray            # raw ids, 0 ends a line
    raw_vocab: int
    vocab: H.Vocab
    stream: np.ndarray         # vocabulary ids incl. 0 for "</s>"
    counts: np.ndarray
    train_words: int
    groups: np.ndarray         # vocab ids: clusters [C, m] (zipf) or -1 padded
    labels: np.ndarray | None  # sbm: community of each vocab id (-1 for "</s>")
    raw_clusters: np.ndarray | None = None  # zipf: raw ids of the planted clusters [C, m]


def make(name: str, scale: float = 1.0) -> Corpus:
    """Generates workload `name`; `scale` shrinks the token count (tests)."""
    w = WORKLOADS[name]
    g = dict(w.gen)
    labels = None
    if w.kind == "zipf":
        g["n_tokens"] = max(10_000, int(g["n_tokens"] * scale))
        raw, clusters = H.synth_zipf(**g)
        raw_vocab = g["raw_vocab"]
    else:
        if scale != 1.0:
            g["n_nodes"] = max(200, int(g["n_nodes"] * scale))
        raw, comm = H.synth_sbm_walks(**g)
        raw_vocab = g["n_nodes"]
        clusters = None
    vocab = H.Vocab.from_raw(raw, raw_vocab, w.min_count)
    stream = H.map_raw_stream(vocab, raw)
    rmap = vocab.raw_map()
    if clusters is not None:
        groups = rmap[clusters]
    else:
        labels = np.full(len(vocab), -1, np.int32)
        ids = rmap[np.arange(1, raw_vocab + 1)]
        ok = ids >= 0
        labels[ids[ok]] = comm[ok]
        groups = np.zeros((0, 0), np.int32)
    return Corpus(w, raw, raw_vocab, vocab, stream, vocab.counts(), vocab.train_words, groups, labels,
                  clusters if w.kind == "zipf" else None)


def heldout_stream(c: Corpus, n_tokens: int, seed: int = 99) -> np.ndarray:
    """Held-out lines of a Zipf workload: fresh tokens from the same rank
    distribution, line lengths and planted clusters as the training corpus,
    mapped through the training vocabulary (words it dropped are skipped).
    The generator's replacement rule -- each token of a clustered word
    replaces each neighbour within +-span on its line by a random word of its
    cluster with probability p -- is applied one offset at a time over all
    tokens.  Returns vocabulary ids, 0 ending each line."""
    w = c.workload
    if w.kind != "zipf":
        raise ValueError("held-out lines are defined for the Zipf workloads")
    g = w.gen
    rng = np.random.default_rng(seed)
    V = g["raw_vocab"]
    cdf = np.cumsum(np.arange(1, V + 1, dtype=np.float64) ** -g["zipf_s"])
    cdf /= cdf[-1]
    lens = []
    while sum(lens) < n_tokens:
        if "lognorm" in g:
            mu, sig = g["lognorm"]
            ln = np.clip(np.round(np.exp(mu + sig * rng.standard_normal(1024))), *g["len_clip"]).astype(np.int64)
        else:
            ln = np.full(1024, g["line_len"], np.int64)
        lens.extend(ln.tolist())
    lens = np.array(lens, np.int64)
    lens = lens[: int(np.searchsorted(np.cumsum(lens), n_tokens)) + 1]
    n = int(lens.sum())
    orig = np.minimum(np.searchsorted(cdf, rng.random(n)), V - 1) + 1
    line_of = np.repeat(np.arange(lens.size), lens)
    cl = c.raw_clusters
    owner = np.full(V + 1, -1, np.int64)
    owner[cl.reshape(-1)] = np.repeat(np.arange(cl.shape[0]), cl.shape[1])
    tok = orig.copy()
    src = np.flatnonzero(owner[orig] >= 0)
    for o in range(-g["span"], g["span"] + 1):
        if o == 0:
            continue
        j = src + o
        ok = (j >= 0) & (j < n)
        ok[ok] &= line_of[j[ok]] == line_of[src[ok]]
        ok &= rng.random(src.size) < g["p_replace"]
        tok[j[ok]] = cl[owner[orig[src[ok]]], rng.integers(0, cl.shape[1], int(ok.sum()))]
    ends = np.cumsum(lens) - 1
    raw = np.insert(tok, ends + 1, 0)  # a 0 after each line
    rmap = c.vocab.raw_map()
    ids = np.where(raw == 0, 0, rmap[raw])
    return ids[ids >= 0].astype(np.int32)
