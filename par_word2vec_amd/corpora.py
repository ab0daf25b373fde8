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
    return Corpus(w, raw, raw_vocab, vocab, stream, vocab.counts(), vocab.train_words, groups, labels)
