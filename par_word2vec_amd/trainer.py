"""Host-side mirror of the reference trainer's interface.

`Options` carries the reference's flags with its defaults
(PAR_Word2Vec_cpu.cpp:64-70, parsed at cpu:787-818; the GPU variant adds
-gamma/-max-num-sen, text.cu:1083-1086, which have no role here because the
sentence CSR is sized dynamically and scheduling is a work queue).
`ParWord2Vec.train()` plays Train_PAR_Word2Vec (cpu:372-711): vocabulary,
InitNet, unigram table, then the epochs -- on the GPU through
libpar_sgns.so.  Output follows saveModel (cpu:721-737)."""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from . import host as H
from .device import SgnsDevice


@dataclass
class Options:
    train: str = ""
    output: str = ""
    size: int = 100
    window: int = 5
    sample: float = 1e-3
    negative: int = 5
    threads: int = 12       # accepted for compatibility; the GPU ignores it
    iter: int = 5
    min_count: int = 5
    alpha: float = 0.025
    debug: int = 2
    binary: int = 0
    save_vocab: str = ""
    read_vocab: str = ""
    batch_size: int = 24
    disk: bool = False      # accepted for compatibility; the corpus is always resident
    # device options (no counterpart in the reference)
    mode: str = "persistent"      # det | plain | persistent
    precision: str = "fp32"       # fp32 | bf16
    seed: int = 1
    device: int = 0
    chunk_words: int = 10_000_000  # raw words per training launch (hipGraph chunk)
    use_graph: bool = True
    compute_loss: bool = False
    autotune: int = 0             # persistent mode: tune the launch shape on the first chunk (runs per shape)
    min_sentences_per_wave: int = 0  # persistent mode: staleness floor of the default grid (0 = 24)


def chunk_bounds(sent_raw: np.ndarray, chunk_words: int) -> list[tuple[int, int]]:
    """Split sentences into consecutive ranges of about chunk_words raw words."""
    S = sent_raw.size
    if S == 0:
        return []
    cuts = np.searchsorted(sent_raw, np.arange(chunk_words, int(sent_raw[-1]) + 1, chunk_words), side="left")
    edges = [0] + [int(c) for c in cuts if 0 < c < S] + [S]
    edges = sorted(set(edges))
    return list(zip(edges[:-1], edges[1:]))


class ParWord2Vec:
    """Vocabulary + id stream + device handle; one object per training run."""

    def __init__(self, opts: Options, vocab: H.Vocab, stream: np.ndarray):
        self.opts = opts
        self.vocab = vocab
        self.stream = np.ascontiguousarray(stream, np.int32)
        self.counts = vocab.counts()
        self.train_words = vocab.train_words
        V = len(vocab)
        self.dev = SgnsDevice(V, opts.size, opts.window, opts.negative, opts.batch_size, opts.precision, opts.device)
        self.dev.set_exp_table(H.exp_table())
        w_in, w_out = H.init_net(V, opts.size)
        self.dev.set_weights(w_in, w_out)
        self.entries = H.unigram_entries(self.counts)
        self.dev.set_sampler(self.entries)
        self.epoch_stats: list[dict] = []

    @classmethod
    def from_text(cls, opts: Options) -> "ParWord2Vec":
        vocab = H.Vocab.read(opts.read_vocab, opts.min_count) if opts.read_vocab else \
            H.Vocab.learn_text(opts.train, opts.min_count)
        if opts.save_vocab:
            vocab.save(opts.save_vocab)
        return cls(opts, vocab, H.load_text_stream(vocab, opts.train))

    # ---------------------------------------------------------------- epochs
    def prepare_epoch(self, epoch: int) -> list[tuple[int, int]]:
        o = self.opts
        if o.mode in ("det", "deterministic"):
            if epoch == 0:
                self.plan = H.build_plan(self.stream, self.counts, self.train_words, o.window, o.negative,
                                         o.batch_size, o.iter, o.alpha, o.sample)
                self.dev.set_plan(self.plan)
            return [(0, self.plan.n_sent)] if epoch == 0 else []
        csr = H.subsample(self.stream, self.counts, self.train_words, o.sample, o.seed * 1000003 + epoch)
        self.dev.set_corpus(csr.tokens, csr.sent_off, csr.sent_raw)
        self.csr = csr
        return chunk_bounds(csr.sent_raw, o.chunk_words)

    def run_chunk(self, epoch: int, b: int, e: int) -> dict:
        o = self.opts
        return self.dev.train(o.mode, b, e, o.alpha, self.train_words, iters=o.iter, epoch=epoch, seed=o.seed,
                              compute_loss=o.compute_loss, use_graph=o.use_graph and o.mode == "persistent",
                              min_sentences_per_wave=o.min_sentences_per_wave)

    def train(self) -> list[dict]:
        o = self.opts
        for ep in range(o.iter):
            chunks = self.prepare_epoch(ep)
            if ep == 0 and o.autotune > 0 and o.mode == "persistent" and chunks:  # untimed
                b, e = chunks[0]
                self.tuning = self.dev.autotune(b, e, o.alpha, self.train_words, iters=o.iter, seed=o.seed,
                                                reps=o.autotune, min_sentences_per_wave=o.min_sentences_per_wave)
                if o.debug > 0:
                    print(f"launch shape: {self.tuning['best']}", flush=True)
            t0 = time.perf_counter()
            agg: dict = {}
            for b, e in chunks:
                st = self.run_chunk(ep, b, e)
                for k, v in st.items():
                    agg[k] = agg.get(k, 0) + v
            agg["wall_s"] = time.perf_counter() - t0
            agg["epoch"] = ep
            self.epoch_stats.append(agg)
            if o.debug > 0:
                print(f"epoch {ep}: {agg.get('tokens', 0)} tokens, {agg['wall_s']:.3f} s", flush=True)
        return self.epoch_stats

    def weights(self) -> tuple[np.ndarray, np.ndarray]:
        return self.dev.get_weights()

    def save(self, path: str | None = None) -> None:
        w_in, _ = self.weights()
        self.vocab.save_embeddings(path or self.opts.output, w_in, bool(self.opts.binary))

    def close(self) -> None:
        self.dev.close()
