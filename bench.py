"""Throughput benchmark of the PAR-SGNS trainer on one MI355X.

Workload (default): the 1B-word-shaped corpus -- 800M Zipf(1.0) tokens over
555,000 words, log-normal(3.0, 0.8) line lengths clipped to [5, 2000], 1000
planted clusters of 20 words (p = 0.2), d 300, window 5, 10 negatives,
sample 1e-5, one epoch cut into ~10M-raw-word chunks.  A step is one chunk:
one launch of the training kernel over the sentences of that chunk (a
captured hipGraph for the persistent variant), inputs resident in HBM.

Both variants train the same epoch from the same InitNet weights:
  fp32 plain      one wavefront per sentence, fp32 tables
  bf16 persistent work-queue kernel, bf16 tables, fp32 math, stochastic
                  rounding, repulsion products on the matrix cores, hipGraph
                  per chunk, launched with the saved launch-shape tuning of the
                  workload (par_word2vec_amd/tuned/; --autotune tunes here,
                  before the warmup)
and both write W_in rows back with atomics and W_out rows with plain stores
(--write-back; the reference CPU trainer updates rows in place, cpu:561-565).
`value` is the bf16 persistent rate in Mwords/s: raw corpus words (every
token incl. those removed by subsampling and "</s>") per second of the timed
steps.  Each variant's tables are scored after the full epoch: planted-cluster
recall@10, the planted-cluster cosine margin (does not saturate) and the FP64
SGNS loss per pair of 65,536 held-out windows (Zipf workloads), or community
kNN accuracy (SBM walks).

The CPU baseline is the oracle's OpenMP Hogwild restatement of the
reference CPU trainer, run as the reference's harness runs it (text file,
per-thread byte-offset slices, loading inside the timed region) on a bounded
prefix of the same corpus, with the host's CPU share of threads (lscpu's
model, sockets and cores are reported beside it)."""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from par_word2vec_amd import autotune, corpora, host as H, profiling  # noqa: E402
from par_word2vec_amd.device import SgnsDevice  # noqa: E402
from par_word2vec_amd.trainer import chunk_bounds  # noqa: E402

HBM_PEAK_GBS = 8000.0

def traffic_file(workload: str) -> Path:
    return ROOT / "profiles" / f"sgns_fast_kernel_{workload}_traffic.json"


def log(msg: str) -> None:
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(tokens: int, negatives: int, dim: int, elem: int) -> float:
    """4 B token id per trained token + each trained token's W_in and W_out
    rows read and written once + each shared negative's W_out row read and
    written once."""
    return 4.0 * tokens + elem * dim * (4.0 * tokens + 2.0 * negatives)


def run_variant(c, mode: str, precision: str, chunks, raw_per_chunk, warmup: int, steps: int, csr, seed: int,
                profile_only: bool = False, repulsion: str = "auto", workload: str = "1b",
                tune: bool = False, write_back: str = "atomic", min_sentences_per_wave: int = 0,
                rounding: str = "stochastic") -> dict:
    w = c.workload
    V = len(c.vocab)
    dev = SgnsDevice(V, w.dim, w.window, w.negative, w.batch_size, precision, rounding=rounding)
    dev.set_exp_table(H.exp_table())
    w_in, w_out = H.init_net(V, w.dim)
    dev.set_weights(w_in, w_out)
    dev.set_sampler(H.unigram_entries(c.counts))
    t_up = time.perf_counter()
    dev.set_corpus(csr.tokens, csr.sent_off, csr.sent_raw)  # host -> HBM, outside the timed region
    upload_s = time.perf_counter() - t_up
    n = len(chunks)
    use_graph = mode == "persistent"
    launch = None
    if mode == "persistent":
        # launch shape: the saved tuning of this workload (no trials in this
        # process), or, with --autotune, the tuner run here on chunk 0
        if tune:
            b0, e0 = chunks[0]
            res = dev.autotune(b0, e0, w.alpha, c.train_words, iters=w.iters, seed=seed, repulsion=repulsion,
                               write_back=write_back)
            launch = dict(res["best"], source="autotuned in this run")
        else:
            saved = autotune.load(workload, w, precision, repulsion, write_back)
            if saved is not None:
                launch = dict(autotune.apply(dev, saved), source=str(
                    autotune.path_for(workload, precision, repulsion, write_back).relative_to(ROOT)))
        log(f"{mode}/{precision}: launch shape {launch or 'built-in default'}")

    def step(i: int, prepare_only: bool = False) -> dict:
        b, e = chunks[i % n]
        return dev.train(mode, b, e, w.alpha, c.train_words, iters=w.iters, epoch=0, seed=seed, use_graph=use_graph,
                         prepare_only=prepare_only, repulsion=repulsion, write_back=write_back,
                         min_sentences_per_wave=min_sentences_per_wave)

    for i in range(n):  # queue orders and graphs are built before anything is timed
        step(i, prepare_only=True)

    for i in range(warmup):
        step(i)
    log(f"{mode}/{precision}: warmup {warmup} steps done")
    t0 = time.perf_counter()
    stats = []
    raw = 0
    for i in range(warmup, warmup + steps):
        stats.append(step(i))
        raw += raw_per_chunk[i % n]
        if (i - warmup) % 20 == 19:
            log(f"{mode}/{precision}: step {i - warmup + 1}/{steps}")
    wall = time.perf_counter() - t0
    done = warmup + steps
    if not profile_only:
        for i in range(done, n):  # finish the epoch (untimed) so quality is scored on a full pass
            step(i)
    kern_ms = sum(s["kernel_ms"] for s in stats)
    tokens = sum(s["tokens"] for s in stats)
    negs = sum(s["negatives"] for s in stats)
    elem = 2 if precision == "bf16" else 4
    byts = algorithmic_bytes(tokens, negs, w.dim, elem)
    kind = "matrix" if precision == "bf16" and repulsion != "vector" else "vector"  # kernel instance
    shape = stats[-1] if stats else {}
    # the same rate when the epoch's corpus upload (host buffers across PCIe) is charged pro rata
    pcie_s = wall + upload_s * raw / max(int(c.stream.size), 1)
    out = dict(mode=mode, precision=precision, kernel=f"{precision}-{kind}", write_back=write_back,
               rounding=rounding if precision == "bf16" else None, wall_s=wall,
               upload_s=upload_s,
               pcie_inclusive_mwords_s=raw / pcie_s / 1e6, kernel_ms=kern_ms, raw_words=raw, tokens=tokens,
               negatives=negs, pairs=sum(s["pairs"] for s in stats), mwords_s=raw / wall / 1e6,
               kernel_mwords_s=raw / (kern_ms / 1e3) / 1e6, alg_bytes=byts, launches=len(stats),
               avg_launch_ms=kern_ms / max(len(stats), 1), achieved_gbs=byts / (kern_ms / 1e3) / 1e9)
    out["grid"] = {k: shape.get(k) for k in ("blocks", "waves_per_block", "sentences_per_dequeue")}
    out["min_sentences_per_wave"] = min_sentences_per_wave
    if mode == "persistent":
        out["launch"] = launch or {"source": "built-in default (most resident waves, one sentence per dequeue)"}
    if not profile_only:
        out["w_in"], out["w_out"] = dev.get_weights()
    dev.close()
    log(f"{mode}/{precision}: {out['mwords_s']:.1f} Mwords/s wall, {out['kernel_mwords_s']:.1f} kernel, "
        f"{out['achieved_gbs']:.0f} GB/s algorithmic")
    return out


def calibration_copy(mib: int) -> None:
    """One elementwise kernel that reads `mib` MiB and writes `mib` MiB (a plain
    copy of known size; torch's copy_ may go to a DMA engine, which counters do
    not see), so a counter run of this command can scale FETCH_SIZE/WRITE_SIZE."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.neg(src, out=dst)
    torch.cuda.synchronize()
    log(f"calibration copy: {mib} MiB read and written by one elementwise kernel")
    del src, dst


def measured_traffic(workload: str, write_back: str, chunk_words: int, variant: str) -> float | None:
    """HBM bytes per launch of the training kernel from the committed counter
    profile of this workload (FETCH_SIZE + WRITE_SIZE, each scaled by the
    calibration copy), or None -- also when the profile was taken with another
    write-back or chunking, or on a different build of the kernel's sources."""
    try:
        t = json.loads(traffic_file(workload).read_text())
        if t.get("workload") != f"{workload}/{write_back}/{chunk_words}":
            return None
        if t.get("kernel_source") != profiling.kernel_source_digest():
            log("counter profile is of another kernel build: traffic not reported")
            return None
        return float(t["variants"][variant]["bytes_per_launch"])
    except (OSError, KeyError, ValueError, TypeError):
        return None


def time_loaders(c, workdir: str) -> dict:
    """Host corpus loading at the workload's size, outside any kernel timing:
    the corpus as text ("w<id>" words, one line per sentence), vocabulary +
    id stream by the text path (LearnVocabFromTrainFile + loadStream,
    cpu:253-292, 359-370), then the binary corpus written and read back.
    Both results are checked against the in-memory corpus.  The files are
    read right after being written, so from the page cache."""
    import shutil
    d = tempfile.mkdtemp(prefix="sgns_load_", dir=workdir)
    out = {}
    try:
        text, binp = os.path.join(d, "corpus.txt"), os.path.join(d, "corpus.bin")
        t = time.perf_counter()
        H.write_raw_text(text, c.raw)
        out["text_write_s"] = round(time.perf_counter() - t, 2)
        out["text_bytes"] = os.path.getsize(text)
        log(f"loader: {out['text_bytes'] / 1e9:.2f} GB of text written in {out['text_write_s']} s")
        t = time.perf_counter()
        v = H.Vocab.learn_text(text, c.workload.min_count)
        out["text_vocab_s"] = round(time.perf_counter() - t, 2)
        t = time.perf_counter()
        s = H.load_text_stream(v, text)
        out["text_stream_s"] = round(time.perf_counter() - t, 2)
        assert len(v) == len(c.vocab) and v.train_words == c.train_words and np.array_equal(s, c.stream)
        log(f"loader: text path {out['text_vocab_s']} s vocabulary + {out['text_stream_s']} s stream")
        t = time.perf_counter()
        H.save_corpus(binp, v, s)
        out["binary_write_s"] = round(time.perf_counter() - t, 2)
        out["binary_bytes"] = os.path.getsize(binp)
        del s, v
        t = time.perf_counter()
        v2, s2 = H.load_corpus(binp)
        out["binary_load_s"] = round(time.perf_counter() - t, 2)
        assert len(v2) == len(c.vocab) and np.array_equal(s2, c.stream)
        n = int(c.stream.size)
        out["ids"] = n
        out["text_mwords_s"] = round(n / (out["text_vocab_s"] + out["text_stream_s"]) / 1e6, 2)
        out["binary_mwords_s"] = round(n / out["binary_load_s"] / 1e6, 2)
        out["threads"] = cpu_threads()
        out["method"] = ("text: byte ranges tokenised on every thread, vocabulary in one pass, the id stream in "
                         "two (count, then fill); binary: one sequential read, ids checked on every thread")
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return out


def quality(c, w_in, w_out, windows) -> dict:
    """Scores of one variant's tables after the epoch (see the module doc)."""
    from par_word2vec_amd import evaluate as E
    out = {}
    try:
        if c.workload.kind == "sbm":
            out["community_knn_accuracy"] = round(E.community_knn_accuracy(w_in, c.labels), 5)
        else:
            out["recall_at_10"] = round(E.planted_cluster_recall(w_in, c.groups), 5)
            out["cluster_margin"] = round(E.cluster_margin(w_in, c.groups), 5)
            out["heldout_loss_per_pair"] = round(E.heldout_loss(w_in, w_out, *windows), 6)
    except Exception as ex:  # noqa: BLE001 - scoring must not hide the throughput line
        log(f"scoring failed: {ex!r}")
    return out


def host_cpu() -> dict:
    """lscpu's model / sockets / cores / threads of this host, and the CPUs
    this process may run on."""
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        info = {k.strip(): v.strip() for k, _, v in (ln.partition(":") for ln in out.splitlines())}
    except (OSError, subprocess.SubprocessError) as ex:
        log(f"lscpu failed: {ex!r}")

    def num(key):
        try:
            return int(info.get(key, ""))
        except ValueError:
            return None
    return {"model": info.get("Model name"), "sockets": num("Socket(s)"), "cores_per_socket": num("Core(s) per socket"),
            "threads_per_core": num("Thread(s) per core"), "online_cpus": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0))}


def cpu_threads() -> int:
    """The CPU share of this process: OMP_NUM_THREADS where the host sets it
    (a GPU's share of a shared host), else every CPU it may run on."""
    return int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))


def cpu_baseline(c, seconds: float, windows=None) -> dict:
    """The oracle's Hogwild restatement of PAR_Word2Vec_cpu.cpp run the way
    the reference's harness runs it (cpu:395-416, 702-706): the sample is
    written as a text file and a "word count" vocabulary file, ReadVocab
    reads the vocabulary, each thread seeks to its byte offset, loads its
    share of words and trains it, and the time covers loading and training.
    The vocabulary keeps the whole corpus's counts, so subsampling and the
    learning-rate schedule are the full corpus's.  When the sample is the
    whole corpus, its tables are scored like the GPU variants' (`windows`:
    the held-out windows of a Zipf workload)."""
    import shutil
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O

    w = c.workload
    threads = cpu_threads()
    # bounded sample: a prefix of the corpus sized to roughly `seconds` of CPU
    # work, calibrated by a short in-memory probe run
    table = O.unigram_table(c.counts)
    probe = min(c.stream.size, 400_000)
    cut = int(np.flatnonzero(c.stream[:probe] == 0)[-1]) + 1
    secs, _, _ = O.train_hogwild(c.stream[:cut], c.counts, c.train_words, w.dim, w.window, w.negative,
                                 w.batch_size, 1, w.alpha, w.sample, threads, table=table)
    del table
    rate = cut / max(secs, 1e-6)
    n = int(min(c.stream.size, max(cut, rate * seconds)))
    ends = np.flatnonzero(c.stream[:n] == 0)
    n = int(ends[-1]) + 1 if ends.size else n
    d = tempfile.mkdtemp(prefix="sgns_cpu_")
    try:
        words = [c.vocab.word(i) for i in range(len(c.vocab))]
        text, vpath = os.path.join(d, "sample.txt"), os.path.join(d, "vocab.txt")
        t0 = time.perf_counter()
        size = O.write_text_corpus(text, c.stream[:n], words)
        O.write_vocab(vpath, words, c.counts)
        log(f"cpu baseline: {n} words, {size / 1e6:.0f} MB of text written in {time.perf_counter() - t0:.1f} s; "
            f"{threads} threads")
        v = O.OracleVocab.read(vpath, text, w.min_count, flavour="fast")
        r = O.train_hogwild_file(v, text, n, v.train_words, w.dim, w.window, w.negative, w.batch_size, 1,
                                 w.alpha, w.sample, threads)
        scores = None
        if n == c.stream.size:  # rows in the oracle vocabulary's order -> the corpus vocabulary's
            rows = np.array([O.lib().orc_vocab_search(v.h, wd.encode()) for wd in words], np.int64)
            if (rows >= 0).all():
                scores = quality(c, r["w_in"][rows], r["w_out"][rows], windows)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    st, secs = r["stats"], r["seconds"]
    return {"value": round(n / secs / 1e6, 3), "unit": "Mwords/s", "cores": threads, "kind": "port",
            "sample": f"first {n} raw words ({st['trained_tokens']} trained tokens) of the same corpus as text, "
                      f"1 epoch, {secs:.1f} s including the per-thread file-offset loads (longest "
                      f"{r['load_max']:.1f} s); OpenMP Hogwild restatement of PAR_Word2Vec_cpu.cpp with naive "
                      f"GEMM loops in place of MKL",
            "load_max_s": round(r["load_max"], 3), "seconds": round(secs, 3),
            "host": host_cpu(), "quality": scores}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="timed chunks (0 = rest of the epoch)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="1b")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--chunk-words", type=int, default=10_000_000)
    ap.add_argument("--variants", default="fp32-plain,bf16-persistent")
    ap.add_argument("--write-back", default="store_out", choices=["atomic", "store_out", "store_in", "store"],
                    help="row write-back of both variants: W_in atomics and W_out plain stores (default), all "
                         "atomics, or plain stores of W_in / both tables (bf16 matrix-core path only)")
    ap.add_argument("--min-sentences-per-wave", type=int, default=0,
                    help="Hogwild staleness bound of both variants: persistent grid floor (0 = 24 sentences per "
                         "wave), plain mode at most n / this sentences in flight (0 = one launch)")
    ap.add_argument("--rounding", default="stochastic", choices=["stochastic", "nearest", "nearest_out", "nearest_in"],
                    help="bf16 write-back rounding: stochastic (default) or round-to-nearest for both tables, or "
                         "nearest for W_out (nearest_out) / W_in (nearest_in) rows only")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-score", action="store_true")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--autotune", action="store_true",
                    help="run the launch-shape tuner in this process instead of using the saved tuning")
    ap.add_argument("--loaders", metavar="DIR", default="",
                    help="time the text and binary corpus loaders on the workload's corpus (files under DIR), "
                         "print one JSON line and exit; no device use")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="run a device copy of this many MiB first (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported by this build "
              f"(the multi-GPU path is a design note, not code)", file=sys.stderr)
        return 2

    if args.loaders:
        c = corpora.make(args.workload, args.scale)
        print(json.dumps({"workload": c.workload.name, "loaders": time_loaders(c, args.loaders)}), flush=True)
        return 0
    if args.calibrate_copy_mb > 0:
        calibration_copy(args.calibrate_copy_mb)
    t0 = time.perf_counter()
    c = corpora.make(args.workload, args.scale)
    w = c.workload
    log(f"corpus {args.workload}: {c.stream.size} raw words, V={len(c.vocab)}, {time.perf_counter() - t0:.1f} s")
    csr = H.subsample(c.stream, c.counts, c.train_words, w.sample, seed=args.seed * 1000003)
    chunks = chunk_bounds(csr.sent_raw, args.chunk_words)
    bounds_raw = [int(csr.sent_raw[b]) for b, _ in chunks] + [int(c.stream.size)]
    raw_per_chunk = [bounds_raw[i + 1] - bounds_raw[i] for i in range(len(chunks))]
    steps = args.steps if args.steps > 0 else max(1, len(chunks) - args.warmup)
    log(f"{csr.n_sent} sentences, {csr.tokens.size} trained tokens, {len(chunks)} chunks; warmup {args.warmup}, "
        f"steps {steps}")

    results = {}
    for v in args.variants.split(","):
        prec, mode, *rep = v.split("-")  # precision-mode[-repulsion]
        results[v] = run_variant(c, mode, prec, chunks, raw_per_chunk, args.warmup, steps, csr, args.seed,
                                 profile_only=args.no_score, repulsion=rep[0] if rep else "auto",
                                 workload=args.workload, tune=args.autotune,
                                 write_back=args.write_back if (prec == "bf16" and not rep) or
                                 args.write_back in ("atomic", "store_out") else "atomic",
                                 min_sentences_per_wave=args.min_sentences_per_wave, rounding=args.rounding)
    for r in results.values():  # counter-measured HBM bytes (the default chunking's profile)
        t = measured_traffic(args.workload, r["write_back"], args.chunk_words, r["kernel"]) if args.scale == 1.0 else None
        if t is not None:
            r["counter_bytes_per_launch"] = t
            r["counter_gbs"] = t / (r["avg_launch_ms"] / 1e3) / 1e9
    scores = {}
    windows = None
    if not args.no_score:
        if w.kind == "zipf":
            from par_word2vec_amd import evaluate as E
            held = corpora.heldout_stream(c, 3_000_000, seed=args.seed + 98)
            windows = E.heldout_windows(held, c.counts, c.train_words, w.sample, w.window, 5, 65_536,
                                        seed=args.seed + 7)
        for v, r in results.items():
            scores[v] = quality(c, r.pop("w_in"), r.pop("w_out"), windows)
            log(f"{v}: {scores[v]}")
    head = results.get("bf16-persistent") or next(iter(results.values()))
    base = results.get("fp32-plain")
    cpu = None if args.no_cpu else cpu_baseline(c, args.cpu_seconds, None if args.no_score else windows)
    line = {
        "metric": "Mwords/sec trained, BF16 persistent vs FP32 plain, at equal planted-cluster score",
        "value": round(head["mwords_s"], 2),
        "unit": "Mwords/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["wall_s"] * 1e3 / steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16" if head["precision"] == "bf16" else "f32",
        "data": "synthetic",
        "config": {"workload": w.name, "raw_words": int(c.stream.size), "vocab": len(c.vocab), "dim": w.dim,
                   "window": w.window, "negative": w.negative, "sample": w.sample, "batch_size": w.batch_size,
                   "chunk_words": args.chunk_words, "chunks": len(chunks), "trained_tokens": int(csr.tokens.size),
                   "variant": f"{head['precision']} {head['mode']}", "write_back": head["write_back"],
                   "parallelism": "single GPU"},
        "fp32_plain_mwords_s": round(base["mwords_s"], 2) if base else None,
        "speedup_vs_fp32_plain": round(head["mwords_s"] / base["mwords_s"], 3) if base else None,
        "quality": scores,
        "variants": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in r.items()}
                     for k, r in results.items()},
        "roofline": {"bound": "hbm", "achieved": round(head["achieved_gbs"], 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(head["achieved_gbs"] / HBM_PEAK_GBS, 4),
                     "traffic": head.get("counter_bytes_per_launch"),
                     "algorithmic_bytes_per_launch": round(head["alg_bytes"] / max(head["launches"], 1)),
                     "kernel": f"sgns_fast_kernel ({head['kernel']} repulsion)",
                     "avg_launch_ms": round(head["avg_launch_ms"], 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
