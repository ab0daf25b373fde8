"""Launch-shape autotuner of the persistent SGNS kernel and its saved tunings.

    python -m par_word2vec_amd.autotune --workload 1b [--precision bf16] [--write-back store_out] [--reps 3]

times every feasible launch shape of the persistent kernel -- waves per block
{1, 2, 4, 8}, resident waves per CU {4, 6, 8} (as many as the LDS a wave needs
allows) and sentences per dequeue {1, 2, 4} -- on the workload's first chunk
(par_sgns_autotune: both tables are restored after every trial, so nothing is
trained) and writes the measurements and the fastest shape to
``tuned/<workload>-<precision>-<repulsion>[-<write-back>].json`` beside this
module (the write-back policy changes the kernel instance and its timing).  The
LDS window itself is the ring of 2W+1 word-keyed slots, fixed by the window;
its bytes per wave decide how many waves per CU the tuner may try.

``load()`` returns a saved tuning whose kernel shape (dim, window, negatives,
batch, precision, repulsion) matches; ``apply()`` makes it a device handle's
default (par_sgns_set_launch_shape).  bench.py uses both, so the timed run
launches no tuning trials.  The reference has no tuner: its launch shape is
fixed (gamma*56 blocks of `dim` threads, PAR_Word2Vec_gpu_text.cu:848-868).
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

TUNED = Path(__file__).resolve().parent / "tuned"


def shape_key(w, precision: str, repulsion: str = "auto", write_back: str = "atomic") -> dict:
    """What a tuning depends on besides the device: the kernel instance and
    its LDS per wave."""
    k = {"dim": int(w.dim), "window": int(w.window), "negative": int(w.negative),
         "batch_size": int(w.batch_size), "precision": precision, "repulsion": repulsion}
    if write_back != "atomic":
        k["write_back"] = write_back
    return k


def path_for(workload: str, precision: str, repulsion: str = "auto", write_back: str = "atomic") -> Path:
    tail = "" if write_back == "atomic" else f"-{write_back}"
    return TUNED / f"{workload}-{precision}-{repulsion}{tail}.json"


def load(workload: str, w, precision: str, repulsion: str = "auto", write_back: str = "atomic") -> dict | None:
    """The saved tuning of this workload, kernel shape and write-back policy, or None."""
    p = path_for(workload, precision, repulsion, write_back)
    if not p.exists():
        return None
    data = json.loads(p.read_text())
    if data.get("key") != shape_key(w, precision, repulsion, write_back):
        return None
    return data


def apply(dev, tuning: dict) -> dict:
    """Make the tuning's best shape the handle's default; returns that shape."""
    b = tuning["best"]
    dev.set_launch_shape(int(b["waves_per_block"]), int(b["waves_per_cu"]), int(b["sentences_per_dequeue"]))
    return b


def device_name() -> dict:
    try:
        import torch
        if torch.cuda.is_available():
            p = torch.cuda.get_device_properties(0)
            return {"name": p.name, "arch": getattr(p, "gcnArchName", ""), "cus": int(p.multi_processor_count)}
    except Exception:  # informational only
        pass
    return {}


def run(workload: str, scale: float, precision: str, repulsion: str, chunk_words: int, reps: int, seed: int,
        out: Path | None, write_back: str = "atomic") -> dict:
    from . import corpora, host as H
    from .device import SgnsDevice
    from .trainer import chunk_bounds
    c = corpora.make(workload, scale)
    w = c.workload
    csr = H.subsample(c.stream, c.counts, c.train_words, w.sample, seed=seed * 1000003)
    b, e = chunk_bounds(csr.sent_raw, chunk_words)[0]
    V = len(c.vocab)
    with SgnsDevice(V, w.dim, w.window, w.negative, w.batch_size, precision) as dev:
        dev.set_exp_table(H.exp_table())
        dev.set_weights(*H.init_net(V, w.dim))
        dev.set_sampler(H.unigram_entries(c.counts))
        dev.set_corpus(csr.tokens, csr.sent_off, csr.sent_raw)
        dev.train("persistent", b, e, w.alpha, c.train_words, iters=w.iters, seed=seed, repulsion=repulsion,
                  write_back=write_back)  # warm
        t0 = time.perf_counter()
        res = dev.autotune(b, e, w.alpha, c.train_words, iters=w.iters, seed=seed, reps=reps, repulsion=repulsion,
                           write_back=write_back)
        secs = time.perf_counter() - t0
    data = {"workload": workload, "scale": scale, "key": shape_key(w, precision, repulsion, write_back),
            "device": device_name(),
            "sample": {"sentences": [int(b), int(e)], "trained_tokens": int(csr.sent_off[e] - csr.sent_off[b]),
                       "reps": reps, "seconds": round(secs, 2)},
            "best": res["best"], "points": res["points"]}
    if out is not None:
        out.parent.mkdir(parents=True, exist_ok=True)
        out.write_text(json.dumps(data, indent=1) + "\n")
    return data


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workload", default="1b")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--precision", default="bf16", choices=["bf16", "fp32"])
    ap.add_argument("--repulsion", default="auto", choices=["auto", "vector", "matrix"])
    ap.add_argument("--write-back", default="atomic", choices=["atomic", "store_out", "store_in", "store"])
    ap.add_argument("--chunk-words", type=int, default=10_000_000)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default="", help="output file (default: the saved-tuning path of the workload)")
    a = ap.parse_args()
    out = Path(a.out) if a.out else path_for(a.workload, a.precision, a.repulsion, a.write_back)
    data = run(a.workload, a.scale, a.precision, a.repulsion, a.chunk_words, a.reps, a.seed, out, a.write_back)
    for p in data["points"]:
        print(f"  waves/block {p['waves_per_block']}  waves/CU {p['waves_per_cu']}  "
              f"sentences/dequeue {p['sentences_per_dequeue']}: {p['kernel_ms']:.3f} ms", flush=True)
    print(f"best {data['best']} -> {out}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
