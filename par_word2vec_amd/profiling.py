"""Summaries of rocprofv3 output for the training kernel (the figures the
committed profile notes and bench.py's roofline use).

    python -m par_word2vec_amd.profiling stats  <kernel_stats.csv>
    python -m par_word2vec_amd.profiling traffic <fetch counter csv> <write counter csv> <copy MiB> <out.json> [workload/write-back/chunk-words]

`stats` prints the per-kernel dispatch count and average duration of a
`rocprofv3 --kernel-trace --stats --output-format csv` run.  `traffic` takes
two `--pmc FETCH_SIZE` / `--pmc WRITE_SIZE` runs (KiB per dispatch) of a
command that first copies `copy MiB` device to device (bench.py
--calibrate-copy-mb), scales each counter by what it reads for that copy
(bytes per counted KiB) and writes bytes per launch for every training-kernel
variant, with a tag naming the workload, write-back and chunk size (default
"1b/store_out/10000000") and a digest of the kernel's sources, so a profile of
another workload, write-back, chunking or kernel build is not used."""
from __future__ import annotations

import csv
import hashlib
import json
import sys
from collections import defaultdict
from pathlib import Path

COPY_KERNEL = "neg_kernel"  # torch.neg of the calibration buffer (bench.py --calibrate-copy-mb)
KERNEL_SOURCES = ("fast_kernel.h", "fast_kernel.hip", "fast_inst_mfma_atomic.hip", "fast_inst_mfma_hot.hip",
                  "fast_inst_mfma_store.hip", "fast_inst_mfma_store_out.hip", "fast_inst_vec_bf16.hip",
                  "fast_inst_vec_fp32.hip", "sgns_device.h", "sgns_internal.h")


def kernel_source_digest() -> str:
    """SHA-256 over the training kernel's source files: a counter profile is
    valid only for the kernel build it was taken on."""
    h = hashlib.sha256()
    d = Path(__file__).resolve().parent / "csrc" / "device"
    for name in KERNEL_SOURCES:
        h.update((d / name).read_bytes())
    return h.hexdigest()[:16]


def variant_of(kernel_name: str) -> str | None:
    """'bf16-persistent' style name of a training-kernel instance, else None
    (mode is not in the symbol; plain and persistent share an instance)."""
    if "sgns_fast_kernel<" not in kernel_name:
        return None
    # template arguments: element type, pairs per lane, negatives per block,
    # matrix-core repulsion, W_in and W_out write-back
    targs = kernel_name.split("sgns_fast_kernel<", 1)[1].split(">", 1)[0].split(",")
    prec = "bf16" if targs[0].strip() == "unsigned short" else "fp32"
    rep = "matrix" if len(targs) > 3 and targs[3].strip() == "true" else "vector"
    return f"{prec}-{rep}"


def kernel_stats(path: str) -> list[dict]:
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            rows.append({"name": r["Name"], "calls": int(r["Calls"]),
                         "avg_ms": float(r["AverageNs"]) / 1e6, "total_ms": float(r["TotalDurationNs"]) / 1e6})
    return rows


def counter_values(path: str, counter: str) -> dict[str, list[float]]:
    """Per kernel name, the counter's value of each dispatch."""
    acc: dict[str, list[float]] = defaultdict(list)
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] == counter:
                acc[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return dict(acc)


def traffic(fetch_csv: str, write_csv: str, copy_mib: float, workload: str = "1b/store_out/10000000") -> dict:
    copy_bytes = copy_mib * (1 << 20)
    out = {"workload": workload, "kernel_source": kernel_source_digest(), "copy_bytes": copy_bytes, "counters": {},
           "variants": {}}
    per_kernel: dict[str, dict] = defaultdict(dict)
    for name, path in (("FETCH_SIZE", fetch_csv), ("WRITE_SIZE", write_csv)):
        vals = counter_values(path, name)
        copy = [x for k, v in vals.items() if COPY_KERNEL in k for x in v]
        if not copy:
            raise SystemExit(f"no {COPY_KERNEL} dispatch in {path}")
        kib_copy = max(copy)  # the calibration copy is the largest such dispatch
        factor = copy_bytes / kib_copy  # bytes per counted KiB
        out["counters"][name] = {"copy_kib": kib_copy, "bytes_per_kib": factor}
        for k, v in vals.items():
            var = variant_of(k)
            if var:
                kib = sum(v) / len(v)
                per_kernel[var][name] = {"dispatches": len(v), "kib": kib, "bytes": kib * factor}
    for v, c in per_kernel.items():
        if "FETCH_SIZE" in c and "WRITE_SIZE" in c:
            out["variants"][v] = {"fetch_bytes": c["FETCH_SIZE"]["bytes"], "write_bytes": c["WRITE_SIZE"]["bytes"],
                                  "bytes_per_launch": c["FETCH_SIZE"]["bytes"] + c["WRITE_SIZE"]["bytes"],
                                  "dispatches": c["FETCH_SIZE"]["dispatches"]}
    return out


def main(argv: list[str]) -> int:
    if len(argv) >= 2 and argv[0] == "stats":
        for r in kernel_stats(argv[1]):
            print(f"{r['calls']:8d} {r['avg_ms']:12.4f} ms avg {r['total_ms']:12.2f} ms total  {r['name'][:110]}")
        return 0
    if len(argv) in (5, 6) and argv[0] == "traffic":
        t = traffic(argv[1], argv[2], float(argv[3]), *argv[5:6])
        with open(argv[4], "w") as f:
            json.dump(t, f, indent=1)
        print(json.dumps(t, indent=1))
        return 0
    print(__doc__)
    return 2


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
