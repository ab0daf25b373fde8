"""The rocprofv3 summaries bench.py's roofline uses: kernel-instance names,
counter calibration and the stale-profile checks.  CPU only (synthetic CSVs)."""
import csv
import json
import sys
from pathlib import Path

import pytest

from par_word2vec_amd import profiling as P

ROOT = Path(__file__).resolve().parent.parent

BF16_MATRIX = "void sgns::fast::sgns_fast_kernel<unsigned short, 3, 6, true, 0, 1>(sgns::FastArgs)"
FP32_VECTOR = "void sgns::fast::sgns_fast_kernel<float, 3, 5, false, 0, 1>(sgns::FastArgs)"
BF16_VECTOR = "void sgns::fast::sgns_fast_kernel<unsigned short, 2, 8, false, 0, 0>(sgns::FastArgs)"
COPY = "void at::native::vectorized_elementwise_kernel<4, at::native::neg_kernel_cuda(...)>"


def test_variant_of_reads_the_template_arguments():
    assert P.variant_of(BF16_MATRIX) == "bf16-matrix"
    assert P.variant_of(FP32_VECTOR) == "fp32-vector"
    assert P.variant_of(BF16_VECTOR) == "bf16-vector"
    assert P.variant_of(COPY) is None
    assert P.variant_of("void sgns::fast::sgns_sample_kernel(...)") is None


def _counters(path, counter, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for name, value in rows:
            w.writerow({"Kernel_Name": name, "Counter_Name": counter, "Counter_Value": value})


def test_traffic_scales_each_counter_by_the_calibration_copy(tmp_path):
    """1 GiB copied: FETCH_SIZE counts half of it (2 bytes per KiB counted),
    WRITE_SIZE all of it; the kernels' means are scaled by those factors."""
    f, w = tmp_path / "f.csv", tmp_path / "w.csv"
    _counters(f, "FETCH_SIZE", [(COPY, 524288.0), (COPY, 1000.0), (BF16_MATRIX, 100.0), (BF16_MATRIX, 300.0),
                                (FP32_VECTOR, 50.0)])
    _counters(w, "WRITE_SIZE", [(COPY, 1048576.0), (BF16_MATRIX, 10.0), (BF16_MATRIX, 30.0), (FP32_VECTOR, 5.0)])
    t = P.traffic(str(f), str(w), 1024, "text8/atomic/20000000")
    assert t["workload"] == "text8/atomic/20000000" and t["kernel_source"] == P.kernel_source_digest()
    assert t["counters"]["FETCH_SIZE"]["bytes_per_kib"] == 2048.0
    assert t["counters"]["WRITE_SIZE"]["bytes_per_kib"] == 1024.0
    m = t["variants"]["bf16-matrix"]
    assert m["fetch_bytes"] == 200.0 * 2048 and m["write_bytes"] == 20.0 * 1024 and m["dispatches"] == 2
    assert m["bytes_per_launch"] == m["fetch_bytes"] + m["write_bytes"]
    assert t["variants"]["fp32-vector"]["bytes_per_launch"] == 50.0 * 2048 + 5.0 * 1024


def test_bench_ignores_a_profile_of_another_write_back_or_build(tmp_path, monkeypatch):
    sys.path.insert(0, str(ROOT))
    import bench
    good = {"workload": "1b/store_out/10000000", "kernel_source": P.kernel_source_digest(),
            "variants": {"bf16-matrix": {"bytes_per_launch": 123.0}}}
    path = tmp_path / "sgns_fast_kernel_1b_traffic.json"
    monkeypatch.setattr(bench, "traffic_file", lambda workload: path)
    path.write_text(json.dumps(good))
    assert bench.measured_traffic("1b", "store_out", 10_000_000, "bf16-matrix") == 123.0
    assert bench.measured_traffic("1b", "atomic", 10_000_000, "bf16-matrix") is None
    assert bench.measured_traffic("1b", "store_out", 5_000_000, "bf16-matrix") is None
    assert bench.measured_traffic("1b", "store_out", 10_000_000, "fp32-vector") is None
    path.write_text(json.dumps(dict(good, kernel_source="0" * 16)))
    assert bench.measured_traffic("1b", "store_out", 10_000_000, "bf16-matrix") is None


@pytest.mark.parametrize("name", ["1b"])
def test_committed_traffic_profile_matches_the_kernel_sources(name):
    """The committed counter profile was taken on the kernel as it is in the
    tree (otherwise bench.py reports no traffic)."""
    t = json.loads((ROOT / "profiles" / f"sgns_fast_kernel_{name}_traffic.json").read_text())
    assert t["kernel_source"] == P.kernel_source_digest()
    assert {"bf16-matrix", "fp32-vector"} <= set(t["variants"])
