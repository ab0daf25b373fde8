"""Builds the native libraries of par_word2vec_amd in-tree.

    libpar_host.so   host C++ (tokenizer, vocabulary, subsampling, schedule,
                     synthetic corpora, writer)            -- g++
    libpar_sgns.so   HIP kernels for gfx950 + the device C-ABI  -- hipcc
    par_word2vec     command-line trainer (pWord2Vec-style flags)

Every call recompiles (the build check runs on a machine without a GPU;
hipcc cross-compiles gfx950 code objects).
"""
from __future__ import annotations

import os
import shlex
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
LIB = PKG / "lib"
OBJ = PKG / "_build"
INC = ROOT / "include"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PAR_SGNS_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-munsafe-fp-atomics", "-fno-gpu-flush-denormals-to-zero",
    # packed fp32 instructions stay disabled on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INC}",
]
DEVICE_SOURCES = [
    ("det_kernel.hip", ["-ffp-contract=off"]),
    ("fast_kernel.hip", []),
    ("fast_inst_vec_fp32.hip", []),
    ("fast_inst_vec_bf16.hip", []),
    ("fast_inst_mfma_atomic.hip", []),
    ("fast_inst_mfma_store_out.hip", []),
    ("fast_inst_mfma_store.hip", []),
    ("fast_inst_mfma_hot.hip", []),
    ("window_kernel.hip", []),
    ("api.hip", ["-fopenmp"]),
]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print("+", " ".join(shlex.quote(c) for c in cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")


def build_host(verbose: bool = False) -> Path:
    LIB.mkdir(exist_ok=True)
    out = LIB / "libpar_host.so"
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-fopenmp", f"-I{INC}",
          "-o", str(out), str(PKG / "csrc" / "host" / "par_host.cpp")], verbose)
    return out


def build_device(verbose: bool = False) -> Path:
    LIB.mkdir(exist_ok=True)
    OBJ.mkdir(exist_ok=True)
    def compile_one(item):
        src, extra = item
        o = OBJ / (Path(src).stem + ".o")
        _run([HIPCC, *HIP_FLAGS, *extra, "-c", str(PKG / "csrc" / "device" / src), "-o", str(o)], verbose)
        return str(o)

    # translation units compile in parallel (the training kernel's instances are split over several)
    with ThreadPoolExecutor(max_workers=min(len(DEVICE_SOURCES), os.cpu_count() or 1)) as ex:
        objs = list(ex.map(compile_one, DEVICE_SOURCES))
    out = LIB / "libpar_sgns.so"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-fopenmp", "-o", str(out), *objs], verbose)
    return out


def build_cli(verbose: bool = False) -> Path:
    out = LIB / "par_word2vec"
    _run(["g++", "-std=c++17", "-O2", "-fopenmp", f"-I{INC}", "-o", str(out),
          str(PKG / "csrc" / "host" / "cli.cpp"),
          f"-L{LIB}", "-lpar_host", "-lpar_sgns", f"-Wl,-rpath,$ORIGIN",
          "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"], verbose)
    return out


def build_all(verbose: bool = False) -> None:
    build_host(verbose)
    build_device(verbose)
    if (PKG / "csrc" / "host" / "cli.cpp").exists():
        build_cli(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
