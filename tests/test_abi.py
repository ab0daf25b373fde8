"""The C-ABI libraries load and export every entry point their headers
declare, and the Python bindings cover all of them.  No compute calls."""
import re
from pathlib import Path

import pytest

from par_word2vec_amd import _native as N

ROOT = Path(__file__).resolve().parent.parent


def declared(header: str) -> set[str]:
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b((?:ph|par_sgns)_[a-z0-9_]+)\s*\(", text))


@pytest.mark.parametrize("header,loader,table", [
    ("par_host.h", N.host, N.HOST_SYMBOLS),
    ("par_sgns.h", N.sgns, N.SGNS_SYMBOLS),
])
def test_every_declared_symbol_is_exported_and_bound(header, loader, table):
    names = declared(header)
    assert names, header
    lib = loader()
    for n in sorted(names):
        assert hasattr(lib, n), f"{n} declared in {header} but not exported"
    assert names == set(table), f"binding table differs: {names ^ set(table)}"


def test_create_rejects_bad_config_without_gpu():
    import ctypes as C
    lib = N.sgns()
    h = C.c_void_p()
    cfg = N.ParSgnsConfig(1, 100, 5, 5, 24, 0, 0, 0)  # vocab_size 1 is invalid
    assert lib.par_sgns_create(C.byref(cfg), C.byref(h)) == -1 and not h.value


def test_launch_shape_calls_reject_a_null_handle_without_gpu():
    import ctypes as C
    lib = N.sgns()
    args = N.ParSgnsTrainArgs()
    best = N.ParSgnsTunePoint()
    n = C.c_int32(0)
    assert lib.par_sgns_autotune(None, C.byref(args), 1, None, 0, C.byref(n), C.byref(best)) == -1
    assert lib.par_sgns_set_launch_shape(None, 1, 8, 1) == -1


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(N, "LIB_DIR", tmp_path)
    with pytest.raises(N.NativeLibraryMissing):
        N._load("libpar_sgns.so")
