"""bench.py's command line on a host without a GPU: it runs one GPU only and
says so for --gpus N > 1 (exit status 2) before touching any device."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_cli_rejects_an_unknown_rounding():
    """The trainer checks -rounding before it touches a file or a device."""
    cli = ROOT / "par_word2vec_amd" / "lib" / "par_word2vec"
    r = subprocess.run([str(cli), "-train", "/nonexistent", "-rounding", "bogus"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "-rounding" in r.stderr
