"""End to end through the drop-in executable: text corpus in, embedding file
out, compared with the oracle's training + saveModel (cpu:721-737)."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from par_word2vec_amd import host as H

pytestmark = pytest.mark.gpu
CLI = Path(__file__).resolve().parent.parent / "par_word2vec_amd" / "lib" / "par_word2vec"


def _corpus(tmp_path):
    raw, _ = H.synth_zipf(80_000, 2000, 1.0, line_len=300, n_clusters=5, cluster_size=10, p_replace=0.3, seed=21)
    p = str(tmp_path / "corpus.txt")
    H.write_raw_text(p, raw)
    return p


def test_cli_det_output_bytes_equal_oracle(tmp_path):
    p = _corpus(tmp_path)
    out = tmp_path / "vec.bin"
    flags = ["-train", p, "-output", str(out), "-size", "64", "-window", "5", "-negative", "5", "-sample", "1e-3",
             "-iter", "2", "-min-count", "5", "-batch-size", "24", "-binary", "1", "-mode", "det",
             "-time-file", str(tmp_path / "t")]
    r = subprocess.run([str(CLI), *flags], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    ov = O.OracleVocab.learn(p, 5)
    st = ov.stream(p)
    res = O.train(st, ov.counts(), ov.train_words, 64, 5, 5, 24, 2, 0.025, 1e-3)
    ov.save_model(str(tmp_path / "ref.bin"), res.w_in, True)
    assert out.read_bytes() == (tmp_path / "ref.bin").read_bytes()
    assert (tmp_path / "t").read_text().startswith("Elapsed ")


def test_cli_persistent_text_output(tmp_path):
    p = _corpus(tmp_path)
    out = tmp_path / "vec.txt"
    r = subprocess.run([str(CLI), "-train", p, "-output", str(out), "-size", "32", "-iter", "2", "-precision",
                        "bf16", "-sample", "1e-4", "-save-vocab", str(tmp_path / "v.txt"),
                        "-time-file", str(tmp_path / "t")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = out.read_text().split("\n")
    V, d = map(int, lines[0].split())
    assert d == 32 and V == len(H.Vocab.learn_text(p, 5))
    assert lines[1].startswith("</s> ")
    vals = np.array(lines[2].split()[1:], dtype=np.float64)
    assert vals.size == 32 and np.isfinite(vals).all()
    assert (tmp_path / "v.txt").read_text().startswith("</s> ")


def test_cli_autotune(tmp_path):
    """-autotune N times the launch shapes on the first chunk before training
    (untimed) and reports the chosen one; the run then trains as usual."""
    p = _corpus(tmp_path)
    out = tmp_path / "vec.txt"
    r = subprocess.run([str(CLI), "-train", p, "-output", str(out), "-size", "64", "-iter", "1", "-precision",
                        "bf16", "-sample", "1e-4", "-autotune", "1", "-time-file", str(tmp_path / "t")],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "launch shape (" in r.stdout and "sentences per dequeue" in r.stdout
    V, d = map(int, out.read_text().split("\n")[0].split())
    assert d == 64 and V > 1


def test_cli_min_sentences_per_wave(tmp_path):
    """-min-sentences-per-wave reaches the persistent grid; a bad value is an error."""
    p = _corpus(tmp_path)
    out = tmp_path / "vec.txt"
    base = [str(CLI), "-train", p, "-output", str(out), "-size", "32", "-iter", "1", "-precision", "bf16",
            "-sample", "1e-4", "-time-file", str(tmp_path / "t")]
    shapes = {}
    for floor in ("1", "16", "100000"):
        r = subprocess.run(base + ["-min-sentences-per-wave", floor, "-debug", "2"], capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stderr
        assert out.read_text().split("\n")[0].split()[1] == "32"
        line = next(x for x in r.stdout.splitlines() if x.startswith("launch shape:"))
        blocks, waves = int(line.split()[2]), int(line.split()[5])
        shapes[floor] = blocks * waves
    # the floor reached the grid: one wavefront at 100000 sentences per wave,
    # more wavefronts at one sentence per wave than at the default 16
    assert shapes["100000"] == 1
    assert shapes["1"] > shapes["16"]
    r = subprocess.run(base + ["-min-sentences-per-wave", "-3"], capture_output=True, text=True, timeout=300)
    assert r.returncode != 0 and "min_sentences_per_wave" in r.stderr


def test_cli_binary_corpus_trains_like_text(tmp_path):
    """-train-bin on the binary corpus that -save-bin wrote from the text gives
    the same embedding file byte for byte (deterministic mode)."""
    p = _corpus(tmp_path)
    b = tmp_path / "corpus.bin"
    r = subprocess.run([str(CLI), "-train", p, "-save-bin", str(b)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    flags = ["-size", "64", "-window", "5", "-negative", "5", "-sample", "1e-3", "-iter", "1", "-binary", "1",
             "-mode", "det", "-time-file", str(tmp_path / "t")]
    outs = []
    for src in (["-train", p], ["-train-bin", str(b)]):
        out = tmp_path / f"vec{len(outs)}.bin"
        r = subprocess.run([str(CLI), *src, "-output", str(out), *flags], capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stderr
        outs.append(out.read_bytes())
    assert outs[0] == outs[1]
