"""The oracle's multi-thread Hogwild trainers, which the CPU baseline times.

  orc_train_hogwild        equal word slices of an in-memory stream; with one
                           thread (seed 0, the whole stream) it must be
                           orc_train bit for bit (cpu:397-416 with -threads 1)
  orc_train_hogwild_file   the reference's harness: per-thread byte-offset
                           slices of a text file, loading inside the timed
                           region (cpu:395-416, 702-706); with one thread it
                           loads the whole file and must again be orc_train on
                           the stream loadStream (cpu:359-370) reads; with
                           several threads the slices must cover the file
"""
import numpy as np
import pytest

import oracle as O
from par_word2vec_amd import host as H

DIM, W, NEG, B = 16, 3, 4, 8


@pytest.fixture(scope="module")
def corpus(tmp_path_factory):
    raw, _ = H.synth_zipf(60_000, 400, 1.0, line_len=150, seed=7)
    d = tmp_path_factory.mktemp("hogwild")
    words = ["</s>"] + [f"w{i:04d}" for i in range(1, 401)]
    text = str(d / "train.txt")
    O.write_text_corpus(text, raw, words)
    learned = O.OracleVocab.learn(text, 5)
    vwords = [learned.word(i) for i in range(len(learned))]
    vpath = str(d / "vocab.txt")
    O.write_vocab(vpath, vwords, learned.counts())
    return text, vpath, learned


def test_hogwild_one_thread_is_orc_train(corpus):
    text, _, v = corpus
    stream, counts, T = v.stream(text), v.counts(), v.train_words
    ref = O.train(stream, counts, T, DIM, W, NEG, B, 1, 0.025, 1e-3)
    wo = []
    secs, st, w_in = O.train_hogwild(stream, counts, T, DIM, W, NEG, B, 1, 0.025, 1e-3, 1, flavour="parity",
                                     w_out_out=wo)
    assert secs >= 0
    assert np.array_equal(w_in.view(np.uint32), ref.w_in.view(np.uint32))
    assert np.array_equal(wo[0].view(np.uint32), ref.w_out.view(np.uint32))
    for k in ("raw_words", "trained_tokens", "sentences", "pairs", "batches", "negatives", "rep_pairs", "draws",
              "final_rng"):
        assert st[k] == ref.stats[k], k
    assert st["final_alpha"] == ref.stats["final_alpha"]


def test_hogwild_file_one_thread_is_orc_train(corpus):
    text, vpath, _ = corpus
    v = O.OracleVocab.read(vpath, text, 5)
    stream, counts, T = v.stream(text), v.counts(), v.train_words
    assert v.file_size == len(open(text, "rb").read())
    ref = O.train(stream, counts, T, DIM, W, NEG, B, 1, 0.025, 1e-3)
    r = O.train_hogwild_file(v, text, stream.size, T, DIM, W, NEG, B, 1, 0.025, 1e-3, 1, flavour="parity")
    assert 0 <= r["load_max"] <= r["seconds"]
    assert np.array_equal(r["w_in"].view(np.uint32), ref.w_in.view(np.uint32))
    assert np.array_equal(r["w_out"].view(np.uint32), ref.w_out.view(np.uint32))
    assert r["stats"]["raw_words"] == ref.stats["raw_words"]
    assert r["stats"]["pairs"] == ref.stats["pairs"]


@pytest.mark.parametrize("threads", [3, 8])
def test_hogwild_file_slices_cover_the_file(corpus, threads):
    """Byte-offset slices: every thread loads about 1/threads of the words;
    the union reads each word about once (a thread may start inside a word
    and stops after ceil(words / threads) words, so a few words at the seams
    are read twice or not at all)."""
    text, vpath, _ = corpus
    v = O.OracleVocab.read(vpath, text, 5)
    n = v.stream(text).size
    r = O.train_hogwild_file(v, text, n, v.train_words, DIM, W, NEG, B, 1, 0.025, 1e-3, threads)
    raw = r["stats"]["raw_words"]
    assert abs(raw - n) <= 0.02 * n, (raw, n)
    assert r["stats"]["pairs"] > 0
    assert np.isfinite(r["w_in"]).all() and np.abs(r["w_in"]).max() > 0


def test_hogwild_file_reports_unreadable_file(corpus):
    _, vpath, _ = corpus
    text = corpus[0]
    v = O.OracleVocab.read(vpath, text, 5)
    with pytest.raises(RuntimeError, match="cannot read"):
        O.train_hogwild_file(v, text + ".missing", 1000, v.train_words, DIM, W, NEG, B, 1, 0.025, 1e-3, 2)
