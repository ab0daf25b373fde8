"""Host library (libpar_host.so) against the reference semantics and the
oracle: tokenizer, vocabulary, id streams, the exact random schedule of the
deterministic mode, subsampling into the sentence CSR, unigram table, writer
and synthetic corpora, the binary corpus format.  CPU only."""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from par_word2vec_amd import corpora, host as H

CLI = Path(__file__).resolve().parent.parent / "par_word2vec_amd" / "lib" / "par_word2vec"


def write(tmp_path, name, data: bytes):
    p = tmp_path / name
    p.write_bytes(data)
    return str(p)


def test_tokenizer_known_answer(tmp_path):
    """ReadWord (cpu:120-144): \\r dropped, tabs/spaces delimit, each newline is
    "</s>", a word cut off by EOF is dropped (feof, cpu:172)."""
    p = write(tmp_path, "t.txt", b"a b\r\nc\t\tb\n\nd a a\nx")
    v = H.Vocab.learn_text(p, 1)
    words = {v.word(i): int(c) for i, c in enumerate(v.counts())}
    assert words == {"</s>": 4, "a": 3, "b": 2, "c": 1, "d": 1}
    assert v.word(0) == "</s>" and v.train_words == 11
    ov = O.OracleVocab.learn(p, 1)
    assert {ov.word(i): int(c) for i, c in enumerate(ov.counts())} == words
    ids = H.load_text_stream(v, p)
    assert [v.word(i) for i in ids] == ["a", "b", "</s>", "c", "b", "</s>", "</s>", "d", "a", "a", "</s>"]


def test_tokenizer_long_words_match_oracle(tmp_path):
    long1 = b"x" * 150
    long2 = b"x" * 98 + b"yz"  # differs after byte 98: same truncated word
    data = b" ".join([long1, b"short", long2, b"short", long1]) + b"\n"
    p = write(tmp_path, "l.txt", data)
    v = H.Vocab.learn_text(p, 1)
    ov = O.OracleVocab.learn(p, 1)
    assert len(v) == len(ov)
    for i in range(len(v)):
        assert v.word(i) == ov.word(i)
    assert np.array_equal(v.counts(), ov.counts())
    assert len(v.word(1)) == 98


def test_sort_ties_first_appearance_and_min_count(tmp_path):
    p = write(tmp_path, "s.txt", b"c b a c b a d d d e\n")
    v = H.Vocab.learn_text(p, 2)
    assert [v.word(i) for i in range(len(v))] == ["</s>", "d", "c", "b", "a"]
    v1 = H.Vocab.learn_text(p, 5)
    assert [v1.word(i) for i in range(len(v1))] == ["</s>"]  # </s> is never pruned


@pytest.mark.parametrize("seed", [1, 2])
def test_raw_path_equals_text_path_and_oracle(tmp_path, seed):
    raw, _ = H.synth_zipf(50_000, 2000, 1.0, line_len=123, n_clusters=4, cluster_size=5, p_replace=0.3, seed=seed)
    v = H.Vocab.from_raw(raw, 2000, 5)
    st = H.map_raw_stream(v, raw)
    p = str(tmp_path / "c.txt")
    H.write_raw_text(p, raw)
    vt = H.Vocab.learn_text(p, 5)
    ov = O.OracleVocab.learn(p, 5)
    assert len(v) == len(vt) == len(ov)
    assert np.array_equal(v.counts(), vt.counts()) and np.array_equal(v.counts(), ov.counts())
    assert all(v.word(i) == vt.word(i) == ov.word(i) for i in range(len(v)))
    assert np.array_equal(st, H.load_text_stream(vt, p))
    assert np.array_equal(st, ov.stream(p))


@pytest.mark.parametrize("W,NEG,B,iters,sample,V,line", [
    (5, 5, 24, 1, 1e-3, 3000, 300),
    (8, 5, 16, 2, 1e-4, 3000, 1500),   # lines longer than 1000 kept tokens are split
    (3, 2, 7, 1, 0.0, 400, 50),
    (2, 3, 5, 3, 1e-2, 60, 40),        # tiny vocabulary: many in-sentence rejections
])
def test_plan_equals_oracle_schedule(W, NEG, B, iters, sample, V, line):
    raw, _ = H.synth_zipf(60_000, V, 1.0, line_len=line, seed=9)
    v = H.Vocab.from_raw(raw, V, 1)
    stream = H.map_raw_stream(v, raw)
    counts, T = v.counts(), v.train_words
    plan = H.build_plan(stream, counts, T, W, NEG, B, iters, 0.025, sample)
    r = O.train(stream, counts, T, 8, W, NEG, B, iters, 0.025, sample, want_trace=True)
    tr = r.trace
    assert np.array_equal(plan.tokens, tr["tokens"])
    assert np.array_equal(plan.b, tr["b"])
    assert np.array_equal(np.diff(plan.sent_off), tr["sent_len"])
    assert np.array_equal(plan.sent_alpha, tr["sent_alpha"])
    assert np.array_equal(plan.negs, tr["negs"])
    assert np.array_equal(np.diff(plan.neg_off) // max(NEG, 1), tr["rep_b"]) or NEG == 0
    assert plan.draws == r.stats["draws"] and plan.final_rng == r.stats["final_rng"]
    assert plan.final_alpha == r.stats["final_alpha"]


def test_plan_drops_unterminated_tail():
    stream = np.array([1, 2, 3, 0, 4, 5, 6], np.int32)  # last line has no "</s>"
    counts = np.array([1, 5, 5, 5, 5, 5, 5], np.uint32)
    plan = H.build_plan(stream, counts, 7, 2, 1, 4, 1, 0.025, 0.0)
    assert list(plan.tokens) == [1, 2, 3]
    csr = H.subsample(stream, counts, 7, 0.0, 1)
    assert list(csr.tokens) == [1, 2, 3] and list(csr.sent_off) == [0, 3]


def test_subsample_csr_properties():
    raw, _ = H.synth_zipf(300_000, 5000, 1.0, line_len=2500, seed=4)
    v = H.Vocab.from_raw(raw, 5000, 1)
    stream = H.map_raw_stream(v, raw)
    counts, T = v.counts(), v.train_words
    full = H.subsample(stream, counts, T, 0.0, 1)
    L = np.diff(full.sent_off)
    assert L.max() == 1000 and full.tokens.size == np.count_nonzero(stream)  # lines of 2500 cut at 1000
    assert np.array_equal(full.tokens, stream[stream != 0])
    assert full.sent_raw[0] == 0 and np.all(np.diff(full.sent_raw) > 0)
    sample = 1e-3
    s = H.subsample(stream, counts, T, sample, 7)
    kept = np.bincount(s.tokens, minlength=counts.size)
    ratio = (np.float32(sample) * np.float32(T)) / counts.astype(np.float32)
    p_keep = np.minimum(1.0, np.sqrt(ratio) + ratio)
    top = np.argsort(-counts)[1:20]
    expect = counts[top] * p_keep[top]
    assert np.all(np.abs(kept[top] - expect) < 5 * np.sqrt(expect) + 5)
    # deterministic given the seed, independent of the thread count
    s2 = H.subsample(stream, counts, T, sample, 7)
    assert np.array_equal(s.tokens, s2.tokens) and np.array_equal(s.sent_off, s2.sent_off)


@pytest.mark.parametrize("seed", [0, 1])
def test_unigram_entries_equal_oracle_table(seed):
    rng = np.random.default_rng(seed)
    counts = np.sort(rng.zipf(1.3, 20_000).astype(np.uint32))[::-1].copy()
    counts[0] = 123
    ent = H.unigram_entries(counts)
    tab = O.unigram_table(counts)
    assert np.array_equal(ent, np.bincount(tab, minlength=counts.size))
    assert ent.sum() == 100_000_000


def test_exp_table_and_init_net_equal_oracle():
    assert np.array_equal(H.exp_table(), O.exp_table())
    a, b = H.init_net(123, 77)
    oa, ob = O.init_net(123, 77)
    assert np.array_equal(a, oa) and np.array_equal(b, ob)


@pytest.mark.parametrize("binary", [0, 1])
def test_writer_bytes_equal_oracle(tmp_path, binary):
    raw, _ = H.synth_zipf(20_000, 300, 1.0, line_len=50, seed=2)
    p = str(tmp_path / "c.txt")
    H.write_raw_text(p, raw)
    v = H.Vocab.learn_text(p, 5)
    ov = O.OracleVocab.learn(p, 5)
    w = (np.random.default_rng(0).standard_normal((len(v), 17)) * 0.1).astype(np.float32)
    v.save_embeddings(str(tmp_path / "a.out"), w, bool(binary))
    ov.save_model(str(tmp_path / "b.out"), w, bool(binary))
    assert (tmp_path / "a.out").read_bytes() == (tmp_path / "b.out").read_bytes()


def test_vocab_save_read_roundtrip(tmp_path):
    raw, _ = H.synth_zipf(20_000, 300, 1.0, line_len=50, seed=3)
    v = H.Vocab.from_raw(raw, 300, 5)
    p = str(tmp_path / "v.txt")
    v.save(p)
    v2 = H.Vocab.read(p, 5)
    assert len(v2) == len(v) and np.array_equal(v2.counts(), v.counts())
    assert all(v.word(i) == v2.word(i) for i in range(len(v)))


def test_synthetic_zipf_structure():
    raw, cw = H.synth_zipf(2_000_000, 10_000, 1.0, lognorm=(3.0, 0.8), len_clip=(5, 2000), n_clusters=50,
                           cluster_size=20, p_replace=0.3, seed=5)
    ends = np.flatnonzero(raw == 0)
    lens = np.diff(np.concatenate([[-1], ends])) - 1
    assert lens[:-1].min() >= 5 and lens.max() <= 2000
    assert abs(np.log(lens[:-1]).mean() - 3.0) < 0.05
    assert len(np.unique(cw)) == 1000
    raw2, cw2 = H.synth_zipf(2_000_000, 10_000, 1.0, lognorm=(3.0, 0.8), len_clip=(5, 2000), n_clusters=50,
                             cluster_size=20, p_replace=0.3, seed=5)
    assert np.array_equal(raw, raw2) and np.array_equal(cw, cw2)
    # planted co-occurrence: a cluster word's neighbours are same-cluster far above chance
    owner = np.full(10_001, -1)
    owner[cw.reshape(-1)] = np.repeat(np.arange(50), 20)
    a, b = raw[:-1], raw[1:]
    m = (owner[a] >= 0) & (b != 0)
    assert (owner[b[m]] == owner[a[m]]).mean() > 0.2


def test_synthetic_sbm_walks():
    raw, comm = H.synth_sbm_walks(n_nodes=2000, n_comm=20, walks_per_node=2, walk_len=40, seed=3)
    lines = np.split(raw, np.flatnonzero(raw == 0) + 1)[:-1]
    assert len(lines) == 4000
    nodes = raw[raw > 0] - 1
    a, b = raw[:-1], raw[1:]
    m = (a > 0) & (b > 0)
    intra = (comm[a[m] - 1] == comm[b[m] - 1]).mean()
    assert 0.75 < intra < 0.95
    assert nodes.max() < 2000


def test_sbm_graph_matches_its_specification():
    """The config-3 graph (10,000 nodes, 20 communities, degree exponent 2.5,
    mean degree 40, 85% intra-community edges): the mean degree is 40, the
    fraction of intra-community edges is 0.85 (less the self-loops the
    generator redraws), and the Hill estimate of the degree tail over the top
    5% of nodes is the power law's exponent minus one (1.5; 2.0 would mean
    exponent 3).  The walks step along intra-community edges at the same rate."""
    deg, m, intra = H.synth_sbm_graph(n_nodes=10_000, n_comm=20, exponent=2.5, mean_degree=40.0, p_intra=0.85,
                                      seed=3)
    assert m == 200_000 and deg.sum() == 2 * m
    assert deg.mean() == pytest.approx(40.0, rel=1e-3)
    assert intra / m == pytest.approx(0.85, abs=0.005)
    d = np.sort(deg)[::-1].astype(np.float64)
    k = 500
    hill = k / np.log(d[:k] / d[k]).sum()
    assert hill == pytest.approx(1.5, abs=0.15)
    other, _, _ = H.synth_sbm_graph(n_nodes=10_000, n_comm=20, exponent=3.5, mean_degree=40.0, p_intra=0.85, seed=3)
    o = np.sort(other)[::-1].astype(np.float64)
    assert k / np.log(o[:k] / o[k]).sum() > 2.0  # a steeper law is told apart
    raw, comm = H.synth_sbm_walks(n_nodes=10_000, n_comm=20, exponent=2.5, mean_degree=40.0, p_intra=0.85,
                                  walks_per_node=2, walk_len=80, seed=3)
    a, b = raw[:-1], raw[1:]
    mk = (a > 0) & (b > 0)
    assert (comm[a[mk] - 1] == comm[b[mk] - 1]).mean() == pytest.approx(intra / m, abs=0.01)


def test_workloads_generate_at_small_scale():
    for name in ("zipf1m", "sbm"):
        c = corpora.make(name, scale=0.05)
        assert c.stream.size > 0 and c.train_words == c.stream.size


# ---------------------------------------------------------------- binary corpus
def _text_corpus(tmp_path, n=1_500_000, seed=5):
    raw, _ = H.synth_zipf(n, 20_000, 1.0, lognorm=(3.0, 0.8), len_clip=(5, 2000), n_clusters=20, cluster_size=10,
                          p_replace=0.2, seed=seed)
    p = str(tmp_path / "corpus.txt")
    H.write_raw_text(p, raw)
    return p


def test_binary_corpus_equals_text_path(tmp_path):
    """ph_corpus_save + ph_corpus_vocab / ph_corpus_ids give back the text
    path's vocabulary (words, order, counts, train_words) and id stream
    (LearnVocabFromTrainFile + loadStream, cpu:253-292, 359-370)."""
    p = _text_corpus(tmp_path)
    v = H.Vocab.learn_text(p, 5)
    s = H.load_text_stream(v, p)
    b = str(tmp_path / "corpus.bin")
    H.save_corpus(b, v, s)
    v2, s2 = H.load_corpus(b)
    assert len(v2) == len(v) and v2.train_words == v.train_words
    assert np.array_equal(v2.counts(), v.counts())
    assert all(v2.word(i) == v.word(i) for i in range(len(v)))
    assert all(v2.lookup(v.word(i)) == i for i in range(0, len(v), 97))
    assert np.array_equal(s2, s)
    assert os.path.getsize(b) == 64 + 4 * len(v) + -(-sum(len(v.word(i)) + 1 for i in range(len(v))) // 8) * 8 \
        + 4 * s.size


def test_binary_corpus_cli_conversion(tmp_path):
    """-save-bin writes the same file as ph_corpus_save on the text path, and
    -train-bin reads it (vocabulary report only: no -output, no device)."""
    p = _text_corpus(tmp_path, 300_000)
    b = tmp_path / "cli.bin"
    r = subprocess.run([str(CLI), "-train", p, "-save-bin", str(b)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    v = H.Vocab.learn_text(p, 5)
    H.save_corpus(str(tmp_path / "py.bin"), v, H.load_text_stream(v, p))
    assert b.read_bytes() == (tmp_path / "py.bin").read_bytes()
    r = subprocess.run([str(CLI), "-train-bin", str(b)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert f"Vocab size: {len(v)}" in r.stdout and f"Words in train file: {v.train_words}" in r.stdout
    r = subprocess.run([str(CLI), "-train", p, "-train-bin", str(b)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "not both" in r.stderr


def test_binary_corpus_rejects_malformed_files(tmp_path):
    p = _text_corpus(tmp_path, 100_000)
    v = H.Vocab.learn_text(p, 5)
    s = H.load_text_stream(v, p)
    b = tmp_path / "c.bin"
    H.save_corpus(str(b), v, s)
    data = bytearray(b.read_bytes())
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"XXXXXXXX" + data[8:])
    with pytest.raises(H.HostError, match="bad magic"):
        H.load_corpus(str(bad))
    bad.write_bytes(data[: len(data) - 4])
    with pytest.raises(H.HostError, match="truncated id stream"):
        H.load_corpus(str(bad))
    ids_off = int(np.frombuffer(bytes(data[40:48]), np.uint64)[0])
    d2 = bytearray(data)
    d2[ids_off + 40: ids_off + 44] = np.int32(len(v)).tobytes()
    bad.write_bytes(bytes(d2))
    with pytest.raises(H.HostError, match="out of vocabulary range"):
        H.load_corpus(str(bad))
    d3 = bytearray(data)
    d3[64:68] = np.uint32(int(np.frombuffer(bytes(data[64:68]), np.uint32)[0]) + 1).tobytes()
    bad.write_bytes(bytes(d3))
    with pytest.raises(H.HostError, match="do not match"):
        H.load_corpus(str(bad))
    with pytest.raises(H.HostError, match="out of vocabulary range"):
        H.save_corpus(str(bad), v, np.array([0, len(v)], np.int32))


def test_parallel_text_loader_equals_oracle(tmp_path):
    """Files above 1 MiB are tokenised in byte ranges on several threads.  On
    a 6 MB file with every tokenizer case at random places (tabs, runs of
    separators, "\\r\\n", blank lines, words longer than 98 bytes, a last
    word cut off by EOF) the vocabulary, counts, first-appearance order and
    id stream equal the oracle's single pass (cpu:120-144, 253-292, 359-370)."""
    rng = np.random.default_rng(11)
    words = [f"w{i}".encode() for i in range(3000)] + [b"L" * 120 + str(i).encode() for i in range(20)]
    p = np.arange(1, len(words) + 1, dtype=np.float64) ** -1.0
    tok = rng.choice(len(words), 1_400_000, p=p / p.sum())
    seps = [b" ", b"\t", b"  ", b"\n", b"\r\n", b" \n\n", b"\t \t"]
    sp = rng.choice(len(seps), tok.size, p=[0.6, 0.1, 0.1, 0.1, 0.05, 0.03, 0.02])
    data = b"".join(words[t] + seps[s] for t, s in zip(tok.tolist(), sp.tolist())) + b"tail"
    path = write(tmp_path, "adv.txt", data)
    assert len(data) > 6_000_000
    v = H.Vocab.learn_text(path, 3)
    ov = O.OracleVocab.learn(path, 3)
    assert len(v) == len(ov)
    assert all(v.word(i) == ov.word(i) for i in range(len(v)))
    assert np.array_equal(v.counts(), ov.counts()) and v.train_words == ov.train_words
    assert np.array_equal(H.load_text_stream(v, path), ov.stream(path))
