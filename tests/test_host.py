"""Host library (libpar_host.so) against the reference semantics and the
oracle: tokenizer, vocabulary, id streams, the exact random schedule of the
deterministic mode, subsampling into the sentence CSR, unigram table, writer
and synthetic corpora.  CPU only."""
import numpy as np
import pytest

import oracle as O
from par_word2vec_amd import corpora, host as H


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


def test_workloads_generate_at_small_scale():
    for name in ("zipf1m", "sbm"):
        c = corpora.make(name, scale=0.05)
        assert c.stream.size > 0 and c.train_words == c.stream.size
