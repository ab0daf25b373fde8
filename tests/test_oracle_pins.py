"""Pins the oracle (CPU restatement of PAR_Word2Vec_cpu.cpp) before it is
trusted as the parity reference.  The reference itself cannot be built here
(it includes "mkl.h" unconditionally, cpu:27), so every part of the oracle is
checked against something the reference or mathematics fixes:

  LCG, InitNet        known answers derived from the closed form (cpu:354-355, 461)
  EXP_RESOLUTION      the float value of 1000 / 12.0f (cpu:70)
  expTable            float64 sigmoid of the same float32 grid, <= 1 ulp (cpu:822-826)
  unigram table       brute-force transcription of cpu:96-117 (pure Python)
  GEMMs               numpy float64 products (cpu:633, 652, 655 shapes)
  one pair            closed form of the attraction update with W_out = 0 (cpu:551-565)
  whole trainer       an independent pure-Python float32 transcription of
                      cpu:430-685 on tiny corpora, bit for bit
  window loss/grad    numpy float64 and central finite differences

The training schedule itself (RNG interleaving, learning-rate timing, batch
rules, rejection, the dropped tail) is pinned by hand-derived known answers in
test_oracle_known_answers.py, which also shows each planted misreading fails.
"""
import math

import numpy as np
import pytest

import oracle as O

f32 = np.float32


def lcg(x):
    return (x * 25214903917 + 11) % (1 << 64)


def test_lcg_known_answers():
    x, seq = 0, []
    for _ in range(5):
        x = int(O.lib().orc_lcg(x))
        seq.append(x)
    assert seq == [11, 277363943098, 2389171320405252413, 9903746776048127236, 1952413064874484415]
    y = 0
    for v in seq:
        y = lcg(y)
        assert y == v


def test_init_net_known_answers():
    w_in, w_out = O.init_net(3, 100)
    np.testing.assert_array_equal(w_in[0, :4], np.array([0.0040026857, 0.0044194032, -0.0038302613, -0.0032780457],
                                                        dtype=np.float32))
    assert not w_out.any()
    # closed form for the whole table
    x = 1
    ref = np.empty(300, np.float32)
    for i in range(300):
        x = lcg(x)
        ref[i] = (f32(x & 0xFFFF) / f32(65536.0) - f32(0.5)) / f32(100)
    np.testing.assert_array_equal(w_in.reshape(-1), ref)


def test_exp_resolution_and_table():
    assert O.lib().orc_exp_resolution() == float(np.float32(1000.0) / np.float32(12.0))
    assert O.lib().orc_exp_resolution() == pytest.approx(83.33333587646484, abs=0)
    t = O.exp_table()
    i = np.arange(1001)
    x = (i.astype(np.float32) / f32(1000) * f32(2) - f32(1)) * f32(6)
    # expf is within one ulp of the float64 exponential; the second step
    # e / (e + 1) is exact float32 arithmetic on that value
    e = np.exp(x.astype(np.float64)).astype(np.float32)
    ok = np.zeros(1001, bool)
    for cand in (e, np.nextafter(e, f32(np.inf)), np.nextafter(e, f32(0))):
        ok |= (cand / (cand + f32(1))).astype(np.float32) == t
    assert ok.all()
    assert t[500] == f32(0.5)
    assert np.all(np.diff(t) >= 0)


def brute_unigram(cn, table_size):
    """cpu:96-117 transcribed: double Z, float d1, float a/table_size."""
    z = 0.0
    for c in cn:
        z += float(c) ** 0.75
    i = 0
    d1 = f32(float(cn[0]) ** 0.75 / z)
    out = np.empty(table_size, np.int32)
    V = len(cn)
    for a in range(table_size):
        out[a] = i
        if f32(a) / f32(table_size) > d1:
            i += 1
            if i < V:
                d1 = f32(float(d1) + float(cn[i]) ** 0.75 / z)
        if i >= V:
            i = V - 1
    return out


@pytest.mark.parametrize("cn,size", [
    ([5, 100, 50, 20, 7, 7, 1], 5000),
    ([0, 3, 3, 3, 2, 1, 1, 1, 1, 1], 20000),
    (list(range(200, 0, -1)), 30000),
])
def test_unigram_table_brute_force(cn, size):
    cn = np.array(cn, np.uint32)
    t = np.empty(size, np.int32)
    O.lib().orc_unigram_table_n(O.P(cn), cn.size, O.P(t), size)
    np.testing.assert_array_equal(t, brute_unigram(cn, size))


def test_gemms_match_float64():
    rng = np.random.default_rng(0)
    m, n, k = 7, 5, 33
    a = rng.standard_normal((m, k)).astype(np.float32)
    b = rng.standard_normal((n, k)).astype(np.float32)
    c = np.empty((m, n), np.float32)
    O.lib().orc_gemm_nt(m, n, k, O.P(a), k, O.P(b), k, O.P(c), n)
    np.testing.assert_allclose(c, a.astype(np.float64) @ b.astype(np.float64).T, rtol=1e-5, atol=1e-5)
    b2 = rng.standard_normal((k, n)).astype(np.float32)
    O.lib().orc_gemm_nn(m, n, k, O.P(a), k, O.P(b2), n, O.P(c), n)
    np.testing.assert_allclose(c, a.astype(np.float64) @ b2.astype(np.float64), rtol=1e-5, atol=1e-5)
    a3 = rng.standard_normal((k, m)).astype(np.float32)
    O.lib().orc_gemm_tn(m, n, k, O.P(a3), m, O.P(b2), n, O.P(c), n)
    np.testing.assert_allclose(c, a3.astype(np.float64).T @ b2.astype(np.float64), rtol=1e-5, atol=1e-5)
    # a transposed operand would fail this: shape (m,n) differs from (n,m) products
    assert not np.allclose(c[:n, :n], (a3.astype(np.float64).T @ b2.astype(np.float64))[:n, :n].T)


def test_single_pair_closed_form():
    """Sentence [a, b], window 1, no subsampling: two attraction pairs with
    W_out = 0, so f = 0, g = (1 - expTable[500]) * alpha = alpha/2; repulsion
    draws rep_b = x % 1 = 0 negatives."""
    V, d, alpha = 3, 4, 0.025
    counts = np.array([1, 1, 1], np.uint32)
    stream = np.array([1, 2, 0], np.int32)
    w_in0, w_out0 = O.init_net(V, d)
    r = O.train(stream, counts, 3, d, 1, 5, 24, 1, alpha, 0.0, table=np.zeros(100_000_000, np.int32))
    g = (f32(1) - f32(0.5)) * f32(alpha)
    np.testing.assert_array_equal(r.w_in, w_in0)
    np.testing.assert_array_equal(r.w_out[1], f32(0) + g * w_in0[2])
    np.testing.assert_array_equal(r.w_out[2], f32(0) + g * w_in0[1])
    assert r.stats["pairs"] == 2 and r.stats["negatives"] == 0 and r.stats["draws"] == 3  # 2 centres + 1 batch


# --------------------------------------------------------- transcription ---
def py_train(stream, cn, T, d, W, NEG, B, iters, alpha0, sample, table, et, w_in, w_out):
    """Independent pure-Python float32 transcription of cpu:430-685 with one
    thread (seed 0) and naive ascending-k products for cblas_sgemm."""
    TABLE = 100_000_000
    V = len(cn)
    RES = f32(1000) / (f32(6) * f32(2.0))
    x = 0
    alpha = f32(alpha0)
    a0 = f32(alpha0)
    local_iter = iters
    wc = last = wca = 0
    L = pos = rpos = 0
    sen = [0] * 1000
    n = len(stream) - 1  # stream carries the 0 sentinel

    def sig_idx(f):
        return int((f + f32(6)) * RES)

    while True:
        if wc - last > 10000:
            wca += wc - last
            last = wc
            alpha = a0 * (f32(1) - f32(wca) / f32(iters * T + 1))
            if alpha < a0 * f32(0.0001):
                alpha = a0 * f32(0.0001)
        if L == 0:
            while True:
                w = int(stream[wc])
                wc += 1
                if w == 0:
                    break
                if sample > 0:
                    ratio = (f32(sample) * f32(T)) / f32(cn[w])
                    ran = np.sqrt(ratio, dtype=np.float32) + ratio
                    x = lcg(x)
                    if ran < f32(x & 0xFFFF) / f32(65536.0):
                        continue
                sen[L] = w
                L += 1
                if L >= 1000:
                    break
            pos = rpos = 0
        if wc > n:
            wca += wc - last
            local_iter -= 1
            if local_iter == 0:
                break
            wc = last = 0
            L = 0
            continue
        target = sen[pos]
        x = lcg(x)
        b = x % W
        for i in range(b, 2 * W + 1 - b):
            if i == W:
                continue
            c = pos - W + i
            if c < 0:
                continue
            if c >= L:
                break
            u, v = w_in[sen[c]], w_out[target]
            f = f32(0)
            for j in range(d):
                f = f32(f + u[j] * v[j])
            if f > 6:
                g = f32(0) * alpha
            elif f < -6:
                g = f32(1) * alpha
            else:
                g = (f32(1) - et[sig_idx(f)]) * alpha
            for j in range(d):
                tmp = g * v[j]
                v[j] = v[j] + g * u[j]
                u[j] = u[j] + tmp
        pos += 1
        if pos >= L:
            rpos = 0
            if L > 0:
                for nb in range((L + B - 1) // B):
                    lo = nb * B
                    isz = min(lo + B - 1, L - 1) - lo + 1
                    x = lcg(x)
                    rep_b = x % W if (lo <= W - 1 or (L - 1) - lo <= W - 1) else x % (2 * W)
                    K = rep_b * NEG
                    negs = []
                    for _ in range(K):
                        x = lcg(x)
                        s = int(table[(x >> 16) % TABLE])
                        if s == 0:
                            s = x % (V - 1) + 1
                        while s in sen[:L]:
                            x = lcg(x)
                            s = x % (V - 1) + 1
                        negs.append(s)
                    Min = np.array([w_in[sen[lo + i]] for i in range(isz)], np.float32)
                    Mout = np.array([w_out[s] for s in negs], np.float32).reshape(K, d)
                    G = np.empty((K, isz), np.float32)
                    for k in range(K):
                        for i in range(isz):
                            acc = f32(0)
                            for j in range(d):
                                acc = f32(acc + Mout[k, j] * Min[i, j])
                            f = acc
                            if f > 6:
                                f = f32(-1) * alpha
                            elif f < -6:
                                f = f32(0) * alpha
                            else:
                                f = (f32(0) - et[sig_idx(f)]) * alpha
                            G[k, i] = f
                    Uo = np.empty((K, d), np.float32)
                    for k in range(K):
                        for j in range(d):
                            acc = f32(0)
                            for i in range(isz):
                                acc = f32(acc + G[k, i] * Min[i, j])
                            Uo[k, j] = acc
                    Ui = np.empty((isz, d), np.float32)
                    for i in range(isz):
                        for j in range(d):
                            acc = f32(0)
                            for k in range(K):
                                acc = f32(acc + G[k, i] * Mout[k, j])
                            Ui[i, j] = acc
                    for i in range(isz):
                        row = w_in[sen[lo + i]]
                        for j in range(d):
                            row[j] = row[j] + Ui[i, j]
                    for k in range(K):
                        row = w_out[negs[k]]
                        for j in range(d):
                            row[j] = row[j] + Uo[k, j]
                    rpos += isz
        if rpos >= L:
            L = 0
            continue
    return x


@pytest.mark.parametrize("n_tok,V,d,W,NEG,B,iters,sample", [
    (400, 25, 3, 2, 2, 4, 1, 0.0),
    (900, 40, 5, 3, 2, 5, 2, 5e-3),
    (11_000, 30, 2, 2, 1, 6, 2, 1e-3),   # crosses the 10000-word alpha update, two epochs
])
def test_trainer_matches_python_transcription(n_tok, V, d, W, NEG, B, iters, sample):
    rng = np.random.default_rng(n_tok)
    lens = rng.integers(1, 60, size=n_tok)
    stream = []
    for L in lens:
        stream += list(rng.integers(1, V, size=int(L))) + [0]
        if len(stream) >= n_tok:
            break
    if n_tok < 1000:  # one over-long line (split at 1000 kept tokens) only makes sense when long
        pass
    stream = np.array(stream, np.int32)
    cn = np.bincount(stream, minlength=V).astype(np.uint32)
    cn[cn == 0] = 1
    T = int(stream.size)
    table = O.unigram_table(cn)
    et = O.exp_table()
    w_in0, w_out0 = O.init_net(V, d)
    r = O.train(stream, cn, T, d, W, NEG, B, iters, 0.025, sample, table=table)
    w_in, w_out = w_in0.copy(), w_out0.copy()
    full = np.append(stream, 0).astype(np.int32)
    x = py_train(full, cn, T, d, W, NEG, B, iters, 0.025, sample, table, et, w_in, w_out)
    assert x == r.stats["final_rng"]
    np.testing.assert_array_equal(r.w_in, w_in)
    np.testing.assert_array_equal(r.w_out, w_out)


def test_long_line_split_and_unterminated_tail():
    """A line of 2500 tokens is cut into 1000 + 1000 + 500 (cpu:467); a final
    line without "</s>" reads the sentinel and is dropped (cpu:472)."""
    V = 5000  # the line must not hold every word, or the rejection loop of cpu:611-615 never ends
    line = (np.arange(2500) % 3000 + 1).astype(np.int32)
    tail = np.array([3, 4, 5], np.int32)
    stream = np.concatenate([line, [0], tail])
    cn = np.bincount(stream, minlength=V).astype(np.uint32)
    cn[cn == 0] = 1
    r = O.train(stream, cn, stream.size, 4, 2, 1, 24, 1, 0.025, 0.0, want_trace=True)
    assert list(r.trace["sent_len"]) == [1000, 1000, 500]
    assert r.stats["trained_tokens"] == 2500


def test_window_loss_and_grad_f64():
    rng = np.random.default_rng(3)
    V, d, nw, nctx, nneg = 20, 6, 9, 4, 3
    w_in = rng.standard_normal((V, d)) * 0.3
    w_out = rng.standard_normal((V, d)) * 0.3
    centre = rng.integers(0, V, nw).astype(np.int32)
    ctx = rng.integers(0, V, (nw, nctx)).astype(np.int32)
    neg = rng.integers(0, V, (nw, nneg)).astype(np.int32)
    L = O.window_loss_f64(centre, ctx, neg, w_in, w_out)
    ref = 0.0
    for w in range(nw):
        Y = np.vstack([w_out[centre[w]], w_out[neg[w]]])
        F = w_in[ctx[w]] @ Y.T
        ref += np.logaddexp(0, -F[:, 0]).sum() + np.logaddexp(0, F[:, 1:]).sum()
    assert L == pytest.approx(ref, rel=1e-12)
    gi, go = O.window_grad_f64(centre, ctx, neg, w_in, w_out)
    h = 1e-6
    for (tab, grad) in ((w_in, gi), (w_out, go)):
        for _ in range(12):
            r_, c_ = rng.integers(0, V), rng.integers(0, d)
            tab[r_, c_] += h
            lp = O.window_loss_f64(centre, ctx, neg, w_in, w_out)
            tab[r_, c_] -= 2 * h
            lm = O.window_loss_f64(centre, ctx, neg, w_in, w_out)
            tab[r_, c_] += h
            assert (lp - lm) / (2 * h) == pytest.approx(grad[r_, c_], rel=1e-5, abs=1e-8)


def test_oracle_test_options_seed_and_bf16():
    """The oracle's test-only options: seed 0 is the reference's thread-0 run;
    another seed gives another run of the same rules; bf16 leaves every table
    entry bf16-representable and, with stochastic rounding, keeps the loss."""
    raw = np.random.default_rng(4).integers(1, 60, 30_000).astype(np.int32)
    stream = np.where(np.arange(raw.size) % 50 == 49, 0, raw).astype(np.int32)
    cn = np.bincount(stream, minlength=60).astype(np.uint32)
    cn[cn == 0] = 1
    args = (stream, cn, int(stream.size), 16, 3, 3, 12, 1, 0.025, 1e-3)
    a, b, c = O.train(*args), O.train(*args, seed=0), O.train(*args, seed=1)
    assert np.array_equal(a.w_in, b.w_in) and a.stats["final_rng"] == b.stats["final_rng"]
    assert not np.array_equal(a.w_in, c.w_in) and a.stats["pairs"] != c.stats["pairs"]
    h = O.train(*args, bf16=True)
    assert not (h.w_in.view(np.uint32) & 0xFFFF).any() and not (h.w_out.view(np.uint32) & 0xFFFF).any()
    la = (a.stats["loss_attr"] + a.stats["loss_rep"]) / (a.stats["pairs"] + a.stats["rep_pairs"])
    lh = (h.stats["loss_attr"] + h.stats["loss_rep"]) / (h.stats["pairs"] + h.stats["rep_pairs"])
    assert h.stats["final_rng"] == a.stats["final_rng"]  # rounding noise has its own stream
    assert abs(lh / la - 1) < 0.01
