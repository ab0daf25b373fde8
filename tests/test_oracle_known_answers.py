"""Known answers for the oracle's training schedule, derived by hand from the
reference's text (PAR_Word2Vec_cpu.cpp, "cpu:" below) and the LCG closed form.

Each scenario is a crafted corpus small enough that every random decision can
be followed draw by draw: x_n is the n-th LCG value from seed 0 (thread 0,
cpu:401), x_{n+1} = 25214903917 x_n + 11 mod 2^64 (cpu:461), pinned by the
known answers x_1..x_5 = 11, 277363943098, ... in test_oracle_pins.  Two
properties of that generator make the draws easy to follow: the parity of x_n
alternates (odd, even, odd, ...: an odd multiplier and an odd increment), so
x_n % 2 = n % 2; and x_n % 4 cycles 3, 2, 1, 0 (multiplier = 1, increment = 3
mod 4).  The order of draws follows the reference's loop: per sentence, one
subsampling draw per non-"</s>" word read (cpu:458-464), then per centre one
window draw (cpu:532-534), then per batch one rep_b draw (cpu:589-601), then
per negative one table draw (cpu:605-608) and one uniform draw per in-sentence
rejection (cpu:609-617).

The expected values are literal numbers with their derivation beside them, not
a second transcription of the trainer.  test_planted_misreadings then switches
on, one at a time, misreadings of the reference planted in the oracle
(ORC_MUT_* in par_sgns_oracle.h) and requires every one to break at least one
known answer here or in test_oracle_pins.
"""
import numpy as np
import pytest

import oracle as O

f32 = np.float32
TS = 100_000_000  # table_size, cpu:75
A0 = f32(0.025)


def xs(n: int) -> list[int]:
    """x_0 .. x_n of the LCG from seed 0."""
    out = [0]
    for _ in range(n):
        out.append((out[-1] * 25214903917 + 11) % (1 << 64))
    return out


X = xs(30_000)


def _train(stream, cn, T, d, W, neg, B, iters, sample, table=None):
    if table is None:
        table = np.zeros(TS, np.int32)  # every slot is "</s>": each negative is uniform x % (V-1) + 1 (cpu:607-608)
    return O.train(np.array(stream, np.int32), np.array(cn, np.uint32), T, d, W, neg, B, iters, float(A0), sample,
                   table=table, want_trace=True)


def _check(got: dict, want: dict) -> list[str]:
    bad = []
    for k, v in want.items():
        g = got[k]
        same = np.array_equal(np.asarray(g), np.asarray(v)) if isinstance(v, (list, np.ndarray)) else g == v
        if not same:
            bad.append(f"{k}: got {g!r}, want {v!r}")
    return bad


def _flat(r) -> dict:
    t = r.trace
    return dict(r.stats, b=t["b"].tolist(), rep_b=t["rep_b"].tolist(), negs=t["negs"].tolist(),
                sent_len=t["sent_len"].tolist(), tokens=t["tokens"].tolist(),
                sent_alpha=t["sent_alpha"].reshape(-1, 2).tolist())


# --------------------------------------------------------------- scenarios
def k1_window_empty_line_rejection() -> list[str]:
    """Stream [1 2 </s> </s> 4 1 3 </s>], V = 11, W = 2, one negative per
    rep_b, no subsampling.  The table is all "</s>" except the slot
    (x_4 >> 16) % 1e8, which holds word 2.

    sentence [1 2]: centre 0 draws x_1 (odd) -> b = 1: window slots i = 1, 3
      (i < 2W+1-b = 4, i != W): c = -1 skipped, c = 1 -> 1 pair.
      centre 1 draws x_2 (even) -> b = 0: c = -1 skipped, c = 0 pair, c = 2 >= L
      stops -> 1 pair.  One batch (min_pos 0 <= W-1): rep_b = x_3 % 2 = 1,
      K = 1.  The negative draws x_4: table slot -> word 2, which is in the
      sentence, so one uniform redraw x_5: x_5 % 10 + 1 = 6.
    sentence [] (second "</s>"): one window draw x_6, no pairs, no batch.
    sentence [4 1 3]: b = x_7 % 2, x_8 % 2, x_9 % 2 = 1, 0, 1 -> 1 + 2 + 1
      pairs; one batch: rep_b = x_10 % 2 = 0, K = 0.
    The sentinel then ends the epoch.  Draws: 6 window + 2 batch + 1 negative
    + 1 redraw = 10, so the generator ends at x_10."""
    table = np.zeros(TS, np.int32)
    table[(X[4] >> 16) % TS] = 2
    assert X[5] % 10 + 1 == 6 and X[4] % 10 + 1 not in (1, 2)  # the uniform fallback would differ
    r = _train([1, 2, 0, 0, 4, 1, 3, 0], [2] + [1] * 10, 11, 3, 2, 1, 24, 1, 0.0, table)
    return _check(_flat(r), dict(b=[1, 0, 1, 0, 1], rep_b=[1, 0], negs=[6], sent_len=[2, 3], pairs=6,
                                 negatives=1, rep_pairs=2, batches=2, trained_tokens=5, sentences=2, draws=10,
                                 final_rng=X[10]))


def k2_learning_rate_schedule() -> list[str]:
    """12 lines of 999 words + "</s>" (T = 12000), 2 epochs, W = 1, no
    negatives.  The check `word_count - last_word_count > 10000` (cpu:431) runs
    at the top of every centre's iteration, before a new sentence is read
    (cpu:447), so the jump of a sentence load is seen from that sentence's
    centre 1 on.  Epoch 1: after 10 lines word_count = 10000 (not > 10000);
    line 10 brings 11000 -> from its centre 1, alpha = a0 (1 - 11000/24001)
    (iter * T + 1 = 24001, float arithmetic, cpu:441).  Reading the sentinel
    makes word_count 12001 > 12000 and the epoch ends, adding 12001 - 11000
    (cpu:515-516): word_count_actual = 12001.  Epoch 2 repeats the pattern:
    from line 10's centre 1, alpha = a0 (1 - 23001/24001).  Draws: per line
    999 window draws + ceil(999/24) = 42 batch draws, 24 lines."""
    line = [(i % 50) + 1 for i in range(999)]
    stream = (line + [0]) * 12
    cn = np.bincount(stream, minlength=51)
    a1 = A0 * (f32(1) - f32(11000) / f32(24001))
    a2 = A0 * (f32(1) - f32(23001) / f32(24001))
    r = _train(stream, cn, 12000, 2, 1, 0, 24, 2, 0.0)
    want_alpha = ([[A0, A0]] * 10 + [[A0, a1], [a1, a1]] + [[a1, a1]] * 10 + [[a1, a2], [a2, a2]])
    got = _flat(r)
    got["sent_alpha"] = np.array(got["sent_alpha"], np.float32)
    return _check(got, dict(sent_alpha=np.array(want_alpha, np.float32), final_alpha=float(a2),
                            draws=24 * (999 + 42), final_rng=X[24 * (999 + 42)], trained_tokens=24 * 999,
                            raw_words=24000))


def k3_batch_rules() -> list[str]:
    """One line of 8 distinct words (ids 11..18), V = 1000, W = 2, B = 3, one
    negative per rep_b.  Window draws x_1..x_8 alternate b = 1, 0, ...: pairs
    per centre 1, 3, 2, 4, 2, 4, 2, 2 = 20.  Batches start at 0, 3, 6
    (cpu:580-587): min_pos 0 <= W-1 -> rep_b = x_9 % 2 = 1; min_pos 3 is an
    interior batch ((L-1) - 3 = 4 > W-1) -> rep_b = x_11 % 4 = 1 (the cycle
    3, 2, 1, 0 at n = 11); min_pos 6 ends the line ((L-1) - 6 = 1 <= W-1) ->
    rep_b = x_13 % 2 = 1.  Each batch draws one uniform negative (x_10, x_12,
    x_14) % 999 + 1 = 286, 108, 932, none in the line.  Draws 8 + 3 + 3."""
    r = _train([11, 12, 13, 14, 15, 16, 17, 18, 0], [1] * 1000, 1000, 2, 2, 1, 3, 1, 0.0)
    return _check(_flat(r), dict(b=[1, 0, 1, 0, 1, 0, 1, 0], pairs=20, rep_b=[1, 1, 1], negs=[286, 108, 932],
                                 batches=3, rep_pairs=3 + 3 + 2, draws=14, final_rng=X[14]))


def k4_subsampling() -> list[str]:
    """Line [1 2 3 4 5 6], each word with count 1000, T = 6001, sample 0.03:
    ratio = 0.03 * 6001 / 1000 = 0.18003, ran = sqrt(ratio) + ratio = 0.6043
    (float, cpu:459-460).  A word is dropped when ran < (x & 0xFFFF) / 65536
    (cpu:462): the draws x_1..x_6 give 0.00017, 0.901, 0.349, 0.801, 0.886,
    0.146, so words 1, 3, 6 stay.  Then 3 window draws and one batch draw
    (rep_b = x_10 % 2 = 0)."""
    r = _train([1, 2, 3, 4, 5, 6, 0], [1] + [1000] * 6, 6001, 2, 2, 1, 24, 1, 0.03)
    return _check(_flat(r), dict(tokens=[1, 3, 6], trained_tokens=3, draws=10, final_rng=X[10]))


def k5_unterminated_tail() -> list[str]:
    """Stream [1 2 </s> 3 4] with no final "</s>" and sample 1 (every word
    kept: ran >= 2, but each still costs a draw).  The second line reads the
    sentinel, word_count = 6 > 5 ends the epoch and the line is dropped
    (cpu:472) after its two subsampling draws.  Draws: 2 subsampling + 2
    window + 1 batch (rep_b = x_5 % 2 = 1) + 1 negative (x_6 % 9 + 1 = 5)
    + 2 subsampling of the dropped line = 8."""
    r = _train([1, 2, 0, 3, 4], [1] * 10, 10, 2, 2, 1, 24, 1, 1.0)
    return _check(_flat(r), dict(trained_tokens=2, sentences=1, sent_len=[2], negs=[5], draws=8, final_rng=X[8],
                                 raw_words=5))


def k6_repulsion_closed_form() -> list[str]:
    """Stream [1 2 </s>], V = 11, W = 2, one negative, the table of k1 (so the
    negative is word 6 after one rejection).  W_out starts at zero, so both
    attraction pairs see f = 0 and g = (1 - expTable[500]) alpha = alpha / 2
    (expTable[500] = 0.5, index (0 + 6) * 83.333336 = 500): W_out[1] =
    g W_in[2], W_out[2] = g W_in[1], and W_in is unchanged (tmp = g * 0,
    cpu:562-564).  The batch (rows 1, 2) against negative 6 (a zero row) has
    logits 0, label-0 gradient (0 - 0.5) alpha = -alpha/2 (cpu:641-647):
    W_out[6] = (-alpha/2) W_in[1] + (-alpha/2) W_in[2] (cpu:652-653, summed
    in row order) and W_in += G^T * 0 = 0.  Every other row is untouched."""
    table = np.zeros(TS, np.int32)
    table[(X[4] >> 16) % TS] = 2
    V, d = 11, 5
    r = O.train(np.array([1, 2, 0], np.int32), np.array([2] + [1] * 10, np.uint32), 11, d, 2, 1, 24, 1, float(A0),
                0.0, table=table)
    w_in0, _ = O.init_net(V, d)
    g = (f32(1) - f32(0.5)) * A0
    gn = (f32(0) - f32(0.5)) * A0
    want_out = np.zeros((V, d), np.float32)
    want_out[1] = f32(0) + g * w_in0[2]
    want_out[2] = f32(0) + g * w_in0[1]
    want_out[6] = f32(0) + ((f32(0) + gn * w_in0[1]) + gn * w_in0[2])
    bad = []
    if not np.array_equal(r.w_in, w_in0):
        bad.append(f"W_in changed in {np.count_nonzero(r.w_in != w_in0)} entries")
    if not np.array_equal(r.w_out, want_out):
        bad.append(f"W_out differs from the closed form in rows {sorted(set(np.nonzero(r.w_out != want_out)[0]))}")
    return bad


def k7_unigram_closed_form() -> list[str]:
    """The table rule of cpu:98-115 gives word i the slots up to the first a
    with a / table_size > d1_i, d1_i the cumulative unigram^0.75 mass of words
    0..i; so the number of slots holding a word <= i is
    C_i = floor(table_size * P_i) + 2, P_i = sum_{k<=i} cn_k^0.75 / sum cn^0.75
    (or C_{i-1} + 1 when word i's mass is below one slot), and the last word
    takes the rest.  d1 is a float sum, so each term may move a boundary by
    table_size * 2^-24 relative: at table_size 2000 that is 1e-4 slot, below
    the distance of every 2000 * P_i here from an integer."""
    bad = []
    for cn in ([5, 100, 50, 20, 7, 7, 1], [1, 3, 3, 3, 2, 1, 1, 1, 1, 1], [40, 9, 30, 2, 2, 2, 500, 60]):
        cn = np.array(cn, np.uint32)
        ts = 2000
        t = np.empty(ts, np.int32)
        O.lib().orc_unigram_table_n(O.P(cn), cn.size, O.P(t), ts)
        p = np.cumsum(cn.astype(np.float64) ** 0.75) / (cn.astype(np.float64) ** 0.75).sum()
        frac = ts * p[:-1] - np.floor(ts * p[:-1])
        assert np.all(np.minimum(frac, 1 - frac) > 1e-3), "a boundary too close to a slot edge for this check"
        want, prev = [], 0
        for i in range(cn.size - 1):
            c = max(int(np.floor(ts * p[i])) + 2, prev + 1)
            want.append(c)
            prev = c
        want.append(ts)
        got = [int(np.count_nonzero(t <= i)) for i in range(cn.size)]
        if got != want:
            bad.append(f"cn={cn.tolist()}: cumulative slots {got}, closed form {want}")
    return bad


def k8_unigram_full_table_bounds() -> list[str]:
    """The 1e8-slot table of a Zipf vocabulary (5,000 words, counts 1e6 / r):
    C_i lies within floor(1e8 P_i) + 2 +- the float error of the running sum
    d1 (at most 2^-24 d1 per addition, so 1e8 * 2^-24 * sum_{k<=i} P_k)."""
    cn = (1e6 / np.arange(1, 5001)).astype(np.uint32)
    t = O.unigram_table(cn)
    w = cn.astype(np.float64) ** 0.75
    p = np.cumsum(w) / w.sum()
    C = np.searchsorted(t, np.arange(cn.size), side="right")  # the table is non-decreasing
    tol = TS * 2.0 ** -24 * np.cumsum(p) + 1
    dev = C[:-1] - (np.floor(TS * p[:-1]) + 2)
    bad = []
    if not np.all(np.diff(t) >= 0):
        bad.append("table not non-decreasing")
    worst = int(np.argmax(np.abs(dev) - tol[:-1]))
    if np.any(np.abs(dev) > tol[:-1]):
        bad.append(f"word {worst}: {C[worst]} slots cumulative, closed form {np.floor(TS * p[worst]) + 2:.0f} "
                   f"+- {tol[worst]:.1f}")
    if C[-1] != TS:
        bad.append("last word does not fill the table")
    return bad


SCENARIOS = {f.__name__: f for f in (k1_window_empty_line_rejection, k2_learning_rate_schedule, k3_batch_rules,
                                     k4_subsampling, k5_unterminated_tail, k6_repulsion_closed_form,
                                     k7_unigram_closed_form, k8_unigram_full_table_bounds)}


@pytest.mark.parametrize("name", list(SCENARIOS))
def test_known_answer(name):
    assert SCENARIOS[name]() == []


# Which known answers a planted misreading must break (at least one of them).
CATCHES = {
    "alpha_after_load": ["k2_learning_rate_schedule"],
    "alpha_ge": ["k2_learning_rate_schedule"],
    "no_rejection": ["k1_window_empty_line_rejection", "k6_repulsion_closed_form"],
    "repb_2w": ["k1_window_empty_line_rejection", "k3_batch_rules"],
    "window_incl": ["k1_window_empty_line_rejection", "k3_batch_rules"],
    "pair_order": ["k6_repulsion_closed_form"],
    "unigram_linear": ["k7_unigram_closed_form", "k8_unigram_full_table_bounds"],
    "subsample_flip": ["k4_subsampling"],
    "table_index": ["k1_window_empty_line_rejection", "k6_repulsion_closed_form"],
    "wca_local": ["k2_learning_rate_schedule"],
    "rep_label": ["k6_repulsion_closed_form"],
    "batch_end": ["k3_batch_rules"],
}


def test_every_planted_misreading_is_listed():
    assert set(CATCHES) == set(O.MUTATIONS)


@pytest.mark.parametrize("mutation", sorted(CATCHES))
def test_planted_misreading_fails_a_known_answer(mutation):
    """With the misreading switched on in the oracle, each listed known answer
    is run; at least one must fail (each list names the scenarios built to
    see that rule), and every listed scenario still passes once it is off."""
    with O.mutation(mutation):
        failed = {n: SCENARIOS[n]() for n in CATCHES[mutation]}
    assert any(failed.values()), f"misreading {mutation} passed every known answer: {failed}"
    for n in CATCHES[mutation]:
        assert SCENARIOS[n]() == []
