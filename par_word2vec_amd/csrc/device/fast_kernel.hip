// fast_kernel.hip -- Hogwild PAR-SGNS trainer for gfx950 (plain and
// persistent scheduling, fp32 or bf16 tables).
//
// One wavefront owns one sentence at a time (PAR_Word2Vec_gpu_text.cu gives
// a thread block a contiguous sentence range, text.cu:385-390; here the unit
// is a wave so every reduction is a DPP butterfly with no barrier).
//
// Attraction (cpu:529-568): the W_in rows of the sliding window [p-W, p+W]
// live in an LDS ring of R = 2W+1 word-keyed slots holding the current fp32
// value and the value that was loaded; a row is fetched once when its word
// enters the window (prefetched one centre ahead) and written back once when
// its last occurrence leaves, as delta = current - loaded.  The centre's
// W_out row stays in registers for all its pairs (and across a repeated
// centre).  Pairs are applied sequentially, exactly as the reference does, so
// later pairs see earlier updates.
//
// Repulsion (cpu:571-679): per batch of B positions, K = rep_b * negative
// shared negatives are drawn from the on-device alias table with the
// reference's in-sentence rejection; the K x B logits, the label-0 gradient
// and both updates are computed in register blocks of NB negatives.
//
// Write-back: fp32 tables take the delta with global float atomics (no lost
// updates, as the reference GPU's atomicAdd, text.cu:438-439, 710-717);
// bf16 tables re-read the current value, add the fp32 delta and store it
// stochastically rounded (a Hogwild read-modify-write, as the reference
// CPU's plain += at cpu:561-565).
//
// Randomness: Philox4x32-10 keyed by the seed; the counter names the
// sentence, epoch, purpose and index of each draw.
#include "sgns_device.h"
#include "sgns_internal.h"

namespace sgns {
namespace {

constexpr int kMaxRedraw = 64;
constexpr int kUnusedSlot = -(1 << 30);

// ---------------------------------------------------------------- stores
__device__ __forceinline__ void wb_pair(float *w, int64_t row, int ld, int q, float2 dlt, uint32_t, int) {
    float *p = w + row * ld + 2 * q;
    if (dlt.x != 0.0f) unsafeAtomicAdd(p, dlt.x);
    if (dlt.y != 0.0f) unsafeAtomicAdd(p + 1, dlt.y);
}
__device__ __forceinline__ void wb_pair(uint16_t *w, int64_t row, int ld, int q, float2 dlt, uint32_t noise,
                                        int nearest) {
    if (dlt.x == 0.0f && dlt.y == 0.0f) return;
    uint32_t *p = reinterpret_cast<uint32_t *>(w + row * ld) + q;
    const uint32_t cur = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const float x = bf16lo(cur) + dlt.x, y = bf16hi(cur) + dlt.y;
    const uint32_t out = nearest ? (bf16_rne(x) | (bf16_rne(y) << 16))
                                 : (bf16_sr(x, noise) | (bf16_sr(y, noise >> 16) << 16));
    __hip_atomic_store(p, out, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T, int NP>
__device__ __forceinline__ void load_row(const T *w, int64_t row, int ld, int nq, int lane,
                                         typename Pair<T>::P (&out)[NP]) {
#pragma unroll
    for (int t = 0; t < NP; t++) {
        const int q = lane + kWave * t;
        if (q < nq) out[t] = Pair<T>::load(w, row, ld, q);
        else out[t] = typename Pair<T>::P{};
    }
}

template <typename T, int NP>
__device__ __forceinline__ void to_f(const typename Pair<T>::P (&in)[NP], float2 (&out)[NP]) {
#pragma unroll
    for (int t = 0; t < NP; t++) out[t] = Pair<T>::tof(in[t]);
}

__device__ __forceinline__ uint32_t noise_of(uint32_t salt, int64_t row, int q) {
    return hash32(salt ^ hash32((uint32_t)row * 0x9E3779B1u + (uint32_t)q));
}

template <typename T, int NP>
__device__ __forceinline__ void store_delta(T *w, int64_t row, int ld, int nq, int lane, const float2 (&cur)[NP],
                                            const float2 (&ref)[NP], uint32_t salt, int nearest) {
#pragma unroll
    for (int t = 0; t < NP; t++) {
        const int q = lane + kWave * t;
        if (q < nq)
            wb_pair(w, row, ld, q, make_float2(cur[t].x - ref[t].x, cur[t].y - ref[t].y), noise_of(salt, row, q), nearest);
    }
}

// One draw from the unigram^0.75 alias table (before in-sentence rejection).
__device__ __forceinline__ int alias_draw(const float *prob, const int32_t *alias, int V, U4 rr) {
    const uint32_t bucket = __umulhi(rr.x, (uint32_t)V);
    const float u2 = (float)(rr.y >> 8) * (1.0f / 16777216.0f);
    return u2 < prob[bucket] ? (int)bucket : alias[bucket];
}

// ------------------------------------------------------ one sentence
// `raw` is the training progress in raw words when the sentence starts: its
// corpus offset in sentence order, or the words dequeued before it when a
// queue reorders sentences -- the reference decays alpha with the words
// processed so far (word_count_actual, cpu:431-445), not with corpus position.
template <typename T, int NP, int NB>
__device__ void train_sentence(const FastArgs &a, int64_t s, int64_t raw, int lane, char *lds, double &loss_a,
                               double &loss_r, uint64_t (&cnt)[kCntSlots], const float *et) {
    using PT = typename Pair<T>::P;
    T *w_in = static_cast<T *>(a.w_in);
    T *w_out = static_cast<T *>(a.w_out);
    const int ld = a.ld, nq = a.nq, W = a.W, R = a.R;
    const int64_t o0 = a.sent_off[s];
    const int L = (int)(a.sent_off[s + 1] - o0);
    if (L <= 0) return;  // an empty sentence trains nothing (cpu:539-542, 575)
    const int32_t *sen = a.tokens + o0;
    const bool planned = a.plan_b != nullptr;
    const float alpha = planned ? a.plan_alpha[2 * s + 1] : sentence_alpha(a.a0, a.epoch, a.T, a.iter, raw);
    const float alpha0 = planned ? a.plan_alpha[2 * s] : alpha;  // centre 0 (cpu:431-447 timing)
    const uint32_t ep = (uint32_t)a.epoch;
    const uint32_t salt0 = hash32((uint32_t)a.seed ^ hash32((uint32_t)s * 0x632BE5ABu ^ (ep << 24)));
    // LDS of this wave: ring values (fp32), ring snapshots (storage type),
    // shared negatives, and the slot tables of the ring
    float2 *ring_u = reinterpret_cast<float2 *>(lds);
    PT *ring_0 = reinterpret_cast<PT *>(lds + (size_t)R * nq * sizeof(float2));
    const int ring_bytes = R * nq * (int)(sizeof(float2) + sizeof(PT));
    const int snap_bytes = a.B * nq * (int)sizeof(PT);
    int32_t *negs = reinterpret_cast<int32_t *>(lds + (((ring_bytes > snap_bytes ? ring_bytes : snap_bytes) + 15) & ~15));
    int32_t *slot_word = negs + a.kmax;  // word held by a slot
    int32_t *slot_last = slot_word + R;  // last window position referencing it
    int32_t *pos_slot = slot_last + R;   // slot of position c, indexed c % R

    cnt[kCntSentences]++;
    cnt[kCntTokens] += (uint64_t)L;

    // Slot of word w among slots still referenced at or after position `from`
    // (-1 if none).  One lane per slot, first match wins.
    auto find_slot = [&](int w, int from) -> int {
        for (int r0 = 0; r0 < R; r0 += kWave) {
            const int sx = r0 + lane;
            const bool m = sx < R && slot_last[sx] >= from && slot_word[sx] == w;
            const uint64_t mask = __builtin_amdgcn_ballot_w64(m);
            if (mask) return r0 + (int)__builtin_ctzll(mask);
        }
        return -1;
    };
    // A slot referenced by no position at or after `from`.
    auto find_free = [&](int from) -> int {
        for (int r0 = 0; r0 < R; r0 += kWave) {
            const int sx = r0 + lane;
            const uint64_t mask = __builtin_amdgcn_ballot_w64(sx < R && slot_last[sx] < from);
            if (mask) return r0 + (int)__builtin_ctzll(mask);
        }
        return 0;  // unreachable: the window references at most R slots
    };
    auto fill = [&](int slot, const PT (&row)[NP]) {
#pragma unroll
        for (int t = 0; t < NP; t++) {
            const int q = lane + kWave * t;
            if (q < nq) {
                ring_0[slot * nq + q] = row[t];
                ring_u[slot * nq + q] = Pair<T>::tof(row[t]);
            }
        }
    };
    auto flush = [&](int slot) {
        const int64_t row = slot_word[slot];
        const uint32_t salt = salt0 ^ hash32((uint32_t)slot_last[slot] + 0x1000u);
#pragma unroll
        for (int t = 0; t < NP; t++) {
            const int q = lane + kWave * t;
            if (q < nq) {
                const float2 u = ring_u[slot * nq + q];
                const float2 u0 = Pair<T>::tof(ring_0[slot * nq + q]);
                wb_pair(w_in, row, ld, q, make_float2(u.x - u0.x, u.y - u0.y), noise_of(salt, row, q), a.rounding);
            }
        }
    };
    auto bind = [&](int c, int slot, int w) {
        if (lane == 0) {
            slot_word[slot] = w;
            slot_last[slot] = c;
            pos_slot[c % R] = slot;
        }
    };

    // ---- attraction
    // The W_in rows of window [p-W, p+W] live in LDS slots keyed by word: a
    // word occurring twice in the window shares one slot, so every pair sees
    // the updates of the pairs before it exactly as cpu:551-565 does in place.
    // A slot is written back when the last position referencing it leaves the
    // window; the next centre's W_out row and entering W_in row are prefetched
    // one centre ahead.
    {
        // unused slots sit below every window start (window starts go down to -W)
        for (int r0 = 0; r0 < R; r0 += kWave)
            if (r0 + lane < R) slot_last[r0 + lane] = kUnusedSlot;
        const int first = W < L - 1 ? W : L - 1;
        for (int c = 0; c <= first; c++) {
            const int w = sen[c];
            int sl = find_slot(w, 0);
            if (sl < 0) {
                sl = find_free(0);
                PT row[NP];
                load_row<T, NP>(w_in, w, ld, nq, lane, row);
                fill(sl, row);
            }
            bind(c, sl, w);
        }
        PT vnext[NP];
        load_row<T, NP>(w_out, sen[0], ld, nq, lane, vnext);
        float2 v[NP], v0[NP];
        bool carried = false;
        uint32_t bwin = 0;
        for (int p = 0; p < L; p++) {
            if ((p & (kWave - 1)) == 0)
                bwin = planned ? (p + lane < L ? a.plan_b[o0 + p + lane] : 0u)
                               : draw(a.seed, s, ep, kTagWindow, (uint32_t)(p + lane)).x % (uint32_t)W;
            const int b = __builtin_amdgcn_readlane((int)bwin, p & (kWave - 1));
            const float alpha_p = p == 0 ? alpha0 : alpha;
            if (!carried) {  // a centre repeated at p-1 keeps its register row
                to_f<T, NP>(vnext, v);
#pragma unroll
                for (int t = 0; t < NP; t++) v0[t] = v[t];
            }
            const bool carry_next = p + 1 < L && sen[p + 1] == sen[p];
            // prefetch for centre p+1: its W_out row and the W_in row entering its window
            const int np = p + 1 + W;
            const int np_slot = np < L ? find_slot(sen[np], p - W) : -1;
            PT pf[NP];
            if (p + 1 < L && !carry_next) load_row<T, NP>(w_out, sen[p + 1], ld, nq, lane, vnext);
            if (np < L && np_slot < 0) load_row<T, NP>(w_in, sen[np], ld, nq, lane, pf);
            for (int i = b; i < 2 * W + 1 - b; i++) {
                if (i == W) continue;
                const int c = p - W + i;
                if (c < 0) continue;
                if (c >= L) break;
                const int slot = pos_slot[c % R];
                float2 u[NP];
                float part = 0.0f;
#pragma unroll
                for (int t = 0; t < NP; t++) {
                    const int q = lane + kWave * t;
                    u[t] = q < nq ? ring_u[slot * nq + q] : make_float2(0.0f, 0.0f);
                    part += u[t].x * v[t].x + u[t].y * v[t].y;
                }
                const float f = wave_sum(part);
                const float g = grad_label1(f, alpha_p, et);
#pragma unroll
                for (int t = 0; t < NP; t++) {
                    const int q = lane + kWave * t;
                    if (q < nq) ring_u[slot * nq + q] = make_float2(u[t].x + g * v[t].x, u[t].y + g * v[t].y);
                    v[t].x += g * u[t].x;
                    v[t].y += g * u[t].y;
                }
                cnt[kCntPairs]++;
                if (a.compute_loss) loss_a += (double)softplusf(-f);
            }
            if (!carry_next)
                store_delta<T, NP>(w_out, sen[p], ld, nq, lane, v, v0, salt0 ^ hash32((uint32_t)p), a.rounding);
            carried = carry_next;
            // position p-W leaves the window; position np enters it
            const int r = p - W;
            const int r_slot = r >= 0 ? pos_slot[r % R] : -1;
            if (np < L && np_slot >= 0) bind(np, np_slot, sen[np]);
            if (r_slot >= 0 && slot_last[r_slot] == r) flush(r_slot);
            if (np < L && np_slot < 0) {
                const int sl = find_free(p + 1 - W);
                fill(sl, pf);
                bind(np, sl, sen[np]);
            }
        }
        for (int r0 = 0; r0 < R; r0 += kWave) {  // slots still referenced by the last window
            const int sx = r0 + lane;
            uint64_t mask = __builtin_amdgcn_ballot_w64(sx < R && slot_last[sx] >= L - W);
            while (mask) {
                const int sl = r0 + (int)__builtin_ctzll(mask);
                mask &= mask - 1;
                flush(sl);
            }
        }
    }

    // ---- repulsion (cpu:571-679)
    // Per batch of B positions: K shared negatives; logits G = N C^T against
    // the batch rows C, label-0 gradient, then W_out[n_k] += sum_i G_ki c_i and
    // W_in[c_i] += sum_k G_ki n_k, all from one snapshot as the reference's
    // three cblas_sgemm calls (cpu:633-656).  The batch rows are snapshotted
    // into the (now idle) ring memory; repeated negatives are merged with a
    // multiplicity so every word's W_out row is read and written by exactly
    // one register block of NB negatives.
    const int B = a.B, NEG = a.NEG;
    const int nbat = (L + B - 1) / B;
    const float alpha_r = L == 1 ? alpha0 : alpha;
    PT *snap = reinterpret_cast<PT *>(lds);
    int32_t *uneg = pos_slot + R;   // unique negatives of the batch
    int32_t *umul = uneg + a.kmax;  // their multiplicities
    for (int bi = 0; bi < nbat; bi++) {
        const int lo = bi * B;
        const int isz = L - lo < B ? L - lo : B;
        int K;
        int64_t n0 = 0;
        if (planned) {
            const int64_t bo = a.plan_batch_off[s] + bi;
            n0 = a.plan_neg_off[bo];
            K = (int)(a.plan_neg_off[bo + 1] - n0);
        } else {
            const uint32_t r = draw(a.seed, s, ep, kTagRepWindow, (uint32_t)bi).x;
            const int rep_b =
                (lo <= W - 1 || (L - 1) - lo <= W - 1) ? (int)(r % (uint32_t)W) : (int)(r % (uint32_t)(2 * W));
            K = rep_b * NEG;
        }
        cnt[kCntBatches]++;
        cnt[kCntNegatives] += (uint64_t)K;
        cnt[kCntRepPairs] += (uint64_t)K * (uint64_t)isz;
        if (K == 0) continue;
        // draw K negatives, 64 at a time, rejecting in-sentence words (cpu:604-622)
        for (int k0 = 0; k0 < K; k0 += kWave) {
            const int k = k0 + lane;
            const bool live = k < K;
            if (planned) {
                if (live) negs[k] = a.plan_negs[n0 + k];
                continue;
            }
            const uint32_t idx = (uint32_t)bi * 8192u + (uint32_t)k;
            int cand = alias_draw(a.alias_prob, a.alias_idx, a.V, draw(a.seed, s, ep, kTagNeg, idx));
            // Redraw uniformly over [1, V-1] while the candidate occurs in the
            // sentence.  The reference loops without bound (cpu:611-615); the
            // bound here only matters when the sentence holds nearly the whole
            // vocabulary, and then the last candidate is kept.
            for (int attempt = 1; attempt <= kMaxRedraw; attempt++) {
                bool hit = false;
                for (int t0 = 0; t0 < L; t0 += kWave) {
                    const int tk = t0 + lane < L ? sen[t0 + lane] : -1;
                    const int n = L - t0 < kWave ? L - t0 : kWave;
                    for (int j = 0; j < n; j++) hit |= (cand == __builtin_amdgcn_readlane(tk, j));
                }
                const bool pending = live && hit;
                if (__builtin_amdgcn_ballot_w64(pending) == 0) break;
                if (pending) {
                    const U4 r2 = draw(a.seed, s, ep, kTagRedraw + ((uint32_t)attempt << 8), idx);
                    cand = 1 + (int)__umulhi(r2.x, (uint32_t)(a.V - 1));
                }
            }
            if (live) negs[k] = cand;
        }
        // merge repeated negatives: first occurrence keeps the word, counts copies
        int U = 0;
        for (int k0 = 0; k0 < K; k0 += kWave) {
            const int k = k0 + lane;
            const int w = k < K ? negs[k] : -1;
            bool first = k < K;
            int m = 0;
            for (int j = 0; j < K; j++) {
                const int wj = negs[j];
                m += wj == w;
                first = first && !(wj == w && j < k);
            }
            const uint64_t fm = __builtin_amdgcn_ballot_w64(first);
            const int before = __builtin_amdgcn_mbcnt_hi((uint32_t)(fm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)fm, 0));
            if (first) {
                uneg[U + before] = w;
                umul[U + before] = m;
            }
            U += __builtin_popcountll(fm);
        }
        // snapshot of the batch rows (after this sentence's attraction)
        for (int i = 0; i < isz; i++) {
            PT row[NP];
            load_row<T, NP>(w_in, sen[lo + i], ld, nq, lane, row);
#pragma unroll
            for (int t = 0; t < NP; t++) {
                const int q = lane + kWave * t;
                if (q < nq) snap[i * nq + q] = row[t];
            }
        }
        const uint32_t salt_b = salt0 ^ hash32((uint32_t)bi * 0x27D4EB2Fu + 0x5000u);
        for (int kb = 0; kb < U; kb += NB) {
            const int nbk = U - kb < NB ? U - kb : NB;
            int32_t nrow[NB];
            float mult[NB];
            float2 nv[NB][NP], dn[NB][NP];
#pragma unroll
            for (int k = 0; k < NB; k++) {
                nrow[k] = k < nbk ? uneg[kb + k] : -1;
                mult[k] = k < nbk ? (float)umul[kb + k] : 0.0f;
                PT tmp[NP];
                if (k < nbk) load_row<T, NP>(w_out, nrow[k], ld, nq, lane, tmp);
                else {
#pragma unroll
                    for (int t = 0; t < NP; t++) tmp[t] = PT{};
                }
                to_f<T, NP>(tmp, nv[k]);
#pragma unroll
                for (int t = 0; t < NP; t++) dn[k][t] = make_float2(0.0f, 0.0f);
            }
            for (int i = 0; i < isz; i++) {
                float2 c[NP];
#pragma unroll
                for (int t = 0; t < NP; t++) {
                    const int q = lane + kWave * t;
                    c[t] = q < nq ? Pair<T>::tof(snap[i * nq + q]) : make_float2(0.0f, 0.0f);
                }
                float g[NB];
#pragma unroll
                for (int k = 0; k < NB; k++) {
                    float part = 0.0f;
#pragma unroll
                    for (int t = 0; t < NP; t++) part += nv[k][t].x * c[t].x + nv[k][t].y * c[t].y;
                    const float f = wave_sum(part);
                    g[k] = k < nbk ? grad_label0(f, alpha_r, et) : 0.0f;
                    if (a.compute_loss && k < nbk) loss_r += (double)mult[k] * (double)softplusf(f);
                }
                float2 dc[NP];
#pragma unroll
                for (int t = 0; t < NP; t++) {
                    dc[t] = make_float2(0.0f, 0.0f);
#pragma unroll
                    for (int k = 0; k < NB; k++) {
                        const float gm = g[k] * mult[k];
                        dc[t].x += gm * nv[k][t].x;
                        dc[t].y += gm * nv[k][t].y;
                        dn[k][t].x += g[k] * c[t].x;
                        dn[k][t].y += g[k] * c[t].y;
                    }
                }
                float2 zero[NP];
#pragma unroll
                for (int t = 0; t < NP; t++) zero[t] = make_float2(0.0f, 0.0f);
                store_delta<T, NP>(w_in, sen[lo + i], ld, nq, lane, dc, zero,
                                   salt_b ^ hash32((uint32_t)(kb * 1024 + i)), a.rounding);
            }
#pragma unroll
            for (int k = 0; k < NB; k++) {
                if (k < nbk) {
                    float2 zero[NP];
#pragma unroll
                    for (int t = 0; t < NP; t++) {
                        zero[t] = make_float2(0.0f, 0.0f);
                        dn[k][t].x *= mult[k];
                        dn[k][t].y *= mult[k];
                    }
                    store_delta<T, NP>(w_out, nrow[k], ld, nq, lane, dn[k], zero,
                                       salt_b ^ hash32((uint32_t)(kb + k) + 0x9000u), a.rounding);
                }
            }
        }
    }
}

template <typename T, int NP, int NB>
__global__ __launch_bounds__(256) void sgns_fast_kernel(FastArgs a) {
    extern __shared__ __align__(16) char smem[];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    char *lds = smem + (size_t)wave * a.lds_per_wave;
    const float *et = a.exp_table;
    double loss_a = 0.0, loss_r = 0.0;
    uint64_t cnt[kCntSlots] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int64_t n = a.s_end - a.s_begin;
    if (a.persistent) {
        for (;;) {
            unsigned long long qi = 0;
            if (lane == 0) qi = atomicAdd(a.queue, 1ull);
            const int64_t idx = readfirst_i64((int64_t)qi);
            if (idx >= n) break;
            const int64_t slot = a.s_begin + idx;
            const int64_t s = a.order ? (int64_t)a.order[slot] : slot;
            const int64_t raw = a.plan_b ? 0 : (a.order ? a.queue_raw[slot] : a.sent_raw[s]);
            train_sentence<T, NP, NB>(a, s, raw, lane, lds, loss_a, loss_r, cnt, et);
        }
    } else {
        const int64_t idx = (int64_t)blockIdx.x * (blockDim.x / kWave) + wave;
        const int64_t s = a.s_begin + idx;
        if (idx < n) train_sentence<T, NP, NB>(a, s, a.plan_b ? 0 : a.sent_raw[s], lane, lds, loss_a, loss_r, cnt, et);
    }
    if (lane == 0) {
        if (a.compute_loss) {
            atomicAdd(&a.loss[0], loss_a);
            atomicAdd(&a.loss[1], loss_r);
        }
#pragma unroll
        for (int k = 0; k < kCntRepPairs + 1; k++)
            if (cnt[k]) atomicAdd(&a.counters[k], (unsigned long long)cnt[k]);
    }
}

// n draws of the sampler for sentence id `sent` (the counter layout of the
// trainer's negatives), for testing the on-device distribution.
__global__ void sgns_sample_kernel(const float *prob, const int32_t *alias, int V, uint64_t seed, int64_t sent,
                                   int64_t n, int32_t *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = alias_draw(prob, alias, V, draw(seed, sent + (i >> 13), 0, kTagNeg, (uint32_t)(i & 8191)));
}

__global__ void sgns_reset_kernel(unsigned long long *counters, int n, double *loss, unsigned long long *queue) {
    const int i = threadIdx.x;
    if (i < n) counters[i] = 0;
    if (i < 2) loss[i] = 0.0;
    if (i == 0 && queue) *queue = 0;
}

template <typename T, int NP>
hipError_t launch_typed(const FastArgs &a, int wpb, int blocks, hipStream_t st) {
    constexpr int NB = NP <= 1 ? 8 : (NP == 2 ? 6 : 4);
    const int64_t n = a.s_end - a.s_begin;
    if (n <= 0) return hipSuccess;
    const size_t lds = (size_t)a.lds_per_wave * (size_t)wpb;
    auto kern = sgns_fast_kernel<T, NP, NB>;
    hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    dim3 grid;
    if (a.persistent) grid = dim3((unsigned)blocks);
    else grid = dim3((unsigned)((n + wpb - 1) / wpb));
    hipLaunchKernelGGL(kern, grid, dim3(kWave * wpb), lds, st, a);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_fast(const FastArgs &a, int bf16, int np, int wpb, int blocks, hipStream_t st) {
    if (bf16) {
        switch (np) {
        case 1: return launch_typed<uint16_t, 1>(a, wpb, blocks, st);
        case 2: return launch_typed<uint16_t, 2>(a, wpb, blocks, st);
        case 3: return launch_typed<uint16_t, 3>(a, wpb, blocks, st);
        case 4: return launch_typed<uint16_t, 4>(a, wpb, blocks, st);
        }
    } else {
        switch (np) {
        case 1: return launch_typed<float, 1>(a, wpb, blocks, st);
        case 2: return launch_typed<float, 2>(a, wpb, blocks, st);
        case 3: return launch_typed<float, 3>(a, wpb, blocks, st);
        case 4: return launch_typed<float, 4>(a, wpb, blocks, st);
        }
    }
    return hipErrorInvalidValue;
}

hipError_t launch_sample(const float *prob, const int32_t *alias, int V, uint64_t seed, int64_t sent, int64_t n,
                         int32_t *out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(sgns_sample_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, prob, alias, V, seed,
                       sent, n, out);
    return hipGetLastError();
}

hipError_t launch_reset(unsigned long long *counters, int n, double *loss, unsigned long long *queue, hipStream_t st) {
    hipLaunchKernelGGL(sgns_reset_kernel, dim3(1), dim3(64), 0, st, counters, n, loss, queue);
    return hipGetLastError();
}

}  // namespace sgns
