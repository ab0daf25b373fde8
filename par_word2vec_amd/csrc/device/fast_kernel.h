// fast_kernel.h -- Hogwild PAR-SGNS trainer for gfx950 (plain and
// persistent scheduling, fp32 or bf16 tables).
//
// One wavefront owns one sentence at a time (PAR_Word2Vec_gpu_text.cu gives
// a thread block a contiguous sentence range, text.cu:385-390; here the unit
// is a wave so every reduction is a DPP butterfly with no barrier).
//
// Attraction (cpu:529-568): the W_in rows of the sliding window [p-W, p+W]
// live in an LDS ring of R = 2W+1 word-keyed slots holding the current fp32
// value and the value that was loaded; a row is fetched once when its word
// enters the window (loaded two centres ahead) and written back once when
// its last occurrence leaves, as delta = current - loaded.  The centre's
// W_out row stays in registers for all its pairs (and across a repeated
// centre).  Pairs are applied in the reference's order with its in-place
// semantics, so later pairs see earlier updates.
//
// Repulsion (cpu:571-679): per batch of B positions, K = rep_b * negative
// shared negatives are drawn from the on-device alias table with the
// reference's in-sentence rejection; the K x B logits, the label-0 gradient
// and both updates come from one snapshot, on the matrix cores (bf16 tables,
// repulsion_mfma) or in register blocks of NB negatives (vector path, which
// sums each batch row's W_in delta over the blocks in LDS).  Either way every
// W_in and W_out row of a batch is written back once.
//
// Write-back: fp32 tables take the delta with global float atomics (no lost
// updates, as the reference GPU's atomicAdd, text.cu:438-439, 710-717);
// bf16 tables add the packed bf16 step that moves the wave's base value to
// the stochastically rounded base + delta (packed-bf16 atomics: a concurrent
// update is kept, never overwritten).  The matrix-core instance can instead
// store the rounded value of rows at or past a "hot row" bound (plain
// stores, the reference CPU trainer's in-place Hogwild updates, cpu:561-565,
// 658-674): a concurrent update of such a row may be overwritten, and the
// frequent rows below the bound, where concurrent updates are common, keep
// their atomics.
//
// Randomness: Philox4x32-10 keyed by the seed; the counter names the
// sentence, epoch, purpose and index of each draw.
#pragma once
#include "sgns_device.h"
#include "sgns_internal.h"

namespace sgns {
namespace fast {

constexpr int kMaxRedraw = 64;
// Open-addressing hash set of a sentence's token ids in LDS, for the
// in-sentence rejection check of sentences longer than the 128 tokens kept in
// registers: slots hold a token id or kTabEmpty; a sentence uses the set when
// it holds at most 3/4 of the slots.
constexpr int32_t kTabEmpty = -1;
__device__ __forceinline__ uint32_t tab_hash(int32_t w, int log2) { return ((uint32_t)w * 0x9E3779B1u) >> (32 - log2); }
constexpr int kUnusedSlot = -(1 << 30);

// Rounding of bf16 write-backs per table (par_sgns_rounding): stochastic for
// both, round-to-nearest for both, or nearest for one table only.
// A sentence's wave-uniform metadata (queue order, offsets, rates), which
// the kernel never writes.  On the matrix-core instances with two or more
// pairs per lane it is loaded through the scalar cache: the scalar counter
// tracks those loads, so waiting for them does not wait for this wave's
// vector atomics to retire (the previous sentence's write-backs).  On the
// other instances the same loads moved the compiler's register allocation
// the wrong way (with one pair per lane it reloaded kernel arguments inside
// the centre loop), so there they stay vector loads.
template <int NP, bool MF, typename X>
__device__ __forceinline__ X load_meta(const X *p) {
    if constexpr (MF && NP >= 2) return *(const __attribute__((address_space(4))) X *)p;
    else return *p;
}
__device__ __forceinline__ int nearest_in(const FastArgs &a) { return a.rounding == 1 || a.rounding == 3; }
__device__ __forceinline__ int nearest_out(const FastArgs &a) { return a.rounding == 1 || a.rounding == 2; }

// ---------------------------------------------------------------- stores
// Every write-back is a no-return atomic: fp32 tables add the fp32 delta;
// bf16 tables decide the rounding of base + delta against the wave's own
// snapshot `base` (stochastic or nearest), and add the resulting bf16 step
// s - base (exact in bf16) with a packed bf16 atomic.  With no concurrent
// writer the row ends exactly at the rounded value; a concurrent update is
// kept, never overwritten.  A step that rounds to zero issues no write.
typedef __bf16 __attribute__((ext_vector_type(2))) bf16x2_t;

// The packed bf16 step s - base from base to s = cur rounded to bf16 (0 if
// the pair does not move).
__device__ __forceinline__ uint32_t bf16_step_to(float2 base, float2 cur, uint32_t &rs, int nearest) {
    // per-lane Weyl stream of rounding noise, seeded per sentence and lane by
    // a hash: stochastic rounding needs each noise value uniform and
    // independent of the value it rounds, which the golden-ratio step gives
    // at one add (an xorshift32 step took six instructions)
    rs += 0x9E3779B9u;
    // both roundings add a value below the kept 16 bits to the fp32 bit
    // pattern and truncate: 16 uniform bits for stochastic, half an ulp for
    // nearest (ties away from zero)
    const uint32_t noise = nearest ? 0x80008000u : rs;
    const uint32_t bx = __builtin_bit_cast(uint32_t, cur.x), by = __builtin_bit_cast(uint32_t, cur.y);
    const float sx = __builtin_bit_cast(float, (bx + (noise & 0xFFFFu)) & 0xFFFF0000u) - base.x;
    const float sy = __builtin_bit_cast(float, (by + (noise >> 16)) & 0xFFFF0000u) - base.y;
    // the steps' high halves: exact when base and the rounded value share a
    // binade (s - base is then a multiple of base's ulp below 2^8 ulps);
    // otherwise the step's low bits are truncated toward zero
    return __builtin_amdgcn_perm(__builtin_bit_cast(uint32_t, sy), __builtin_bit_cast(uint32_t, sx), 0x07060302u);
}
// cur rounded to bf16 (the same rounding and noise stream as bf16_step_to),
// packed: x in the low half, y in the high half.
__device__ __forceinline__ uint32_t bf16_round_to(float2 cur, uint32_t &rs, int nearest) {
    rs += 0x9E3779B9u;
    const uint32_t noise = nearest ? 0x80008000u : rs;
    const uint32_t bx = __builtin_bit_cast(uint32_t, cur.x), by = __builtin_bit_cast(uint32_t, cur.y);
    return __builtin_amdgcn_perm(by + (noise >> 16), bx + (noise & 0xFFFFu), 0x07060302u);
}
// The packed bf16 step from base to the packed bf16 value v (as bf16_step_to).
__device__ __forceinline__ uint32_t bf16_step_from(float2 base, uint32_t v) {
    const float sx = __builtin_bit_cast(float, v << 16) - base.x;
    const float sy = __builtin_bit_cast(float, v & 0xFFFF0000u) - base.y;
    return __builtin_amdgcn_perm(__builtin_bit_cast(uint32_t, sy), __builtin_bit_cast(uint32_t, sx), 0x07060302u);
}
// The same step for base + dlt.
__device__ __forceinline__ uint32_t bf16_step(float2 base, float2 dlt, uint32_t &rs, int nearest) {
    return bf16_step_to(base, make_float2(base.x + dlt.x, base.y + dlt.y), rs, nearest);
}
__device__ __forceinline__ void wb_pair(uint16_t *w, int64_t row, int ld, int q, float2 base, float2 dlt,
                                        uint32_t &rs, int nearest) {
    if (dlt.x == 0.0f && dlt.y == 0.0f) return;
    const uint32_t step = bf16_step(base, dlt, rs, nearest);
    if (step == 0u) return;
    bf16x2_t *p = reinterpret_cast<bf16x2_t *>(w + row * ld) + q;
    __builtin_amdgcn_global_atomic_fadd_v2bf16(p, __builtin_bit_cast(bf16x2_t, step));
}

template <typename T, int NP>
__device__ __forceinline__ void load_row(const T *w, int64_t row, int ld, int nq, int lane,
                                         typename Pair<T>::P (&out)[NP]) {
#pragma unroll
    for (int t = 0; t < NP; t++) {
        const int q = lane + kWave * t;
        if (t < NP - 1 || q < nq) out[t] = Pair<T>::load(w, row, ld, q);
        else out[t] = typename Pair<T>::P{};
    }
}

template <typename T, int NP>
__device__ __forceinline__ void to_f(const typename Pair<T>::P (&in)[NP], float2 (&out)[NP]) {
#pragma unroll
    for (int t = 0; t < NP; t++) out[t] = Pair<T>::tof(in[t]);
}


// STORE: a plain store of base + dlt (fp32) or of its rounding (bf16)
// instead of an atomic add of the delta -- the table's plain-store write-back
// policy (a concurrent update of the row may be overwritten).
template <typename T, int NP, bool STORE = false>
__device__ __forceinline__ void store_delta(T *w, int64_t row, int ld, int nq, int lane, const float2 (&base)[NP],
                                            const float2 (&dlt)[NP], uint32_t &rs, int nearest) {
    if constexpr (sizeof(T) == 4 && STORE) {
        // the same one-float-per-lane layout as the atomic path below
        float *p = reinterpret_cast<float *>(w) + row * ld;
        const int s0 = lane >> 1, s1 = 32 + (lane >> 1), n = 2 * nq;
        const bool odd = lane & 1;
#pragma unroll
        for (int t = 0; t < NP; t++) {
            const float a0 = __shfl(dlt[t].x, s0), a1 = __shfl(dlt[t].y, s0);
            const float b0 = __shfl(dlt[t].x, s1), b1 = __shfl(dlt[t].y, s1);
            const float c0 = __shfl(base[t].x, s0), c1 = __shfl(base[t].y, s0);
            const float e0 = __shfl(base[t].x, s1), e1 = __shfl(base[t].y, s1);
            const float va = odd ? a1 : a0, vb = odd ? b1 : b0;
            const float ba = odd ? c1 : c0, bb = odd ? e1 : e0;
            const int f = 128 * t + lane;
            if (f < n && va != 0.0f) __hip_atomic_store(p + f, ba + va, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (f + 64 < n && vb != 0.0f)
                __hip_atomic_store(p + f + 64, bb + vb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    } else if constexpr (STORE) {
#pragma unroll
        for (int t = 0; t < NP; t++) {
            const int q = lane + kWave * t;
            if ((t < NP - 1 || q < nq) && (dlt[t].x != 0.0f || dlt[t].y != 0.0f)) {
                const uint32_t v = bf16_round_to(make_float2(base[t].x + dlt[t].x, base[t].y + dlt[t].y), rs, nearest);
                uint32_t *p = reinterpret_cast<uint32_t *>(w + row * ld) + q;
                __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    } else if constexpr (sizeof(T) == 4) {
        // fp32: one float per lane and atomic.  Lane l takes float l of a
        // 128-float block from lane l/2 and float 64 + l from lane 32 + l/2
        // (shuffles: every lane of the wave takes part), so an instruction adds
        // 64 consecutive floats -- four 64-byte requests, where a pair per lane
        // (two atomics 8 bytes apart) costs eight.  Pairs past the row are zero.
        float *p = reinterpret_cast<float *>(w) + row * ld;
        const int s0 = lane >> 1, s1 = 32 + (lane >> 1), n = 2 * nq;
        const bool odd = lane & 1;
#pragma unroll
        for (int t = 0; t < NP; t++) {
            const float a0 = __shfl(dlt[t].x, s0), a1 = __shfl(dlt[t].y, s0);
            const float b0 = __shfl(dlt[t].x, s1), b1 = __shfl(dlt[t].y, s1);
            const float va = odd ? a1 : a0, vb = odd ? b1 : b0;
            const int f = 128 * t + lane;
            if (f < n && va != 0.0f) unsafeAtomicAdd(p + f, va);
            if (f + 64 < n && vb != 0.0f) unsafeAtomicAdd(p + f + 64, vb);
        }
    } else {
#pragma unroll
        for (int t = 0; t < NP; t++) {
            const int q = lane + kWave * t;
            if (t < NP - 1 || q < nq) wb_pair(w, row, ld, q, base[t], dlt[t], rs, nearest);
        }
    }
}

// One draw from the unigram^0.75 alias table (before in-sentence rejection).
// A bucket's entry is {keep probability (float bits), alias}: one 8-byte load.
__device__ __forceinline__ int alias_draw(const uint2 *tab, int V, U4 rr) {
    const uint32_t bucket = __umulhi(rr.x, (uint32_t)V);
    const float u2 = (float)(rr.y >> 8) * (1.0f / 16777216.0f);
    const uint2 e = tab[bucket];
    return u2 < __builtin_bit_cast(float, e.x) ? (int)bucket : (int)e.y;
}

// ------------------------------------------- repulsion on the matrix cores
// One batch of the repulsion for bf16 tables.  C = the batch's W_in rows
// (snapshot in LDS, zero rows up to a multiple of 16, zero columns up to
// ld), N = the W_out rows of the unique negatives, m = their multiplicities:
//   G    = N C^T                v_mfma_f32_16x16x32_bf16 (bf16 operands are
//                               the table values, products exact, fp32 sums)
//   Gm   = m (0 - sigmoid(G)) alpha   (exp-table sigmoid, cpu:641-645), bf16
//   dC   = Gm^T N               v_mfma_f32_16x16x16_bf16
//   dN   = Gm C                 v_mfma_f32_16x16x16_bf16
// the three cblas_sgemm calls of cpu:633-656, from one snapshot.  An output
// tile of dC or dN holds one dimension of four rows per lane (D layout of a
// 16x16 MFMA: lane l, register r -> row 4(l/16)+r, column l%16), written back
// a step pair at a time (write_pair): each row of C takes one write-back per
// group of kMfU*16 unique negatives (one per batch up to 32), each row of N
// one.
// The batch holds at most kMfI*16 = 32 positions (the host keeps B <= 32 on
// this path), so every product reads the one snapshot the reference uses.
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef short s16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2_t __attribute__((ext_vector_type(2)));

// Table reads of the matrix-core path go through buffer resources: 32-bit
// byte offsets (the host enables the path only for tables below 4 GiB - 4 KiB)
// and the agent-coherent cache policy (sc1) of the row loads.  A lane with no
// row reads at kNoRow (+ less than 4 KiB), past the resource's size: the
// hardware range check returns zero.  So every load is unconditional: a load
// under a branch, or whose result is merged with a default, makes the
// compiler wait for it (and, since a wave's memory operations retire in
// order, for everything issued before it) where the values merge, and loads
// issued ahead of their use would gain nothing.  Operand buffers for steps
// ahead alternate between two register sets (no rotating copies, which would
// wait for the load just issued), and a scheduling barrier keeps each step's
// loads ahead of its write-back atomics (buffer loads and global atomics are
// otherwise free to be reordered).
constexpr int kSc1 = 16;
constexpr int kRsrcWord3 = 0x00020000;
constexpr uint32_t kNoRow = 0xFFFFF000u;

// LLVM's packed-bf16 buffer atomic add (buffer_atomic_pk_add_bf16), bound
// by name: clang has no builtin for it.  Its result is never used (the
// no-return form is emitted).
__device__ bf16x2_t buffer_atomic_pk_add_bf16(bf16x2_t v, __amdgpu_buffer_rsrc_t r, int off, int soff, int aux)
    __asm("llvm.amdgcn.raw.ptr.buffer.atomic.fadd.v2bf16");

// A dropped atomic (offset kNoRow): issued where a loop's entry must have
// the same operations in flight as its later iterations.
__device__ __forceinline__ void drop_atomic(__amdgpu_buffer_rsrc_t rw) {
    buffer_atomic_pk_add_bf16(__builtin_bit_cast(bf16x2_t, 0u), rw, (int)kNoRow, 0, 0);
}

// Write back a step pair of one 16 x 16 output tile whose lane (c, g) holds
// column c (a dimension) of rows 4g..4g+3 (one per register): x = step jt
// (dimensions jt*16 + c), y = step jt+1.  For each register, lanes c and c^1
// exchange one value (DPP quad_perm [1,0,3,2]): an even lane then holds
// dimensions (c, c+1) of step jt, an odd lane (c+15, c+16) of step jt+1 --
// the pair starting at dimp -- so one packed-bf16 atomic writes four rows x
// 64 contiguous bytes (the atomic unit's granularity: four requests per
// instruction).  base = the values the rounding is decided against (the same
// pair), off = the rows' byte offsets (kNoRow: no row).  Every lane issues all
// four atomics; a pair with nothing to add (a step that rounds to zero, no
// row, an inactive tile) lands past the resource's size, which the range
// check drops without a memory request, so the compiler's wait counts
// include every atomic.
// Write-back policy of a table (template argument WB): kWbAtomic adds the
// rounding step with packed-bf16 atomics; kWbStore stores the rounded value
// (value unchanged: no write); kWbHot stores rows whose byte offset is at or
// past hot_bytes and adds the step to the others -- every lane issues both
// the atomic and the store, one of them dropped.
enum { kWbAtomic = 0, kWbStore = 1, kWbHot = 2 };

template <int WB>
__device__ __forceinline__ void write_pair(__amdgpu_buffer_rsrc_t rw, bool active, const uint32_t (&off)[4],
                                           const uint32_t (&base)[4], f32x4_t x, f32x4_t y, int dimp, int odd,
                                           uint32_t &rs, int nearest, uint32_t hot_bytes) {
    uint32_t st[4] = {0u, 0u, 0u, 0u}, o[4] = {kNoRow, kNoRow, kNoRow, kNoRow};
    uint32_t sv[4] = {0u, 0u, 0u, 0u}, os[4] = {kNoRow, kNoRow, kNoRow, kNoRow};
    if (active) {  // uniform
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const float sent = odd ? x[r] : y[r];
            const float got = __builtin_bit_cast(
                float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, sent), 0xB1, 0xF, 0xF, false));
            const float2 dl = odd ? make_float2(got, y[r]) : make_float2(x[r], got);
            const float2 b = Pair<uint16_t>::tof(base[r]);
            if constexpr (WB == kWbStore) {
                sv[r] = bf16_round_to(make_float2(b.x + dl.x, b.y + dl.y), rs, nearest);
                os[r] = sv[r] != base[r] ? off[r] + 2u * (uint32_t)dimp : kNoRow;
            } else if constexpr (WB == kWbHot) {
                sv[r] = bf16_round_to(make_float2(b.x + dl.x, b.y + dl.y), rs, nearest);
                st[r] = bf16_step_from(b, sv[r]);
                // (off = kNoRow is never hot, and kNoRow + 2 dimp is past the resource)
                const bool hot = off[r] < hot_bytes;
                o[r] = hot && st[r] != 0u ? off[r] + 2u * (uint32_t)dimp : kNoRow;
                os[r] = !hot && sv[r] != base[r] ? off[r] + 2u * (uint32_t)dimp : kNoRow;
            } else {
                st[r] = bf16_step(b, dl, rs, nearest);
                // (kNoRow + 2 dimp is past the resource too; pairs at or past d
                // hold zero padding and a zero update, so their step is 0)
                o[r] = st[r] != 0u ? off[r] + 2u * (uint32_t)dimp : kNoRow;
            }
        }
    }
    if constexpr (WB != kWbStore) {
#pragma unroll
        for (int r = 0; r < 4; r++) buffer_atomic_pk_add_bf16(__builtin_bit_cast(bf16x2_t, st[r]), rw, (int)o[r], 0, 0);
    }
    if constexpr (WB != kWbAtomic) {
#pragma unroll
        for (int r = 0; r < 4; r++) __builtin_amdgcn_raw_buffer_store_b32(sv[r], rw, (int)os[r], 0, 0);
    }
}

// One row's write-back on a matrix-core instance (bf16 tables below 4 GiB -
// 4 KiB, addressed through a buffer resource): the step from the loaded
// values `base` to the current values `cur` rounded to bf16.  Every lane
// issues its NP atomics and a pair past the row (q >= nq) or with a zero
// step lands past the resource, dropped by the range check -- no branch per
// pair.
// WB: the table's write-back policy (see write_pair).
template <int NP, int WB>
__device__ __forceinline__ void store_value_rsrc(__amdgpu_buffer_rsrc_t rw, uint32_t row, uint32_t row_bytes, int nq,
                                                 int lane, const float2 (&base)[NP], const float2 (&cur)[NP],
                                                 uint32_t &rs, int nearest, uint32_t hot_bytes) {
#pragma unroll
    for (int t = 0; t < NP; t++) {
        const int q = lane + kWave * t;
        const uint32_t at = row * row_bytes + 4u * (uint32_t)q;
        if constexpr (WB == kWbStore) {
            const uint32_t v = bf16_round_to(cur[t], rs, nearest);
            const uint32_t bp = __builtin_amdgcn_perm(__builtin_bit_cast(uint32_t, base[t].y),
                                                      __builtin_bit_cast(uint32_t, base[t].x), 0x07060302u);
            __builtin_amdgcn_raw_buffer_store_b32(v, rw, (int)(q < nq && v != bp ? at : kNoRow), 0, 0);
        } else if constexpr (WB == kWbHot) {
            const uint32_t v = bf16_round_to(cur[t], rs, nearest);
            const uint32_t st = bf16_step_from(base[t], v);
            const uint32_t bp = __builtin_amdgcn_perm(__builtin_bit_cast(uint32_t, base[t].y),
                                                      __builtin_bit_cast(uint32_t, base[t].x), 0x07060302u);
            const bool hot = at < hot_bytes;
            const uint32_t o = q < nq && hot && st != 0u ? at : kNoRow;
            const uint32_t os = q < nq && !hot && v != bp ? at : kNoRow;
            buffer_atomic_pk_add_bf16(__builtin_bit_cast(bf16x2_t, st), rw, (int)o, 0, 0);
            __builtin_amdgcn_raw_buffer_store_b32(v, rw, (int)os, 0, 0);
        } else {
            const uint32_t st = bf16_step_to(base[t], cur[t], rs, nearest);
            const uint32_t o = q < nq && st != 0u ? at : kNoRow;
            buffer_atomic_pk_add_bf16(__builtin_bit_cast(bf16x2_t, st), rw, (int)o, 0, 0);
        }
    }
}

template <int NP, int WIN, int WOUT>
__device__ __forceinline__ void repulsion_mfma(const FastArgs &a, const int32_t *bat, int bat_tok, bool bat_in_reg,
                                               int isz, int U, const int32_t *uneg, const int32_t *umul, char *lds,
                                               int lane, float alpha_r, double &loss_r, uint32_t &rs) {
    uint16_t *w_in = static_cast<uint16_t *>(a.w_in);
    uint16_t *w_out = static_cast<uint16_t *>(a.w_out);
    const int ld = a.ld, d = a.d, SC = mf_snap_stride(ld);
    const uint32_t row_bytes = (uint32_t)ld * 2u;
    const __amdgpu_buffer_rsrc_t rout =
        __builtin_amdgcn_make_buffer_rsrc(w_out, (short)0, (int)((uint32_t)a.V * row_bytes), kRsrcWord3);
    const __amdgpu_buffer_rsrc_t rin =
        __builtin_amdgcn_make_buffer_rsrc(w_in, (short)0, (int)((uint32_t)a.V * row_bytes), kRsrcWord3);
    uint16_t *snapC = reinterpret_cast<uint16_t *>(lds);
    uint16_t *tr = snapC + (a.B + 1) * SC;
    uint16_t *stg = tr + 16 * kMfTS;  // two staging tiles of 16 x kMfSS
    const int r16 = lane & 15, g4 = lane >> 4;
    const int KS = ld / 32, JT = (d + 15) / 16;
    const int ni = isz, IT = (ni + 15) >> 4;
    auto srow = [&](int i) { return i < ni ? i : a.B; };  // tile rows past the batch read the zero row
    double lr = 0.0;  // this lane's share of the batch loss (summed over the wave at the end)
    {
        // snapshot of the batch rows (after this sentence's attraction) and the
        // zero row.  The rows go from the table straight into LDS (buffer loads
        // with an LDS destination: lane l's 4 bytes land at the instruction's
        // LDS address + 4 l), 256 bytes of a row per instruction, every row of
        // the batch in flight at once and no registers staged; the lanes past
        // the row are switched off, so nothing lands in the row's padding or
        // the next row.  One wait for all of them follows (such loads count
        // in vmcnt like any other).
        // (the batch's row ids are taken first, lane i holding row i: a load
        // inside the loop would wait for every row already in flight.  For a
        // sentence of at most 128 tokens they come from the registers that
        // hold its tokens, bat_tok; a load would wait for the sentence's last
        // write-back atomics, which retire before it)
        for (int q = lane; 2 * q < ld; q += kWave) *reinterpret_cast<uint32_t *>(snapC + a.B * SC + 2 * q) = 0u;
        const int bat_id = bat_in_reg ? bat_tok : (lane < ni ? bat[lane] : 0);  // ni <= kMfI * 16 <= 64
        for (int i = 0; i < ni; i++) {
            const uint32_t off = (uint32_t)__builtin_amdgcn_readlane(bat_id, i) * row_bytes + 4u * (uint32_t)lane;
#pragma unroll
            for (int t = 0; t < NP; t++) {
                auto *dst = (__attribute__((address_space(3))) void *)(snapC + i * SC + 128 * t);
                if (t < NP - 1 || 2 * (lane + kWave * t) < ld)
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(rin, dst, 4, off + 256u * t, 0, 0, kSc1);
            }
        }
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the snapshot has landed

        for (int u0 = 0; u0 < U; u0 += 16 * kMfU) {
            const int nu = U - u0 < 16 * kMfU ? U - u0 : 16 * kMfU;
            const int UT = (nu + 15) >> 4;
            s16x4_t gd[kMfU][kMfI];
            int32_t rowN[kMfU];
            // ---- G = N C^T and the gradient Gm (bf16, D layout)
#pragma unroll
            for (int ut = 0; ut < kMfU; ut++) {
                rowN[ut] = -1;
#pragma unroll
                for (int it = 0; it < kMfI; it++) gd[ut][it] = s16x4_t{0, 0, 0, 0};
                if (ut >= UT) continue;
                const int un = u0 + ut * 16 + r16;
                rowN[ut] = un < U ? uneg[un] : -1;
                const uint32_t offN = (rowN[ut] >= 0 ? (uint32_t)rowN[ut] * row_bytes : kNoRow) + 16u * (uint32_t)g4;
                // this tile's 16 rows of N, all of ld: 16 bytes per lane per 32 dimensions
                u32x4_t nv[4 * NP];
#pragma unroll
                for (int ks = 0; ks < 4 * NP; ks++)
                    nv[ks] = __builtin_amdgcn_raw_buffer_load_b128(rout, ks < KS ? offN + 64u * ks : kNoRow, 0, kSc1);
                f32x4_t acc[kMfI];
#pragma unroll
                for (int it = 0; it < kMfI; it++) acc[it] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int ks = 0; ks < 4 * NP; ks++) {
                    if (ks >= KS) continue;
                    const bf16x8_t av = __builtin_bit_cast(bf16x8_t, nv[ks]);
#pragma unroll
                    for (int it = 0; it < kMfI; it++) {
                        if (it >= IT) continue;
                        const uint16_t *q = snapC + srow(it * 16 + r16) * SC + ks * 32 + 8 * g4;
                        const uint2 b0 = *reinterpret_cast<const uint2 *>(q);
                        const uint2 b1 = *reinterpret_cast<const uint2 *>(q + 4);
                        const bf16x8_t bv = __builtin_bit_cast(bf16x8_t, u32x4_t{b0.x, b0.y, b1.x, b1.y});
                        acc[it] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[it], 0, 0, 0);
                    }
                }
#pragma unroll
                for (int it = 0; it < kMfI; it++) {
                    if (it >= IT) continue;
                    uint32_t gb[4];
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int ul = u0 + ut * 16 + 4 * g4 + r;  // D row: negative
                        const int il = it * 16 + r16;              // D column: position
                        float gm = 0.0f;
                        if (ul < U && il < ni) {
                            const float f = acc[it][r];
                            const float m = (float)umul[ul];
                            gm = grad_label0_r(f, alpha_r) * m;
                            if (a.compute_loss) lr += (double)m * (double)softplusf(f);
                        }
                        gb[r] = bf16_rne(gm);
                    }
                    gd[ut][it] = s16x4_t{(short)gb[0], (short)gb[1], (short)gb[2], (short)gb[3]};
                }
                // keep the next tile's row prefetch (4*NP x 16 bytes per lane) out of this one
                __builtin_amdgcn_sched_barrier(0);
            }
            // ---- Gm^T: each 16 x 16 tile transposed through LDS (the A operand of dN)
            s16x4_t gt[kMfU][kMfI];
#pragma unroll
            for (int ut = 0; ut < kMfU; ut++) {
#pragma unroll
                for (int it = 0; it < kMfI; it++) {
                    gt[ut][it] = s16x4_t{0, 0, 0, 0};
                    if (ut >= UT || it >= IT) continue;
#pragma unroll
                    for (int r = 0; r < 4; r++) tr[(4 * g4 + r) * kMfTS + r16] = (uint16_t)gd[ut][it][r];
                    const uint2 t2 = *reinterpret_cast<const uint2 *>(tr + r16 * kMfTS + 4 * g4);
                    gt[ut][it] = __builtin_bit_cast(s16x4_t, t2);
                }
            }
            // ---- dC = Gm^T N and dN = Gm C, one pass over step pairs (32
            // dimensions).  With Gm^T / Gm as the A operand an output tile holds
            // one dimension per lane and four rows per register, and one atomic
            // per register writes four rows x 32 dimensions (write_pair).  N
            // enters as one 128-bit load per lane and tile per step pair (16
            // rows x 32 dimensions, one pair ahead), staged through LDS to give
            // dC its B operands and dN its rounding bases.  A pair's slices are
            // loaded before that pair's write-backs, so dC and the bases see N
            // as G did (no dimension of a W_out row of the group is written
            // before it is read).  Every loop iteration issues the same loads
            // and atomics (dropped ones included) and is entered with the same
            // operations in flight, so each wait is for the previous pair's loads.
            const int odd = r16 & 1;
            uint32_t offT[kMfU];      // this lane's row of tile ut (row r16), its 8-dimension chunk g4
            uint32_t offN[kMfU][4];   // dN output rows: negatives ut*16 + 4*g4 + r (kNoRow past U)
#pragma unroll
            for (int ut = 0; ut < kMfU; ut++) {
                offT[ut] = (rowN[ut] >= 0 ? (uint32_t)rowN[ut] * row_bytes : kNoRow) + 16u * (uint32_t)g4;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int ua = u0 + ut * 16 + 4 * g4 + r;
                    offN[ut][r] = ut < UT && ua < U ? (uint32_t)uneg[ua] * row_bytes : kNoRow;
                }
            }
            uint32_t offC[kMfI][4];   // dC output rows: positions it*16 + 4*g4 + r
#pragma unroll
            for (int it = 0; it < kMfI; it++) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int ip = it * 16 + 4 * g4 + r;
                    // (every lane shuffles, then selects: a shuffle inside the
                    // select would run in the lanes with ip < ni only and read
                    // lanes outside that set)
                    const uint32_t id = (uint32_t)__shfl(bat_id, ip & (kWave - 1), kWave);
                    offC[it][r] = ip < ni ? id * row_bytes : kNoRow;
                }
            }
            // (slices past the last pair read the next row's start or zeros, unused)
            auto load_slice = [&](int jt, u32x4_t (&sl)[kMfU]) {
#pragma unroll
                for (int ut = 0; ut < kMfU; ut++)
                    sl[ut] = __builtin_amdgcn_raw_buffer_load_b128(rout, offT[ut] + 32u * (uint32_t)jt, 0, kSc1);
            };
            const int dpair = (r16 & ~1) + (odd ? 16 : 0);  // first dimension of this lane's pair in a step pair
            auto step_pair = [&](int jt, const u32x4_t (&sl)[kMfU], u32x4_t (&nxt)[kMfU]) {
                load_slice(jt + 2, nxt);
                __builtin_amdgcn_sched_barrier(0);
                // C: positions it*16 + 4*g4 + r at dimensions jt*16 + r16 and +16
                s16x4_t cvX[kMfI], cvY[kMfI];
#pragma unroll
                for (int it = 0; it < kMfI; it++) {
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const uint16_t *q = snapC + srow(it * 16 + 4 * g4 + r) * SC + jt * 16 + r16;
                        cvX[it][r] = (short)q[0];
                        cvY[it][r] = (short)q[16];
                    }
                }
                f32x4_t cX[kMfI], cY[kMfI];
#pragma unroll
                for (int it = 0; it < kMfI; it++) cX[it] = cY[it] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
                const int dimp = jt * 16 + dpair;
#pragma unroll
                for (int ut = 0; ut < kMfU; ut++) {
                    uint16_t *st = stg + (ut & 1) * 16 * kMfSS;
                    // stage this lane's 8 dimensions of row r16
                    uint32_t *sw = reinterpret_cast<uint32_t *>(st + r16 * kMfSS + 8 * g4);
                    *reinterpret_cast<uint2 *>(sw) = make_uint2(sl[ut].x, sl[ut].y);
                    *reinterpret_cast<uint2 *>(sw + 2) = make_uint2(sl[ut].z, sl[ut].w);
                    // dC's B operands (rows 4*g4 + r at dimensions r16 and 16 + r16)
                    // and dN's rounding bases (the pair this lane writes)
                    s16x4_t bX, bY;
                    uint32_t base[4];
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const uint16_t *row = st + (4 * g4 + r) * kMfSS;
                        bX[r] = (short)row[r16];
                        bY[r] = (short)row[16 + r16];
                        base[r] = *reinterpret_cast<const uint32_t *>(row + dpair);
                    }
                    f32x4_t nX = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f}, nY = nX;
                    if (ut < UT) {  // uniform
                        // dN's products first: this tile writes them back next,
                        // and the dC products issued after cover their latency
#pragma unroll
                        for (int it = 0; it < kMfI; it++) {
                            if (it >= IT) continue;
                            nX = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(gt[ut][it], cvX[it], nX, 0, 0, 0);
                            nY = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(gt[ut][it], cvY[it], nY, 0, 0, 0);
                        }
#pragma unroll
                        for (int it = 0; it < kMfI; it++) {
                            if (it >= IT) continue;
                            cX[it] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(gd[ut][it], bX, cX[it], 0, 0, 0);
                            cY[it] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(gd[ut][it], bY, cY[it], 0, 0, 0);
                        }
                    }
                    write_pair<WOUT>(rout, ut < UT, offN[ut], base, nX, nY, dimp, odd, rs, nearest_out(a), a.hot_bytes);
                }
                // dC write-back, rounding against the snapshot
#pragma unroll
                for (int it = 0; it < kMfI; it++) {
                    uint32_t base[4];
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        base[r] = *reinterpret_cast<const uint32_t *>(snapC + srow(it * 16 + 4 * g4 + r) * SC + dimp);
                    write_pair<WIN>(rin, it < IT, offC[it], base, cX[it], cY[it], dimp, odd, rs, nearest_in(a),
                                    a.hot_bytes);
                }
            };
            u32x4_t sX[kMfU], sY[kMfU];
            load_slice(0, sX);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int k = 0; k < 4 * (kMfU + kMfI); k++) drop_atomic(rout);  // a step pair's atomics
            for (int jt = 0; jt < JT; jt += 4) {  // pairs of step pairs: dimensions up to ld
                step_pair(jt, sX, sY);
                step_pair(jt + 2, sY, sX);
            }
        }
    }
    loss_r += lr;
}

// Counts of one sentence (scalars); the kernel folds them into per-lane
// 64-bit totals so no 64-bit counter stays live in scalar registers.
struct SentCount {
    uint32_t tokens, pairs, batches, negatives;
    uint64_t rep_pairs;
};

// ------------------------------------------------------ one sentence
// `raw` is the training progress in raw words when the sentence starts: its
// corpus offset in sentence order, or the words dequeued before it when a
// queue reorders sentences -- the reference decays alpha with the words
// processed so far (word_count_actual, cpu:431-445), not with corpus position.
// MF: matrix-core repulsion (bf16 tables); WIN / WOUT: write-back policy of
// W_in / W_out on the matrix-core instance (kWbAtomic, kWbStore, kWbHot).
template <typename T, int NP, int NB, bool MF, int WIN, int WOUT>
__device__ __forceinline__ void train_sentence(const FastArgs &a, int64_t s, int64_t raw, int lane, char *lds, double &loss_a,
                               double &loss_r, SentCount &cnt) {
    using PT = typename Pair<T>::P;
    T *w_in = static_cast<T *>(a.w_in);
    T *w_out = static_cast<T *>(a.w_out);
    const int ld = a.ld, nq = a.nq, W = a.W, R = a.R;
    // matrix-core instances address the tables through buffer resources
    // (the host enables them only below 4 GiB - 4 KiB)
    const uint32_t row_bytes = (uint32_t)ld * (uint32_t)sizeof(T);
    const uint32_t tab_bytes = (uint32_t)a.V * row_bytes;
    const __amdgpu_buffer_rsrc_t rin_w = __builtin_amdgcn_make_buffer_rsrc(a.w_in, (short)0, (int)tab_bytes, kRsrcWord3);
    const __amdgpu_buffer_rsrc_t rout_w = __builtin_amdgcn_make_buffer_rsrc(a.w_out, (short)0, (int)tab_bytes, kRsrcWord3);
    // the host picks NP = ceil(nq / 64): only a lane's last pair can be past the row
    __builtin_assume(nq > kWave * (NP - 1) && nq <= kWave * NP);
    const int64_t o0 = load_meta<NP, MF>(a.sent_off + s);
    const int L = (int)(load_meta<NP, MF>(a.sent_off + s + 1) - o0);
    if (L <= 0) return;  // an empty sentence trains nothing (cpu:539-542, 575)
    const int32_t *sen = a.tokens + o0;
    const bool planned = a.plan_b != nullptr;
    const float alpha =
        planned ? load_meta<NP, MF>(a.plan_alpha + 2 * s + 1) : sentence_alpha(a.a0, a.epoch, a.T, a.iter, raw);
    const float alpha0 = planned ? load_meta<NP, MF>(a.plan_alpha + 2 * s) : alpha;  // centre 0 (cpu:431-447 timing)
    const uint32_t ep = (uint32_t)a.epoch;
    // start of this lane's rounding-noise stream for this sentence, hashed
    // from the seed, the sentence, the epoch and the lane
    uint32_t rs = hash32((uint32_t)a.seed ^ hash32((uint32_t)s * 0x632BE5ABu ^ (ep << 24)) ^ ((uint32_t)lane << 8)) | 1u;
    // LDS of this wave: ring values (fp32), ring snapshots (storage type),
    // shared negatives
    float2 *ring_u = reinterpret_cast<float2 *>(lds);
    PT *ring_0 = reinterpret_cast<PT *>(lds + (size_t)R * nq * sizeof(float2));
    int32_t *negs = reinterpret_cast<int32_t *>(lds + a.lds_negs);
    // Slot tables of the ring, one entry per lane (R = 2W+1 <= 64): lane k
    // holds the word of slot k, the last window position referencing slot k,
    // and the slot of the positions c with (c & pmask) == k.
    const int pmask = pow2_ceil(R) - 1;
    int slot_word = -1, slot_last = kUnusedSlot, pos_slot = 0;
    // byte stride of a ring slot, in a vector register: the pair loop's
    // address is then one multiply-add with no scalar reload of nq
    // (nq is moved into a vector register and the stride derived from it, so
    // the stride's low three bits stay known: the ring addresses are then
    // 8-byte aligned 32-bit LDS addresses, one 24-bit multiply-add each)
    uint32_t nq_v = (uint32_t)nq;
    asm volatile("" : "+v"(nq_v));
    const uint32_t slot_bytes = nq_v * (uint32_t)sizeof(float2);
    char *ring_lane = reinterpret_cast<char *>(ring_u + lane);
    // A lane whose last pair lies past the row (q >= nq) points that pair at a
    // zero pair of its own in the slack after the ring, for every slot: its
    // centre-row value is zero too, so it reads 0 and writes 0 there, and the
    // pair loop needs no lane predicate or mask.
    const bool last_in = lane + kWave * (NP - 1) < nq;
    float2 *zero_pair = reinterpret_cast<float2 *>(lds + (size_t)R * nq * (sizeof(float2) + sizeof(PT))) + lane;
    const uint32_t slot_bytes_last = (last_in ? nq_v : 0u) * (uint32_t)sizeof(float2);
    char *ring_last = last_in ? ring_lane + kWave * sizeof(float2) * (NP - 1) : reinterpret_cast<char *>(zero_pair);
    typedef __attribute__((address_space(3))) char lchar;
    typedef float __attribute__((ext_vector_type(2))) vf2;
    typedef __attribute__((address_space(3))) vf2 lfloat2;
    lchar *ring_lane_l = (lchar *)ring_lane;
    lchar *ring_last_l = (lchar *)ring_last;

    cnt.tokens = (uint32_t)L;
    // the sentence's first 128 tokens, lane k holding tokens k and 64 + k (-1
    // past the end), kept for the repulsion: a sentence of at most 128 tokens
    // is then checked against its negatives and gives its batch rows without
    // reloading them (a per-lane load retires after every atomic before it)
    const bool short_sen = L <= 2 * kWave;
    int tok_lo = -1, tok_hi = -1;

    // Slot of word w among slots still referenced at or after position `from`
    // (-1 if none), first match wins.
    auto find_slot = [&](int w, int from) -> int {
        const uint64_t mask = __builtin_amdgcn_ballot_w64(lane < R && slot_last >= from && slot_word == w);
        return mask ? (int)__builtin_ctzll(mask) : -1;
    };
    // A slot referenced by no position at or after `from` (the window
    // references at most R slots, so one exists).
    auto find_free = [&](int from) -> int {
        const uint64_t mask = __builtin_amdgcn_ballot_w64(lane < R && slot_last < from);
        return (int)__builtin_ctzll(mask);
    };
    auto slot_of = [&](int c) -> int { return __builtin_amdgcn_readlane(pos_slot, c & pmask); };
    auto last_of = [&](int slot) -> int { return __builtin_amdgcn_readlane(slot_last, slot); };
    auto word_of = [&](int slot) -> int { return __builtin_amdgcn_readlane(slot_word, slot); };
    auto fill = [&](int slot, const PT (&row)[NP]) {
#pragma unroll
        for (int t = 0; t < NP; t++) {
            const int q = lane + kWave * t;
            if (t < NP - 1 || q < nq) {
                ring_0[slot * nq + q] = row[t];
                ring_u[slot * nq + q] = Pair<T>::tof(row[t]);
            }
        }
    };
    // a slot's current and loaded values, then their write-back to W_in
    auto read_slot = [&](int slot, float2 (&u)[NP], float2 (&u0)[NP]) {
#pragma unroll
        for (int t = 0; t < NP; t++) {
            const int q = lane + kWave * t;
            u[t] = u0[t] = make_float2(0.0f, 0.0f);
            if (t < NP - 1 || q < nq) {
                u[t] = ring_u[slot * nq + q];
                u0[t] = Pair<T>::tof(ring_0[slot * nq + q]);
            }
        }
    };
    auto write_slot = [&](int64_t row, const float2 (&u)[NP], const float2 (&u0)[NP]) {
        if constexpr (MF) {
            store_value_rsrc<NP, WIN>(rin_w, (uint32_t)row, row_bytes, nq, lane, u0, u, rs, nearest_in(a), a.hot_bytes);
        } else {
            float2 du[NP];
#pragma unroll
            for (int t = 0; t < NP; t++) du[t] = make_float2(u[t].x - u0[t].x, u[t].y - u0[t].y);
            store_delta<T, NP>(w_in, row, ld, nq, lane, u0, du, rs, nearest_in(a));
        }
    };
    auto flush = [&](int slot) {
        float2 u[NP], u0[NP];
        read_slot(slot, u, u0);
        write_slot(word_of(slot), u, u0);
    };
    auto bind = [&](int c, int slot, int w) {
        if (lane == slot) {
            slot_word = w;
            slot_last = c;
        }
        if (lane == (c & pmask)) pos_slot = slot;
    };

    // The attraction is a chain of dependent pairs (latency bound); the
    // repulsion has independent matrix, LDS and atomic work.  The attraction
    // runs at wave priority 1, so when both waves of a SIMD can issue, its
    // chain goes first and the other wave's repulsion fills the gaps.
    __builtin_amdgcn_s_setprio(1);
    // ---- attraction
    // The W_in rows of window [p-W, p+W] live in LDS slots keyed by word: a
    // word occurring twice in the window shares one slot, so every pair sees
    // the updates of the pairs before it exactly as cpu:551-565 does in place.
    // A slot is written back when the last position referencing it leaves the
    // window.  At the end of centre p the slot bookkeeping and the fill of the
    // entering row come first (their operands were loaded a centre earlier),
    // then this centre's write-backs, then the loads for two centres ahead:
    // the W_out row of centre p+2 and the W_in row entering at p+2+W.  Loaded
    // after every write-back this wave makes before using them, they hold the
    // values of the reference's sequential order; and since a wave's memory
    // operations retire in issue order, every wait in the loop is for
    // operations issued a full centre earlier.
    {
        *zero_pair = make_float2(0.0f, 0.0f);  // (the repulsion may have used these bytes)
        // the sentence's tokens [tb, tb+128) in two registers, lane k holding
        // tokens tb+k and tb+64+k; every token the centre loop needs lies there
        int tb = 0;
        auto tok_block = [&](int base) -> int { return base + lane < L ? sen[base + lane] : -1; };
        int tw0 = tok_block(0), tw1 = tok_block(kWave);
        tok_lo = tw0;
        tok_hi = tw1;
        auto tok = [&](int x) -> int {
            const int o = x - tb;
            return o < kWave ? __builtin_amdgcn_readlane(tw0, o) : __builtin_amdgcn_readlane(tw1, o - kWave);
        };
        // unused slots (slot_last = kUnusedSlot) sit below every window start
        // (window starts go down to -W)
        const int first = W < L - 1 ? W : L - 1;
        for (int c = 0; c <= first; c++) {  // bind the first window's positions to slots
            const int w = tok(c);
            int sl = find_slot(w, 0);
            if (sl < 0) sl = find_free(0);
            bind(c, sl, w);
        }
        // vq0/vq1: W_out rows of centres p and p+1; pq: the W_in row entering at
        // p+1+W, loaded only if its word then holds no slot (np_slot < 0).  With
        // one pair per lane they go out before the first window's rows, which
        // follow eight at a time, so a window of up to eight positions is set up
        // in one round trip; wider rows keep four rows in flight and load the
        // centre rows after them (at two waves per SIMD the other wave hides
        // those round trips, and eight wide rows cost registers).
        PT vq0[NP], vq1[NP], pq[NP];
        int np_slot = -1;
        if constexpr (NP == 1) {
            load_row<T, NP>(w_out, tok(0), ld, nq, lane, vq0);
            if (1 < L && tok(1) != tok(0)) load_row<T, NP>(w_out, tok(1), ld, nq, lane, vq1);
            np_slot = 1 + W < L ? find_slot(tok(1 + W), -W) : -1;
            if (1 + W < L && np_slot < 0) load_row<T, NP>(w_in, tok(1 + W), ld, nq, lane, pq);
        }
        {  // fetch every bound slot
            constexpr int kFetch = NP == 1 ? 8 : 4;
            uint64_t mask = __builtin_amdgcn_ballot_w64(lane < R && slot_last >= 0);
            while (mask) {
                int slk[kFetch];
#pragma unroll
                for (int k = 0; k < kFetch; k++) {
                    slk[k] = mask ? (int)__builtin_ctzll(mask) : -1;
                    mask &= mask ? mask - 1 : 0;
                }
                PT rows[kFetch][NP];
#pragma unroll
                for (int k = 0; k < kFetch; k++)
                    if (slk[k] >= 0) load_row<T, NP>(w_in, word_of(slk[k]), ld, nq, lane, rows[k]);
#pragma unroll
                for (int k = 0; k < kFetch; k++)
                    if (slk[k] >= 0) fill(slk[k], rows[k]);
            }
        }
        if constexpr (NP != 1) {
            load_row<T, NP>(w_out, tok(0), ld, nq, lane, vq0);
            if (1 < L && tok(1) != tok(0)) load_row<T, NP>(w_out, tok(1), ld, nq, lane, vq1);
            np_slot = 1 + W < L ? find_slot(tok(1 + W), -W) : -1;
            if (1 + W < L && np_slot < 0) load_row<T, NP>(w_in, tok(1 + W), ld, nq, lane, pq);
        }
        // centre 0's row, settled here: every iteration then enters the loop
        // with the same loads in flight (the next centres' rows only)
#pragma unroll
        for (int t = 0; t < NP; t++) vq0[t] = Pair<T>::copy(vq0[t]);
        float2 v[NP], v0[NP];
        bool carried = false;
        uint32_t bwin = 0;
        for (int p = 0; p < L; p++) {
            if ((p & (kWave - 1)) == 0) {
                bwin = planned ? (p + lane < L ? a.plan_b[o0 + p + lane] : 0u)
                               : draw(a.seed, s, ep, kTagWindow, (uint32_t)(p + lane)).x % (uint32_t)W;
                if (p > 0) {  // slide the token window: [p, p+128)
                    tb = p;
                    tw0 = tw1;
                    tw1 = (int)copy_u32((uint32_t)tok_block(p + kWave));
                }
                bwin = copy_u32(bwin);
            }
            const int b = __builtin_amdgcn_readlane((int)bwin, p & (kWave - 1));
            const float alpha_p = p == 0 ? alpha0 : alpha;
            const int wp = tok(p);
            if (!carried) {  // a centre repeated at p-1 keeps its register row
                to_f<T, NP>(vq0, v);
#pragma unroll
                for (int t = 0; t < NP; t++) v0[t] = v[t];
            }
            const int wn = p + 1 < L ? tok(p + 1) : -1;
            const bool carry_next = wn == wp;
            const int np = p + 1 + W;
            const int wnp = np < L ? tok(np) : -1;
            // Pairs of centre p: positions lo..hi except p (cpu:547-550), in
            // order.  Pair k's logit f_k = u_k . v_k, with u_k its ring row and
            // v_k the centre row after k pairs, is assembled as
            //   f_k = u_k . v_{k-1} + g_{k-1} (u_k . u_{k-1})
            // (v_k = v_{k-1} + g_{k-1} u_{k-1}), so both products and their
            // reductions run while g_{k-1} is still being computed; only a
            // pair reading the slot the previous pair just wrote (the same
            // word at adjacent positions) takes its logit after that write.
            const int lo = p - W + b > 0 ? p - W + b : 0;
            const int hi = p + W - b < L - 1 ? p + W - b : L - 1;
            const int npair = hi - lo;
            float fbuf = 0.0f;  // lane k: logit of the k-th pair (loss only)
            if (npair > 0) {
                auto pos = [&](int k) { return lo + k + (lo + k >= p ? 1 : 0); };
                // ring entry t of this lane in a slot (the last one: see zero_pair)
                auto ring_at = [&](int slot, int t) {
                    return t == NP - 1 ? (lfloat2 *)(ring_last_l + __umul24((uint32_t)slot, slot_bytes_last))
                                       : (lfloat2 *)(ring_lane_l + __umul24((uint32_t)slot, slot_bytes)) + kWave * t;
                };
                auto read_row = [&](int slot, float2 (&u)[NP]) {
#pragma unroll
                    for (int t = 0; t < NP; t++) {
                        const vf2 x = *ring_at(slot, t);
                        u[t] = make_float2(x.x, x.y);
                    }
                };
                auto dot = [&](const float2 (&x)[NP], const float2 (&y)[NP]) {  // one FMA chain
                    float s = x[0].x * y[0].x;
                    s = __builtin_fmaf(x[0].y, y[0].y, s);
#pragma unroll
                    for (int t = 1; t < NP; t++) {
                        s = __builtin_fmaf(x[t].x, y[t].x, s);
                        s = __builtin_fmaf(x[t].y, y[t].y, s);
                    }
                    return s;
                };
                // Pair k: on entry v = v_k, cur = u_k and f = f_k.  The next
                // pair's row goes into nxt; the two row buffers alternate over
                // an unrolled step pair, so no row is ever copied.
                float2 uA[NP], uB[NP];
                int slot = slot_of(pos(0));
                read_row(slot, uA);
                float f = wave_sum(dot(uA, v));
                auto step = [&](int k, float2 (&cur)[NP], float2 (&nxt)[NP], auto more, auto loss) {
                    int slot_n = slot;
                    float An = 0.0f, Bn = 0.0f;
                    // the next row's LDS read goes out first and the sigmoid
                    // of f_k covers its latency; the scheduling barriers
                    // keep that order (left free, the compiler waits for the
                    // read and then interleaves the sigmoid with the sums)
                    if constexpr (decltype(more)::value) {
                        slot_n = slot_of(pos(k + 1));
                        read_row(slot_n, nxt);
                    }
                    __builtin_amdgcn_sched_barrier(0);
                    const float g = grad_label1_r(f, alpha_p);
                    __builtin_amdgcn_sched_barrier(0);
                    if constexpr (decltype(more)::value)  // u_{k+1}.v_k and u_{k+1}.u_k
                        wave_sum2(dot(nxt, v), dot(nxt, cur), An, Bn);
                    float2 nw[NP];
#pragma unroll
                    for (int t = 0; t < NP; t++) {
                        nw[t] = make_float2(cur[t].x + g * v[t].x, cur[t].y + g * v[t].y);
                        *ring_at(slot, t) = vf2{nw[t].x, nw[t].y};
                    }
#pragma unroll
                    for (int t = 0; t < NP; t++) {  // v_{k+1}
                        v[t].x += g * cur[t].x;
                        v[t].y += g * cur[t].y;
                    }
                    if constexpr (decltype(loss)::value) fbuf = lane == k ? f : fbuf;  // pair k's logit in lane k
                    if constexpr (decltype(more)::value) {
                        if (slot_n == slot) {  // pair k+1 reads the row pair k just wrote
#pragma unroll
                            for (int t = 0; t < NP; t++) nxt[t] = nw[t];
                            f = wave_sum(dot(nxt, v));
                        } else {
                            f = An + g * Bn;
                        }
                        slot = slot_n;
                    }
                };
                auto pairs = [&](auto loss) {
                    int k = 0;
                    for (; k + 2 < npair; k += 2) {
                        step(k, uA, uB, std::true_type{}, loss);
                        step(k + 1, uB, uA, std::true_type{}, loss);
                    }
                    if (k + 1 < npair) {
                        step(k, uA, uB, std::true_type{}, loss);
                        step(k + 1, uB, uA, std::false_type{}, loss);
                    } else {
                        step(k, uA, uB, std::false_type{}, loss);
                    }
                };
                if (a.compute_loss) pairs(std::true_type{});
                else pairs(std::false_type{});
            }
            cnt.pairs += (uint32_t)npair;
            if (a.compute_loss && lane < npair) loss_a += (double)softplusf(-fbuf);  // this lane's share
            // position p-W leaves the window; position np enters it (a word
            // still in the window keeps its slot, so bind before the flush test)
            const int r = p - W;
            const int r_slot = r >= 0 ? slot_of(r) : -1;
            if (np < L && np_slot >= 0) bind(np, np_slot, wnp);
            const bool leave = r_slot >= 0 && last_of(r_slot) == r;
            float2 fu[NP], fu0[NP];
            int64_t frow = 0;
            if (leave) {  // read before the fill, which may reuse the slot
                frow = word_of(r_slot);
                read_slot(r_slot, fu, fu0);
            }
            if (np < L && np_slot < 0) {
                const int sl = find_free(p + 1 - W);
                fill(sl, pq);
                bind(np, sl, wnp);
            }
#pragma unroll
            for (int t = 0; t < NP; t++) vq0[t] = Pair<T>::copy(vq1[t]);
            // this centre's write-backs
            if (!carry_next) {
                if constexpr (MF) {
                    store_value_rsrc<NP, WOUT>(rout_w, (uint32_t)wp, row_bytes, nq, lane, v0, v, rs, nearest_out(a),
                                               a.hot_bytes);
                } else {
                    float2 dv[NP];
#pragma unroll
                    for (int t = 0; t < NP; t++) dv[t] = make_float2(v[t].x - v0[t].x, v[t].y - v0[t].y);
                    store_delta<T, NP, WOUT == kWbStore>(w_out, wp, ld, nq, lane, v0, dv, rs, nearest_out(a));
                }
            }
            if (leave) write_slot(frow, fu, fu0);
            carried = carry_next;
            // loads for two centres ahead: the W_out row of centre p+2 (unless
            // centre p+1 carries its row into it) and the W_in row entering at
            // p+2+W, unless its word holds a slot at the start of centre p+1
            const int w2 = p + 2 < L ? tok(p + 2) : -1;
            if (w2 >= 0 && w2 != wn) load_row<T, NP>(w_out, w2, ld, nq, lane, vq1);
            const int e2 = np + 1 < L ? tok(np + 1) : -1;
            np_slot = e2 >= 0 ? find_slot(e2, p + 1 - W) : -1;
            if (e2 >= 0 && np_slot < 0) load_row<T, NP>(w_in, e2, ld, nq, lane, pq);
        }
        {  // slots still referenced by the last window
            uint64_t mask = __builtin_amdgcn_ballot_w64(lane < R && slot_last >= L - W);
            while (mask) {
                const int sl = (int)__builtin_ctzll(mask);
                mask &= mask - 1;
                flush(sl);
            }
        }
    }

    __builtin_amdgcn_s_setprio(0);
    // ---- repulsion (cpu:571-679)
    // Per batch of B positions: K shared negatives; logits G = N C^T against
    // the batch rows C, label-0 gradient, then W_out[n_k] += sum_i G_ki c_i and
    // W_in[c_i] += sum_k G_ki n_k, as the reference's three cblas_sgemm calls
    // (cpu:633-656).  Repeated negatives are merged with a multiplicity so
    // every word's W_out row is read and written by exactly one register block
    // of NB negatives.  Vector path: each block re-reads the batch rows C
    // (cache-resident) and adds its share of W_in[c_i]'s delta to an fp32
    // accumulator in the (now idle) ring memory; the rows are written once per
    // batch, after the last block (the reference's one scatter per batch,
    // cpu:658-674).  With no concurrent writer -- one wave -- every block
    // reads the same C, the reference's snapshot.
    const int B = a.B, NEG = a.NEG;
    const int nbat = (L + B - 1) / B;
    const float alpha_r = L == 1 ? alpha0 : alpha;
    int32_t *uneg = negs + a.kmax;  // unique negatives of the batch
    int32_t *umul = uneg + a.kmax;  // their multiplicities
    // A sentence longer than its token registers is checked against its
    // negatives through a hash set of its tokens, built here once: a probe or
    // two per candidate instead of a pass over every token of the sentence
    // per check.  The set holds exactly the sentence's words, so the check
    // rejects what the pass does.
    // (one value stays live over the batches: log2 of the slots, 0 when the
    // sentence takes the token pass instead)
    const int tab_lg = !short_sen && !planned && a.tab_log2 > 0 && 4 * L <= 3 << a.tab_log2 ? a.tab_log2 : 0;
    if (tab_lg) {
        int32_t *tab = reinterpret_cast<int32_t *>(lds + a.tab_off);
        const uint32_t tab_mask = (1u << tab_lg) - 1u;
        int4 *t4 = reinterpret_cast<int4 *>(tab);
        for (int i = lane; i < (1 << tab_lg) / 4; i += kWave) t4[i] = make_int4(kTabEmpty, kTabEmpty, kTabEmpty, kTabEmpty);
        __builtin_amdgcn_wave_barrier();
        for (int t0 = 0; t0 < L; t0 += kWave) {
            const int w = t0 + lane < L ? sen[t0 + lane] : kTabEmpty;
            uint32_t h = tab_hash(w, tab_lg);
            bool pend = w != kTabEmpty;
            while (__builtin_amdgcn_ballot_w64(pend)) {  // linear probing; a word already there is done
                if (pend) {
                    const int old = atomicCAS(tab + h, kTabEmpty, w);
                    if (old == kTabEmpty || old == w) pend = false;
                    else h = (h + 1) & tab_mask;
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
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
        cnt.batches++;
        cnt.negatives += (uint32_t)K;
        cnt.rep_pairs += (uint64_t)K * (uint64_t)isz;
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
            int cand = alias_draw(a.alias, a.V, draw(a.seed, s, ep, kTagNeg, idx));
            // Redraw uniformly over [1, V-1] while the candidate occurs in the
            // sentence.  The reference loops without bound (cpu:611-615); the
            // bound here only matters when the sentence holds nearly the whole
            // vocabulary, and then the last candidate is kept.
            for (int attempt = 1; attempt <= kMaxRedraw; attempt++) {
                bool hit = false;
                if (tab_lg) {
                    const int32_t *tab = reinterpret_cast<const int32_t *>(lds + a.tab_off);
                    const uint32_t tab_mask = (1u << tab_lg) - 1u;
                    uint32_t h = tab_hash(cand, tab_lg);
                    bool pend = live;
                    while (__builtin_amdgcn_ballot_w64(pend)) {
                        if (pend) {
                            const int v = tab[h];
                            if (v == cand) hit = true;
                            if (v == cand || v == kTabEmpty) pend = false;
                            else h = (h + 1) & tab_mask;
                        }
                    }
                } else for (int t0 = 0; t0 < L; t0 += kWave) {
                    const int tk = short_sen ? (t0 == 0 ? tok_lo : tok_hi) : (t0 + lane < L ? sen[t0 + lane] : -1);
                    const int n = L - t0 < kWave ? L - t0 : kWave;
                    // runs of 16 constant lanes with one uniform test per run: a
                    // loop per lane costs a branch and two scalar-vector round
                    // trips per token; lanes past the sentence hold -1, which
                    // matches no candidate
#pragma unroll
                    for (int j0 = 0; j0 < kWave; j0 += 16) {
                        if (j0 < n) {
#pragma unroll
                            for (int j = 0; j < 16; j++) hit |= (cand == __builtin_amdgcn_readlane(tk, j0 + j));
                        }
                    }
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
        if constexpr (MF) {  // matrix-core path (bf16 tables, its own kernel instance)
            int bat_tok = 0;  // lane i: token lo + i (sentences of at most 128 tokens)
            if (short_sen) {
                const int x = lo + lane;
                const int t_lo = __shfl(tok_lo, x & (kWave - 1), kWave), t_hi = __shfl(tok_hi, x & (kWave - 1), kWave);
                bat_tok = x < kWave ? t_lo : t_hi;
            }
            repulsion_mfma<NP, WIN, WOUT>(a, sen + lo, bat_tok, short_sen, isz, U, uneg, umul, lds, lane, alpha_r,
                                          loss_r, rs);
            continue;
        }
        // the batch's row ids, lane i holding row i (isz <= B <= 64): from the
        // registers that hold a short sentence's tokens, else one load per
        // batch -- a per-row load would put a dependent round trip, behind the
        // wave's last atomics, before every row
        int bat_id;
        if (short_sen) {
            const int x = lo + lane;
            const int t_lo = __shfl(tok_lo, x & (kWave - 1), kWave), t_hi = __shfl(tok_hi, x & (kWave - 1), kWave);
            bat_id = x < kWave ? t_lo : t_hi;
        } else {
            bat_id = lane < isz ? sen[lo + lane] : 0;
        }
        auto row_id = [&](int i) { return __builtin_amdgcn_readlane(bat_id, i); };
        float2 *acc = reinterpret_cast<float2 *>(lds);  // [isz][nq] fp32 W_in deltas
        for (int i = 0; i < isz; i++) {
#pragma unroll
            for (int t = 0; t < NP; t++) {
                const int q = lane + kWave * t;
                if (t < NP - 1 || q < nq) acc[i * nq + q] = make_float2(0.0f, 0.0f);
            }
        }
        // W_out rows of a register block; the next block's rows are loaded
        // before this block's write-back atomics (vector memory operations
        // retire in issue order, so waiting for them then skips the atomics)
        PT pre[NB][NP];
        auto load_block = [&](int kb) {
#pragma unroll
            for (int k = 0; k < NB; k++) {
                if (kb + k < U) load_row<T, NP>(w_out, uneg[kb + k], ld, nq, lane, pre[k]);
                else {
#pragma unroll
                    for (int t = 0; t < NP; t++) pre[k][t] = PT{};
                }
            }
        };
        load_block(0);
        for (int kb = 0; kb < U; kb += NB) {
            const int nbk = U - kb < NB ? U - kb : NB;
            int32_t nrow[NB];
            float mult[NB];
            float2 nv[NB][NP], dn[NB][NP];
#pragma unroll
            for (int k = 0; k < NB; k++) {
                nrow[k] = k < nbk ? uneg[kb + k] : -1;
                mult[k] = k < nbk ? (float)umul[kb + k] : 0.0f;
                to_f<T, NP>(pre[k], nv[k]);
#pragma unroll
                for (int t = 0; t < NP; t++) dn[k][t] = make_float2(0.0f, 0.0f);
            }
            PT ra[NP], rb[NP];
            load_row<T, NP>(w_in, row_id(0), ld, nq, lane, ra);
            if (kb + NB < U) load_block(kb + NB);
            const float mlane = lane < nbk ? (float)umul[kb + lane] : 0.0f;  // loss: lane k <-> negative k
            for (int i = 0; i < isz; i++) {
                float2 c[NP];
                if ((i & 1) == 0) {
                    if (i + 1 < isz) load_row<T, NP>(w_in, row_id(i + 1), ld, nq, lane, rb);
                    to_f<T, NP>(ra, c);
                } else {
                    if (i + 1 < isz) load_row<T, NP>(w_in, row_id(i + 1), ld, nq, lane, ra);
                    to_f<T, NP>(rb, c);
                }
                float g[NB], part[NB], fk[NB];
                float fbuf = 0.0f;  // lane k: logit against negative k (loss only)
#pragma unroll
                for (int k = 0; k < NB; k++) {
                    part[k] = 0.0f;
#pragma unroll
                    for (int t = 0; t < NP; t++) part[k] += nv[k][t].x * c[t].x + nv[k][t].y * c[t].y;
                }
                wave_sum_n<NB>(part, fk);  // the block's logits in one reduction tree
#pragma unroll
                for (int k = 0; k < NB; k++) {
                    g[k] = k < nbk ? grad_label0_r(fk[k], alpha_r) : 0.0f;
                    fbuf = lane == k ? fk[k] : fbuf;
                }
                if (a.compute_loss && lane < nbk) loss_r += (double)mlane * (double)softplusf(fbuf);
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
#pragma unroll
                for (int t = 0; t < NP; t++) {
                    const int q = lane + kWave * t;
                    if (t < NP - 1 || q < nq) {
                        float2 v = acc[i * nq + q];
                        v.x += dc[t].x;
                        v.y += dc[t].y;
                        acc[i * nq + q] = v;
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < NB; k++) {
                if (k < nbk) {
#pragma unroll
                    for (int t = 0; t < NP; t++) {
                        dn[k][t].x *= mult[k];
                        dn[k][t].y *= mult[k];
                    }
                    store_delta<T, NP, WOUT == kWbStore>(w_out, nrow[k], ld, nq, lane, nv[k], dn[k], rs, nearest_out(a));
                }
            }
        }
        for (int i = 0; i < isz; i++) {
            float2 c[NP], dc[NP];
            if constexpr (sizeof(T) == 2) {
                PT r[NP];
                load_row<T, NP>(w_in, row_id(i), ld, nq, lane, r);
                to_f<T, NP>(r, c);
            } else {
#pragma unroll
                for (int t = 0; t < NP; t++) c[t] = make_float2(0.0f, 0.0f);
            }
#pragma unroll
            for (int t = 0; t < NP; t++) {
                const int q = lane + kWave * t;
                dc[t] = t < NP - 1 || q < nq ? acc[i * nq + q] : make_float2(0.0f, 0.0f);
            }
            store_delta<T, NP>(w_in, row_id(i), ld, nq, lane, c, dc, rs, nearest_in(a));
        }
    }
}

// At most 256 VGPRs + AGPRs, so two waves fit on a SIMD: the kernel is bound
// by the latency of dependent row loads and needs resident waves to hide it.
template <typename T, int NP, int NB, bool MF, int WIN, int WOUT>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2))) void sgns_fast_kernel(FastArgs a) {
    static_assert(!MF || sizeof(T) == 2, "the matrix-core repulsion runs on bf16 tables");
    extern __shared__ __align__(16) char smem[];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    char *lds = smem + (size_t)wave * a.lds_per_wave;
    double loss_a = 0.0, loss_r = 0.0;
    uint64_t total = 0;  // lane k: counter k of this wave (Counter order)
    auto fold = [&](const SentCount &c) {
        uint64_t v = 0;
        v = lane == kCntTokens ? (uint64_t)c.tokens : v;
        v = lane == kCntSentences ? (uint64_t)(c.tokens > 0) : v;
        v = lane == kCntPairs ? (uint64_t)c.pairs : v;
        v = lane == kCntBatches ? (uint64_t)c.batches : v;
        v = lane == kCntNegatives ? (uint64_t)c.negatives : v;
        v = lane == kCntRepPairs ? c.rep_pairs : v;
        total += v;
    };
    const int64_t n = a.s_end - a.s_begin;
    if (a.persistent) {
        // each dequeue takes `grab` consecutive queue slots (one atomic per
        // grab sentences); every wave leaves once the queue is past n
        const int grab = a.grab;
        for (;;) {
            unsigned long long qi = 0;
            if (lane == 0) qi = atomicAdd(a.queue, (unsigned long long)grab);
            const int64_t idx0 = readfirst_i64((int64_t)qi);
            if (idx0 >= n) break;
            const int64_t idx1 = idx0 + grab < n ? idx0 + grab : n;
            for (int64_t idx = idx0; idx < idx1; idx++) {
                const int64_t slot = a.s_begin + idx;
                const int64_t s = a.order ? (int64_t)load_meta<NP, MF>(a.order + slot) : slot;
                const int64_t raw =
                    a.plan_b ? 0 : (a.order ? load_meta<NP, MF>(a.queue_raw + slot) : load_meta<NP, MF>(a.sent_raw + s));
                SentCount c{0u, 0u, 0u, 0u, 0ull};
                train_sentence<T, NP, NB, MF, WIN, WOUT>(a, s, raw, lane, lds, loss_a, loss_r, c);
                fold(c);
            }
        }
    } else {
        const int64_t idx = (int64_t)blockIdx.x * (blockDim.x / kWave) + wave;
        const int64_t s = a.s_begin + idx;
        if (idx < n) {
            SentCount c{0u, 0u, 0u, 0u, 0ull};
            train_sentence<T, NP, NB, MF, WIN, WOUT>(a, s, a.plan_b ? 0 : a.sent_raw[s], lane, lds, loss_a, loss_r, c);
            fold(c);
        }
    }
    if (a.compute_loss) {  // every lane holds a share of the wave's losses
#pragma unroll
        for (int m = 1; m < kWave; m <<= 1) {
            loss_a += __shfl_xor(loss_a, m, kWave);
            loss_r += __shfl_xor(loss_r, m, kWave);
        }
        if (lane == 0) {
            atomicAdd(&a.loss[0], loss_a);
            atomicAdd(&a.loss[1], loss_r);
        }
    }
    if (lane <= kCntRepPairs && total) atomicAdd(&a.counters[lane], (unsigned long long)total);
}

template <typename T, int NP, bool MF, int WIN = kWbAtomic, int WOUT = kWbAtomic>
hipError_t launch_typed(const FastArgs &a, int wpb, int blocks, hipStream_t st) {
    // negatives per register block: as many as fit without spilling at each width
    constexpr int NB = NP <= 2 ? 8 : (NP == 3 ? (sizeof(T) == 2 ? 6 : 5) : 2);
    const int64_t n = a.s_end - a.s_begin;
    if (n <= 0) return hipSuccess;
    const size_t lds = (size_t)a.lds_per_wave * (size_t)wpb;
    auto kern = sgns_fast_kernel<T, NP, NB, MF, WIN, WOUT>;
    hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    dim3 grid;
    if (a.persistent) grid = dim3((unsigned)blocks);
    else grid = dim3((unsigned)((n + wpb - 1) / wpb));
    hipLaunchKernelGGL(kern, grid, dim3(kWave * wpb), lds, st, a);
    return hipGetLastError();
}

// Vector-repulsion instances with one W_out write-back policy (W_in: atomics).
template <typename T, int WOUT>
hipError_t launch_vector(const FastArgs &a, int np, int wpb, int blocks, hipStream_t st) {
    switch (np) {
    case 1: return launch_typed<T, 1, false, kWbAtomic, WOUT>(a, wpb, blocks, st);
    case 2: return launch_typed<T, 2, false, kWbAtomic, WOUT>(a, wpb, blocks, st);
    case 3: return launch_typed<T, 3, false, kWbAtomic, WOUT>(a, wpb, blocks, st);
    case 4: return launch_typed<T, 4, false, kWbAtomic, WOUT>(a, wpb, blocks, st);
    }
    return hipErrorInvalidValue;
}

// Matrix-core instances for one write-back policy pair.
template <int WIN, int WOUT>
hipError_t launch_mfma(const FastArgs &a, int np, int wpb, int blocks, hipStream_t st) {
    switch (np) {
    case 1: return launch_typed<uint16_t, 1, true, WIN, WOUT>(a, wpb, blocks, st);
    case 2: return launch_typed<uint16_t, 2, true, WIN, WOUT>(a, wpb, blocks, st);
    case 3: return launch_typed<uint16_t, 3, true, WIN, WOUT>(a, wpb, blocks, st);
    case 4: return launch_typed<uint16_t, 4, true, WIN, WOUT>(a, wpb, blocks, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace fast
}  // namespace sgns
