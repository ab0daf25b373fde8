// sgns_device.h -- gfx950 device helpers shared by the PAR-SGNS kernels.
//
// Row layout on the device: W_in / W_out are [V][ld] in float or bf16, ld a
// multiple of 32 (fp32) or 64 (bf16) so every row starts on a 128-byte line.
// A wavefront covers a row in element PAIRS: lane l owns pairs l, l+64, ...
// (NP pairs per lane), so each load/store instruction of the wave touches
// 64 consecutive pairs (512 B fp32, 256 B bf16) and LDS copies are
// bank-conflict free.  Pairs at or beyond nq = ceil(d/2) are idle; padding
// elements are zero and stay zero under every update (g*0 = 0).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sgns {

constexpr int kWave = 64;
constexpr int kExpTable = 1000;
constexpr int kMaxExp = 6;
constexpr int kMaxSentence = 1000;

// ------------------------------------------------------------- reductions
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}

__device__ __forceinline__ float readlane_f(float v, int lane) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}

// Sum over the 64 lanes, result uniform.  Butterfly inside each row of 16
// lanes (quad_perm xor1, xor2, row_half_mirror, row_mirror), then gfx9's row
// broadcasts: row_bcast:15 adds lane 15 of rows 0/2 into rows 1/3, row_bcast:31
// adds lane 31 into rows 2/3, and lane 63 holds (r3 + r2) + (r1 + r0).  The
// two broadcast steps are single DPP adds whose masked-out rows are left
// unwritten (the compiler emits a DPP move, a move of the old value and an
// add for each); the s_nop gives the two wait states a DPP read needs after
// a VALU write of its source.
__device__ __forceinline__ float add_row_bcast15(float v) {
    asm("s_nop 1\n\tv_add_f32_dpp %0, %0, %0 row_bcast:15 row_mask:0xa bank_mask:0xf" : "+v"(v));
    return v;
}
__device__ __forceinline__ float add_row_bcast31(float v) {
    asm("s_nop 1\n\tv_add_f32_dpp %0, %0, %0 row_bcast:31 row_mask:0xc bank_mask:0xf" : "+v"(v));
    return v;
}
__device__ __forceinline__ float wave_sum(float v) {
    v += dpp<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp<0x141>(v);  // row_half_mirror
    v += dpp<0x140>(v);  // row_mirror
    v = add_row_bcast15(v);  // rows 1 and 3
    v = add_row_bcast31(v);  // rows 2 and 3
    return readlane_f(v, 63);
}
// Two sums over the 64 lanes in one tree.  v_permlane32_swap trades the upper
// half of a with the lower half of b, so after one add lanes 0-31 hold a's
// half-sums and lanes 32-63 b's; the same row butterfly then reduces both,
// and one row_bcast:15 step leaves a's total in row 1 and b's in row 3.
// (The swap's two results are read into scalars first: clang's bit_cast of an
// ext-vector element reads element 0 whatever the index.)
__device__ __forceinline__ void wave_sum2(float a, float b, float &sa, float &sb) {
    const auto r = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(uint32_t, a), __builtin_bit_cast(uint32_t, b),
                                                    false, false);
    const uint32_t lo = r[0], hi = r[1];
    float v = __builtin_bit_cast(float, lo) + __builtin_bit_cast(float, hi);
    v += dpp<0xB1>(v);
    v += dpp<0x4E>(v);
    v += dpp<0x141>(v);
    v += dpp<0x140>(v);
    v = add_row_bcast15(v);
    sa = readlane_f(v, 31);
    sb = readlane_f(v, 63);
}

// N sums over the 64 lanes in one tree (N <= 8).  permlane32_swap pairs the
// values (each register then holds two values' half-sums, one per 32-lane
// half), permlane16_swap pairs those registers (each 16-lane row then holds
// one value's quarter-sums), and the row butterfly finishes every row; the
// rows of register j hold values 4j, 4j+2, 4j+1, 4j+3.  A value without a
// partner is paired with zeros.
__device__ __forceinline__ float2 swap32(float a, float b) {
    const auto r = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(uint32_t, a), __builtin_bit_cast(uint32_t, b),
                                                    false, false);
    const uint32_t x = r[0], y = r[1];
    return make_float2(__builtin_bit_cast(float, x), __builtin_bit_cast(float, y));
}
__device__ __forceinline__ float2 swap16(float a, float b) {
    const auto r = __builtin_amdgcn_permlane16_swap(__builtin_bit_cast(uint32_t, a), __builtin_bit_cast(uint32_t, b),
                                                    false, false);
    const uint32_t x = r[0], y = r[1];
    return make_float2(__builtin_bit_cast(float, x), __builtin_bit_cast(float, y));
}
template <int N>
__device__ __forceinline__ void wave_sum_n(const float (&v)[N], float (&out)[N]) {
    static_assert(N >= 1 && N <= 8, "wave_sum_n: 1..8 values");
    constexpr int M1 = (N + 1) / 2, M2 = (M1 + 1) / 2;
    float h[M1];  // half-sums: value 2j in lanes 0-31, 2j+1 in lanes 32-63
#pragma unroll
    for (int j = 0; j < M1; j++) {
        const float2 p = swap32(v[2 * j], 2 * j + 1 < N ? v[2 * j + 1] : 0.0f);
        h[j] = p.x + p.y;
    }
    float q[M2];  // quarter-sums: rows 0..3 hold values 4j, 4j+2, 4j+1, 4j+3
#pragma unroll
    for (int j = 0; j < M2; j++) {
        const float2 p = swap16(h[2 * j], 2 * j + 1 < M1 ? h[2 * j + 1] : 0.0f);
        float t = p.x + p.y;
        t += dpp<0xB1>(t);
        t += dpp<0x4E>(t);
        t += dpp<0x141>(t);
        t += dpp<0x140>(t);
        q[j] = t;
    }
#pragma unroll
    for (int k = 0; k < N; k++) {
        const int j = k / 4, r = k % 4;
        out[k] = readlane_f(q[j], r == 0 ? 0 : r == 1 ? 32 : r == 2 ? 16 : 48);
    }
}

__device__ __forceinline__ int64_t readfirst_i64(int64_t v) {
    int lo = __builtin_amdgcn_readfirstlane((int)(uint32_t)(uint64_t)v);
    int hi = __builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// --------------------------------------------------------------- Philox
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as
// easy as 1, 2, 3", SC'11): 10 rounds of two 32x32->64 multiplies with the
// published multipliers and Weyl key increments.
struct U4 { uint32_t x, y, z, w; };

__device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint32_t hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
        c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// Counter layout of one random draw: (index, purpose tag, sentence lo,
// sentence hi ^ epoch<<20).  Every draw of the trainer is a pure function of
// the key, the sentence, the epoch and its index, independent of scheduling.
enum : uint32_t { kTagWindow = 1, kTagRepWindow = 2, kTagNeg = 3, kTagRedraw = 4, kTagRound = 5 };

__device__ __forceinline__ U4 draw(uint64_t seed, int64_t sent, uint32_t epoch, uint32_t tag, uint32_t idx) {
    return philox4x32_10(U4{idx, tag, (uint32_t)(uint64_t)sent, (uint32_t)((uint64_t)sent >> 32) ^ (epoch << 20)},
                         (uint32_t)seed, (uint32_t)(seed >> 32));
}

// 32-bit integer hash (lowbias32 by C. Wellons) for rounding noise.
__device__ __forceinline__ uint32_t hash32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

// ---------------------------------------------------------------- bf16
__device__ __forceinline__ float bf16lo(uint32_t p) { return __builtin_bit_cast(float, p << 16); }
__device__ __forceinline__ float bf16hi(uint32_t p) { return __builtin_bit_cast(float, p & 0xFFFF0000u); }

// Stochastic rounding to bf16: add 16 uniform bits below the kept mantissa
// and truncate.  Rounds up with probability equal to the discarded fraction.
__device__ __forceinline__ uint32_t bf16_sr(float x, uint32_t noise16) {
    uint32_t b = __builtin_bit_cast(uint32_t, x);
    return (b + (noise16 & 0xFFFFu)) >> 16;
}
__device__ __forceinline__ uint32_t bf16_rne(float x) {
    uint32_t b = __builtin_bit_cast(uint32_t, x);
    return (b + 0x7FFFu + ((b >> 16) & 1u)) >> 16;
}

// ------------------------------------------------------------- row pairs
// Storage traits: a "pair" is two adjacent elements of a row.
template <typename T> struct Pair;
// A register copy made where it is written.  A loop that rotates prefetch
// buffers (next <- after-next, then a new load into after-next) otherwise
// gets the copy at its latch, after the new load was issued, and the copy
// waits for that load.  Also used to take a value loaded inside a rarely
// taken branch out of the loop's pending loads, so the loop header does not
// wait for every load in flight.
__device__ __forceinline__ uint32_t copy_u32(uint32_t p) {
    uint32_t r;
    asm volatile("v_mov_b32 %0, %1" : "=v"(r) : "v"(p));
    return r;
}
// Table rows are read with agent-coherent loads (sc1 on gfx950): every
// update reaches memory through the L2 (float atomics or coherent stores), so
// an ordinary load could keep returning a row's stale copy from the CU's
// vector L1 for as long as that row stays hot there.
template <> struct Pair<float> {
    using P = float2;
    __device__ static __forceinline__ float2 load(const float *w, int64_t row, int ld, int q) {
        const uint64_t *p = reinterpret_cast<const uint64_t *>(w + row * ld) + q;
        return __builtin_bit_cast(float2, __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    }
    __device__ static __forceinline__ float2 tof(float2 p) { return p; }
    __device__ static __forceinline__ float2 copy(float2 p) {
        float2 r;
        asm volatile("v_mov_b32 %0, %2\n\tv_mov_b32 %1, %3" : "=&v"(r.x), "=&v"(r.y) : "v"(p.x), "v"(p.y));
        return r;
    }
};
template <> struct Pair<uint16_t> {  // bf16 bit patterns
    using P = uint32_t;
    __device__ static __forceinline__ uint32_t load(const uint16_t *w, int64_t row, int ld, int q) {
        const uint32_t *p = reinterpret_cast<const uint32_t *>(w + row * ld) + q;
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ static __forceinline__ float2 tof(uint32_t p) { return make_float2(bf16lo(p), bf16hi(p)); }
    __device__ static __forceinline__ uint32_t copy(uint32_t p) { return copy_u32(p); }
};

// sigmoid-table gradient of a pair, cpu:554-559 (label 1) and cpu:642-647
// (label 0); table index (int)((f + 6) * 83.333336f).
__device__ __forceinline__ float grad_label1(float f, float alpha, const float *et) {
    if (f > kMaxExp) return 0.0f;
    if (f < -kMaxExp) return alpha;
    return (1.0f - et[(int)((f + (float)kMaxExp) * (1000.0f / 12.0f))]) * alpha;
}
__device__ __forceinline__ float grad_label0(float f, float alpha, const float *et) {
    if (f > kMaxExp) return -alpha;
    if (f < -kMaxExp) return 0.0f;
    return (0.0f - et[(int)((f + (float)kMaxExp) * (1000.0f / 12.0f))]) * alpha;
}

// The same rule with the table entry computed in registers instead of read
// from LDS: entry i is e/(e+1) = 1/(1 + exp(-x_i)) at x_i = (i/1000*2-1)*6
// (cpu:822-826).  The label-1 gradient takes 1 - sigma(x_i) = 1/(1 + e^x_i)
// directly (no cancellation near x = 6), each exponential is one exp2 of
// i*(0.012 log2 e) -/+ 6 log2 e, so results can differ from the table's
// bits in the last places.  Used by the Hogwild kernels, where the lookup's
// LDS round trip sits in every pair's dependent chain.  The index needs no
// clamp: outside [-6, 6] the select discards it (and exp2 of a large
// argument is inf or 0, never a trap).
__device__ __forceinline__ float sig_index(float f) {  // (float)(int) of the table index, f in [-6, 6]
    return __builtin_truncf((f + (float)kMaxExp) * (1000.0f / 12.0f));
}
constexpr float kLog2e = 1.44269504088896341f;
constexpr float kGridStep2 = (2.0f * kMaxExp / kExpTable) * kLog2e;  // grid step, in log2 units
constexpr float kGridLo2 = kMaxExp * kLog2e;                          // 6 log2 e
// (selects, no branches: the clipped cases are rare and a branch would end
// the basic block the scheduler interleaves independent work in)
__device__ __forceinline__ float grad_label1_r(float f, float alpha) {
    const float e = __builtin_amdgcn_exp2f(__builtin_fmaf(sig_index(f), kGridStep2, -kGridLo2));  // e^x_i
    const float g = alpha * __builtin_amdgcn_rcpf(1.0f + e);
    return f > kMaxExp ? 0.0f : (f < -kMaxExp ? alpha : g);
}
__device__ __forceinline__ float grad_label0_r(float f, float alpha) {
    const float e = __builtin_amdgcn_exp2f(__builtin_fmaf(sig_index(f), -kGridStep2, kGridLo2));  // e^-x_i
    const float g = -alpha * __builtin_amdgcn_rcpf(1.0f + e);
    return f > kMaxExp ? -alpha : (f < -kMaxExp ? 0.0f : g);
}

__device__ __forceinline__ float softplusf(float x) {  // log(1 + e^x)
    return x > 0.0f ? x + log1pf(__expf(-x)) : log1pf(__expf(x));
}

// Learning rate of a sentence that starts `raw` words into the epoch: the
// schedule of cpu:443-445 evaluated at the sentence start.
__device__ __forceinline__ float sentence_alpha(float a0, int32_t epoch, uint64_t T, int32_t iter, int64_t raw) {
    float wca = (float)((uint64_t)epoch * T + (uint64_t)raw);
    float a = a0 * (1.0f - wca / (float)((uint64_t)iter * T + 1));
    float fl = a0 * 0.0001f;
    return a < fl ? fl : a;
}

}  // namespace sgns
