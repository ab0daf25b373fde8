// window_kernel.hip -- separately callable SGNS loss-forward and adjoint over
// a batch of windows (centre, nctx contexts, nneg shared negatives).
//
// Pair semantics are those of the trainer: contexts are W_in rows paired with
// the centre's W_out row under label 1 (cpu:546-559) and with every shared
// negative's W_out row under label 0 (cpu:633-647); here with the exact
// sigmoid.  Per window the logits are F = X Y^T with X = nctx context rows and
// Y = [centre, negatives]; loss = sum_i softplus(-F_i0) + sum_{i,k>0}
// softplus(F_ik); the adjoint is dF = sigmoid(F) - label, dX = dF Y,
// dY = dF^T X, scattered into dense gradients with float atomics.
//
// One wavefront per window (grid-stride).  The Y rows stay in registers for
// all contexts of the window.  The FP64 variant evaluates every logit and
// loss term in double (finite-difference checks); the FP32 variant reduces
// per-wave partial losses in double.
#include "sgns_device.h"
#include "sgns_internal.h"

namespace sgns {
namespace {

constexpr int kMaxY = 8;  // 1 centre + up to 7 shared negatives

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
    return v;
}

// Adds a row held as pairs (lane l: pairs l, l+64, ...) to a dense fp32
// gradient row with one float per lane and atomic: lane l takes float l of a
// 128-float block from lane l/2 and float 64 + l from lane 32 + l/2 (shuffles
// through ds_bpermute), so each atomic instruction adds 64 consecutive floats
// -- four 64-byte requests, where a pair per lane (two atomics 8 bytes apart)
// costs eight per instruction.  Pairs past the row hold zeros.
template <int NP>
__device__ __forceinline__ void add_row_f32(float *row, const float2 (&v)[NP], int d, int lane) {
    const int s0 = lane >> 1, s1 = 32 + (lane >> 1);
    const bool odd = lane & 1;
#pragma unroll
    for (int t = 0; t < NP; t++) {
        const float a0 = __shfl(v[t].x, s0), a1 = __shfl(v[t].y, s0);
        const float b0 = __shfl(v[t].x, s1), b1 = __shfl(v[t].y, s1);
        const float va = odd ? a1 : a0, vb = odd ? b1 : b0;
        const int f = 128 * t + lane;
        if (f < d && va != 0.f) unsafeAtomicAdd(row + f, va);
        if (f + 64 < d && vb != 0.f) unsafeAtomicAdd(row + f + 64, vb);
    }
}

template <typename T, int NP, typename ACC, bool ADJ>
__global__ __launch_bounds__(256) void sgns_window_kernel(WindowArgs a) {
    using PT = typename Pair<T>::P;
    const T *w_in = static_cast<const T *>(a.w_in);
    const T *w_out = static_cast<const T *>(a.w_out);
    const int lane = threadIdx.x % kWave;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
    const int nq = (a.d + 1) / 2, ld = a.ld, ny = 1 + a.nneg;
    double loss = 0.0;
    for (int64_t w = wave; w < a.n_windows; w += nwaves) {
        int32_t yrow[kMaxY];
        float2 y[kMaxY][NP], dy[kMaxY][NP];
#pragma unroll
        for (int k = 0; k < kMaxY; k++) {
            yrow[k] = k == 0 ? a.centre[w] : (k < ny ? a.neg[w * a.nneg + (k - 1)] : -1);
#pragma unroll
            for (int t = 0; t < NP; t++) {
                const int q = lane + kWave * t;
                y[k][t] = (k < ny && q < nq) ? Pair<T>::tof(Pair<T>::load(w_out, yrow[k], ld, q)) : make_float2(0.f, 0.f);
                dy[k][t] = make_float2(0.f, 0.f);
            }
        }
        for (int i = 0; i < a.nctx; i++) {
            const int32_t xrow = a.ctx[w * a.nctx + i];
            float2 x[NP];
#pragma unroll
            for (int t = 0; t < NP; t++) {
                const int q = lane + kWave * t;
                x[t] = q < nq ? Pair<T>::tof(Pair<T>::load(w_in, xrow, ld, q)) : make_float2(0.f, 0.f);
            }
            float2 dx[NP];
#pragma unroll
            for (int t = 0; t < NP; t++) dx[t] = make_float2(0.f, 0.f);
#pragma unroll
            for (int k = 0; k < kMaxY; k++) {
                if (k >= ny) continue;
                ACC part = 0;
#pragma unroll
                for (int t = 0; t < NP; t++) part += (ACC)x[t].x * (ACC)y[k][t].x + (ACC)x[t].y * (ACC)y[k][t].y;
                double f;
                if constexpr (sizeof(ACC) == 8) f = wave_sum_d(part);
                else f = (double)wave_sum(part);
                const double z = k == 0 ? -f : f;  // softplus argument
                if (lane == 0) loss += z > 0 ? z + log1p(exp(-z)) : log1p(exp(z));
                if (ADJ) {
                    const float sg = (float)(1.0 / (1.0 + exp(-f)));
                    const float gf = k == 0 ? sg - 1.0f : sg;
#pragma unroll
                    for (int t = 0; t < NP; t++) {
                        dx[t].x += gf * y[k][t].x;
                        dx[t].y += gf * y[k][t].y;
                        dy[k][t].x += gf * x[t].x;
                        dy[k][t].y += gf * x[t].y;
                    }
                }
            }
            if (ADJ) add_row_f32<NP>(a.g_in + (int64_t)xrow * ld, dx, a.d, lane);
        }
        if (ADJ) {
#pragma unroll
            for (int k = 0; k < kMaxY; k++) {
                if (k >= ny) continue;
                add_row_f32<NP>(a.g_out + (int64_t)yrow[k] * ld, dy[k], a.d, lane);
            }
        }
    }
    if (lane == 0 && loss != 0.0) atomicAdd(a.loss, loss);
}

template <typename T, int NP>
hipError_t launch_np(const WindowArgs &a, int adjoint, int fp64, hipStream_t st) {
    const int64_t waves = a.n_windows;
    int64_t blocks = (waves + 3) / 4;
    if (blocks > 8192) blocks = 8192;
    if (blocks < 1) blocks = 1;
    dim3 grid((unsigned)blocks), block(256);
    if (fp64) {
        if (adjoint) hipLaunchKernelGGL((sgns_window_kernel<T, NP, double, true>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((sgns_window_kernel<T, NP, double, false>), grid, block, 0, st, a);
    } else {
        if (adjoint) hipLaunchKernelGGL((sgns_window_kernel<T, NP, float, true>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((sgns_window_kernel<T, NP, float, false>), grid, block, 0, st, a);
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_t(const WindowArgs &a, int np, int adjoint, int fp64, hipStream_t st) {
    switch (np) {
    case 1: return launch_np<T, 1>(a, adjoint, fp64, st);
    case 2: return launch_np<T, 2>(a, adjoint, fp64, st);
    case 3: return launch_np<T, 3>(a, adjoint, fp64, st);
    case 4: return launch_np<T, 4>(a, adjoint, fp64, st);
    }
    return hipErrorInvalidValue;
}

// ------------------------------------------------------- matrix-core path
// bf16 tables, one wavefront per window.  The window's rows are staged in LDS
// (X: contexts, zero rows up to a multiple of 16; Y: centre and negatives,
// zero rows up to 16), then
//   F    = X Y^T            v_mfma_f32_16x16x32_bf16 (exact bf16 products, fp32 sums)
//   G    = sigmoid(F) - label, split into bf16 hi + lo (about 16 significant bits)
//   dY   = G^T X            v_mfma_f32_16x16x16_bf16, twice (hi, lo)
//   dX   = G Y              idem, G's A-operand layout through a 16 x 16 LDS tile
// With the gradient as the A operand an output tile holds, per lane, one
// dimension of four rows, so each float atomic instruction adds four rows x
// 16 consecutive floats (four 64-byte requests).
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef short s16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
constexpr int kWinTS = 20;  // transpose tile row stride (bf16 elements)

inline int win_mfma_stride(int ld) { return ld + 4; }

template <int MT, bool ADJ>
__global__ __launch_bounds__(256) void sgns_window_mfma_kernel(WindowArgs a, int lds_per_wave) {
    extern __shared__ __align__(16) char smem[];
    const uint16_t *w_in = static_cast<const uint16_t *>(a.w_in);
    const uint16_t *w_out = static_cast<const uint16_t *>(a.w_out);
    const int lane = threadIdx.x % kWave, r16 = lane & 15, g4 = lane >> 4;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
    const int ld = a.ld, d = a.d, SC = ld + 4, KS = ld / 32, JT = (d + 15) / 16;
    const int nctx = a.nctx, ny = 1 + a.nneg, NX = MT * 16;
    uint16_t *xs = reinterpret_cast<uint16_t *>(smem + (size_t)(threadIdx.x / kWave) * lds_per_wave);
    uint16_t *ys = xs + NX * SC;
    uint16_t *tr = ys + 16 * SC;  // two tiles: hi, lo
    double loss = 0.0;
    for (int64_t w = wave; w < a.n_windows; w += nwaves) {
        // ---- stage the rows (four in flight); zero rows past nctx / ny
        for (int r0 = 0; r0 < NX + 16; r0 += 4) {
            uint32_t v[4][8];
            int32_t id[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int rr = r0 + k;
                const bool isx = rr < NX;
                const int yi = rr - NX;
                id[k] = isx ? (rr < nctx ? a.ctx[w * nctx + rr] : -1)
                            : (yi < ny ? (yi == 0 ? a.centre[w] : a.neg[w * a.nneg + (yi - 1)]) : -1);
                const uint16_t *tab = isx ? w_in : w_out;
#pragma unroll
                for (int t = 0; t < 8; t++) {
                    const int q = lane + kWave * t;  // pair index, ld <= 1024
                    v[k][t] = id[k] >= 0 && 2 * q < ld ? Pair<uint16_t>::load(tab, id[k], ld, q) : 0u;
                }
            }
#pragma unroll
            for (int k = 0; k < 4; k++) {
                uint16_t *dst = (r0 + k < NX ? xs + (r0 + k) * SC : ys + (r0 + k - NX) * SC);
#pragma unroll
                for (int t = 0; t < 8; t++) {
                    const int q = lane + kWave * t;
                    if (2 * q < ld) *reinterpret_cast<uint32_t *>(dst + 2 * q) = v[k][t];
                }
            }
        }
        // ---- F = X Y^T
        f32x4_t acc[MT];
#pragma unroll
        for (int mt = 0; mt < MT; mt++) acc[mt] = f32x4_t{0.f, 0.f, 0.f, 0.f};
        for (int ks = 0; ks < KS; ks++) {
            const uint16_t *qb = ys + r16 * SC + ks * 32 + 8 * g4;
            const uint2 b0 = *reinterpret_cast<const uint2 *>(qb), b1 = *reinterpret_cast<const uint2 *>(qb + 4);
            const bf16x8_t bv = __builtin_bit_cast(bf16x8_t, u32x4_t{b0.x, b0.y, b1.x, b1.y});
#pragma unroll
            for (int mt = 0; mt < MT; mt++) {
                const uint16_t *qa = xs + (mt * 16 + r16) * SC + ks * 32 + 8 * g4;
                const uint2 a0 = *reinterpret_cast<const uint2 *>(qa), a1 = *reinterpret_cast<const uint2 *>(qa + 4);
                const bf16x8_t av = __builtin_bit_cast(bf16x8_t, u32x4_t{a0.x, a0.y, a1.x, a1.y});
                acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[mt], 0, 0, 0);
            }
        }
        // acc[mt][r] = F[context mt*16 + 4*g4 + r][y r16]
        s16x4_t gh[MT], gl[MT];
#pragma unroll
        for (int mt = 0; mt < MT; mt++) {
            uint32_t hb[4], lb[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = mt * 16 + 4 * g4 + r;
                const bool ok = i < nctx && r16 < ny;
                const double f = (double)acc[mt][r];
                const double z = r16 == 0 ? -f : f;
                if (ok) loss += z > 0 ? z + log1p(exp(-z)) : log1p(exp(z));
                float g = 0.f;
                if (ADJ && ok) {
                    const float sg = (float)(1.0 / (1.0 + exp(-f)));
                    g = r16 == 0 ? sg - 1.0f : sg;
                }
                hb[r] = bf16_rne(g);
                lb[r] = bf16_rne(g - __builtin_bit_cast(float, hb[r] << 16));
            }
            gh[mt] = s16x4_t{(short)hb[0], (short)hb[1], (short)hb[2], (short)hb[3]};
            gl[mt] = s16x4_t{(short)lb[0], (short)lb[1], (short)lb[2], (short)lb[3]};
        }
        if constexpr (ADJ) {
            // ---- dY = G^T X: lane (r16 = dimension, g4) gets rows y = 4*g4 + r
            int32_t yrow[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int y = 4 * g4 + r;
                yrow[r] = y < ny ? (y == 0 ? a.centre[w] : a.neg[w * a.nneg + (y - 1)]) : -1;
            }
            for (int jt = 0; jt < JT; jt++) {
                f32x4_t dy = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int mt = 0; mt < MT; mt++) {
                    s16x4_t xv;
#pragma unroll
                    for (int r = 0; r < 4; r++) xv[r] = (short)xs[(mt * 16 + 4 * g4 + r) * SC + jt * 16 + r16];
                    dy = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(gh[mt], xv, dy, 0, 0, 0);
                    dy = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(gl[mt], xv, dy, 0, 0, 0);
                }
                const int j = jt * 16 + r16;
#pragma unroll
                for (int r = 0; r < 4; r++)
                    if (yrow[r] >= 0 && j < d && dy[r] != 0.f) unsafeAtomicAdd(a.g_out + (int64_t)yrow[r] * ld + j, dy[r]);
            }
            // ---- dX = G Y: lane (r16 = dimension, g4) gets contexts mt*16 + 4*g4 + r
#pragma unroll
            for (int mt = 0; mt < MT; mt++) {
                // G^T tile of this context block (hi and lo) through LDS
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    tr[(4 * g4 + r) * kWinTS + r16] = (uint16_t)gh[mt][r];
                    tr[16 * kWinTS + (4 * g4 + r) * kWinTS + r16] = (uint16_t)gl[mt][r];
                }
                const s16x4_t th = __builtin_bit_cast(s16x4_t, *reinterpret_cast<const uint2 *>(tr + r16 * kWinTS + 4 * g4));
                const s16x4_t tl =
                    __builtin_bit_cast(s16x4_t, *reinterpret_cast<const uint2 *>(tr + 16 * kWinTS + r16 * kWinTS + 4 * g4));
                int32_t xrow[4];
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int i = mt * 16 + 4 * g4 + r;
                    xrow[r] = i < nctx ? a.ctx[w * nctx + i] : -1;
                }
                for (int jt = 0; jt < JT; jt++) {
                    s16x4_t yv;
#pragma unroll
                    for (int r = 0; r < 4; r++) yv[r] = (short)ys[(4 * g4 + r) * SC + jt * 16 + r16];
                    f32x4_t dx = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(th, yv, f32x4_t{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
                    dx = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(tl, yv, dx, 0, 0, 0);
                    const int j = jt * 16 + r16;
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        if (xrow[r] >= 0 && j < d && dx[r] != 0.f) unsafeAtomicAdd(a.g_in + (int64_t)xrow[r] * ld + j, dx[r]);
                }
            }
        }
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) loss += __shfl_xor(loss, m, kWave);
    if (lane == 0 && loss != 0.0) atomicAdd(a.loss, loss);
}

template <int MT, bool ADJ>
hipError_t launch_mfma_mt(const WindowArgs &a, hipStream_t st) {
    const int lds = (MT * 16 + 16) * win_mfma_stride(a.ld) * 2 + 2 * 16 * kWinTS * 2;
    const int wpb = 2;
    auto kern = sgns_window_mfma_kernel<MT, ADJ>;
    hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, lds * wpb);
    if (e != hipSuccess) return e;
    int64_t blocks = (a.n_windows + wpb - 1) / wpb;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(kWave * wpb), (size_t)lds * wpb, st, a, lds);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_window(const WindowArgs &a, int bf16, int np, int adjoint, int flags, hipStream_t st) {
    const int fp64 = (flags & 2) != 0;
    if (flags & 1) {  // matrix cores: bf16 tables, fp32 accumulation, ld <= 1024
        if (!bf16 || fp64 || a.ld > 1024) return hipErrorNotSupported;
        if (a.nctx <= 16) return adjoint ? launch_mfma_mt<1, true>(a, st) : launch_mfma_mt<1, false>(a, st);
        return adjoint ? launch_mfma_mt<2, true>(a, st) : launch_mfma_mt<2, false>(a, st);
    }
    return bf16 ? launch_t<uint16_t>(a, np, adjoint, fp64, st) : launch_t<float>(a, np, adjoint, fp64, st);
}

}  // namespace sgns
