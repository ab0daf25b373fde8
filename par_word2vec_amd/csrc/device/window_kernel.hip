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
// bf16 tables below 4 GiB - 4 KiB (addressed through buffer resources), one
// wavefront per window, grid-stride, the next window's row indices loaded
// while the current window's rows are in flight.
//   F    = X Y^T            v_mfma_f32_16x16x32_bf16 (exact bf16 products, fp32 sums)
//   G    = sigmoid(F) - label, split into bf16 hi + lo (about 16 significant bits)
//   dY   = G^T X            v_mfma_f32_16x16x16_bf16, twice (hi, lo)
//   dX   = G Y              idem, G's A-operand layout through a 16 x 16 LDS tile
// F's operands come straight from the tables in the 16x16x32 operand layout:
// lane (r, g) of a 16-row tile holds row r at dimensions 32 ks + 8 g .. + 7,
// one 128-bit load per 32 dimensions.  X tiles are the contexts (rows past
// nctx read at kWinNoRow, past the resource: the range check returns zeros
// without a memory request), the Y tile the centre and the negatives (rows
// past ny likewise), so every row of a window is loaded at once and the
// forward needs no LDS.  The adjoint's products read X and Y with one
// dimension per lane, so it also writes the window's rows to LDS (only the
// nctx + ny real ones, and one zero row that the padding rows read).  With
// the gradient as the A operand an output tile holds, per lane, one dimension
// of four rows, so each float atomic instruction adds four rows x 16
// consecutive floats (four 64-byte requests).
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef short s16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
constexpr int kWinTS = 20;                   // transpose tile row stride (bf16 elements)
constexpr int kWinRsrcWord3 = 0x00020000;    // buffer resource: raw 32-bit data
constexpr uint32_t kWinNoRow = 0xFFFFF000u;  // past every table the path accepts

__host__ __device__ inline int win_mfma_stride(int ld) { return ld + 4; }

// softplus(z) = log(1 + e^z) in fp32 without overflow: max(z, 0) + log1p(e^-|z|)
__device__ __forceinline__ float softplus_stable(float z) {
    return fmaxf(z, 0.0f) + log1pf(expf(-fabsf(z)));
}

template <int MT, int KSM, bool ADJ>
__global__ __launch_bounds__(256) void sgns_window_mfma_kernel(WindowArgs a, int lds_per_wave) {
    extern __shared__ __align__(16) char smem[];
    const int lane = threadIdx.x % kWave, r16 = lane & 15, g4 = lane >> 4;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
    const int64_t nw = a.n_windows;
    const int ld = a.ld, d = a.d, SC = win_mfma_stride(ld), KS = ld / 32, JT = (d + 15) / 16;
    const int nctx = a.nctx, nneg = a.nneg, ny = 1 + nneg;
    const uint32_t row_bytes = (uint32_t)ld * 2u;
    const __amdgpu_buffer_rsrc_t rin = __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(a.w_in), (short)0,
                                                                         (int)((uint32_t)a.V * row_bytes), kWinRsrcWord3);
    const __amdgpu_buffer_rsrc_t rout = __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(a.w_out), (short)0,
                                                                          (int)((uint32_t)a.V * row_bytes), kWinRsrcWord3);
    // LDS of this wave (adjoint only): X rows [0, nctx), Y rows [nctx, nctx + ny),
    // the zero row, then the two transpose tiles (hi, lo)
    uint16_t *xs = reinterpret_cast<uint16_t *>(smem + (size_t)(threadIdx.x / kWave) * lds_per_wave);
    const int zrow = nctx + ny;
    uint16_t *tr = xs + (zrow + 1) * SC;
    if constexpr (ADJ) {
        for (int q = lane; 2 * q < ld; q += kWave) *reinterpret_cast<uint32_t *>(xs + zrow * SC + 2 * q) = 0u;
    }
    auto xmap = [&](int i) { return i < nctx ? i : zrow; };       // LDS row of context i
    auto ymap = [&](int k) { return k < ny ? nctx + k : zrow; };  // LDS row of Y row k
    // this lane's rows of window w: context mt*16 + r16 and Y row r16 (-1: none).
    // The index loads are unconditional (clamped in range) and masked after
    // with bit operations: a select of a loaded value against a default
    // becomes a branch around the load, whose wait would cover every row load
    // in flight.
    auto load_ids = [&](int64_t w, int32_t (&xi)[MT], int32_t &yi) {
        const int64_t wc = w < nw ? w : nw - 1;
        const int32_t past = w < nw ? 0 : -1;  // all bits set: no window
#pragma unroll
        for (int mt = 0; mt < MT; mt++) {
            const int i = mt * 16 + r16;
            xi[mt] = a.ctx[wc * nctx + (i < nctx ? i : 0)] | (i < nctx ? past : -1);
        }
        const int32_t c = a.centre[wc];
        const int32_t *pn = nneg > 0 ? a.neg + wc * nneg + (r16 >= 1 && r16 <= nneg ? r16 - 1 : 0) : a.centre + wc;
        const int32_t n = *pn;
        const int32_t m0 = r16 == 0 ? -1 : 0;
        yi = ((c & m0) | (n & ~m0)) | (r16 < ny ? past : -1);
    };
    double loss = 0.0;
    if (nw <= 0) return;
    int32_t xi[MT], yi;
    load_ids(wave, xi, yi);
    for (int64_t w = wave; w < nw; w += nwaves) {
        // ---- every row of the window in flight: X tiles and the Y tile
        u32x4_t xv[MT][KSM], yv[KSM];
        {
            uint32_t ox[MT];
#pragma unroll
            for (int mt = 0; mt < MT; mt++) ox[mt] = (xi[mt] >= 0 ? (uint32_t)xi[mt] * row_bytes : kWinNoRow) + 16u * g4;
            const uint32_t oy = (yi >= 0 ? (uint32_t)yi * row_bytes : kWinNoRow) + 16u * g4;
#pragma unroll
            for (int ks = 0; ks < KSM; ks++) {
#pragma unroll
                for (int mt = 0; mt < MT; mt++)
                    xv[mt][ks] = __builtin_amdgcn_raw_buffer_load_b128(rin, ks < KS ? ox[mt] + 64u * ks : kWinNoRow, 0, 0);
                yv[ks] = __builtin_amdgcn_raw_buffer_load_b128(rout, ks < KS ? oy + 64u * ks : kWinNoRow, 0, 0);
            }
        }
        // the row ids this lane writes gradients to (adjoint), then the next window's ids
        int32_t yrow[4], xrow[MT][4];
        if constexpr (ADJ) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                yrow[r] = __shfl(yi, 4 * g4 + r, kWave);
#pragma unroll
                for (int mt = 0; mt < MT; mt++) xrow[mt][r] = __shfl(xi[mt], 4 * g4 + r, kWave);
            }
        }
        load_ids(w + nwaves, xi, yi);
        // ---- F = X Y^T: acc[mt][r] = F[context mt*16 + 4*g4 + r][y r16]
        f32x4_t acc[MT];
#pragma unroll
        for (int mt = 0; mt < MT; mt++) acc[mt] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < KSM; ks++) {
            if (ks >= KS) continue;
            const bf16x8_t bv = __builtin_bit_cast(bf16x8_t, yv[ks]);
#pragma unroll
            for (int mt = 0; mt < MT; mt++)
                acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, xv[mt][ks]), bv, acc[mt],
                                                                  0, 0, 0);
        }
        if constexpr (ADJ) {  // stage the real rows for the adjoint's transposed reads
#pragma unroll
            for (int ks = 0; ks < KSM; ks++) {
                if (ks >= KS) continue;
#pragma unroll
                for (int mt = 0; mt < MT; mt++) {
                    const int i = mt * 16 + r16;
                    if (i < nctx) {
                        uint2 *p = reinterpret_cast<uint2 *>(xs + i * SC + ks * 32 + 8 * g4);
                        p[0] = make_uint2(xv[mt][ks].x, xv[mt][ks].y);
                        p[1] = make_uint2(xv[mt][ks].z, xv[mt][ks].w);
                    }
                }
                if (r16 < ny) {
                    uint2 *p = reinterpret_cast<uint2 *>(xs + (nctx + r16) * SC + ks * 32 + 8 * g4);
                    p[0] = make_uint2(yv[ks].x, yv[ks].y);
                    p[1] = make_uint2(yv[ks].z, yv[ks].w);
                }
            }
        }
        // ---- loss and the logit gradient (label 1 for the centre column)
        s16x4_t gh[MT], gl[MT];
#pragma unroll
        for (int mt = 0; mt < MT; mt++) {
            uint32_t hb[4], lb[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = mt * 16 + 4 * g4 + r;
                const bool ok = i < nctx && r16 < ny;
                const float f = acc[mt][r];
                if (ok) loss += (double)softplus_stable(r16 == 0 ? -f : f);
                float g = 0.f;
                if (ADJ && ok) {
                    const float sg = 1.0f / (1.0f + expf(-f));
                    g = r16 == 0 ? sg - 1.0f : sg;
                }
                hb[r] = bf16_rne(g);
                lb[r] = bf16_rne(g - __builtin_bit_cast(float, hb[r] << 16));
            }
            gh[mt] = s16x4_t{(short)hb[0], (short)hb[1], (short)hb[2], (short)hb[3]};
            gl[mt] = s16x4_t{(short)lb[0], (short)lb[1], (short)lb[2], (short)lb[3]};
        }
        if constexpr (ADJ) {
            // ---- dY = G^T X: lane (r16 = dimension, g4) gets Y rows 4*g4 + r
            for (int jt = 0; jt < JT; jt++) {
                f32x4_t dy = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int mt = 0; mt < MT; mt++) {
                    s16x4_t xt;
#pragma unroll
                    for (int r = 0; r < 4; r++) xt[r] = (short)xs[xmap(mt * 16 + 4 * g4 + r) * SC + jt * 16 + r16];
                    dy = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(gh[mt], xt, dy, 0, 0, 0);
                    dy = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(gl[mt], xt, dy, 0, 0, 0);
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
                for (int jt = 0; jt < JT; jt++) {
                    s16x4_t yt;
#pragma unroll
                    for (int r = 0; r < 4; r++) yt[r] = (short)xs[ymap(4 * g4 + r) * SC + jt * 16 + r16];
                    f32x4_t dx = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(th, yt, f32x4_t{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
                    dx = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(tl, yt, dx, 0, 0, 0);
                    const int j = jt * 16 + r16;
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        if (xrow[mt][r] >= 0 && j < d && dx[r] != 0.f)
                            unsafeAtomicAdd(a.g_in + (int64_t)xrow[mt][r] * ld + j, dx[r]);
                }
            }
        }
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) loss += __shfl_xor(loss, m, kWave);
    if (lane == 0 && loss != 0.0) atomicAdd(a.loss, loss);
}

// One-tile variant, for windows with nctx + 1 + nneg <= 16 rows: the
// contexts (rows 0..nctx-1) and Y (rows nctx..nctx+ny-1) share one 16-row
// tile T, the rows past them repeat context 0.  F = T T^T (A and B operands
// are the same registers: lane (r, g) holds row r at dimensions 32 ks + 8 g
// .. + 7 in both layouts), of which the context x Y block is used.  With
// M[i][j] = dL/dF[i][j - nctx] on that block and 0 elsewhere, S = M + M^T is
// symmetric, so the 16x16x16 A operand of S is its D layout: each lane
// computes the gradient of its own F entry, which is the logit of pair
// (i, j) or, transposed, of (j, i).  Then S T gives every gradient row at
// once: row i < nctx is sum_j M[i][j] T[j] = dX_i, row nctx + k is
// sum_i M[i][nctx+k] T[i] = dY_k -- one product pair (gradient hi, lo) per
// 16 dimensions instead of four, and no transpose tile.  Rows are loaded
// with per-lane 64-bit addresses (a lane's row lives in either table), KS =
// ld / 32 loads of 16 bytes each, all in flight at once.
template <int KS, bool ADJ>
__global__ __launch_bounds__(256) void sgns_window_tile_kernel(WindowArgs a, int lds_per_wave) {
    extern __shared__ __align__(16) char smem[];
    const int lane = threadIdx.x % kWave, r16 = lane & 15, g4 = lane >> 4;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
    const int64_t nw = a.n_windows;
    const int ld = a.ld, d = a.d, SC = win_mfma_stride(ld), JT = (d + 15) / 16;
    const int nctx = a.nctx, nneg = a.nneg, nr = nctx + 1 + nneg;
    if (nw <= 0) return;
    const uint16_t *w_in = static_cast<const uint16_t *>(a.w_in);
    const uint16_t *w_out = static_cast<const uint16_t *>(a.w_out);
    // LDS of this wave (adjoint only): the nr rows of T, then a zero row
    uint16_t *ts = reinterpret_cast<uint16_t *>(smem + (size_t)(threadIdx.x / kWave) * lds_per_wave);
    if constexpr (ADJ) {
        for (int q = lane; 2 * q < ld; q += kWave) *reinterpret_cast<uint32_t *>(ts + nr * SC + 2 * q) = 0u;
    }
    // the row of tile row r16 in window w: a context, the centre, a negative,
    // or (past nr) context 0 again.  The index's address is chosen by lane,
    // so the load is unconditional.
    const bool is_y = r16 >= nctx && r16 < nr;
    const int k = r16 - nctx;
    auto id_of = [&](int64_t w) -> int32_t {
        const int64_t wc = w < nw ? w : nw - 1;
        const int32_t *p = !is_y ? a.ctx + wc * nctx + (r16 < nctx ? r16 : 0)
                                 : (k == 0 ? a.centre + wc : a.neg + wc * nneg + (k - 1));
        return *p;
    };
    const uint16_t *tab = is_y ? w_out : w_in;
    double loss = 0.0;
    int32_t id = id_of(wave);
    for (int64_t w = wave; w < nw; w += nwaves) {
        // the next window's row id first (loads retire in order: it is then
        // in hand when this window's rows are), then this window's rows
        const int32_t id_next = id_of(w + nwaves);
        u32x4_t tv[KS];
        {
            const u32x4_t *p = reinterpret_cast<const u32x4_t *>(tab + (int64_t)id * ld + 8 * g4);
#pragma unroll
            for (int ks = 0; ks < KS; ks++) tv[ks] = p[4 * ks];
        }
        // every load issued before the first product: otherwise the scheduler
        // may interleave them, two registers sets, one round trip per step
        __builtin_amdgcn_sched_barrier(0);
        // (adjoint) the table and row of the gradient rows this lane adds: 4 g4 + r
        int32_t orow[4];
        if constexpr (ADJ) {
#pragma unroll
            for (int r = 0; r < 4; r++) orow[r] = __shfl(id, 4 * g4 + r, kWave);
        }
        id = id_next;
        f32x4_t acc = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            const bf16x8_t v = __builtin_bit_cast(bf16x8_t, tv[ks]);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(v, v, acc, 0, 0, 0);
        }
        if constexpr (ADJ) {  // stage the real rows for the transposed reads
            if (r16 < nr) {
#pragma unroll
                for (int ks = 0; ks < KS; ks++) {
                    uint2 *q = reinterpret_cast<uint2 *>(ts + r16 * SC + ks * 32 + 8 * g4);
                    q[0] = make_uint2(tv[ks].x, tv[ks].y);
                    q[1] = make_uint2(tv[ks].z, tv[ks].w);
                }
            }
        }
        // acc[r] = F[4 g4 + r][r16]: the loss once per pair (context row, Y
        // column), the gradient in both orientations (S is symmetric)
        uint32_t hb[4], lb[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int i = 4 * g4 + r;
            const bool xy = i < nctx && is_y;                    // (context i, Y r16)
            const bool yx = r16 < nctx && i >= nctx && i < nr;   // (Y i, context r16)
            const int yc = xy ? k : i - nctx;                    // Y row of the pair: 0 = centre
            const float f = acc[r];
            if (xy) loss += (double)softplus_stable(yc == 0 ? -f : f);
            float g = 0.f;
            if (ADJ && (xy || yx)) {
                const float sg = 1.0f / (1.0f + expf(-f));
                g = yc == 0 ? sg - 1.0f : sg;
            }
            hb[r] = bf16_rne(g);
            lb[r] = bf16_rne(g - __builtin_bit_cast(float, hb[r] << 16));
        }
        if constexpr (ADJ) {
            const s16x4_t sh = s16x4_t{(short)hb[0], (short)hb[1], (short)hb[2], (short)hb[3]};
            const s16x4_t sl = s16x4_t{(short)lb[0], (short)lb[1], (short)lb[2], (short)lb[3]};
            float *dst[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = 4 * g4 + r;
                dst[r] = i < nr ? (i < nctx ? a.g_in : a.g_out) + (int64_t)orow[r] * ld : nullptr;
            }
            for (int jt = 0; jt < JT; jt++) {
                s16x4_t tt;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int i = 4 * g4 + r;
                    tt[r] = (short)ts[(i < nr ? i : nr) * SC + jt * 16 + r16];
                }
                f32x4_t o = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(sh, tt, f32x4_t{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
                o = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(sl, tt, o, 0, 0, 0);
                const int j = jt * 16 + r16;
#pragma unroll
                for (int r = 0; r < 4; r++)
                    if (dst[r] && j < d && o[r] != 0.f) unsafeAtomicAdd(dst[r] + j, o[r]);
            }
        }
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) loss += __shfl_xor(loss, m, kWave);
    if (lane == 0 && loss != 0.0) atomicAdd(a.loss, loss);
}

template <int KS, bool ADJ>
hipError_t launch_tile_ks(const WindowArgs &a, hipStream_t st) {
    if (a.n_windows <= 0) return hipSuccess;  // (the kernel reads the last window's indices ahead)
    const int nr = a.nctx + 1 + a.nneg;
    const int lds = ADJ ? ((nr + 1) * win_mfma_stride(a.ld) * 2 + 15) & ~15 : 0;
    const int wpb = 4;  // at most 4 x 17 rows of 516 bf16: 70 KB
    auto kern = sgns_window_tile_kernel<KS, ADJ>;
    if (ADJ) {
        hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, lds * wpb);
        if (e != hipSuccess) return e;
    }
    int64_t blocks = (a.n_windows + wpb - 1) / wpb;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(kWave * wpb), (size_t)lds * wpb, st, a, lds);
    return hipGetLastError();
}

template <bool ADJ>
hipError_t launch_tile(const WindowArgs &a, hipStream_t st) {
    switch (a.ld / 32) {  // ld is a multiple of 64 on bf16 tables, at most 512
    case 2: return launch_tile_ks<2, ADJ>(a, st);
    case 4: return launch_tile_ks<4, ADJ>(a, st);
    case 6: return launch_tile_ks<6, ADJ>(a, st);
    case 8: return launch_tile_ks<8, ADJ>(a, st);
    case 10: return launch_tile_ks<10, ADJ>(a, st);
    case 12: return launch_tile_ks<12, ADJ>(a, st);
    case 14: return launch_tile_ks<14, ADJ>(a, st);
    case 16: return launch_tile_ks<16, ADJ>(a, st);
    }
    return hipErrorNotSupported;
}

// The one-tile kernel on fp32 tables, on the fp32 matrix cores
// (v_mfma_f32_16x16x4_f32: A is 16 x 4, lane (r, g) holding A[r][g]; B is
// 4 x 16, lane (r, g) holding B[g][r]).  Lane (r, g) loads row r of T in
// 16-byte pieces: piece kb covers dimensions 16 kb + 4 g .. + 3.  A k-step
// may take the dimensions in any order as long as A and B agree, and in
// F = T T^T they are the same registers, so component c of piece kb feeds
// step (kb, c) with k = 16 kb + 4 g + c: four steps per piece, KB = ld / 16
// pieces.  The adjoint S T takes its k-steps over tile rows the same way:
// step kk pairs lane group g with tile row 4 g + kk, whose S entry this lane
// holds in accumulator register kk (S is symmetric), and whose T entry it
// reads from LDS.  The gradient stays in fp32 (no hi/lo split).
template <int KB, bool ADJ>
__global__ __launch_bounds__(256) void sgns_window_tile_f32_kernel(WindowArgs a, int lds_per_wave) {
    extern __shared__ __align__(16) char smem[];
    const int lane = threadIdx.x % kWave, r16 = lane & 15, g4 = lane >> 4;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
    const int64_t nw = a.n_windows;
    const int ld = a.ld, d = a.d, SC = win_mfma_stride(ld), JT = (d + 15) / 16;
    const int nctx = a.nctx, nneg = a.nneg, nr = nctx + 1 + nneg;
    if (nw <= 0) return;
    const float *w_in = static_cast<const float *>(a.w_in);
    const float *w_out = static_cast<const float *>(a.w_out);
    float *ts = reinterpret_cast<float *>(smem + (size_t)(threadIdx.x / kWave) * lds_per_wave);
    if constexpr (ADJ) {
        for (int q = lane; q < ld; q += kWave) ts[nr * SC + q] = 0.f;
    }
    const bool is_y = r16 >= nctx && r16 < nr;
    const int k = r16 - nctx;
    auto id_of = [&](int64_t w) -> int32_t {
        const int64_t wc = w < nw ? w : nw - 1;
        const int32_t *p = !is_y ? a.ctx + wc * nctx + (r16 < nctx ? r16 : 0)
                                 : (k == 0 ? a.centre + wc : a.neg + wc * nneg + (k - 1));
        return *p;
    };
    const float *tab = is_y ? w_out : w_in;
    double loss = 0.0;
    int32_t id = id_of(wave);
    for (int64_t w = wave; w < nw; w += nwaves) {
        const int32_t id_next = id_of(w + nwaves);  // (first: see the bf16 kernel)
        f32x4_t tv[KB];
        {
            const f32x4_t *p = reinterpret_cast<const f32x4_t *>(tab + (int64_t)id * ld + 4 * g4);
#pragma unroll
            for (int kb = 0; kb < KB; kb++) tv[kb] = p[4 * kb];
        }
        __builtin_amdgcn_sched_barrier(0);  // (every load before the first product)
        int32_t orow[4];
        if constexpr (ADJ) {
#pragma unroll
            for (int r = 0; r < 4; r++) orow[r] = __shfl(id, 4 * g4 + r, kWave);
        }
        id = id_next;
        f32x4_t acc = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < KB; kb++) {
#pragma unroll
            for (int c = 0; c < 4; c++) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(tv[kb][c], tv[kb][c], acc, 0, 0, 0);
        }
        if constexpr (ADJ) {
            if (r16 < nr) {
#pragma unroll
                for (int kb = 0; kb < KB; kb++) *reinterpret_cast<f32x4_t *>(ts + r16 * SC + 16 * kb + 4 * g4) = tv[kb];
            }
        }
        float sg4[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int i = 4 * g4 + r;
            const bool xy = i < nctx && is_y;
            const bool yx = r16 < nctx && i >= nctx && i < nr;
            const int yc = xy ? k : i - nctx;
            const float f = acc[r];
            if (xy) loss += (double)softplus_stable(yc == 0 ? -f : f);
            float g = 0.f;
            if (ADJ && (xy || yx)) {
                const float sg = 1.0f / (1.0f + expf(-f));
                g = yc == 0 ? sg - 1.0f : sg;
            }
            sg4[r] = g;
        }
        if constexpr (ADJ) {
            float *dst[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = 4 * g4 + r;
                dst[r] = i < nr ? (i < nctx ? a.g_in : a.g_out) + (int64_t)orow[r] * ld : nullptr;
            }
            for (int jt = 0; jt < JT; jt++) {
                f32x4_t o = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kk = 0; kk < 4; kk++) {
                    const int i = 4 * g4 + kk;
                    const float t = ts[(i < nr ? i : nr) * SC + jt * 16 + r16];
                    o = __builtin_amdgcn_mfma_f32_16x16x4f32(sg4[kk], t, o, 0, 0, 0);
                }
                const int j = jt * 16 + r16;
#pragma unroll
                for (int r = 0; r < 4; r++)
                    if (dst[r] && j < d && o[r] != 0.f) unsafeAtomicAdd(dst[r] + j, o[r]);
            }
        }
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) loss += __shfl_xor(loss, m, kWave);
    if (lane == 0 && loss != 0.0) atomicAdd(a.loss, loss);
}

template <int KB, bool ADJ>
hipError_t launch_tile_f32_kb(const WindowArgs &a, hipStream_t st) {
    if (a.n_windows <= 0) return hipSuccess;
    const int nr = a.nctx + 1 + a.nneg;
    const int lds = ADJ ? ((nr + 1) * win_mfma_stride(a.ld) * 4 + 15) & ~15 : 0;
    const int wpb = 4;  // at most 4 x 17 rows of 516 floats: 140 KB
    auto kern = sgns_window_tile_f32_kernel<KB, ADJ>;
    if (ADJ) {
        hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, lds * wpb);
        if (e != hipSuccess) return e;
    }
    int64_t blocks = (a.n_windows + wpb - 1) / wpb;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(kWave * wpb), (size_t)lds * wpb, st, a, lds);
    return hipGetLastError();
}

template <bool ADJ>
hipError_t launch_tile_f32(const WindowArgs &a, hipStream_t st) {
    switch (a.ld / 32) {  // ld is a multiple of 32 on fp32 tables, at most 512
    case 1: return launch_tile_f32_kb<2, ADJ>(a, st);
    case 2: return launch_tile_f32_kb<4, ADJ>(a, st);
    case 3: return launch_tile_f32_kb<6, ADJ>(a, st);
    case 4: return launch_tile_f32_kb<8, ADJ>(a, st);
    case 5: return launch_tile_f32_kb<10, ADJ>(a, st);
    case 6: return launch_tile_f32_kb<12, ADJ>(a, st);
    case 7: return launch_tile_f32_kb<14, ADJ>(a, st);
    case 8: return launch_tile_f32_kb<16, ADJ>(a, st);
    case 9: return launch_tile_f32_kb<18, ADJ>(a, st);
    case 10: return launch_tile_f32_kb<20, ADJ>(a, st);
    case 11: return launch_tile_f32_kb<22, ADJ>(a, st);
    case 12: return launch_tile_f32_kb<24, ADJ>(a, st);
    case 13: return launch_tile_f32_kb<26, ADJ>(a, st);
    case 14: return launch_tile_f32_kb<28, ADJ>(a, st);
    case 15: return launch_tile_f32_kb<30, ADJ>(a, st);
    case 16: return launch_tile_f32_kb<32, ADJ>(a, st);
    }
    return hipErrorNotSupported;
}

// LDS bytes per wave of the matrix-core adjoint (the forward uses none).
inline int win_mfma_lds(const WindowArgs &a) {
    return (a.nctx + 1 + a.nneg + 1) * win_mfma_stride(a.ld) * 2 + 2 * 16 * kWinTS * 2;
}

template <int MT, int KSM, bool ADJ>
hipError_t launch_mfma_mt(const WindowArgs &a, hipStream_t st) {
    if (a.n_windows <= 0) return hipSuccess;  // (the kernel reads the last window's indices ahead)
    const int lds = ADJ ? (win_mfma_lds(a) + 15) & ~15 : 0;
    // four waves per block, fewer where their adjoint LDS would pass 160 KB
    const int wpb = ADJ ? (4 * lds <= 160 * 1024 ? 4 : (2 * lds <= 160 * 1024 ? 2 : 1)) : 4;
    auto kern = sgns_window_mfma_kernel<MT, KSM, ADJ>;
    if (ADJ) {
        hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, lds * wpb);
        if (e != hipSuccess) return e;
    }
    // grid-stride over the windows: up to 8 blocks of 4 waves per CU (256 CUs)
    int64_t blocks = (a.n_windows + wpb - 1) / wpb;
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(kWave * wpb), (size_t)lds * wpb, st, a, lds);
    return hipGetLastError();
}

template <int MT, bool ADJ>
hipError_t launch_mfma_ks(const WindowArgs &a, hipStream_t st) {
    const int ks = a.ld / 32;  // ld is a multiple of 64 on bf16 tables, at most 512 (dim <= 512)
    if (ks <= 4) return launch_mfma_mt<MT, 4, ADJ>(a, st);
    if (ks <= 8) return launch_mfma_mt<MT, 8, ADJ>(a, st);
    if (ks <= 12) return launch_mfma_mt<MT, 12, ADJ>(a, st);
    if (ks <= 16) return launch_mfma_mt<MT, 16, ADJ>(a, st);
    return hipErrorNotSupported;
}

}  // namespace

hipError_t launch_window(const WindowArgs &a, int bf16, int np, int adjoint, int flags, hipStream_t st) {
    const int fp64 = (flags & 2) != 0;
    if ((flags & 1) && !bf16) {  // fp32 matrix cores: windows of at most 16 rows, ld <= 512
        if (fp64 || a.ld > 512 || a.nctx + 1 + a.nneg > 16) return hipErrorNotSupported;
        return adjoint ? launch_tile_f32<true>(a, st) : launch_tile_f32<false>(a, st);
    }
    if (flags & 1) {  // matrix cores: bf16 tables below 4 GiB - 4 KiB, fp32 accumulation, ld <= 512
        if (fp64 || a.ld > 512 || (uint64_t)a.V * (uint64_t)a.ld * 2u > (uint64_t)kWinNoRow)
            return hipErrorNotSupported;
        if (a.nctx + 1 + a.nneg <= 16) return adjoint ? launch_tile<true>(a, st) : launch_tile<false>(a, st);
        if (a.nctx <= 16) return adjoint ? launch_mfma_ks<1, true>(a, st) : launch_mfma_ks<1, false>(a, st);
        return adjoint ? launch_mfma_ks<2, true>(a, st) : launch_mfma_ks<2, false>(a, st);
    }
    return bf16 ? launch_t<uint16_t>(a, np, adjoint, fp64, st) : launch_t<float>(a, np, adjoint, fp64, st);
}

}  // namespace sgns
