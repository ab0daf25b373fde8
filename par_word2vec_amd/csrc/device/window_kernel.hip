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
            if (ADJ) {
#pragma unroll
                for (int t = 0; t < NP; t++) {
                    const int q = lane + kWave * t;
                    if (q < nq) {
                        float *p = a.g_in + (int64_t)xrow * ld + 2 * q;
                        if (dx[t].x != 0.f) unsafeAtomicAdd(p, dx[t].x);
                        if (dx[t].y != 0.f) unsafeAtomicAdd(p + 1, dx[t].y);
                    }
                }
            }
        }
        if (ADJ) {
#pragma unroll
            for (int k = 0; k < kMaxY; k++) {
                if (k >= ny) continue;
#pragma unroll
                for (int t = 0; t < NP; t++) {
                    const int q = lane + kWave * t;
                    if (q < nq) {
                        float *p = a.g_out + (int64_t)yrow[k] * ld + 2 * q;
                        if (dy[k][t].x != 0.f) unsafeAtomicAdd(p, dy[k][t].x);
                        if (dy[k][t].y != 0.f) unsafeAtomicAdd(p + 1, dy[k][t].y);
                    }
                }
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

}  // namespace

hipError_t launch_window(const WindowArgs &a, int bf16, int np, int adjoint, int flags, hipStream_t st) {
    const int fp64 = (flags & 2) != 0;
    if (flags & 1) return hipErrorNotSupported;  // matrix-core path: see window_mfma
    return bf16 ? launch_t<uint16_t>(a, np, adjoint, fp64, st) : launch_t<float>(a, np, adjoint, fp64, st);
}

}  // namespace sgns
