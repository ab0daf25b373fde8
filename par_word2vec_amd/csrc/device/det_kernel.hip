// det_kernel.hip -- deterministic single-wavefront PAR-SGNS executor.
//
// Executes an uploaded schedule (kept tokens, window shrink b per centre,
// shared negatives per repulsion batch, learning rate per centre) with the
// exact arithmetic of the reference CPU trainer, PAR_Word2Vec_cpu.cpp:
//   attraction cpu:536-567: f = sequential float sum of products over j,
//     g from expTable, then per element tmp = g*W_out; W_out += g*W_in;
//     W_in += tmp;
//   repulsion cpu:624-674: gather, G = M_out M_in^T, label-0 gradient,
//     M_out_upd = G M_in, M_in' = G^T M_out, scatter in ascending order;
//     the three products use an ascending-k sum from 0.0f (the oracle's
//     substitute for MKL cblas_sgemm).
// One wavefront, no atomics: results are bit-identical to the oracle built
// without FMA contraction.  This translation unit is compiled with
// -ffp-contract=off so no multiply-add pair is fused.
#include "sgns_device.h"
#include "sgns_internal.h"

namespace sgns {
namespace {

__device__ __forceinline__ void wave_fence() {
    // scratch rows written by one lane are read by another lane of the same
    // wave: wait for the stores and order them at work-group scope
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
}

__device__ __forceinline__ double softplus_d(double x) { return x > 0 ? x + log1p(exp(-x)) : log1p(exp(x)); }

__global__ __launch_bounds__(64) void sgns_det_kernel(DetArgs a) {
#pragma clang fp contract(off)
    __shared__ float et[kExpTable + 1];
    const int lane = threadIdx.x;
    for (int i = lane; i <= kExpTable; i += kWave) et[i] = a.exp_table[i];
    __syncthreads();

    const int d = a.d, ld = a.ld, W = a.W, B = a.B;
    float *__restrict__ w_in = a.w_in;
    float *__restrict__ w_out = a.w_out;
    double loss_a = 0.0, loss_r = 0.0;
    uint64_t n_tok = 0, n_sent = 0, n_pairs = 0, n_bat = 0, n_neg = 0, n_rep = 0;

    for (int64_t s = a.s_begin; s < a.s_end; s++) {
        const int64_t o0 = a.sent_off[s];
        const int L = (int)(a.sent_off[s + 1] - o0);
        const int32_t *sen = a.tokens + o0;
        const uint8_t *bb = a.b + o0;
        const float a_first = a.sent_alpha[2 * s], a_rest = a.sent_alpha[2 * s + 1];
        n_sent++;
        n_tok += (uint64_t)L;

        // ---- attraction, one centre at a time (cpu:529-568)
        for (int p = 0; p < L; p++) {
            const float alpha = p == 0 ? a_first : a_rest;
            const int b = bb[p];
            const int64_t l2 = (int64_t)sen[p] * ld;
            for (int i = b; i < 2 * W + 1 - b; i++) {
                if (i == W) continue;
                const int c = p - W + i;
                if (c < 0) continue;
                if (c >= L) break;
                const int64_t l1 = (int64_t)sen[c] * ld;
                float f = 0.0f;
                for (int j0 = 0; j0 < d; j0 += kWave) {
                    const int j = j0 + lane;
                    const float pr = j < d ? w_in[l1 + j] * w_out[l2 + j] : 0.0f;
                    const int n = d - j0 < kWave ? d - j0 : kWave;
                    for (int q = 0; q < n; q++) f = f + readlane_f(pr, q);
                }
                const float g = grad_label1(f, alpha, et);
                for (int j = lane; j < d; j += kWave) {
                    const float wo = w_out[l2 + j], wi = w_in[l1 + j];
                    const float tmp = g * wo;
                    w_out[l2 + j] = wo + g * wi;
                    w_in[l1 + j] = wi + tmp;
                }
                n_pairs++;
                if (a.compute_loss && lane == 0) loss_a += softplus_d(-(double)f);
            }
        }

        // ---- repulsion over all batches (cpu:571-679)
        const float alpha = L == 1 ? a_first : a_rest;
        const int nb = (L + B - 1) / B;
        const int64_t bo = a.batch_off[s];
        for (int bi = 0; bi < nb; bi++) {
            const int lo = bi * B;
            const int isz = L - lo < B ? L - lo : B;
            const int64_t n0 = a.neg_off[bo + bi];
            const int K = (int)(a.neg_off[bo + bi + 1] - n0);
            const int32_t *neg = a.negs + n0;
            n_bat++;
            n_neg += (uint64_t)K;
            n_rep += (uint64_t)K * (uint64_t)isz;
            if (K == 0) continue;  // SGEMMs with output_size 0 change nothing
            for (int i = 0; i < isz; i++)
                for (int j = lane; j < d; j += kWave) a.m_in[(int64_t)i * ld + j] = w_in[(int64_t)sen[lo + i] * ld + j];
            for (int k = 0; k < K; k++)
                for (int j = lane; j < d; j += kWave) a.m_out[(int64_t)k * ld + j] = w_out[(int64_t)neg[k] * ld + j];
            wave_fence();
            for (int idx = lane; idx < K * isz; idx += kWave) {  // G = M_out M_in^T
                const int k = idx / isz, i = idx - k * isz;
                const float *x = a.m_out + (int64_t)k * ld, *y = a.m_in + (int64_t)i * ld;
                float acc = 0.0f;
                for (int j = 0; j < d; j++) acc = acc + x[j] * y[j];
                if (a.compute_loss) loss_r += softplus_d((double)acc);
                a.m_grad[k * B + i] = grad_label0(acc, alpha, et);
            }
            wave_fence();
            for (int idx = lane; idx < K * d; idx += kWave) {  // M_out_upd = G M_in
                const int k = idx / d, j = idx - k * d;
                float acc = 0.0f;
                for (int i = 0; i < isz; i++) acc = acc + a.m_grad[k * B + i] * a.m_in[(int64_t)i * ld + j];
                a.m_out_upd[(int64_t)k * ld + j] = acc;
            }
            for (int idx = lane; idx < isz * d; idx += kWave) {  // M_in' = G^T M_out
                const int i = idx / d, j = idx - i * d;
                float acc = 0.0f;
                for (int k = 0; k < K; k++) acc = acc + a.m_grad[k * B + i] * a.m_out[(int64_t)k * ld + j];
                a.m_in_new[(int64_t)i * ld + j] = acc;
            }
            wave_fence();
            for (int i = 0; i < isz; i++) {
                const int64_t des = (int64_t)sen[lo + i] * ld;
                for (int j = lane; j < d; j += kWave) w_in[des + j] = w_in[des + j] + a.m_in_new[(int64_t)i * ld + j];
            }
            for (int k = 0; k < K; k++) {
                const int64_t des = (int64_t)neg[k] * ld;
                for (int j = lane; j < d; j += kWave) w_out[des + j] = w_out[des + j] + a.m_out_upd[(int64_t)k * ld + j];
            }
            wave_fence();
        }
    }

    // FP64 loss: per-lane partials summed by lane 0 (loss_a lives in lane 0)
    if (a.compute_loss) {
        __shared__ double red[kWave];
        red[lane] = loss_r;
        __syncthreads();
        if (lane == 0) {
            double t = 0.0;
            for (int i = 0; i < kWave; i++) t += red[i];
            atomicAdd(&a.loss[0], loss_a);
            atomicAdd(&a.loss[1], t);
        }
    }
    if (lane == 0) {
        atomicAdd(&a.counters[kCntTokens], (unsigned long long)n_tok);
        atomicAdd(&a.counters[kCntSentences], (unsigned long long)n_sent);
        atomicAdd(&a.counters[kCntPairs], (unsigned long long)n_pairs);
        atomicAdd(&a.counters[kCntBatches], (unsigned long long)n_bat);
        atomicAdd(&a.counters[kCntNegatives], (unsigned long long)n_neg);
        atomicAdd(&a.counters[kCntRepPairs], (unsigned long long)n_rep);
    }
}

}  // namespace

hipError_t launch_det(const DetArgs &a, hipStream_t st) {
    hipLaunchKernelGGL(sgns_det_kernel, dim3(1), dim3(kWave), 0, st, a);
    return hipGetLastError();
}

}  // namespace sgns
