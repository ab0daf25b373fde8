// fast_kernel.hip -- launch dispatch of the Hogwild trainer (kernel templates in
// fast_kernel.h; each group of instances is compiled in its own translation
// unit, fast_inst_*.hip), the sampler probe and the counter reset.
#include "fast_kernel.h"

namespace sgns {
namespace {

// n draws of the sampler for sentence id `sent` (the counter layout of the
// trainer's negatives), for testing the on-device distribution.
__global__ void sgns_sample_kernel(const uint2 *alias, int V, uint64_t seed, int64_t sent, int64_t n, int32_t *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = fast::alias_draw(alias, V, draw(seed, sent + (i >> 13), 0, kTagNeg, (uint32_t)(i & 8191)));
}

__global__ void sgns_reset_kernel(unsigned long long *counters, int n, double *loss, unsigned long long *queue) {
    const int i = threadIdx.x;
    if (i < n) counters[i] = 0;
    if (i < 2) loss[i] = 0.0;
    if (i == 0 && queue) *queue = 0;
}

}  // namespace

hipError_t launch_fast(const FastArgs &a, int bf16, int np, int wpb, int blocks, hipStream_t st) {
    if (!bf16) return fast_vector_fp32(a, a.write_back == kWbPolicyStoreOut, np, wpb, blocks, st);
    if (!a.rep_mfma) return fast_vector_bf16(a, a.write_back == kWbPolicyStoreOut, np, wpb, blocks, st);
    switch (a.write_back) {
    case kWbPolicyAtomic: return fast_mfma_atomic(a, np, wpb, blocks, st);
    case kWbPolicyStoreOut: return fast_mfma_store_out(a, np, wpb, blocks, st);
    case kWbPolicyStoreIn:
    case kWbPolicyStore: return fast_mfma_store(a, a.write_back == kWbPolicyStoreIn, np, wpb, blocks, st);
    case kWbPolicyHot: return fast_mfma_hot(a, np, wpb, blocks, st);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_sample(const uint2 *alias, int V, uint64_t seed, int64_t sent, int64_t n, int32_t *out,
                         hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(sgns_sample_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, alias, V, seed, sent,
                       n, out);
    return hipGetLastError();
}

hipError_t launch_reset(unsigned long long *counters, int n, double *loss, unsigned long long *queue, hipStream_t st) {
    hipLaunchKernelGGL(sgns_reset_kernel, dim3(1), dim3(64), 0, st, counters, n, loss, queue);
    return hipGetLastError();
}

}  // namespace sgns
