// fast_inst_mfma_hot.hip -- training-kernel instances: bf16 tables, matrix-core repulsion, atomics below the hot row bound and plain stores past it.
// (One group of template instances per translation unit, so they compile in
// parallel; the templates are in fast_kernel.h.)
#include "fast_kernel.h"

namespace sgns {

hipError_t fast_mfma_hot(const FastArgs &a, int np, int wpb, int blocks, hipStream_t st) {
    return fast::launch_mfma<fast::kWbHot, fast::kWbHot>(a, np, wpb, blocks, st);
}

}  // namespace sgns
