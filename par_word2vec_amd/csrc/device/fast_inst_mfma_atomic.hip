// fast_inst_mfma_atomic.hip -- training-kernel instances: bf16 tables, matrix-core repulsion, atomic write-back of both tables.
// (One group of template instances per translation unit, so they compile in
// parallel; the templates are in fast_kernel.h.)
#include "fast_kernel.h"

namespace sgns {

hipError_t fast_mfma_atomic(const FastArgs &a, int np, int wpb, int blocks, hipStream_t st) {
    return fast::launch_mfma<fast::kWbAtomic, fast::kWbAtomic>(a, np, wpb, blocks, st);
}

}  // namespace sgns
