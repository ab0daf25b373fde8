// fast_inst_mfma_store_out.hip -- training-kernel instances: bf16 tables, matrix-core repulsion, W_in atomics and W_out plain stores.
// (One group of template instances per translation unit, so they compile in
// parallel; the templates are in fast_kernel.h.)
#include "fast_kernel.h"

namespace sgns {

hipError_t fast_mfma_store_out(const FastArgs &a, int np, int wpb, int blocks, hipStream_t st) {
    return fast::launch_mfma<fast::kWbAtomic, fast::kWbStore>(a, np, wpb, blocks, st);
}

}  // namespace sgns
