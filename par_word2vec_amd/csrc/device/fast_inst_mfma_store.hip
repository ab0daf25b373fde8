// fast_inst_mfma_store.hip -- training-kernel instances: bf16 tables, matrix-core repulsion, W_in plain stores (W_out atomics or stores).
// (One group of template instances per translation unit, so they compile in
// parallel; the templates are in fast_kernel.h.)
#include "fast_kernel.h"

namespace sgns {

hipError_t fast_mfma_store(const FastArgs &a, int store_in_only, int np, int wpb, int blocks, hipStream_t st) {
    return store_in_only ? fast::launch_mfma<fast::kWbStore, fast::kWbAtomic>(a, np, wpb, blocks, st)
                         : fast::launch_mfma<fast::kWbStore, fast::kWbStore>(a, np, wpb, blocks, st);
}

}  // namespace sgns
