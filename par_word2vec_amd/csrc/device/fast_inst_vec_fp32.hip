// fast_inst_vec_fp32.hip -- training-kernel instances: fp32 tables, vector repulsion.
// (One group of template instances per translation unit, so they compile in
// parallel; the templates are in fast_kernel.h.)
#include "fast_kernel.h"

namespace sgns {

hipError_t fast_vector_fp32(const FastArgs &a, int store_out, int np, int wpb, int blocks, hipStream_t st) {
    return store_out ? fast::launch_vector<float, fast::kWbStore>(a, np, wpb, blocks, st)
                     : fast::launch_vector<float, fast::kWbAtomic>(a, np, wpb, blocks, st);
}

}  // namespace sgns
