// The ensemble kernels built with 256-thread workgroups (namespace fi256), two per CU: members
// up to 256 cells wide, whose rows phase gives one thread per column at this width. Each CU then
// interleaves two workgroups' phase chains (load rounds, LDS transform passes), which a single
// 512-thread workgroup per CU leaves waiting on one another. The parameter blocks, control blocks
// and launch arguments have the layout of namespace fi (same source).
#define FI_NT 256
#define FI_NS fi256
#define FI_BATCH_BOUNDS __launch_bounds__(256, 2)
#include "fi_loop.hpp"

namespace fi256 {

const void* batch_kernel_nt256(bool dist, bool elra) {
    if (elra) return (const void*)fi_batch_persistent_elra;
    return dist ? (const void*)fi_batch_persistent<double, true> : (const void*)fi_batch_persistent<double, false>;
}

}  // namespace fi256
