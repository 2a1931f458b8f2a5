// Batched-launch instances for N = 6 dofs per direction (degree 5),
// Q in {N, N+1, N+2}, FP64 and FP32, cell-local and gathered.
#include "apply_kernel.hpp"

namespace femtc
{
int launch_batch_n6(const BatchReq& r) { return launch_batch_n<6>(r); }
} // namespace femtc
