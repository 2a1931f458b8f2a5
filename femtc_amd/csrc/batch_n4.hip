// Batched-launch instances for N = 4 dofs per direction (degree 3),
// Q in {N, N+1, N+2}, FP64 and FP32, cell-local and gathered.
#include "apply_kernel.hpp"

namespace femtc
{
int launch_batch_n4(const BatchReq& r) { return launch_batch_n<4>(r); }
} // namespace femtc
