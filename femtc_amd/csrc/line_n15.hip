// Line-kernel instances for N = 15 dofs per direction (degree 14),
// Q in {N, N+1, N+2}, mass / stiffness / K+M, FP64 and FP32.
#include "apply_line.hpp"

namespace femtc
{
int launch_line_n15(const ApplyReq& r, int max_lds)
{
  return line_dispatch<15>(r, max_lds);
}
} // namespace femtc
