// Line-kernel instances for N = 13 dofs per direction (degree 12),
// Q in {N, N+1, N+2}, mass / stiffness / K+M, FP64 and FP32.
#include "apply_line.hpp"

namespace femtc
{
int launch_line_n13(const ApplyReq& r, int max_lds)
{
  return line_dispatch<13>(r, max_lds);
}
} // namespace femtc
