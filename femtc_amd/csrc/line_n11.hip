// Line-kernel instances for N = 11 dofs per direction (degree 10),
// Q in {N, N+1, N+2}, mass / stiffness / K+M, FP64 and FP32.
#include "apply_line.hpp"

namespace femtc
{
int launch_line_n11(const ApplyReq& r, int max_lds)
{
  return line_dispatch<11>(r, max_lds);
}
} // namespace femtc
