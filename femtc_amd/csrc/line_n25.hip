// Line-kernel instances for N = 25 dofs per direction (degree 24),
// Q in {N, N+1, N+2}, mass / stiffness / K+M, FP64 and FP32, where they fit
// (line_fits; the rest run on the general kernel).
#include "apply_line.hpp"

namespace femtc
{
int launch_line_n25(const ApplyReq& r, int max_lds)
{
  return line_dispatch<25>(r, max_lds);
}
} // namespace femtc
