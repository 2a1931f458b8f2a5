// Randomised runs of the server's admission code through
// femtc_admission_simulate, for AddressSanitizer / UndefinedBehaviorSanitizer
// builds (tests/test_sanitizers.py): many problems, removals while requests
// run, oversized requests, tight budgets.  Checks the invariants too.
#include <cstdio>
#include <random>
#include <vector>

#include "femtc.h"

int main()
{
  std::mt19937_64 rng(7);
  for (int run = 0; run < 200; ++run)
  {
    const int n = 1 + (int)(rng() % 300), np = 1 + (int)(rng() % 9);
    const int slots = 1 + (int)(rng() % 6);
    const size_t budget = (size_t)(1 + rng() % 1000) << 20;
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::vector<femtc_sim_request> q(n);
    double t = 0.0;
    for (auto& r : q)
    {
      t += U(rng) * 1e-3;
      r.t_arrival = t;
      r.kind = U(rng) < 0.5 ? FEMTC_REQUEST_APPLY : FEMTC_REQUEST_SOLVE;
      r.problem = (int)(rng() % np);
      r.inner = (int)(rng() % 4);
      r.bytes = (size_t)(1 + rng() % 400) << 20;
      r.actual = r.bytes / (1 + rng() % 3);
      r.service = U(rng) * 2e-3;
      r.create = U(rng) * 1e-3;
    }
    std::vector<double> remove_at(np, -1.0);
    for (double& x : remove_at)
      if (U(rng) < 0.3)
        x = U(rng) * t;
    std::vector<femtc_sim_result> out(n);
    femtc_sim_stats st{};
    if (femtc_admission_simulate(q.data(), n, np, remove_at.data(), slots,
                                 budget, out.data(), &st) != FEMTC_OK)
      return 1;
    if (st.high_water > budget || st.left_in_queue != 0)
    {
      std::printf("run %d: budget or queue invariant broken\n", run);
      return 2;
    }
    for (int i = 0; i < n; ++i)
      if (out[i].status == FEMTC_OK && out[i].t_start < q[i].t_arrival - 1e-12)
        return 3;
  }
  std::printf("admission driver done\n");
  return 0;
}
