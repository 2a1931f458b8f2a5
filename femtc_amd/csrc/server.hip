// Request server: apply and solve requests on resident problems, served by
// a fixed number of slots (host threads, one HIP stream each) so that work
// admitted while other requests are running overlaps on the device
// (continuous batching).  Solve requests keep their solver (Jacobi diagonal,
// Krylov and refinement vectors, lower-precision geometry copy) resident per
// (problem, inner precision, slot); admission counts those workspaces
// against an HBM budget and evicts idle ones least-recently-used first.  A
// request that does not fit waits while later ones that do fit go ahead.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "assemble.hpp"
#include "common.hpp"

namespace femtc
{
namespace
{

using Clock = std::chrono::steady_clock;

size_t solver_bytes(const femtc_problem& p, int inner)
{
  const size_t n = (size_t)p.ndofs;
  const size_t q3 = (size_t)p.nq * p.nq * p.nq;
  const size_t C = (size_t)p.num_cells;
  size_t b = 4 * 8 * n;                           // diag, dinv, r, A u (FP64)
  const size_t vs = inner == FEMTC_F64 ? 8 : 4;
  b += 4 * vs * n;                                // rhs, e, p, A p
  if (p.J)
    b += 8 * CELLGEO * C;                         // per-cell geometry, FP64
  if (p.grid[0] > 0)                              // assembled-scatter lattice
    b += vs * (size_t)lattice_size(p.grid[0], p.grid[1], p.grid[2],
                                   p.degree);
  if (inner != FEMTC_F64)
  {
    b += 2 * 4 * n;                               // r, D^-1 (FP32)
    if (p.J)
      b += 4 * CELLGEO * C;                       // per-cell geometry, FP32
    else
    {
      if (p.G)
        b += 4 * 6 * q3 * C;
      if (p.detJw)
        b += 4 * q3 * C;
    }
  }
  return b + 64 * 1024; // partial sums, scalars, tables
}

} // namespace
} // namespace femtc

using namespace femtc;

struct femtc_server
{
  struct Pending
  {
    femtc_request r;
    double t_arrival;
  };
  struct Cached
  {
    const femtc_problem* problem;
    int inner;
    int slot;
    femtc_solver* solver;
    size_t bytes;
    uint64_t last_use;
    bool busy;
  };
  int slots = 1;
  int device = 0;
  size_t budget = 0;
  size_t used = 0;
  uint64_t tick = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Pending> queue;
  std::vector<femtc_request_result> done;
  std::vector<Cached> cache;
  int inflight = 0;
  bool stop = false;
  std::vector<std::thread> workers;
  Clock::time_point t0;

  double now() const
  {
    return std::chrono::duration<double>(Clock::now() - t0).count();
  }

  // Index of a cached solver for (p, inner, slot), or -1.
  int find(const femtc_problem* p, int inner, int slot) const
  {
    for (size_t i = 0; i < cache.size(); ++i)
      if (cache[i].problem == p && cache[i].inner == inner &&
          cache[i].slot == slot)
        return (int)i;
    return -1;
  }

  // Make room for `need` bytes by evicting idle solvers (LRU).  Called with
  // the lock held; returns false if the budget cannot be met now.
  bool make_room(size_t need)
  {
    while (used + need > budget)
    {
      int victim = -1;
      for (size_t i = 0; i < cache.size(); ++i)
        if (!cache[i].busy &&
            (victim < 0 || cache[i].last_use < cache[victim].last_use))
          victim = (int)i;
      if (victim < 0)
        return false;
      femtc_solver_destroy(cache[victim].solver);
      used -= cache[victim].bytes;
      cache.erase(cache.begin() + victim);
    }
    return true;
  }

  void worker(int slot)
  {
    (void)hipSetDevice(device);
    hipStream_t st = nullptr;
    (void)hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    for (;;)
    {
      Pending job{};
      femtc_solver* cached = nullptr; // resident solver taken by this job
      size_t reserve = 0;             // budget reserved for a new solver
      {
        std::unique_lock<std::mutex> lk(mu);
        bool got = false;
        for (;;)
        {
          if (stop && queue.empty())
            break;
          // first admissible request in arrival order
          for (auto it = queue.begin(); it != queue.end(); ++it)
          {
            if (it->r.kind == FEMTC_REQUEST_SOLVE)
            {
              const int c = find(it->r.problem, it->r.inner_dtype, slot);
              if (c >= 0)
              {
                cached = cache[c].solver;
                cache[c].busy = true;
                cache[c].last_use = ++tick;
              }
              else
              {
                const size_t need =
                    solver_bytes(*it->r.problem, it->r.inner_dtype);
                if (!make_room(need))
                  continue; // does not fit now: try a later request
                used += need;
                reserve = need;
              }
            }
            job = *it;
            queue.erase(it);
            got = true;
            break;
          }
          if (got)
            break;
          cv_work.wait(lk);
        }
        if (!got)
          break;
        ++inflight;
      }
      femtc_request_result res{};
      res.tag = job.r.tag;
      res.kind = job.r.kind;
      res.slot = slot;
      res.t_arrival = job.t_arrival;
      res.t_start = now();
      const femtc_problem& p = *job.r.problem;
      femtc_solver* solver = nullptr;
      if (job.r.kind == FEMTC_REQUEST_APPLY)
      {
        res.status = femtc_global_apply(p.op, FEMTC_F64, p.degree, p.nq,
                                        p.tables, job.r.input, job.r.output,
                                        p.ndofs, p.dofmap, p.G, p.detJw,
                                        p.num_cells, st);
        if (res.status == FEMTC_OK && hipStreamSynchronize(st) != hipSuccess)
          res.status = FEMTC_ERR_HIP;
      }
      else
      {
        if (cached)
          solver = cached;
        else
          res.status = femtc_solver_create(&solver, &p, job.r.inner_dtype, st);
        if (res.status == FEMTC_OK)
        {
          femtc_solve_info info{};
          res.status = femtc_solver_solve(
              solver, (const double*)job.r.input, (double*)job.r.output,
              job.r.rtol, job.r.inner_rtol, job.r.maxit, &info);
          res.iterations = info.inner_iterations;
          res.outer_iterations = info.outer_iterations;
          res.relres = info.relres;
        }
      }
      res.t_end = now();
      {
        std::lock_guard<std::mutex> lk(mu);
        if (job.r.kind == FEMTC_REQUEST_SOLVE)
        {
          if (cached)
          {
            // busy entries are never evicted, so it is still there
            const int c = find(job.r.problem, job.r.inner_dtype, slot);
            if (c >= 0)
              cache[c].busy = false;
          }
          else if (solver)
            cache.push_back(Cached{job.r.problem, job.r.inner_dtype, slot,
                                   solver, reserve, ++tick, false});
          else
            used -= reserve; // creation failed
        }
        done.push_back(res);
        --inflight;
      }
      cv_done.notify_all();
      cv_work.notify_all();
    }
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
  }
};

extern "C" {

int femtc_server_create(femtc_server** out, int slots, size_t hbm_budget)
{
  if (!out || slots < 1 || slots > 16)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "server needs 1..16 slots");
  femtc_server* s = new femtc_server;
  s->slots = slots;
  s->budget = hbm_budget;
  s->t0 = Clock::now();
  if (hipGetDevice(&s->device) != hipSuccess)
  {
    delete s;
    return fail(FEMTC_ERR_NO_DEVICE, "no current HIP device");
  }
  for (int i = 0; i < slots; ++i)
    s->workers.emplace_back([s, i] { s->worker(i); });
  *out = s;
  return FEMTC_OK;
}

int femtc_server_submit(femtc_server* s, const femtc_request* r)
{
  if (!s || !r || !r->problem || !r->input || !r->output)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "bad request");
  if (r->kind != FEMTC_REQUEST_APPLY && r->kind != FEMTC_REQUEST_SOLVE)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "kind must be apply or solve");
  if (r->kind == FEMTC_REQUEST_SOLVE &&
      (!(r->rtol > 0.0) || r->maxit < 1 ||
       (r->inner_dtype != FEMTC_F64 && r->inner_dtype != FEMTC_F32)))
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "bad solve parameters");
  // a request that could never be admitted would wait forever
  if (r->kind == FEMTC_REQUEST_SOLVE &&
      solver_bytes(*r->problem, r->inner_dtype) > s->budget)
    return fail(FEMTC_ERR_OUT_OF_MEMORY,
                "the solve workspace exceeds the server's HBM budget");
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop)
      return fail(FEMTC_ERR_INVALID_ARGUMENT, "server is stopping");
    s->queue.push_back(femtc_server::Pending{*r, s->now()});
  }
  s->cv_work.notify_all();
  return FEMTC_OK;
}

int femtc_server_drain(femtc_server* s)
{
  if (!s)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "server is NULL");
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [s] { return s->queue.empty() && s->inflight == 0; });
  return FEMTC_OK;
}

int femtc_server_results(femtc_server* s, femtc_request_result* out,
                         int max, int* n)
{
  if (!s || !n || (max > 0 && !out))
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "bad arguments");
  std::lock_guard<std::mutex> lk(s->mu);
  const int k = std::min<int>(max, (int)s->done.size());
  for (int i = 0; i < k; ++i)
    out[i] = s->done[i];
  s->done.erase(s->done.begin(), s->done.begin() + k);
  *n = k;
  return FEMTC_OK;
}

int femtc_server_memory(const femtc_server* s, size_t* used, int* cached)
{
  if (!s)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "server is NULL");
  std::lock_guard<std::mutex> lk(const_cast<femtc_server*>(s)->mu);
  if (used)
    *used = s->used;
  if (cached)
    *cached = (int)s->cache.size();
  return FEMTC_OK;
}

int femtc_server_destroy(femtc_server* s)
{
  if (!s)
    return FEMTC_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto& t : s->workers)
    t.join();
  for (auto& c : s->cache)
    femtc_solver_destroy(c.solver);
  delete s;
  return FEMTC_OK;
}

} // extern "C"
