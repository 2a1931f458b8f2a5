// Request server: apply and solve requests on registered problems, served by
// a fixed number of slots (host threads, one HIP stream each) so that work
// admitted while other requests are running overlaps on the device
// (continuous batching).  Solve requests keep their solver (Jacobi diagonal,
// Krylov and refinement vectors, lower-precision geometry copy) resident per
// (problem, inner precision, slot).  Admission, queueing and eviction are the
// host-only Admission class (admission.hpp), run here under the server's
// lock with real solvers and the real clock: a request that does not fit
// waits while later ones that do fit go ahead, idle solvers are evicted
// least-recently-used first, and they are destroyed after the lock is
// released, so a slot's bookkeeping never waits on another's hipFree.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "admission.hpp"
#include "assemble.hpp"
#include "common.hpp"

namespace femtc
{
namespace
{

using Clock = std::chrono::steady_clock;

// Estimated device bytes of a solver of problem p (femtc_solver_create's
// allocations); the admission reserves this much until the solver exists
// and then charges what femtc_solver_bytes reports.
size_t solver_bytes(const femtc_problem& p, int inner)
{
  const size_t n = (size_t)p.ndofs;
  const size_t q3 = (size_t)p.nq * p.nq * p.nq;
  const size_t C = (size_t)p.num_cells;
  const size_t N3 = (size_t)(p.degree + 1) * (p.degree + 1) * (p.degree + 1);
  size_t b = 4 * 8 * n;                    // diag, D^-1, r, A u (FP64)
  const size_t vs = inner == FEMTC_F64 ? 8 : 4;
  b += 4 * vs * n;                         // rhs, e, p, A p
  b += 8 * C * 4;                          // energy partials (<= 4 per cell)
  if (p.J)
    b += 8 * CELLGEO * C;                  // per-cell geometry, FP64
  if (p.grid[0] > 0)                       // assembled-scatter lattice
    b += vs * C * N3;
  if (inner != FEMTC_F64)
  {
    b += (4 + 2) * n;                      // r, D^-1 (bfloat16)
    if (p.J)
      b += 4 * CELLGEO * C;                // per-cell geometry, FP32
    else
    {
      if (p.G)
        b += 4 * 6 * q3 * C;
      if (p.detJw)
        b += 4 * q3 * C;
    }
  }
  return b + (1 << 20); // partial sums, scalars, tables, operand sets
}

} // namespace
} // namespace femtc

using namespace femtc;

struct femtc_server
{
  struct Job
  {
    femtc_request r;
    femtc_problem p; // the registered copy
    uint64_t key;
    double t_arrival;
  };
  struct Registered
  {
    uint64_t key;
    femtc_problem copy;
  };
  int slots = 1;
  int device = 0;
  Admission adm{0};
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::unordered_map<int64_t, Job> jobs; // queued and running requests
  int64_t next_id = 0;
  std::unordered_map<const femtc_problem*, Registered> reg;
  std::unordered_map<uint64_t, int> running; // requests running per problem
  std::vector<femtc_request_result> done;
  int inflight = 0;
  bool stop = false;
  std::vector<std::thread> workers;
  Clock::time_point t0;

  double now() const
  {
    return std::chrono::duration<double>(Clock::now() - t0).count();
  }

  static void destroy_all(std::vector<void*>& v)
  {
    for (void* h : v)
      femtc_solver_destroy(static_cast<femtc_solver*>(h));
    v.clear();
  }

  // Results of requests dropped with their problem (lock held).
  void drop(const std::vector<int64_t>& ids)
  {
    for (int64_t id : ids)
    {
      auto it = jobs.find(id);
      if (it == jobs.end())
        continue;
      femtc_request_result res{};
      res.tag = it->second.r.tag;
      res.kind = it->second.r.kind;
      res.status = FEMTC_ERR_INVALID_ARGUMENT;
      res.slot = -1;
      res.t_arrival = it->second.t_arrival;
      res.t_start = res.t_end = now();
      done.push_back(res);
      jobs.erase(it);
    }
  }

  void worker(int slot)
  {
    (void)hipSetDevice(device);
    hipStream_t st = nullptr;
    (void)hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    std::vector<void*> victims;
    for (;;)
    {
      AdmTake t;
      Job job{};
      bool got = false;
      {
        std::unique_lock<std::mutex> lk(mu);
        for (;;)
        {
          if (adm.take(slot, &t, &victims))
          {
            job = jobs.at(t.req.id);
            ++running[job.key];
            ++inflight;
            got = true;
            break;
          }
          if (stop && adm.queue_empty())
            break;
          cv_work.wait(lk);
        }
      }
      destroy_all(victims); // evicted for this admission, outside the lock
      if (!got)
        break;
      femtc_request_result res{};
      res.tag = job.r.tag;
      res.kind = job.r.kind;
      res.slot = slot;
      res.t_arrival = job.t_arrival;
      res.t_start = now();
      const femtc_problem& p = job.p;
      femtc_solver* solver = static_cast<femtc_solver*>(t.cached);
      size_t actual = 0;
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
        if (!solver)
        {
          res.status = femtc_solver_create(&solver, &p, job.r.inner_dtype, st);
          if (res.status == FEMTC_OK)
            (void)femtc_solver_bytes(solver, &actual);
          else
            solver = nullptr;
        }
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
      void* dead = nullptr;
      {
        std::lock_guard<std::mutex> lk(mu);
        if (job.r.kind == FEMTC_REQUEST_SOLVE)
        {
          if (t.cached)
            dead = adm.release(t, slot);
          else if (solver)
            dead = adm.adopt(t, slot, solver, actual);
          else
            adm.cancel(t);
        }
        jobs.erase(t.req.id);
        --running[job.key];
        done.push_back(res);
        --inflight;
      }
      if (dead) // its problem was unregistered meanwhile
        femtc_solver_destroy(static_cast<femtc_solver*>(dead));
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
  s->adm = Admission(hbm_budget);
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

int femtc_solver_workspace_bytes(const femtc_problem* p, int inner_dtype,
                                 size_t* bytes)
{
  if (!p || !bytes)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "NULL argument");
  *bytes = solver_bytes(*p, inner_dtype);
  return FEMTC_OK;
}

int femtc_server_register(femtc_server* s, const femtc_problem* p)
{
  if (!s || !p)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->reg.count(p))
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "problem is registered already");
  s->reg[p] = femtc_server::Registered{s->adm.add_problem(), *p};
  return FEMTC_OK;
}

int femtc_server_unregister(femtc_server* s, const femtc_problem* p)
{
  if (!s || !p)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "NULL argument");
  std::vector<void*> victims;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reg.find(p);
    if (it == s->reg.end())
      return fail(FEMTC_ERR_INVALID_ARGUMENT, "problem is not registered");
    const uint64_t key = it->second.key;
    s->reg.erase(it);
    std::vector<int64_t> dropped;
    s->adm.remove_problem(key, &victims, &dropped);
    s->drop(dropped);
    // its running requests finish (their solvers are destroyed by their
    // slots) before the caller may free the problem's buffers
    s->cv_done.wait(lk, [&] { return s->running[key] == 0; });
    s->running.erase(key);
  }
  femtc_server::destroy_all(victims);
  s->cv_done.notify_all();
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
       (r->inner_dtype != FEMTC_F64 && r->inner_dtype != FEMTC_F32 &&
        r->inner_dtype != FEMTC_BF16 && r->inner_dtype != FEMTC_BF16_ROUNDED)))
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "bad solve parameters");
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop)
      return fail(FEMTC_ERR_INVALID_ARGUMENT, "server is stopping");
    auto it = s->reg.find(r->problem);
    if (it == s->reg.end())
      return fail(FEMTC_ERR_INVALID_ARGUMENT,
                  "the request's problem is not registered "
                  "(femtc_server_register)");
    const size_t bytes = r->kind == FEMTC_REQUEST_SOLVE
                             ? solver_bytes(it->second.copy, r->inner_dtype)
                             : 0;
    // a request that could never be admitted would wait forever
    if (bytes > s->adm.budget())
      return fail(FEMTC_ERR_OUT_OF_MEMORY,
                  "the solve workspace exceeds the server's HBM budget");
    const int64_t id = ++s->next_id;
    const double t = s->now();
    s->jobs[id] = femtc_server::Job{*r, it->second.copy, it->second.key, t};
    s->adm.push(AdmReq{id, r->kind, it->second.key, r->inner_dtype, bytes, t});
  }
  s->cv_work.notify_all();
  return FEMTC_OK;
}

int femtc_server_drain(femtc_server* s)
{
  if (!s)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "server is NULL");
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk,
                  [s] { return s->adm.queue_empty() && s->inflight == 0; });
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
    *used = s->adm.used();
  if (cached)
    *cached = s->adm.resident();
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
  std::vector<void*> victims;
  s->adm.drain(&victims);
  femtc_server::destroy_all(victims);
  delete s;
  return FEMTC_OK;
}

} // extern "C"
