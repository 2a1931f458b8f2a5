// Batched operator applies: many independent problems (meshes) of uneven
// size, degree, operator and precision in a few launches.  Problems that
// share (op, dtype, degree, nq, gather) form a group and run as ONE launch
// of apply_batch (each workgroup looks its problem up by block index) or as
// one launch per problem, as femtc_batch_autotune finds faster; the
// launches are spread over four forked streams, longest first, so that
// small ones overlap; the whole sequence can be captured once into a
// hipGraph and replayed.
#include <algorithm>
#include <map>
#include <tuple>
#include <vector>

#include "assemble.hpp"
#include "common.hpp"

namespace femtc
{
namespace
{

struct FillDesc
{
  void* p;
  long long n;
  int dtype;
};

__global__ void __launch_bounds__(256)
    k_fill_batch(const FillDesc* __restrict__ d)
{
  const FillDesc f = d[blockIdx.y];
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < f.n;
       i += (long long)gridDim.x * 256)
  {
    if (f.dtype == FEMTC_F64)
      ((double*)f.p)[i] = 0.0;
    else
      ((float*)f.p)[i] = 0.0f;
  }
}

BatchLauncher batch_launcher(int N)
{
  switch (N)
  {
  case 2: return launch_batch_n2;
  case 3: return launch_batch_n3;
  case 4: return launch_batch_n4;
  case 5: return launch_batch_n5;
  case 6: return launch_batch_n6;
  case 7: return launch_batch_n7;
  case 8: return launch_batch_n8;
  case 9: return launch_batch_n9;
  }
  return nullptr;
}

int cells_per_wg(int nq) { return direct_cells(nq); }

constexpr int NSTREAMS = 4;

// Per problem: its verified structured grid (0 if none) and, once the
// assembled scatter is selected for its group, its lattice buffer.
struct Extra
{
  int grid[3] = {0, 0, 0};
  void* lattice = nullptr;
};

struct Group
{
  BatchReq req;
  size_t first = 0; // offset of its descriptors in the device array
  std::vector<BatchDesc> host;
  std::vector<Extra> extra; // parallel to host
  std::vector<FillDesc> fill; // continuous groups: y of each item
  // the group's clears in the device fill array while the autotuner runs
  size_t fill_first = 0;
  int nfill = 0;
  long long max_fill = 0;
  long long cells = 0;
  bool structured = true; // every problem has a verified grid
  // 0: one fused apply_batch launch; 1: one launch per problem; 2: the
  // assembled scatter per problem; 3: the structured z-run scatter per
  // problem; 4: one fused z-run launch over the group's problems
  int per_problem = 1;
  // the fused z-run launch's descriptors (segment offsets, grids), built
  // when that mode is first selected or timed
  BatchDesc* zdev = nullptr;
  long long zseg_total = 0;
  // time of one launch of the group in its mode, as the autotuner measured
  // it (0: not measured; the scheduler then estimates it from the bytes)
  double est_ms = 0.0;
};

// One unit of the batch's schedule: a group's fused launch (item < 0) or
// one problem of a group that runs one launch per problem.
struct Task
{
  int g;
  int item;
};

// Lattice buffers of a group's problems (allocated once).
int ensure_lattices(Group& g)
{
  if (!g.structured)
    return fail(FEMTC_ERR_INVALID_ARGUMENT,
                "assembled scatter: a problem of the group has no verified "
                "structured grid");
  const size_t es = g.req.dtype == FEMTC_F64 ? 8 : 4;
  for (Extra& e : g.extra)
  {
    if (e.lattice)
      continue;
    const size_t bytes =
        es * (size_t)lattice_size(e.grid[0], e.grid[1], e.grid[2],
                                  g.req.N - 1);
    if (hipMalloc(&e.lattice, bytes) != hipSuccess)
    {
      (void)hipGetLastError();
      e.lattice = nullptr;
      return fail(FEMTC_ERR_OUT_OF_MEMORY,
                  "assembled scatter buffer allocation failed");
    }
  }
  return FEMTC_OK;
}

// Descriptors of the fused z-run launch: per problem its grid and its
// first z-segment (the segment length of the group's instance).
int ensure_zrun(Group& g)
{
  if (g.zdev)
    return FEMTC_OK;
  if (!g.structured || !g.req.gather || g.req.Q > 8)
    return fail(FEMTC_ERR_INVALID_ARGUMENT,
                "z-run scatter: every problem of the group needs a verified "
                "structured grid, and nq <= 8");
  BatchReq q = g.req;
  q.zrun = 1;
  int shape[2] = {0, 0};
  q.zrun_shape = shape;
  BatchLauncher f = batch_launcher(q.N);
  int st = f ? f(q) : FEMTC_ERR_DEGREE_NOT_SUPPORTED;
  if (st != FEMTC_OK)
    return fail(st, "z-run scatter: no instance for the group");
  std::vector<BatchDesc> zh = g.host;
  long long total = 0;
  for (size_t i = 0; i < zh.size(); ++i)
  {
    const Extra& e = g.extra[i];
    for (int k = 0; k < 3; ++k)
      zh[i].grid[k] = e.grid[k];
    zh[i].blk_begin = total;
    total += (long long)e.grid[0] * e.grid[1] *
             ((e.grid[2] + shape[0] - 1) / shape[0]);
  }
  if (hipMalloc((void**)&g.zdev, zh.size() * sizeof(BatchDesc)) != hipSuccess)
  {
    (void)hipGetLastError();
    g.zdev = nullptr;
    return fail(FEMTC_ERR_OUT_OF_MEMORY, "batch descriptor allocation failed");
  }
  if (hipMemcpy(g.zdev, zh.data(), zh.size() * sizeof(BatchDesc),
                hipMemcpyHostToDevice) != hipSuccess)
  {
    (void)hipFree(g.zdev);
    g.zdev = nullptr;
    return fail(FEMTC_ERR_HIP, "batch descriptor upload failed");
  }
  g.zseg_total = total;
  return FEMTC_OK;
}

void free_lattices(Group& g)
{
  for (Extra& e : g.extra)
    if (e.lattice)
    {
      (void)hipFree(e.lattice);
      e.lattice = nullptr;
    }
}

int launch_item(const Group& g, size_t i, hipStream_t s);

// Launch a group: fused, or one launch per problem (atomic, assembled or
// z-run scatter).
int launch_group(const Group& g, hipStream_t s)
{
  if (!g.per_problem || g.per_problem == 4)
  {
    BatchReq r = g.req;
    r.stream = s;
    if (g.per_problem == 4)
    {
      if (!g.zdev)
        return FEMTC_ERR_INVALID_ARGUMENT;
      r.zrun = 1;
      r.items = g.zdev;
      r.nblk_total = g.zseg_total;
    }
    BatchLauncher f = batch_launcher(r.N);
    return f ? f(r) : FEMTC_ERR_DEGREE_NOT_SUPPORTED;
  }
  for (size_t i = 0; i < g.host.size(); ++i)
  {
    const int st = launch_item(g, i, s);
    if (st != FEMTC_OK)
      return st;
  }
  return FEMTC_OK;
}

// Problem i of a group that runs one launch per problem.
int launch_item(const Group& g, size_t i, hipStream_t s)
{
  {
    const BatchDesc& d = g.host[i];
    ApplyReq a;
    if (g.per_problem == 2)
    {
      const Extra& e = g.extra[i];
      a.assembled = 1;
      a.lattice = e.lattice;
      a.ncx = e.grid[0];
      a.ncy = e.grid[1];
      a.ncz = e.grid[2];
      a.overwrite = 1; // y = A x, whatever y held
    }
    else if (g.per_problem == 3)
    {
      const Extra& e = g.extra[i];
      a.zrun = 1; // atomics into y, cleared by the batch's fill
      a.ncx = e.grid[0];
      a.ncy = e.grid[1];
      a.ncz = e.grid[2];
    }
    a.op = g.req.op;
    a.dtype = g.req.dtype;
    a.N = g.req.N;
    a.Q = g.req.Q;
    a.gather = g.req.gather;
    a.tab = d.tab;
    a.x = d.x;
    a.y = d.y;
    a.G = d.G;
    a.detJw = d.detJw;
    a.dofmap = d.dofmap;
    a.num_cells = d.num_cells;
    a.stream = s;
    return run(a);
  }
}

} // namespace
} // namespace femtc

using namespace femtc;

struct femtc_batch
{
  std::vector<Group> groups;
  BatchDesc* d_desc = nullptr;
  FillDesc* d_fill = nullptr; // room for every continuous item
  // continuous items without cells: y = 0 is all they need (their clears
  // come first in d_fill)
  std::vector<FillDesc> empty_fill;
  long long max_empty_fill = 0;
  hipStream_t stream = nullptr;           // the caller's stream
  hipStream_t side[NSTREAMS] = {};        // forked streams
  hipEvent_t ev_fork = nullptr;
  hipEvent_t ev_join[NSTREAMS] = {};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int64_t total_cells = 0;
  int64_t total_dofs = 0;
  // the schedule: each side stream's tasks in launch order, and the range
  // of d_fill it clears before its first task
  std::vector<Task> tasks[NSTREAMS];
  size_t sfill_first[NSTREAMS] = {};
  int sfill_n[NSTREAMS] = {};
  long long sfill_max[NSTREAMS] = {};
};

namespace
{

void release(femtc_batch* b)
{
  if (!b)
    return;
  if (b->exec)
    (void)hipGraphExecDestroy(b->exec);
  if (b->graph)
    (void)hipGraphDestroy(b->graph);
  for (int i = 0; i < NSTREAMS; ++i)
  {
    if (b->side[i])
      (void)hipStreamDestroy(b->side[i]);
    if (b->ev_join[i])
      (void)hipEventDestroy(b->ev_join[i]);
  }
  if (b->ev_fork)
    (void)hipEventDestroy(b->ev_fork);
  if (b->d_desc)
    (void)hipFree(b->d_desc);
  if (b->d_fill)
    (void)hipFree(b->d_fill);
  for (Group& g : b->groups)
  {
    free_lattices(g);
    if (g.zdev)
      (void)hipFree(g.zdev);
  }
  delete b;
}

int check_item(const femtc_batch_item& it, int i)
{
  const std::string where = "batch item " + std::to_string(i) + ": ";
  int st = check_shape_fused(it.degree, it.nq, FEMTC_MODE_CORRECT);
  if (st != FEMTC_OK)
    return st;
  if ((st = check_dtype(it.dtype)) != FEMTC_OK)
    return st;
  if (it.op < FEMTC_OP_MASS || it.op > FEMTC_OP_KPM)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, where + "op must be a femtc_op");
  if (it.num_cells < 0 || !it.tables || !it.x || !it.y)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, where + "missing buffer");
  if (it.op != FEMTC_OP_MASS && !it.G)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, where + "G is required");
  if (it.op != FEMTC_OP_STIFFNESS && !it.detJw)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, where + "detJw is required");
  const uintptr_t galign = 2 * (it.dtype == FEMTC_F64 ? 8 : 4);
  if (it.G && (uintptr_t)it.G % galign != 0)
    return fail(FEMTC_ERR_INVALID_ARGUMENT,
                where + "G must be aligned to two elements");
  if (it.dofmap && it.ndofs < 1)
    return fail(FEMTC_ERR_INVALID_ARGUMENT,
                where + "a continuous item needs ndofs >= 1");
  const bool nogrid = !it.grid[0] && !it.grid[1] && !it.grid[2];
  if (!nogrid && (!it.dofmap || it.grid[0] < 1 || it.grid[1] < 1 ||
                  it.grid[2] < 1 ||
                  (int64_t)it.grid[0] * it.grid[1] * it.grid[2] !=
                      it.num_cells))
    return fail(FEMTC_ERR_INVALID_ARGUMENT,
                where + "grid must be all zero, or positive with "
                        "grid[0]*grid[1]*grid[2] == num_cells on a "
                        "continuous item");
  return FEMTC_OK;
}

// Clear n outputs (descriptors d) on stream s.
void launch_fill(const FillDesc* d, int n, long long max_n, hipStream_t s)
{
  if (n <= 0)
    return;
  const long long blocks =
      std::min<long long>(std::max<long long>((max_n + 255) / 256, 1), 1024);
  hipLaunchKernelGGL(k_fill_batch, dim3((unsigned)blocks, n), dim3(256), 0,
                     s, d);
}

// Launch every group once: the schedule's tasks on the side streams, forked
// from and joined back into b->stream; each side stream first clears the
// continuous outputs of its own tasks, so the clears overlap the other
// streams' applies.
int run_all(femtc_batch* b, hipStream_t s)
{
  launch_fill(b->d_fill, (int)b->empty_fill.size(), b->max_empty_fill, s);
  // only the side streams that have tasks are forked; with one such
  // stream its work runs on s itself (a fork and join cost tens of
  // microseconds in a replayed graph)
  int used = 0;
  for (int i = 0; i < NSTREAMS; ++i)
    used += b->tasks[i].empty() ? 0 : 1;
  const bool fork = used > 1;
  if (fork)
  {
    if (hipEventRecord(b->ev_fork, s) != hipSuccess)
      return fail(FEMTC_ERR_HIP, "batch: event record failed");
    for (int i = 0; i < NSTREAMS; ++i)
      if (!b->tasks[i].empty() &&
          hipStreamWaitEvent(b->side[i], b->ev_fork, 0) != hipSuccess)
        return fail(FEMTC_ERR_HIP, "batch: stream wait failed");
  }
  for (int i = 0; i < NSTREAMS; ++i)
  {
    if (b->tasks[i].empty())
      continue;
    hipStream_t si = fork ? b->side[i] : s;
    launch_fill(b->d_fill + b->sfill_first[i], b->sfill_n[i],
                b->sfill_max[i], si);
    for (const Task& t : b->tasks[i])
    {
      const Group& grp = b->groups[t.g];
      const int st = t.item < 0 ? launch_group(grp, si)
                                : launch_item(grp, (size_t)t.item, si);
      if (st != FEMTC_OK)
        return fail(st, std::string("batch launch: ") +
                            hipGetErrorString(hipGetLastError()));
    }
  }
  for (int i = 0; fork && i < NSTREAMS; ++i)
  {
    if (b->tasks[i].empty())
      continue;
    if (hipEventRecord(b->ev_join[i], b->side[i]) != hipSuccess ||
        hipStreamWaitEvent(s, b->ev_join[i], 0) != hipSuccess)
      return fail(FEMTC_ERR_HIP, "batch: join failed");
  }
  return FEMTC_OK;
}

int upload(femtc_batch* b, const std::vector<FillDesc>& fills)
{
  if (!fills.empty() &&
      hipMemcpy(b->d_fill, fills.data(), fills.size() * sizeof(FillDesc),
                hipMemcpyHostToDevice) != hipSuccess)
    return fail(FEMTC_ERR_HIP, "batch descriptor upload failed");
  return FEMTC_OK;
}

// Every continuous item's clear, group by group (fill_first, nfill), for the
// autotuner, which times each candidate with the clears it needs.
// Synchronous.
int upload_group_fills(femtc_batch* b)
{
  std::vector<FillDesc> fills = b->empty_fill;
  for (Group& g : b->groups)
  {
    g.fill_first = fills.size();
    g.max_fill = 0;
    for (const FillDesc& f : g.fill)
    {
      fills.push_back(f);
      g.max_fill = std::max(g.max_fill, f.n);
    }
    g.nfill = (int)g.fill.size();
  }
  return upload(b, fills);
}

// Time of one launch of a group in its mode: the autotuner's measurement,
// else its algorithmic bytes at 4 TB/s.
double group_estimate(const Group& g)
{
  if (g.est_ms > 0.0)
    return g.est_ms;
  const double es = g.req.dtype == FEMTC_F64 ? 8.0 : 4.0;
  const double N3 = (double)g.req.N * g.req.N * g.req.N;
  const double Q3 = (double)g.req.Q * g.req.Q * g.req.Q;
  const double geo = (g.req.op == FEMTC_OP_MASS ? 1.0
                      : g.req.op == FEMTC_OP_STIFFNESS ? 6.0
                                                       : 7.0) * Q3 * es;
  double bytes = 0.0;
  for (size_t i = 0; i < g.host.size(); ++i)
  {
    const double C = (double)g.host[i].num_cells;
    bytes += C * geo;
    bytes += g.req.gather ? 2.0 * es * (double)g.fill[i].n + 4.0 * C * N3
                          : 2.0 * es * C * N3;
  }
  return bytes / 4e9;
}

// The schedule.  A group that runs one launch per problem becomes one task
// per problem (its estimate shared out by cells), a fused group one task;
// tasks go longest first to the side stream with the least estimated work.
// Then the clears: the items without cells first (on b->stream), then each
// side stream's continuous outputs, except those of assembled-scatter
// tasks, whose second pass overwrites y.  Synchronous; called whenever a
// group's mode changes.
int schedule(femtc_batch* b)
{
  std::vector<std::pair<double, Task>> all;
  for (size_t g = 0; g < b->groups.size(); ++g)
  {
    const Group& grp = b->groups[g];
    const double est = group_estimate(grp);
    if (grp.per_problem >= 1 && grp.per_problem <= 3 && grp.cells > 0)
      for (size_t i = 0; i < grp.host.size(); ++i)
        all.push_back({est * (double)grp.host[i].num_cells / (double)grp.cells,
                       Task{(int)g, (int)i}});
    else
      all.push_back({est, Task{(int)g, -1}});
  }
  std::stable_sort(all.begin(), all.end(),
                   [](const std::pair<double, Task>& a,
                      const std::pair<double, Task>& c) {
                     return a.first > c.first;
                   });
  double load[NSTREAMS] = {};
  for (int i = 0; i < NSTREAMS; ++i)
    b->tasks[i].clear();
  for (const auto& e : all)
  {
    int k = 0;
    for (int i = 1; i < NSTREAMS; ++i)
      if (load[i] < load[k])
        k = i;
    b->tasks[k].push_back(e.second);
    load[k] += e.first;
  }
  std::vector<FillDesc> fills = b->empty_fill;
  b->max_empty_fill = 0;
  for (const FillDesc& f : fills)
    b->max_empty_fill = std::max(b->max_empty_fill, f.n);
  for (int i = 0; i < NSTREAMS; ++i)
  {
    b->sfill_first[i] = fills.size();
    b->sfill_max[i] = 0;
    for (const Task& t : b->tasks[i])
    {
      const Group& grp = b->groups[t.g];
      if (grp.per_problem == 2 || grp.fill.empty())
        continue;
      const size_t lo = t.item < 0 ? 0 : (size_t)t.item;
      const size_t hi = t.item < 0 ? grp.fill.size() : lo + 1;
      for (size_t j = lo; j < hi; ++j)
      {
        fills.push_back(grp.fill[j]);
        b->sfill_max[i] = std::max(b->sfill_max[i], grp.fill[j].n);
      }
    }
    b->sfill_n[i] = (int)(fills.size() - b->sfill_first[i]);
  }
  return upload(b, fills);
}

} // namespace

extern "C" {

int femtc_batch_create(femtc_batch** out, const femtc_batch_item* items,
                       int nitems, void* stream)
{
  if (!out || (nitems > 0 && !items) || nitems < 0)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "bad batch arguments");
  *out = nullptr;
  for (int i = 0; i < nitems; ++i)
  {
    const int st = check_item(items[i], i);
    if (st != FEMTC_OK)
      return st;
  }
  // every dofmap entry in range before a kernel scatters through it
  for (int i = 0; i < nitems; ++i)
  {
    const femtc_batch_item& it = items[i];
    if (!it.dofmap || it.num_cells == 0)
      continue;
    const int64_t n3 = (int64_t)(it.degree + 1) * (it.degree + 1) *
                       (it.degree + 1);
    int ok = 0;
    const int st = check_dofmap_range(it.dofmap, it.num_cells * n3, it.ndofs,
                                      (hipStream_t)stream, &ok);
    if (st != FEMTC_OK)
      return st;
    if (!ok)
      return fail(FEMTC_ERR_INVALID_ARGUMENT,
                  "batch item " + std::to_string(i) +
                      ": a dofmap entry is outside [0, ndofs)");
  }
  femtc_batch* b = new femtc_batch;
  b->stream = (hipStream_t)stream;
  // group by (op, dtype, N, Q, gather)
  std::map<std::tuple<int, int, int, int, int>, size_t> index;
  size_t ncont = 0;
  for (int i = 0; i < nitems; ++i)
  {
    const femtc_batch_item& it = items[i];
    const int N = it.degree + 1, gather = it.dofmap ? 1 : 0;
    auto key = std::make_tuple(it.op, it.dtype, N, it.nq, gather);
    auto f = index.find(key);
    if (f == index.end() ||
        b->groups[f->second].host.size() >= (size_t)BATCH_MAX_ITEMS)
    {
      Group g;
      g.req.op = it.op;
      g.req.dtype = it.dtype;
      g.req.N = N;
      g.req.Q = it.nq;
      g.req.gather = gather;
      index[key] = b->groups.size();
      b->groups.push_back(g);
      f = index.find(key);
    }
    Group& g = b->groups[f->second];
    const int cpw = cells_per_wg(it.nq);
    BatchDesc d{};
    d.tab = it.tables;
    d.x = it.x;
    d.y = it.y;
    d.G = it.G;
    d.detJw = it.detJw;
    d.dofmap = it.dofmap;
    d.num_cells = it.num_cells;
    d.blk_begin = g.req.nblk_total;
    if (it.num_cells > 0)
    {
      Extra e;
      int ok = 0;
      if (it.grid[0] > 0)
      {
        const int vs = verify_structured(it.dofmap, it.grid[0], it.grid[1],
                                         it.grid[2], it.degree, it.ndofs,
                                         (hipStream_t)stream, &ok);
        if (vs != FEMTC_OK)
        {
          release(b);
          return vs;
        }
      }
      if (ok)
        for (int k = 0; k < 3; ++k)
          e.grid[k] = it.grid[k];
      else
        g.structured = false;
      g.host.push_back(d);
      g.extra.push_back(e);
      if (it.dofmap)
        g.fill.push_back(FillDesc{it.y, (long long)it.ndofs, it.dtype});
      g.req.nblk_total += (it.num_cells + cpw - 1) / cpw;
      g.cells += it.num_cells;
    }
    b->total_cells += it.num_cells;
    if (it.dofmap)
    {
      ++ncont;
      if (it.num_cells == 0)
        b->empty_fill.push_back(FillDesc{it.y, (long long)it.ndofs, it.dtype});
      b->total_dofs += it.ndofs;
    }
    else
      b->total_dofs += it.num_cells * (int64_t)N * N * N;
  }
  // drop empty groups; largest groups first
  b->groups.erase(std::remove_if(b->groups.begin(), b->groups.end(),
                                 [](const Group& g) { return g.host.empty(); }),
                  b->groups.end());
  std::stable_sort(b->groups.begin(), b->groups.end(),
                   [](const Group& a, const Group& c) {
                     return a.req.nblk_total > c.req.nblk_total;
                   });
  size_t ndesc = 0;
  for (auto& g : b->groups)
  {
    g.first = ndesc;
    ndesc += g.host.size();
  }
  int st = FEMTC_OK;
  if ((ndesc && hipMalloc((void**)&b->d_desc, ndesc * sizeof(BatchDesc)) !=
                    hipSuccess) ||
      (ncont &&
       hipMalloc((void**)&b->d_fill, ncont * sizeof(FillDesc)) !=
           hipSuccess))
    st = fail(FEMTC_ERR_OUT_OF_MEMORY, "batch descriptor allocation failed");
  for (auto& g : b->groups)
  {
    if (st != FEMTC_OK)
      break;
    g.req.items = b->d_desc + g.first;
    g.req.nitems = (int)g.host.size();
    if (hipMemcpy(b->d_desc + g.first, g.host.data(),
                  g.host.size() * sizeof(BatchDesc),
                  hipMemcpyHostToDevice) != hipSuccess)
      st = fail(FEMTC_ERR_HIP, "batch descriptor upload failed");
  }
  if (st == FEMTC_OK && ncont > 65535)
    st = fail(FEMTC_ERR_INVALID_ARGUMENT, "at most 65535 continuous items");
  if (st == FEMTC_OK)
    st = schedule(b);
  for (int i = 0; i < NSTREAMS && st == FEMTC_OK; ++i)
    if (hipStreamCreateWithFlags(&b->side[i], hipStreamNonBlocking) !=
            hipSuccess ||
        hipEventCreateWithFlags(&b->ev_join[i], hipEventDisableTiming) !=
            hipSuccess)
      st = fail(FEMTC_ERR_HIP, "batch stream creation failed");
  if (st == FEMTC_OK &&
      hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) !=
          hipSuccess)
    st = fail(FEMTC_ERR_HIP, "batch event creation failed");
  if (st != FEMTC_OK)
  {
    release(b);
    return st;
  }
  *out = b;
  return FEMTC_OK;
}

int femtc_batch_run(femtc_batch* b)
{
  if (!b)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "batch is NULL");
  return run_all(b, b->stream);
}

int femtc_batch_graph_launch(femtc_batch* b)
{
  if (!b)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "batch is NULL");
  if (!b->exec)
  {
    // capture on a private stream (the legacy default stream cannot be
    // captured), then replay on the caller's stream
    hipStream_t cs = nullptr;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess)
      return fail(FEMTC_ERR_HIP, "graph: stream creation failed");
    int st = FEMTC_OK;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) !=
        hipSuccess)
      st = fail(FEMTC_ERR_HIP, "graph: begin capture failed");
    if (st == FEMTC_OK)
    {
      const int rs = run_all(b, cs);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(cs, &g);
      if (rs != FEMTC_OK)
        st = rs;
      else if (e != hipSuccess)
        st = fail(FEMTC_ERR_HIP, std::string("graph: end capture failed: ") +
                                     hipGetErrorString(e));
      else
        b->graph = g;
      if (st != FEMTC_OK && g)
        (void)hipGraphDestroy(g);
    }
    (void)hipStreamDestroy(cs);
    if (st != FEMTC_OK)
      return st;
    if (hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) !=
        hipSuccess)
    {
      b->exec = nullptr;
      return fail(FEMTC_ERR_HIP, "graph: instantiate failed");
    }
  }
  if (hipGraphLaunch(b->exec, b->stream) != hipSuccess)
    return fail(FEMTC_ERR_HIP, std::string("graph launch failed: ") +
                                   hipGetErrorString(hipGetLastError()));
  return FEMTC_OK;
}

int femtc_batch_autotune(femtc_batch* b, int reps)
{
  if (!b)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "batch is NULL");
  if (reps < 1)
    reps = 3;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return fail(FEMTC_ERR_HIP, "event creation failed");
  // the whole batch a few times first: a GPU coming out of idle runs its
  // first launches slower, which would favour the candidates timed last
  int st = FEMTC_OK;
  for (int i = 0; i < 4 && st == FEMTC_OK; ++i)
    st = run_all(b, b->stream);
  // then every group's clears on the device, so that a candidate that
  // needs them (all but the assembled scatter) is timed with them
  if (st == FEMTC_OK && hipStreamSynchronize(b->stream) != hipSuccess)
    st = fail(FEMTC_ERR_HIP, "batch autotune: warm-up failed");
  if (st == FEMTC_OK)
    st = upload_group_fills(b);
  for (auto& g : b->groups)
  {
    if (st != FEMTC_OK)
      break;
    double best = 1e300;
    int bestk = 1, best_pp = 1;
    // k = 0: one launch per problem; k > 0: fused, k blocks per workgroup;
    // k = -1: the assembled scatter per problem (structured groups, when
    // its buffers fit); k = -2: the z-run scatter per problem (structured
    // groups, nq <= 8)
    // k = -3: one fused z-run launch over the group's problems (the same
    // condition).  Every candidate is timed in two rounds, the faster round
    // kept, so a candidate is not judged on one disturbed measurement.
    const bool try_asm = g.structured && ensure_lattices(g) == FEMTC_OK;
    const bool try_zrun = g.structured && g.req.gather && g.req.Q <= 8;
    const bool try_fzrun = try_zrun && ensure_zrun(g) == FEMTC_OK;
    for (int round = 0; round < 2 && st == FEMTC_OK; ++round)
      for (int k : {0, 1, 2, 4, 8, -1, -2, -3})
      {
        if ((k == -1 && !try_asm) || (k == -2 && !try_zrun) ||
            (k == -3 && !try_fzrun))
          continue;
        Group t = g;
        t.per_problem = k == 0    ? 1
                        : k == -1 ? 2
                        : k == -2 ? 3
                        : k == -3 ? 4
                                  : 0;
        t.req.per_wg = k > 0 ? k : 1;
        // the candidate as the batch runs it: y cleared first unless the
        // assembled scatter overwrites it
        auto launch = [&]() {
          if (t.per_problem != 2)
            launch_fill(b->d_fill + t.fill_first, t.nfill, t.max_fill,
                        b->stream);
          return launch_group(t, b->stream);
        };
        st = launch(); // warm-up
        (void)hipEventRecord(e0, b->stream);
        for (int i = 0; i < reps && st == FEMTC_OK; ++i)
          st = launch();
        (void)hipEventRecord(e1, b->stream);
        if (st != FEMTC_OK || hipEventSynchronize(e1) != hipSuccess)
          break;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms < best)
        {
          best = ms;
          bestk = t.req.per_wg;
          best_pp = t.per_problem;
        }
      }
    if (st != FEMTC_OK)
      break;
    g.req.per_wg = bestk;
    g.per_problem = best_pp;
    g.est_ms = best / reps;
    if (g.per_problem != 2)
      free_lattices(g);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  // a captured graph holds the old launch shapes
  if (b->exec)
  {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  if (b->graph)
  {
    (void)hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
  // the schedule and the clears for the modes the groups now have, also
  // after a failure (d_fill holds the autotuner's layout until then)
  const int ss = schedule(b);
  if (st == FEMTC_OK)
    st = ss;
  if (st != FEMTC_OK)
    return fail(st, "batch autotune failed");
  return FEMTC_OK;
}

int femtc_batch_stats(const femtc_batch* b, int* ngroups, int64_t* cells,
                      int64_t* dofs)
{
  if (!b)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "batch is NULL");
  if (ngroups)
    *ngroups = (int)b->groups.size();
  if (cells)
    *cells = b->total_cells;
  if (dofs)
    *dofs = b->total_dofs;
  return FEMTC_OK;
}

int femtc_batch_group(const femtc_batch* b, int g, int* op, int* dtype,
                      int* degree, int* nq, int* gather, int* nitems,
                      int* per_wg)
{
  if (!b || g < 0 || g >= (int)b->groups.size())
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "no such group");
  const BatchReq& r = b->groups[g].req;
  if (op)
    *op = r.op;
  if (dtype)
    *dtype = r.dtype;
  if (degree)
    *degree = r.N - 1;
  if (nq)
    *nq = r.Q;
  if (gather)
    *gather = r.gather;
  if (nitems)
    *nitems = r.nitems;
  if (per_wg)
  {
    const int pp = b->groups[g].per_problem;
    *per_wg = pp == 4 ? -3 : pp == 3 ? -2 : pp == 2 ? -1 : pp ? 0 : r.per_wg;
  }
  return FEMTC_OK;
}

int femtc_batch_set_group(femtc_batch* b, int g, int per_wg)
{
  if (!b || g < 0 || g >= (int)b->groups.size() || per_wg < -3 ||
      per_wg > 64)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "no such group or bad per_wg");
  Group& grp = b->groups[g];
  if (per_wg == -1)
  {
    const int st = ensure_lattices(grp);
    if (st != FEMTC_OK)
      return st;
  }
  if (per_wg == -2 && (!grp.structured || !grp.req.gather || grp.req.Q > 8))
    return fail(FEMTC_ERR_INVALID_ARGUMENT,
                "z-run scatter: every problem of the group needs a verified "
                "structured grid, and nq <= 8");
  if (per_wg == -3)
  {
    const int st = ensure_zrun(grp);
    if (st != FEMTC_OK)
      return st;
  }
  grp.per_problem = per_wg == 0    ? 1
                    : per_wg == -1 ? 2
                    : per_wg == -2 ? 3
                    : per_wg == -3 ? 4
                                   : 0;
  grp.req.per_wg = per_wg > 0 ? per_wg : 1;
  grp.est_ms = 0.0; // estimated from the bytes until timed again
  if (grp.per_problem != 2)
    free_lattices(grp);
  if (b->exec)
  {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  if (b->graph)
  {
    (void)hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
  return schedule(b);
}

int femtc_batch_destroy(femtc_batch* b)
{
  release(b);
  return FEMTC_OK;
}

} // extern "C"
