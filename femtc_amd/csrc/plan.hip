// Scatter plans for the continuous operator over a structured cell
// numbering: the direct atomic scatter, an LDS-assembled scatter over 3-D
// blocks of cells, or the assembled scatter (a cell pass storing every
// cell-dof value without atomics, then a pass summing each global dof;
// offered when the dofmap is verified to be the structured numbering at
// plan creation), chosen per (nq, dtype, op) by default tables or by timing
// the candidates on the caller's own data (femtc_plan_autotune).
#include <cfloat>
#include <cstdio>
#include <vector>

#include "assemble.hpp"
#include "common.hpp"
#include "even_odd.hpp"

namespace femtc
{
namespace
{

constexpr int RB = 256;

__global__ void __launch_bounds__(RB)
    k_touch_count(const int32_t* __restrict__ dofmap, int64_t n,
                  int64_t ndofs, uint32_t* __restrict__ cnt,
                  int* __restrict__ bad)
{
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * RB)
  {
    const int32_t g = dofmap[i];
    if (g < 0 || g >= ndofs)
      bad[0] = 1;
    else
      atomicAdd(cnt + g, 1u);
  }
}

__global__ void __launch_bounds__(RB)
    k_touch_narrow(const uint32_t* __restrict__ cnt, uint8_t* __restrict__ t,
                   int64_t n)
{
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * RB)
  {
    const uint32_t c = cnt[i];
    t[i] = (uint8_t)(c < 255u ? c : 255u);
  }
}

__global__ void __launch_bounds__(RB)
    k_zero_u32(uint32_t* __restrict__ p, int64_t n)
{
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * RB)
    p[i] = 0u;
}

unsigned grid_for(int64_t n)
{
  const int64_t b = (n + RB - 1) / RB;
  return (unsigned)(b < 4096 ? (b > 0 ? b : 1) : 4096);
}

// Structured-numbering check: every entry equals the lexicographic global
// index of its cell's local dof (the numbering of femtc_dofmap_structured).
__global__ void __launch_bounds__(RB)
    k_check_structured(const int32_t* __restrict__ dofmap, int ny, int nz,
                       int P, int64_t n, int* __restrict__ bad)
{
  const int N = P + 1;
  const int64_t NY = (int64_t)ny * P + 1, NZ = (int64_t)nz * P + 1;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * RB)
  {
    const int64_t cell = i / (N * N * N);
    const int l = (int)(i - cell * N * N * N);
    const int a = l / (N * N), b = (l / N) % N, c = l % N;
    const int64_t ck = cell % nz;
    const int64_t cj = (cell / nz) % ny;
    const int64_t ci = cell / ((int64_t)nz * ny);
    const int64_t g = ((ci * P + a) * NY + cj * P + b) * NZ + ck * P + c;
    if ((int64_t)dofmap[i] != g)
      bad[0] = 1;
  }
}

// cells per one-wave group of the cell kernel (64 / nq^2, nq <= 8)
int cells_per_wave(int nq) { return nq * nq <= 64 ? 64 / (nq * nq) : 0; }

struct Shape
{
  int x, y, z;
};

// Candidate block shapes for nq (cbz a multiple of the cells per wave, at
// most 8 waves), timed by femtc_plan_autotune.
std::vector<Shape> candidates(int nq)
{
  const int cpw = cells_per_wave(nq);
  std::vector<Shape> v;
  if (cpw == 0)
    return v;
  if (cpw == 1)
    v = {{2, 2, 2}, {1, 2, 2}, {1, 2, 4}, {2, 2, 1}, {1, 1, 4}, {2, 4, 1}};
  else if (cpw == 2)
    v = {{2, 2, 2}, {2, 2, 4}, {2, 4, 2}, {1, 2, 4}, {4, 2, 2}};
  else if (cpw == 4)
    v = {{2, 2, 4}, {2, 4, 4}, {4, 2, 4}, {1, 2, 4}, {2, 2, 8}};
  else if (cpw == 7)
    v = {{2, 2, 7}, {2, 4, 7}, {4, 2, 7}, {1, 4, 7}, {2, 2, 14}};
  else
    v = {{2, 2, cpw}, {2, 4, cpw}, {4, 2, cpw}, {1, 2, cpw}};
  return v;
}

int key_of(int nq, int dtype, int op) { return (nq * 4 + dtype) * 4 + op; }

// Default block shape where the blocked scatter measured faster than the
// direct one on MI355X (structured meshes of 2.4e7-3.6e7 cell-dofs, nq =
// degree+1); everything else defaults to the direct scatter.  At degree >= 5
// and for most stiffness cases the one-wave direct kernel wins: its higher
// occupancy hides more latency than the blocked LDS assembly saves in
// atomics.
Shape default_shape(int degree, int nq, int dtype, int op)
{
  if (nq != degree + 1)
    return Shape{0, 0, 0};
  const bool f32 = dtype == FEMTC_F32;
  switch (degree)
  {
  case 1: return Shape{1, 2, 16};
  case 2: return op == FEMTC_OP_MASS ? Shape{1, 4, 7} : Shape{0, 0, 0};
  case 3:
    if (f32)
      return Shape{2, 2, 4};
    return op == FEMTC_OP_KPM ? Shape{0, 0, 0} : Shape{2, 2, 8};
  case 4:
    if (op == FEMTC_OP_MASS || f32)
      return Shape{1, 2, 4};
    return Shape{0, 0, 0};
  }
  return Shape{0, 0, 0};
}

} // namespace

int verify_structured(const int32_t* dofmap, int nx, int ny, int nz,
                      int degree, int64_t ndofs, hipStream_t s, int* ok)
{
  *ok = 0;
  if (!dofmap || nx < 1 || ny < 1 || nz < 1 || degree < 1 ||
      degree > FEMTC_MAX_FUSED_DEGREE)
    return FEMTC_ERR_INVALID_ARGUMENT;
  const int P = degree;
  if (ndofs != ((int64_t)nx * P + 1) * ((int64_t)ny * P + 1) *
                   ((int64_t)nz * P + 1))
    return FEMTC_OK; // not the structured dof count
  const int n3 = (P + 1) * (P + 1) * (P + 1);
  const int64_t n = (int64_t)nx * ny * nz * n3;
  int* bad = nullptr;
  if (hipMalloc((void**)&bad, sizeof(int)) != hipSuccess)
  {
    (void)hipGetLastError();
    return fail(FEMTC_ERR_OUT_OF_MEMORY, "dofmap check allocation failed");
  }
  int hbad = 1;
  int st = FEMTC_OK;
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(RB), 0, s, (uint32_t*)bad,
                     (int64_t)1);
  hipLaunchKernelGGL(k_check_structured, dim3(grid_for(n)), dim3(RB), 0, s,
                     dofmap, ny, nz, P, n, bad);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    st = fail(FEMTC_ERR_HIP, std::string("dofmap check: ") +
                                 hipGetErrorString(hipGetLastError()));
  (void)hipFree(bad);
  if (st == FEMTC_OK)
    *ok = hbad ? 0 : 1;
  return st;
}

namespace
{
__global__ void __launch_bounds__(RB)
    k_dofmap_range(const int32_t* __restrict__ dofmap, int64_t n,
                   int64_t ndofs, int* __restrict__ bad)
{
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * RB)
  {
    const int32_t g = dofmap[i];
    if (g < 0 || g >= ndofs)
      bad[0] = 1;
  }
}
} // namespace

int check_dofmap_range(const int32_t* dofmap, int64_t n, int64_t ndofs,
                       hipStream_t s, int* ok)
{
  *ok = 0;
  if (n <= 0)
  {
    *ok = 1;
    return FEMTC_OK;
  }
  if (!dofmap)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "dofmap is NULL");
  int* bad = nullptr;
  if (hipMalloc((void**)&bad, sizeof(int)) != hipSuccess)
  {
    (void)hipGetLastError();
    return fail(FEMTC_ERR_OUT_OF_MEMORY, "dofmap check allocation failed");
  }
  int hbad = 1;
  int st = FEMTC_OK;
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(RB), 0, s, (uint32_t*)bad,
                     (int64_t)1);
  hipLaunchKernelGGL(k_dofmap_range, dim3(grid_for(n)), dim3(RB), 0, s,
                     dofmap, n, ndofs, bad);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    st = fail(FEMTC_ERR_HIP, std::string("dofmap range check: ") +
                                 hipGetErrorString(hipGetLastError()));
  (void)hipFree(bad);
  if (st == FEMTC_OK)
    *ok = hbad ? 0 : 1;
  return st;
}

int make_even_odd(const void* tables, int dtype, int N, int Q, const double* w,
                  int out_dtype, hipStream_t s, void** eo)
{
  *eo = nullptr;
  const bool d64 = dtype == FEMTC_F64;
  const int nt = 2 * N * Q;
  std::vector<double> t(nt);
  std::vector<float> t32(d64 ? 0 : nt);
  void* host = d64 ? (void*)t.data() : (void*)t32.data();
  if (hipMemcpyAsync(host, tables, (size_t)nt * (d64 ? 8 : 4),
                     hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return fail(FEMTC_ERR_HIP, std::string("table read: ") +
                                   hipGetErrorString(hipGetLastError()));
  if (!d64)
    for (int i = 0; i < nt; ++i)
      t[i] = t32[i];
  if (!eo_symmetric(t.data(), N, Q, d64 ? DBL_EPSILON : FLT_EPSILON))
    return FEMTC_OK;
  const int ne = eo_table_size(N, Q);
  std::vector<double> e(ne);
  eo_build(t.data(), w, N, Q, e.data());
  const bool o64 = out_dtype == FEMTC_F64;
  std::vector<float> e32(o64 ? 0 : ne);
  for (int i = 0; i < (o64 ? 0 : ne); ++i)
    e32[i] = (float)e[i];
  const size_t bytes = (size_t)ne * (o64 ? 8 : 4);
  void* buf = nullptr;
  if (hipMalloc(&buf, bytes) != hipSuccess)
  {
    (void)hipGetLastError();
    return fail(FEMTC_ERR_OUT_OF_MEMORY, "even-odd table allocation failed");
  }
  if (hipMemcpyAsync(buf, o64 ? (const void*)e.data() : (const void*)e32.data(),
                     bytes, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
  {
    (void)hipFree(buf);
    return fail(FEMTC_ERR_HIP, std::string("even-odd table upload: ") +
                                   hipGetErrorString(hipGetLastError()));
  }
  *eo = buf;
  return FEMTC_OK;
}

} // namespace femtc

using namespace femtc;

struct femtc_plan
{
  int nx = 0, ny = 0, nz = 0, degree = 0;
  int64_t num_cells = 0, ndofs = 0;
  const int32_t* dofmap = nullptr;
  uint8_t* touch = nullptr;
  // the dofmap is the structured numbering (the assembled scatter applies)
  int structured = 0;
  // lattice buffers of the assembled scatter, per dtype (allocated when the
  // scatter is first selected or timed)
  void* lattice[2] = {nullptr, nullptr};
  // chosen block shape per (nq, dtype, op); x == 0 means direct scatter,
  // x == -1 the assembled scatter, x == -2 the structured direct scatter,
  // x == -3 the structured z-run scatter
  std::vector<std::pair<int, Shape>> shapes;
  // tables registered by femtc_plan_set_tables, per (dtype, nq): a copy,
  // and its even-odd table set when the tables are symmetric (else null)
  struct Tables
  {
    int dtype, nq;
    void* tab;
    void* eo;
  };
  std::vector<Tables> tables;

  const Tables* find_tables(int dtype, int nq) const
  {
    for (auto& t : tables)
      if (t.dtype == dtype && t.nq == nq)
        return &t;
    return nullptr;
  }

  Shape shape(int nq, int dtype, int op) const
  {
    const int k = key_of(nq, dtype, op);
    for (auto& s : shapes)
      if (s.first == k)
        return s.second;
    return default_shape(degree, nq, dtype, op);
  }
  void set(int nq, int dtype, int op, Shape s)
  {
    const int k = key_of(nq, dtype, op);
    for (auto& e : shapes)
      if (e.first == k)
      {
        e.second = s;
        return;
      }
    shapes.push_back({k, s});
  }
};

namespace
{

// The tables of an apply: the caller's, or (tables == NULL) the copy that
// femtc_plan_set_tables registered, with its even-odd set (*eo).
int resolve_tables(const femtc_plan* pl, int dtype, int nq, const void** tables,
                   const void** eo)
{
  *eo = nullptr;
  if (*tables)
    return FEMTC_OK;
  const femtc_plan::Tables* t = pl->find_tables(dtype, nq);
  if (!t)
    return fail(FEMTC_ERR_INVALID_ARGUMENT,
                "tables is NULL and none are registered for this dtype and nq "
                "(femtc_plan_set_tables)");
  *tables = t->tab;
  *eo = t->eo;
  return FEMTC_OK;
}

int plan_request(const femtc_plan* pl, int op, int dtype, int nq,
                 const void* tables, const void* eo, const void* x, void* y,
                 const void* G, const void* detJw, int overwrite, Shape s,
                 hipStream_t st, ApplyReq* out)
{
  ApplyReq r;
  r.op = op;
  r.dtype = dtype;
  r.N = pl->degree + 1;
  r.Q = nq;
  r.gather = 1;
  r.tab = tables;
  // the even-odd instances cover the direct and assembled scatters
  r.eo_tab = (s.x == 0 || s.x == -1) && nq == r.N ? eo : nullptr;
  r.x = x;
  r.y = y;
  r.G = G;
  r.detJw = detJw;
  r.dofmap = pl->dofmap;
  r.num_cells = pl->num_cells;
  r.stream = st;
  if (s.x == -2 || s.x == -3)
  {
    r.structured_direct = s.x == -2;
    r.zrun = s.x == -3;
    r.ncx = pl->nx;
    r.ncy = pl->ny;
    r.ncz = pl->nz;
  }
  else if (s.x < 0)
  {
    r.assembled = 1;
    r.lattice = pl->lattice[dtype == FEMTC_F64 ? 0 : 1];
    r.ncx = pl->nx;
    r.ncy = pl->ny;
    r.ncz = pl->nz;
    r.overwrite = overwrite;
  }
  else if (s.x > 0)
  {
    r.blocked = 1;
    r.ncx = pl->nx;
    r.ncy = pl->ny;
    r.ncz = pl->nz;
    r.cbx = s.x;
    r.cby = s.y;
    r.cbz = s.z;
    r.touch = pl->touch;
    r.overwrite = overwrite;
  }
  *out = r;
  return FEMTC_OK;
}

int check_apply_args(const femtc_plan* pl, int op, int dtype, int nq,
                     const void* tables, const void* x, void* y,
                     const void* G, const void* detJw)
{
  if (!pl)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "plan is NULL");
  int st = check_shape_fused(pl->degree, nq, FEMTC_MODE_CORRECT);
  if (st != FEMTC_OK)
    return st;
  if ((st = check_dtype(dtype)) != FEMTC_OK)
    return st;
  if (op < FEMTC_OP_MASS || op > FEMTC_OP_KPM)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "op must be a femtc_op");
  if (!x || !y)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "missing buffer");
  if (op != FEMTC_OP_MASS && !G)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "G is required for this op");
  if (op != FEMTC_OP_STIFFNESS && !detJw)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "detJw is required for this op");
  return FEMTC_OK;
}

// The assembled scatter's lattice buffer for dtype (allocated once).
int ensure_lattice(femtc_plan* pl, int dtype)
{
  if (!pl->structured)
    return fail(FEMTC_ERR_INVALID_ARGUMENT,
                "assembled scatter: the dofmap is not the structured numbering "
                "of the plan's grid");
  void*& buf = pl->lattice[dtype == FEMTC_F64 ? 0 : 1];
  if (buf)
    return FEMTC_OK;
  const size_t bytes = (size_t)lattice_size(pl->nx, pl->ny, pl->nz,
                                            pl->degree) *
                       (dtype == FEMTC_F64 ? 8 : 4);
  if (hipMalloc(&buf, bytes) != hipSuccess)
  {
    (void)hipGetLastError();
    buf = nullptr;
    return fail(FEMTC_ERR_OUT_OF_MEMORY,
                "assembled scatter buffer allocation failed");
  }
  return FEMTC_OK;
}

int launch_plan(const femtc_plan* pl, int op, int dtype, int nq,
                const void* tables, const void* eo, const void* x, void* y,
                const void* G, const void* detJw, int accumulate, Shape s,
                hipStream_t st)
{
  if (s.x == -1 && !pl->lattice[dtype == FEMTC_F64 ? 0 : 1])
    return fail(FEMTC_ERR_INVALID_ARGUMENT,
                "assembled scatter selected without its buffer");
  if ((s.x == -2 || s.x == -3) && !pl->structured)
    return fail(FEMTC_ERR_INVALID_ARGUMENT,
                "structured direct scatter: the dofmap is not the structured "
                "numbering of the plan's grid");
  if (!accumulate && s.x != -1) // the assembly pass overwrites by itself
  {
    int r0 = fill_zero(y, pl->ndofs, dtype, st);
    if (r0 != FEMTC_OK)
      return r0;
  }
  ApplyReq r;
  plan_request(pl, op, dtype, nq, tables, eo, x, y, G, detJw, !accumulate, s,
               st, &r);
  return run(r);
}

} // namespace

extern "C" {

int femtc_plan_create(femtc_plan** out, int nx, int ny, int nz, int degree,
                      const int32_t* dofmap, int64_t ndofs, void* stream)
{
  if (!out)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  if (nx < 1 || ny < 1 || nz < 1 || ndofs < 1 || !dofmap)
    return fail(FEMTC_ERR_INVALID_ARGUMENT,
                "plan needs nx, ny, nz >= 1, ndofs >= 1 and a dofmap");
  if (degree < 1 || degree > FEMTC_MAX_FUSED_DEGREE)
    return fail(FEMTC_ERR_DEGREE_NOT_SUPPORTED, "degree not supported");
  if (ndofs > 0x7fffffffLL)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "ndofs exceeds int32 indexing");
  hipStream_t s = (hipStream_t)stream;
  const int N = degree + 1;
  const int64_t C = (int64_t)nx * ny * nz;
  const int64_t n = C * N * N * N;
  femtc_plan* pl = new femtc_plan;
  pl->nx = nx;
  pl->ny = ny;
  pl->nz = nz;
  pl->degree = degree;
  pl->num_cells = C;
  pl->ndofs = ndofs;
  pl->dofmap = dofmap;
  uint32_t* cnt = nullptr;
  int* bad = nullptr;
  int hbad = 0;
  int st = FEMTC_OK;
  if (hipMalloc((void**)&pl->touch, ndofs) != hipSuccess ||
      hipMalloc((void**)&cnt, 4 * ndofs) != hipSuccess ||
      hipMalloc((void**)&bad, sizeof(int)) != hipSuccess)
  {
    (void)hipGetLastError();
    st = fail(FEMTC_ERR_OUT_OF_MEMORY, "plan allocation failed");
  }
  if (st == FEMTC_OK)
  {
    hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(ndofs)), dim3(RB), 0, s,
                       cnt, ndofs);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(RB), 0, s, (uint32_t*)bad,
                       (int64_t)1);
    hipLaunchKernelGGL(k_touch_count, dim3(grid_for(n)), dim3(RB), 0, s,
                       dofmap, n, ndofs, cnt, bad);
    hipLaunchKernelGGL(k_touch_narrow, dim3(grid_for(ndofs)), dim3(RB), 0, s,
                       (const uint32_t*)cnt, pl->touch, ndofs);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      st = fail(FEMTC_ERR_HIP, std::string("plan setup: ") +
                                   hipGetErrorString(hipGetLastError()));
    else if (hbad)
      st = fail(FEMTC_ERR_INVALID_ARGUMENT,
                "dofmap holds an index outside [0, ndofs)");
  }
  if (st == FEMTC_OK)
    st = verify_structured(dofmap, nx, ny, nz, degree, ndofs, s,
                           &pl->structured);
  if (cnt)
    (void)hipFree(cnt);
  if (bad)
    (void)hipFree(bad);
  if (st != FEMTC_OK)
  {
    if (pl->touch)
      (void)hipFree(pl->touch);
    delete pl;
    return st;
  }
  *out = pl;
  return FEMTC_OK;
}

int femtc_plan_destroy(femtc_plan* plan)
{
  if (plan)
  {
    if (plan->touch)
      (void)hipFree(plan->touch);
    for (void* b : plan->lattice)
      if (b)
        (void)hipFree(b);
    for (auto& t : plan->tables)
    {
      (void)hipFree(t.tab);
      if (t.eo)
        (void)hipFree(t.eo);
    }
    delete plan;
  }
  return FEMTC_OK;
}

int femtc_plan_set_block(femtc_plan* plan, int nq, int dtype, int op,
                         int cbx, int cby, int cbz)
{
  if (!plan)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "plan is NULL");
  if (cbx == 0 && cby == 0 && cbz == 0)
  {
    plan->set(nq, dtype, op, Shape{0, 0, 0});
    return FEMTC_OK;
  }
  if (cbx == -1 && cby == -1 && cbz == -1)
  {
    int st = check_dtype(dtype);
    if (st == FEMTC_OK)
      st = ensure_lattice(plan, dtype);
    if (st != FEMTC_OK)
      return st;
    plan->set(nq, dtype, op, Shape{-1, -1, -1});
    return FEMTC_OK;
  }
  if (cbx == -2 && cby == -2 && cbz == -2)
  {
    if (!plan->structured)
      return fail(FEMTC_ERR_INVALID_ARGUMENT,
                  "structured direct scatter: the dofmap is not the "
                  "structured numbering of the plan's grid");
    plan->set(nq, dtype, op, Shape{-2, -2, -2});
    return FEMTC_OK;
  }
  if (cbx == -3 && cby == -3 && cbz == -3)
  {
    if (!plan->structured)
      return fail(FEMTC_ERR_INVALID_ARGUMENT,
                  "structured z-run scatter: the dofmap is not the "
                  "structured numbering of the plan's grid");
    if (nq > 8)
      return fail(FEMTC_ERR_INVALID_ARGUMENT,
                  "structured z-run scatter: nq <= 8 only");
    plan->set(nq, dtype, op, Shape{-3, -3, -3});
    return FEMTC_OK;
  }
  const int cpw = cells_per_wave(nq);
  const int cells = cbx * cby * cbz;
  if (cpw == 0 || cbx < 1 || cby < 1 || cbz < 1 || cbz % cpw != 0 ||
      (cells + cpw - 1) / cpw > 8)
    return fail(FEMTC_ERR_INVALID_ARGUMENT,
                "block shape: cbz must be a multiple of 64/nq^2 cells and the "
                "block at most 8 waves (nq <= 8)");
  plan->set(nq, dtype, op, Shape{cbx, cby, cbz});
  return FEMTC_OK;
}

int femtc_plan_get_block(const femtc_plan* plan, int nq, int dtype, int op,
                         int* cbx, int* cby, int* cbz)
{
  if (!plan || !cbx || !cby || !cbz)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "NULL argument");
  const Shape s = plan->shape(nq, dtype, op);
  *cbx = s.x;
  *cby = s.y;
  *cbz = s.z;
  return FEMTC_OK;
}

int femtc_plan_apply(const femtc_plan* plan, int op, int dtype, int nq,
                     const void* tables, const void* x, void* y,
                     const void* G, const void* detJw, int accumulate,
                     void* stream)
{
  int st = check_apply_args(plan, op, dtype, nq, tables, x, y, G, detJw);
  const void* eo = nullptr;
  if (st == FEMTC_OK)
    st = resolve_tables(plan, dtype, nq, &tables, &eo);
  if (st != FEMTC_OK)
    return st;
  return launch_plan(plan, op, dtype, nq, tables, eo, x, y, G, detJw,
                     accumulate, plan->shape(nq, dtype, op),
                     (hipStream_t)stream);
}

int femtc_plan_set_tables(femtc_plan* plan, int dtype, int nq,
                          const void* tables, void* stream, int* even_odd)
{
  if (even_odd)
    *even_odd = 0;
  if (!plan || !tables)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "plan and tables are required");
  int st = check_shape_fused(plan->degree, nq, FEMTC_MODE_CORRECT);
  if (st == FEMTC_OK)
    st = check_dtype(dtype);
  if (st != FEMTC_OK)
    return st;
  hipStream_t s = (hipStream_t)stream;
  const int N = plan->degree + 1;
  const size_t bytes = (size_t)2 * N * nq * (dtype == FEMTC_F64 ? 8 : 4);
  void* copy = nullptr;
  void* eo = nullptr;
  if (hipMalloc(&copy, bytes) != hipSuccess)
  {
    (void)hipGetLastError();
    return fail(FEMTC_ERR_OUT_OF_MEMORY, "table copy allocation failed");
  }
  if (hipMemcpyAsync(copy, tables, bytes, hipMemcpyDeviceToDevice, s) !=
      hipSuccess)
    st = fail(FEMTC_ERR_HIP, "table copy failed");
  if (st == FEMTC_OK && nq == N)
    st = make_even_odd(copy, dtype, N, nq, nullptr, dtype, s, &eo);
  else if (st == FEMTC_OK && hipStreamSynchronize(s) != hipSuccess)
    st = fail(FEMTC_ERR_HIP, "table copy failed");
  if (st != FEMTC_OK)
  {
    (void)hipFree(copy);
    return st;
  }
  for (auto it = plan->tables.begin(); it != plan->tables.end(); ++it)
    if (it->dtype == dtype && it->nq == nq)
    {
      (void)hipFree(it->tab);
      if (it->eo)
        (void)hipFree(it->eo);
      plan->tables.erase(it);
      break;
    }
  plan->tables.push_back({dtype, nq, copy, eo});
  if (even_odd)
    *even_odd = eo ? 1 : 0;
  return FEMTC_OK;
}

int femtc_plan_autotune(femtc_plan* plan, int op, int dtype, int nq,
                        const void* tables, const void* x, void* y,
                        const void* G, const void* detJw, int reps,
                        void* stream, double* best_us)
{
  int st = check_apply_args(plan, op, dtype, nq, tables, x, y, G, detJw);
  const void* eo = nullptr;
  if (st == FEMTC_OK)
    st = resolve_tables(plan, dtype, nq, &tables, &eo);
  if (st != FEMTC_OK)
    return st;
  if (reps < 1)
    reps = 5;
  hipStream_t s = (hipStream_t)stream;
  std::vector<Shape> cand = candidates(nq);
  cand.push_back(Shape{0, 0, 0}); // direct scatter
  if (plan->structured)
    cand.push_back(Shape{-2, -2, -2}); // structured direct scatter
  if (plan->structured && nq <= 8)
    cand.push_back(Shape{-3, -3, -3}); // structured z-run scatter
  if (plan->structured && ensure_lattice(plan, dtype) == FEMTC_OK)
    cand.push_back(Shape{-1, -1, -1}); // assembled scatter
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return fail(FEMTC_ERR_HIP, "event creation failed");
  double best = 1e300;
  Shape bs{0, 0, 0};
  for (const Shape& c : cand)
  {
    // warm-up, then time reps overwriting applies (each scatter with the
    // clearing of y it needs: a fill kernel for the atomic and blocked
    // ones, none for the assembled one)
    st = launch_plan(plan, op, dtype, nq, tables, eo, x, y, G, detJw, 0, c, s);
    if (st == FEMTC_ERR_INVALID_ARGUMENT)
      continue; // shape does not fit (LDS) for this op/dtype
    if (st != FEMTC_OK)
      break;
    (void)hipEventRecord(e0, s);
    for (int i = 0; i < reps && st == FEMTC_OK; ++i)
      st = launch_plan(plan, op, dtype, nq, tables, eo, x, y, G, detJw, 0, c, s);
    (void)hipEventRecord(e1, s);
    if (st != FEMTC_OK || hipEventSynchronize(e1) != hipSuccess)
      break;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    const double us = 1e3 * ms / reps;
    if (us < best)
    {
      best = us;
      bs = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st != FEMTC_OK && st != FEMTC_ERR_INVALID_ARGUMENT)
    return st;
  plan->set(nq, dtype, op, bs);
  if (best_us)
    *best_us = best;
  // y holds accumulated garbage from the timing runs: leave it cleared
  return fill_zero(y, plan->ndofs, dtype, s);
}

} // extern "C"
