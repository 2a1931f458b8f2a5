// C ABI of libfemtc.so: argument validation, dispatch to the per-degree
// kernel launchers, and the reference-shaped synchronous host entry points.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

namespace femtc
{
namespace
{
thread_local std::string g_last_error;

ApplyLauncher launcher_for(int N)
{
  switch (N)
  {
  case 2: return launch_apply_n2;
  case 3: return launch_apply_n3;
  case 4: return launch_apply_n4;
  case 5: return launch_apply_n5;
  case 6: return launch_apply_n6;
  case 7: return launch_apply_n7;
  case 8: return launch_apply_n8;
  case 9: return launch_apply_n9;
  }
  return nullptr;
}
} // namespace

int fail(int status, const std::string& msg)
{
  g_last_error = msg;
  return status;
}

int check_shape(int degree, int nq, int mode)
{
  if (degree < 1 || degree > FEMTC_MAX_DEGREE)
    return fail(FEMTC_ERR_DEGREE_NOT_SUPPORTED,
                "degree not supported: the GPU kernels cover degrees 1.." +
                    std::to_string(FEMTC_MAX_DEGREE));
  if (nq < degree + 1 || nq > degree + 3)
    return fail(FEMTC_ERR_INVALID_ARGUMENT,
                "nq must be in [degree+1, degree+3], got " +
                    std::to_string(nq));
  if (mode != FEMTC_MODE_CORRECT && mode != FEMTC_MODE_REFERENCE)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "mode must be 0 or 1");
  if (mode == FEMTC_MODE_REFERENCE && nq != degree + 1)
    return fail(FEMTC_ERR_INVALID_ARGUMENT,
                "reference mode requires nq == degree+1 (the reference "
                "hard-wires Q = P+1)");
  return FEMTC_OK;
}

int check_shape_fused(int degree, int nq, int mode)
{
  if (degree > FEMTC_MAX_FUSED_DEGREE)
    return fail(FEMTC_ERR_DEGREE_NOT_SUPPORTED,
                "degree not supported: plans, solvers and batches cover "
                "degrees 1.." +
                    std::to_string(FEMTC_MAX_FUSED_DEGREE));
  return check_shape(degree, nq, mode);
}

int check_dtype(int dtype)
{
  if (dtype != FEMTC_F64 && dtype != FEMTC_F32)
    return fail(FEMTC_ERR_INVALID_ARGUMENT,
                "dtype must be FEMTC_F64 or FEMTC_F32 for operator apply");
  return FEMTC_OK;
}

int run(const ApplyReq& r)
{
  // The kernels read a cell's geometry block as pairs of elements.
  const uintptr_t galign = 2 * (r.dtype == FEMTC_F64 ? 8 : 4);
  if ((r.op == OP_STIFF || r.op == OP_KPM || r.op == OP_STIFF_LIT) && r.G &&
      (uintptr_t)r.G % galign != 0)
    return fail(FEMTC_ERR_INVALID_ARGUMENT,
                "G must be aligned to two elements (16 bytes for FP64, 8 for "
                "FP32)");
  ApplyLauncher f = launcher_for(r.N);
  if (!f && r.N <= FEMTC_MAX_DEGREE + 1)
    f = launch_apply_general; // degrees above the fused kernels
  if (!f)
    return fail(FEMTC_ERR_DEGREE_NOT_SUPPORTED, "degree not supported");
  const int st = f(r);
  if (st == FEMTC_ERR_HIP)
    return fail(st, std::string("kernel launch failed: ") +
                        hipGetErrorString(hipGetLastError()));
  if (st != FEMTC_OK)
    return fail(st, "operator variant not available for these arguments");
  return st;
}

namespace
{
size_t dsize(int dtype) { return dtype == FEMTC_F64 ? 8 : 4; }

// RAII device buffer for the synchronous host entry points.
struct DevBuf
{
  void* p = nullptr;
  hipError_t alloc(size_t n)
  {
    if (n == 0)
      n = 1;
    return hipMalloc(&p, n);
  }
  ~DevBuf()
  {
    if (p)
      (void)hipFree(p);
  }
};
} // namespace

void set_last_error(const char* msg) { g_last_error = msg; }

} // namespace femtc

using namespace femtc;

extern "C" {

const char* femtc_status_string(int status)
{
  switch (status)
  {
  case FEMTC_OK: return "ok";
  case FEMTC_ERR_INVALID_ARGUMENT: return "invalid argument";
  case FEMTC_ERR_DEGREE_NOT_SUPPORTED: return "degree not supported";
  case FEMTC_ERR_ORDER_NOT_SUPPORTED: return "order not supported";
  case FEMTC_ERR_BLOCK_NOT_SUPPORTED: return "Block size not supported";
  case FEMTC_ERR_INVALID_ORDER_STRING: return "Invalid loop order";
  case FEMTC_ERR_HIP: return "HIP runtime error";
  case FEMTC_ERR_OUT_OF_MEMORY: return "out of device memory";
  case FEMTC_ERR_NOT_CONVERGED: return "solver did not converge";
  case FEMTC_ERR_BREAKDOWN: return "solver breakdown";
  case FEMTC_ERR_NO_DEVICE: return "no HIP device";
  }
  return "unknown status";
}

const char* femtc_last_error(void) { return g_last_error.c_str(); }

int femtc_version(void) { return FEMTC_VERSION; }

int femtc_max_degree(void) { return FEMTC_MAX_DEGREE; }

int femtc_max_fused_degree(void) { return FEMTC_MAX_FUSED_DEGREE; }

int femtc_string2order(const char* name, int* order)
{
  static const char* names[6] = {"ijk", "ikj", "jik", "jki", "kij", "kji"};
  if (!name || !order)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "null argument");
  for (int i = 0; i < 6; ++i)
    if (std::strcmp(name, names[i]) == 0)
    {
      *order = i;
      return FEMTC_OK;
    }
  return fail(FEMTC_ERR_INVALID_ORDER_STRING, "Invalid loop order");
}

int femtc_check_ladder(int degree, int order, int Mb, int Nb)
{
  // Same nesting as the reference ladder: Nb (src/mass.hpp:196-219), then
  // Mb (:171-194), then order (:140-169), then degree (:52-138).
  auto block_ok = [](int b) { return b == 0 || b == 4 || b == 8 || b == 16; };
  if (!block_ok(Nb) || !block_ok(Mb))
    return fail(FEMTC_ERR_BLOCK_NOT_SUPPORTED, "Block size not supported");
  if (order < 0 || order > 5)
    return fail(FEMTC_ERR_ORDER_NOT_SUPPORTED, "order not supported");
  if (degree < 1 || degree > FEMTC_MAX_DEGREE)
    return fail(FEMTC_ERR_DEGREE_NOT_SUPPORTED, "degree not supported");
  return FEMTC_OK;
}

int femtc_mass_apply(int dtype, int degree, int nq, int mode,
                     const void* tables, const void* U, void* W,
                     const void* detJw, int64_t num_cells, void* stream)
{
  int st = check_shape(degree, nq, mode);
  if (st || (st = check_dtype(dtype)))
    return st;
  if (num_cells < 0)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "num_cells < 0");
  if (num_cells == 0)
    return FEMTC_OK;
  if (!tables || !U || !W || !detJw)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "null buffer");
  ApplyReq r;
  r.op = OP_MASS;
  r.dtype = dtype;
  r.N = degree + 1;
  r.Q = nq;
  r.reference_mode = mode;
  r.tab = tables;
  r.x = U;
  r.y = W;
  r.detJw = detJw;
  r.num_cells = num_cells;
  r.stream = (hipStream_t)stream;
  return run(r);
}

int femtc_stiffness_apply(int dtype, int degree, int nq, int mode,
                          const void* tables, const void* U, void* W,
                          const void* G, int64_t num_cells, void* stream)
{
  int st = check_shape(degree, nq, mode);
  if (st || (st = check_dtype(dtype)))
    return st;
  if (num_cells < 0)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "num_cells < 0");
  if (num_cells == 0)
    return FEMTC_OK;
  if (!tables || !U || !W || !G)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "null buffer");
  ApplyReq r;
  r.op = OP_STIFF;
  r.dtype = dtype;
  r.N = degree + 1;
  r.Q = nq;
  r.reference_mode = mode;
  r.tab = tables;
  r.x = U;
  r.y = W;
  r.G = G;
  r.num_cells = num_cells;
  r.stream = (hipStream_t)stream;
  return run(r);
}

int femtc_kpm_apply(int dtype, int degree, int nq, const void* tables,
                    const void* U, void* W, const void* G, const void* detJw,
                    int64_t num_cells, void* stream)
{
  int st = check_shape(degree, nq, FEMTC_MODE_CORRECT);
  if (st || (st = check_dtype(dtype)))
    return st;
  if (num_cells < 0)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "num_cells < 0");
  if (num_cells == 0)
    return FEMTC_OK;
  if (!tables || !U || !W || !G || !detJw)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "null buffer");
  ApplyReq r;
  r.op = OP_KPM;
  r.dtype = dtype;
  r.N = degree + 1;
  r.Q = nq;
  r.tab = tables;
  r.x = U;
  r.y = W;
  r.G = G;
  r.detJw = detJw;
  r.num_cells = num_cells;
  r.stream = (hipStream_t)stream;
  return run(r);
}

static int global_apply_impl(int op, int dtype, int degree, int nq,
                             const void* tables, const void* x, void* y,
                             int64_t ndofs, const int32_t* dofmap,
                             const void* G, const void* detJw,
                             int64_t num_cells, void* stream, bool accumulate)
{
  int st = check_shape(degree, nq, FEMTC_MODE_CORRECT);
  if (st || (st = check_dtype(dtype)))
    return st;
  if (op != FEMTC_OP_MASS && op != FEMTC_OP_STIFFNESS && op != FEMTC_OP_KPM)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "op must be a femtc_op");
  if (num_cells < 0 || ndofs < 0)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "negative size");
  if (ndofs > 0 && (!x || !y))
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "null vector");
  if (num_cells > 0 && (!tables || !dofmap))
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "null tables or dofmap");
  if (num_cells > 0 && op != FEMTC_OP_MASS && !G)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "stiffness needs G");
  if (num_cells > 0 && op != FEMTC_OP_STIFFNESS && !detJw)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "mass needs detJw");
  hipStream_t s = (hipStream_t)stream;
  if (!accumulate && (st = fill_zero(y, ndofs, dtype, s)))
    return fail(st, "fill kernel launch failed");
  if (num_cells == 0)
    return FEMTC_OK;
  ApplyReq r;
  r.op = op == FEMTC_OP_MASS ? OP_MASS
                             : (op == FEMTC_OP_STIFFNESS ? OP_STIFF : OP_KPM);
  r.dtype = dtype;
  r.N = degree + 1;
  r.Q = nq;
  r.gather = 1;
  r.tab = tables;
  r.x = x;
  r.y = y;
  r.G = G;
  r.detJw = detJw;
  r.dofmap = dofmap;
  r.num_cells = num_cells;
  r.stream = s;
  return run(r);
}

int femtc_global_apply(int op, int dtype, int degree, int nq,
                       const void* tables, const void* x, void* y,
                       int64_t ndofs, const int32_t* dofmap, const void* G,
                       const void* detJw, int64_t num_cells, void* stream)
{
  return global_apply_impl(op, dtype, degree, nq, tables, x, y, ndofs, dofmap,
                           G, detJw, num_cells, stream, false);
}

int femtc_global_apply_add(int op, int dtype, int degree, int nq,
                           const void* tables, const void* x, void* y,
                           int64_t ndofs, const int32_t* dofmap, const void* G,
                           const void* detJw, int64_t num_cells, void* stream)
{
  return global_apply_impl(op, dtype, degree, nq, tables, x, y, ndofs, dofmap,
                           G, detJw, num_cells, stream, true);
}

int femtc_global_apply_affine(int op, int dtype, int degree, int nq,
                              const void* tables, const void* x, void* y,
                              int64_t ndofs, const int32_t* dofmap,
                              const void* cellgeo, int64_t num_cells,
                              int accumulate, void* stream)
{
  int st = check_shape(degree, nq, FEMTC_MODE_CORRECT);
  if (st || (st = check_dtype(dtype)))
    return st;
  if (op != FEMTC_OP_MASS && op != FEMTC_OP_STIFFNESS && op != FEMTC_OP_KPM)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "op must be a femtc_op");
  if (num_cells < 0 || ndofs < 0)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "negative size");
  if (ndofs > 0 && (!x || !y))
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "null vector");
  if (num_cells > 0 && (!tables || !dofmap || !cellgeo))
    return fail(FEMTC_ERR_INVALID_ARGUMENT,
                "null tables, dofmap or cell geometry");
  hipStream_t s = (hipStream_t)stream;
  if (!accumulate && (st = fill_zero(y, ndofs, dtype, s)))
    return fail(st, "fill kernel launch failed");
  if (num_cells == 0)
    return FEMTC_OK;
  ApplyReq r;
  r.op = op == FEMTC_OP_MASS
             ? OP_MASS_AFF
             : (op == FEMTC_OP_STIFFNESS ? OP_STIFF_AFF : OP_KPM_AFF);
  r.dtype = dtype;
  r.N = degree + 1;
  r.Q = nq;
  r.gather = 1;
  r.tab = tables;
  r.x = x;
  r.y = y;
  r.G = cellgeo;
  r.dofmap = dofmap;
  r.num_cells = num_cells;
  r.stream = s;
  return run(r);
}

int femtc_basis_tables_weights(int degree, int nq, double* tables)
{
  if (!tables)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "tables is NULL");
  int st = femtc_basis_tables(degree, nq, tables);
  if (st != FEMTC_OK)
    return st;
  std::vector<double> xq(nq);
  return femtc_gauss_rule(nq, xq.data(), tables + 2 * (degree + 1) * nq);
}

int femtc_fill_zero(void* y, int64_t n, int dtype, void* stream)
{
  if (n < 0 || (n > 0 && !y))
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "invalid fill arguments");
  const int st = fill_zero(y, n, dtype, (hipStream_t)stream);
  return st ? fail(st, "fill kernel launch failed") : FEMTC_OK;
}

// Synchronous host-buffer entry points used by the reference-signature shims.
static int host_apply(int op, int dtype, const void* phi, size_t phi_len,
                      const void* U, void* W, const void* geom,
                      long long num_cells, int degree, int order, int Mb,
                      int Nb, int mode)
{
  int st = femtc_check_ladder(degree, order, Mb, Nb);
  if (st)
    return st;
  if ((st = check_dtype(dtype)))
    return st;
  if (mode != FEMTC_MODE_CORRECT && mode != FEMTC_MODE_REFERENCE)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "mode must be 0 or 1");
  if (num_cells < 0)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "num_cells < 0");
  const size_t N = degree + 1, nd = N * N * N;
  const size_t need = (op == OP_MASS ? 1 : 2) * N * N;
  if (phi_len < need)
    return fail(FEMTC_ERR_INVALID_ARGUMENT,
                op == OP_MASS
                    ? "phi must hold (P+1)^2 values"
                    : "phi must hold [phi | dphi] = 2 (P+1)^2 values");
  if (num_cells == 0)
    return FEMTC_OK;
  if (!phi || !U || !W || !geom)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "null buffer");
  const size_t es = dsize(dtype);
  const size_t nU = (size_t)num_cells * nd;
  const size_t nG = (size_t)num_cells * nd * (op == OP_MASS ? 1 : 6);
  DevBuf dphi, dU, dW, dG;
  if (dphi.alloc(need * es) || dU.alloc(nU * es) || dW.alloc(nU * es) ||
      dG.alloc(nG * es))
    return fail(FEMTC_ERR_OUT_OF_MEMORY, "hipMalloc failed");
  if (hipMemcpy(dphi.p, phi, need * es, hipMemcpyHostToDevice) ||
      hipMemcpy(dU.p, U, nU * es, hipMemcpyHostToDevice) ||
      hipMemcpy(dG.p, geom, nG * es, hipMemcpyHostToDevice))
    return fail(FEMTC_ERR_HIP, "host to device copy failed");
  st = op == OP_MASS
           ? femtc_mass_apply(dtype, degree, degree + 1, mode, dphi.p, dU.p,
                              dW.p, dG.p, num_cells, nullptr)
           : femtc_stiffness_apply(dtype, degree, degree + 1, mode, dphi.p,
                                   dU.p, dW.p, dG.p, num_cells, nullptr);
  if (st)
    return st;
  if (hipMemcpy(W, dW.p, nU * es, hipMemcpyDeviceToHost))
    return fail(FEMTC_ERR_HIP, "device to host copy failed");
  return FEMTC_OK;
}

int femtc_mass_operator(int dtype, const void* phi, size_t phi_len,
                        const void* U, void* W, const void* detJ,
                        size_t num_cells, int degree, int order, int Mb,
                        int Nb, int mode)
{
  return host_apply(OP_MASS, dtype, phi, phi_len, U, W, detJ,
                    (long long)num_cells, degree, order, Mb, Nb, mode);
}

int femtc_stiffness_operator(int dtype, const void* phi, size_t phi_len,
                             const void* U, void* W, const void* G,
                             int num_cells, int degree, int order, int Mb,
                             int Nb, int mode)
{
  return host_apply(OP_STIFF, dtype, phi, phi_len, U, W, G, num_cells, degree,
                    order, Mb, Nb, mode);
}

} // extern "C"
