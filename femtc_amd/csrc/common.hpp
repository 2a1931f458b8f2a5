// Shared definitions for the femtc HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "femtc.h"

namespace femtc
{

// Operator selectors used by the fused cell kernel.
enum Op : int
{
  OP_MASS = 0,       // (Phi^T x Phi^T x Phi^T) diag(detJw) (Phi x Phi x Phi)
  OP_STIFF = 1,      // sum_ij B_i^T S_ij B_j
  OP_KPM = 2,        // stiffness + mass
  OP_STIFF_LIT = 3,  // the reference's literal stiffness (src/stiffness.hpp:34-99)
  // the same operators on affine cells with the geometry computed in the
  // kernel from 7 values per cell (detJ J^-1 J^-T packed as G, and detJ)
  // times the tensor-product quadrature weights
  OP_MASS_AFF = 4,
  OP_STIFF_AFF = 5,
  OP_KPM_AFF = 6,
};

// Values per cell of the affine geometry: G's packing, then detJ.
constexpr int CELLGEO = 7;

// One operator application request, filled by the C-ABI layer and consumed
// by the per-degree launchers.
struct ApplyReq
{
  int op = OP_MASS;
  int dtype = FEMTC_F64;
  int N = 0, Q = 0;
  int reference_mode = 0;  // backward tables read untransposed (literal)
  int gather = 0;          // x/y are global vectors addressed through dofmap
  const void* tab = nullptr;
  const void* x = nullptr;
  void* y = nullptr;
  const void* G = nullptr;
  const void* detJw = nullptr;
  const int32_t* dofmap = nullptr;
  int64_t num_cells = 0;
  // optional device flag: when *skip != 0 at launch time the kernel does
  // nothing (lets a solver queue iterations past convergence for free)
  const int* skip = nullptr;
  // optional: per-wave partial sums of x^T A x (the quadrature-point
  // energy), one double per wave, [ceil(C/CPB)][waves per workgroup];
  // unblocked launches only
  double* energy = nullptr;
  hipStream_t stream = nullptr;
  // LDS-assembled scatter (gather only).  Cells are taken as numbered
  // (ci*ncy + cj)*ncz + ck on an ncx x ncy x ncz grid and processed in
  // cbx x cby x cbz blocks, one workgroup each; touch[g] = how many
  // cell-dofs map to global dof g (saturated at 255).  Correct for any
  // dofmap; fast when it is the structured one.
  int blocked = 0;
  int ncx = 0, ncy = 0, ncz = 0;
  int cbx = 0, cby = 0, cbz = 0;
  const uint8_t* touch = nullptr;
  int overwrite = 0; // y was cleared: dofs owned by one block are stored
  // Assembled scatter (gather only, ncx..ncz as above, the dofmap verified to
  // be the structured one by verify_structured): the cell pass stores every
  // cell-dof value into `lattice` (lattice_size() values, assemble.hpp)
  // without atomics; assembled = 1 then sums them into y (y = sum when
  // overwrite, else y += sum), assembled = 2 stops after the cell pass.
  int assembled = 0;
  void* lattice = nullptr;
  // Structured direct scatter (gather only, ncx..ncz as above, the dofmap
  // verified to be the structured one): closed-form gather indices and the
  // atomic scatter at the same indices; the dofmap is not read.
  int structured_direct = 0;
  // Structured z-run scatter (gather only, as structured_direct): each
  // workgroup sums a segment of cells along z in LDS and adds the segment's
  // dof box to y with atomics in runs of consecutive global dofs.
  int zrun = 0;
  // z-run scatter only: when set, the launcher stores the instance's
  // (cells per segment, waves per workgroup) here and launches nothing
  int* zrun_shape = nullptr;
  // Even-odd table set (even_odd.hpp) of symmetric tables, used instead of
  // tab when set: Q == N, gather, correct mode, direct or assembled scatter.
  const void* eo_tab = nullptr;
};

// BW template value of the assembled scatter's cell-pass instances.
constexpr int SCATTER_ASSEMBLED = -1;
// BW template value of the structured direct scatter: the dofmap verified to
// be the structured numbering, the gather and the atomic scatter use the
// global indices in closed form instead of reading the dofmap.
constexpr int SCATTER_STRUCT = -2;
// BW template value of the structured z-run scatter (ApplyReq::zrun).
constexpr int SCATTER_ZRUN = -3;

// Cells per workgroup of the fused cell kernel (Q^2 threads per cell):
// one wave of whole cells up to Q = 8.  At Q = 9 one cell in a 2-wave
// workgroup would leave 47 of its 128 lanes idle; three cells take 243 of
// 4 waves' 256.  At Q = 10, 11 one cell keeps 100 and 121 of 128 busy.
constexpr int direct_cells(int Q)
{
  return Q * Q <= 64 ? 64 / (Q * Q) : Q == 9 ? 3 : 1;
}

// Cells per workgroup of the assembled scatter's cell pass (esize: bytes
// per value).  One wave of whole cells leaves lanes idle at Q = 6 and 7 (36
// and 49 of 64 busy); for FP32 the workgroup there takes 7 and 5 cells over
// 4 waves (252/256, 245/256 lanes), which made the ALU-bound FP32 cell pass
// faster.  FP64 (and the direct and blocked scatters, whose atomics
// measured 4-19 % slower from multi-wave workgroups) keep one-wave
// workgroups there.
constexpr int assembled_cells(int Q, int esize)
{
  return esize == 4 && Q == 6   ? 7
         : esize == 4 && Q == 7 ? 5
                                : direct_cells(Q);
}

// One problem of a batched launch (device-resident descriptor).
struct BatchDesc
{
  const void* tab;
  const void* x;
  void* y;
  const void* G;
  const void* detJw;
  const int32_t* dofmap; // NULL for cell-local problems
  long long num_cells;
  long long blk_begin; // first workgroup-sized block of this problem
  int grid[3];         // the z-run launch: the problem's structured grid
};

// Problems per batched launch (their first blocks are searched in LDS).
constexpr int BATCH_MAX_ITEMS = 64;

// One batched launch: every problem shares (op, dtype, N, Q, gather).
struct BatchReq
{
  int op = OP_MASS;
  int dtype = FEMTC_F64;
  int N = 0, Q = 0;
  int gather = 0;
  const BatchDesc* items = nullptr; // device array, sorted by blk_begin
  int nitems = 0;
  long long nblk_total = 0;
  int per_wg = 1;
  // 1: the z-run scatter (items with their grids; blocks are z-segments,
  // blk_begin the problem's first segment); with zrun_shape set the
  // launcher stores (cells per segment, waves) there and launches nothing
  int zrun = 0;
  int* zrun_shape = nullptr;
  hipStream_t stream = nullptr;
};

using BatchLauncher = int (*)(const BatchReq&);
int launch_batch_n2(const BatchReq&);
int launch_batch_n3(const BatchReq&);
int launch_batch_n4(const BatchReq&);
int launch_batch_n5(const BatchReq&);
int launch_batch_n6(const BatchReq&);
int launch_batch_n7(const BatchReq&);
int launch_batch_n8(const BatchReq&);
int launch_batch_n9(const BatchReq&);

// Kernel-side view of the blocked-scatter parameters.
struct BlockArgs
{
  int ncx, ncy, ncz; // mesh cells per axis
  int cbx, cby, cbz; // cells per block per axis
  int nby, nbz;      // blocks along the two fastest axes
  const uint8_t* touch;
  int overwrite;
  // assembled scatter: cell blocks; the grid is 8 * ceil(nblk / 8)
  // workgroups taken XCD-chunked (see apply_cells)
  int nblk;
};

using ApplyLauncher = int (*)(const ApplyReq&);

// Launchers, one translation unit per N = P+1 (apply_n<N>.hip).
int launch_apply_n2(const ApplyReq&);
int launch_apply_n3(const ApplyReq&);
int launch_apply_n4(const ApplyReq&);
int launch_apply_n5(const ApplyReq&);
int launch_apply_n6(const ApplyReq&);
int launch_apply_n7(const ApplyReq&);
int launch_apply_n8(const ApplyReq&);
int launch_apply_n9(const ApplyReq&);
// N = 10..26 (apply_general.hip); also accepts smaller N
int launch_apply_general(const ApplyReq&);
// Line kernels, N = 10..26 (apply_line.hpp); max_lds = the device's LDS
// bytes per workgroup.  FEMTC_ERR_INVALID_ARGUMENT when the variant does not
// exist or does not fit.
int launch_line_n10(const ApplyReq&, int max_lds);
int launch_line_n11(const ApplyReq&, int max_lds);
int launch_line_n12(const ApplyReq&, int max_lds);
int launch_line_n13(const ApplyReq&, int max_lds);
int launch_line_n14(const ApplyReq&, int max_lds);
int launch_line_n15(const ApplyReq&, int max_lds);
int launch_line_n16(const ApplyReq&, int max_lds);
int launch_line_n17(const ApplyReq&, int max_lds);
int launch_line_n18(const ApplyReq&, int max_lds);
int launch_line_n19(const ApplyReq&, int max_lds);
int launch_line_n20(const ApplyReq&, int max_lds);
int launch_line_n21(const ApplyReq&, int max_lds);
int launch_line_n22(const ApplyReq&, int max_lds);
int launch_line_n23(const ApplyReq&, int max_lds);
int launch_line_n24(const ApplyReq&, int max_lds);
int launch_line_n25(const ApplyReq&, int max_lds);
int launch_line_n26(const ApplyReq&, int max_lds);

// The K+M cell pass with bfloat16 operands on the matrix cores and FP32
// accumulation (mfma_cell.hip): affine cells of a structured ncx x ncy x ncz
// grid, Q == N <= 8, FP32 x in, cell values stored into the assembled
// scatter's lattice (assemble.hpp), one x_e^T A_e x_e partial per cell.
// ks and km scale the stiffness and the mass term (1 or 0: K+M, K or M).
struct MfmaReq
{
  int N = 0, Q = 0;
  int split = 1;               // 1: values split hi + lo as well as tables
  const void* amat = nullptr;  // operand set of mfma_build_tables (device)
  const float* w = nullptr;    // the Q Gauss weights (device)
  const float* x = nullptr;
  float* lattice = nullptr;
  const float* cg = nullptr;   // [C][7]: G's packing of detJ J^-1 J^-T, detJ
  float ks = 1.f, km = 1.f;
  int ncx = 0, ncy = 0, ncz = 0;
  const int* skip = nullptr;
  double* energy = nullptr;
  hipStream_t stream = nullptr;
};
int launch_mfma(const MfmaReq& r);
// Matrices of the operand set, and the set itself: nmat * 64 lanes * 8
// bfloat16 values from the FP64 host tables [phi | phi'] ([2][N][Q]).
int mfma_table_matrices(int split);
int mfma_build_tables(const double* tab, int N, int Q, int split,
                      uint16_t* out);

inline int hip_status(hipError_t e)
{
  return e == hipSuccess ? FEMTC_OK : FEMTC_ERR_HIP;
}

// Shared validation and dispatch (api.cpp).
int fail(int status, const std::string& msg);
int check_shape(int degree, int nq, int mode);
// check_shape, and a degree the fused kernels cover (plans, solvers, batches)
int check_shape_fused(int degree, int nq, int mode);
int check_dtype(int dtype);
int run(const ApplyReq& r); // validated request -> per-degree launcher
int fill_zero(void* y, int64_t n, int dtype, hipStream_t s);

// *ok = 1 when dofmap [C][N^3] is the structured numbering of the
// nx x ny x nz cell grid (femtc_dofmap_structured) and ndofs its dof count,
// which the assembled scatter requires.  Synchronous on `s` (plan.hip).
int verify_structured(const int32_t* dofmap, int nx, int ny, int nz,
                      int degree, int64_t ndofs, hipStream_t s, int* ok);

// *ok = 1 when every one of the n dofmap entries is in [0, ndofs), before
// any kernel indexes through them.  Synchronous on `s` (plan.hip).
int check_dofmap_range(const int32_t* dofmap, int64_t n, int64_t ndofs,
                       hipStream_t s, int* ok);

// Reads the device tables [phi | phi'] ([2][N][Q] in `dtype`) on `s`; when
// they pass eo_symmetric, allocates and fills *eo with their even-odd table
// set (even_odd.hpp) in `out_dtype`, with the weights w (may be null),
// else leaves *eo null.  Synchronous on `s` (plan.hip).
int make_even_odd(const void* tables, int dtype, int N, int Q, const double* w,
                  int out_dtype, hipStream_t s, void** eo);

} // namespace femtc
