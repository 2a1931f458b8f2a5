/*
 * femtc.h -- C ABI of the MI355X (gfx950) sum-factorized finite element
 * operator library (libfemtc.so).
 *
 * Plain pointers and sizes only; no HIP or C++ types cross this boundary.
 * "stream" arguments are hipStream_t values passed as void* (NULL = the
 * default stream).  Device entry points are asynchronous and stream-ordered;
 * host entry points are synchronous.  Every function returns a femtc_status;
 * nothing is thrown across the boundary.  femtc_last_error() returns a
 * human-readable detail string for the calling thread's last failure.
 *
 * Layouts (the reference's, generalised to Q quadrature points per
 * direction; src/mass.hpp:25-29, src/stiffness.hpp:44-57):
 *   N = degree+1, nd = N^3 dofs per cell, nq3 = Q^3 points per cell.
 *   U, W      [num_cells][nd], row-major [a][b][c], a slowest.
 *   tables    [phi | dphi], each [N][Q] row-major: phi[d*Q+q] = phi_d(x_q).
 *   detJw     [num_cells][nq3].
 *   G         [num_cells][nq3][6] packed (S22,S12,S02,S11,S01,S00) of the
 *             symmetric 3x3 tensor acting on (d/da, d/db, d/dc)
 *             (src/stiffness.hpp:25-27).
 *   dofmap    int32 [num_cells][nd], global dof of each cell dof.
 */
#ifndef FEMTC_H
#define FEMTC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FEMTC_VERSION 1
/* Operator apply (cell-local and through a dofmap) and the reference
 * ladder: degrees 1..25 (src/mass.hpp:52-138).  Plans, the solver, batches
 * and the even-odd tables run degrees 1..FEMTC_MAX_FUSED_DEGREE. */
#define FEMTC_MAX_DEGREE 25
#define FEMTC_MAX_FUSED_DEGREE 8

typedef enum femtc_status
{
  FEMTC_OK = 0,
  FEMTC_ERR_INVALID_ARGUMENT = 1,
  FEMTC_ERR_DEGREE_NOT_SUPPORTED = 2, /* src/mass.hpp:134-136 */
  FEMTC_ERR_ORDER_NOT_SUPPORTED = 3,  /* src/mass.hpp:165-167 */
  FEMTC_ERR_BLOCK_NOT_SUPPORTED = 4,  /* src/mass.hpp:190-192, 215-217 */
  FEMTC_ERR_INVALID_ORDER_STRING = 5, /* src/linalg.hpp:36 */
  FEMTC_ERR_HIP = 6,
  FEMTC_ERR_OUT_OF_MEMORY = 7,
  FEMTC_ERR_NOT_CONVERGED = 8,
  FEMTC_ERR_BREAKDOWN = 9,
  FEMTC_ERR_NO_DEVICE = 10
} femtc_status;

typedef enum femtc_dtype
{
  FEMTC_F64 = 0,
  FEMTC_F32 = 1,
  /* Inner precisions of a refined solve only (femtc_solver_create): the
   * operator with bfloat16 operands on the matrix cores and FP32
   * accumulation (femtc_global_apply_mfma), FP32 Krylov vectors.
   * FEMTC_BF16 splits the tables and the values (split = 1, about 1e-5);
   * FEMTC_BF16_ROUNDED rounds the values (split = 0, about 3e-3: pair it
   * with an inner tolerance near 3e-3, tighter solves of the perturbed
   * operator buy nothing). */
  FEMTC_BF16 = 2,
  FEMTC_BF16_ROUNDED = 3
} femtc_dtype;

/* linalg::Order, same enumerators in the same order (src/linalg.hpp:11-19).
 * The GPU kernels do not depend on it; it is validated and accepted. */
typedef enum femtc_order
{
  FEMTC_ORDER_IJK = 0,
  FEMTC_ORDER_IKJ = 1,
  FEMTC_ORDER_JIK = 2,
  FEMTC_ORDER_JKI = 3,
  FEMTC_ORDER_KIJ = 4,
  FEMTC_ORDER_KJI = 5
} femtc_order;

/* FEMTC_MODE_CORRECT: the mathematically intended operator (true transposed
 * tables, all gradient chains summed).  FEMTC_MODE_REFERENCE: what the
 * reference computes, defects included (the backward table is the forward
 * one, src/mass.hpp:20-21; only the z backward chain of the stiffness
 * operator survives and its forward "z" field is not differentiated,
 * src/stiffness.hpp:76-97).  Reference mode requires nq == degree+1. */
typedef enum femtc_mode
{
  FEMTC_MODE_CORRECT = 0,
  FEMTC_MODE_REFERENCE = 1
} femtc_mode;

typedef enum femtc_op
{
  FEMTC_OP_MASS = 0,
  FEMTC_OP_STIFFNESS = 1,
  FEMTC_OP_KPM = 2 /* stiffness + mass */
} femtc_op;

/* ---------------------------------------------------------------------- */
/* Status and version                                                      */
/* ---------------------------------------------------------------------- */
const char* femtc_status_string(int status);
const char* femtc_last_error(void);
int femtc_version(void);
/* Highest degree of the operator apply entry points (FEMTC_MAX_DEGREE), for
 * every quadrature offset nq - degree - 1 in {0,1,2}. */
int femtc_max_degree(void);
/* Highest degree of plans, solvers, batches, the server's solves and the
 * even-odd tables (FEMTC_MAX_FUSED_DEGREE). */
int femtc_max_fused_degree(void);

/* linalg::string2order (src/linalg.hpp:21-37): "ijk".."kji" -> *order;
 * anything else returns FEMTC_ERR_INVALID_ORDER_STRING. */
int femtc_string2order(const char* name, int* order);

/* Validate the CPU blocking knobs of the reference dispatch ladder
 * (src/mass.hpp:140-219): order in 0..5, Mb and Nb in {0,4,8,16}. */
int femtc_check_ladder(int degree, int order, int Mb, int Nb);

/* ---------------------------------------------------------------------- */
/* Device memory helpers (current HIP device)                              */
/* ---------------------------------------------------------------------- */
int femtc_device_count(int* n);
int femtc_device_alloc(void** ptr, size_t bytes);
int femtc_device_free(void* ptr);
int femtc_memcpy_htod(void* dst, const void* src, size_t bytes);
int femtc_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int femtc_device_synchronize(void);

/* ---------------------------------------------------------------------- */
/* 1-D tables on the reference interval [0,1] (host, FP64)                 */
/* ---------------------------------------------------------------------- */
/* Gauss-Lobatto-Legendre nodes, degree+1 values, increasing. */
int femtc_gll_points(int degree, double* x);
/* Gauss-Legendre points and weights (weights sum to 1), nq values each. */
int femtc_gauss_rule(int nq, double* x, double* w);
/* [phi | dphi] of the GLL Lagrange basis at the Gauss points, 2*N*nq values. */
int femtc_basis_tables(int degree, int nq, double* tables);
/* The even-odd table set that plans (femtc_plan_set_tables) and solvers
 * build from [phi | dphi] (2*N*nq values) and the weights (nq values, may
 * be NULL: zeros).  *size = its length in values; *symmetric = 1 when the
 * tables pass the symmetry check (phi symmetric, dphi antisymmetric under
 * b -> N-1-b, q -> nq-1-q, to 64 ulps of the largest entry); out (may be
 * NULL to query the size) receives it.  Layout: four terms, phi forward
 * (dof -> point), dphi forward, phi backward (point -> dof), dphi backward;
 * a term with input length K, output length M and sign S (+1 phi, -1 dphi)
 * holds block A, ceil(K/2) rows (S = +1; floor(K/2) for S = -1) of
 * round4(ceil(M/2)) values, then block B, floor(K/2) rows (S = +1;
 * ceil(K/2) for S = -1) of round4(floor(M/2)) values; then the weights.
 * No counterpart in the reference (host, FP64). */
int femtc_even_odd_tables(int degree, int nq, const double* tables,
                          const double* weights, double* out, int64_t* size,
                          int* symmetric);

/* ---------------------------------------------------------------------- */
/* Cell-local operators on device buffers (stream-ordered)                 */
/* dtype FEMTC_F64 or FEMTC_F32: every array is of that type.              */
/* ---------------------------------------------------------------------- */
/* W = M U per cell. Replaces operators::mass_operator (src/mass.hpp:10-49). */
int femtc_mass_apply(int dtype, int degree, int nq, int mode,
                     const void* tables, const void* U, void* W,
                     const void* detJw, int64_t num_cells, void* stream);
/* W = K U per cell. Replaces operators::stiffness_operator
 * (src/stiffness.hpp:34-99). */
int femtc_stiffness_apply(int dtype, int degree, int nq, int mode,
                          const void* tables, const void* U, void* W,
                          const void* G, int64_t num_cells, void* stream);
/* W = (K + M) U per cell (correct mode only). */
int femtc_kpm_apply(int dtype, int degree, int nq, const void* tables,
                    const void* U, void* W, const void* G, const void* detJw,
                    int64_t num_cells, void* stream);

/* ---------------------------------------------------------------------- */
/* Continuous operator: y = A x through the dofmap (y overwritten).        */
/* op is a femtc_op; G is needed for stiffness/K+M, detJw for mass/K+M.    */
/* Contributions to shared dofs are summed with device atomics, so the     */
/* last bits of y may vary from run to run.                                */
/* ---------------------------------------------------------------------- */
int femtc_global_apply(int op, int dtype, int degree, int nq,
                       const void* tables, const void* x, void* y,
                       int64_t ndofs, const int32_t* dofmap, const void* G,
                       const void* detJw, int64_t num_cells, void* stream);
/* y += A x: the same without clearing y first (y must be initialised). */
int femtc_global_apply_add(int op, int dtype, int degree, int nq,
                           const void* tables, const void* x, void* y,
                           int64_t ndofs, const int32_t* dofmap, const void* G,
                           const void* detJw, int64_t num_cells, void* stream);

/* y[0..n) = 0 with a kernel (usable under graph capture). */
int femtc_fill_zero(void* y, int64_t n, int dtype, void* stream);

/* Kernel choice for degrees above FEMTC_MAX_FUSED_DEGREE (no counterpart
 * in the reference, whose ladder compiles one CPU loop per degree).
 * FEMTC_HIGH_DEGREE_AUTO: the line kernels where they apply (correct mode,
 * mass / stiffness / K+M with per-point or affine geometry, N = degree+1 in
 * 10..26 where a cell's registers and LDS fit: FP64 stiffness and K+M up to
 * N = 21..22, the rest up to 26), else the general kernel;
 * FEMTC_HIGH_DEGREE_GENERAL: the general kernel always.  Process-wide;
 * the previous setting goes to *previous (may be NULL).  Another mode:
 * FEMTC_ERR_INVALID_ARGUMENT, the setting unchanged. */
#define FEMTC_HIGH_DEGREE_AUTO 0
#define FEMTC_HIGH_DEGREE_GENERAL 1
int femtc_set_high_degree_kernel(int mode, int* previous);

/* Which kernel family the operator apply entry points run for (op
 * FEMTC_OP_*, dtype, degree, nq, mode FEMTC_MODE_*) under the current
 * setting above (no counterpart in the reference): *kernel =
 * FEMTC_KERNEL_FUSED (degrees 1..FEMTC_MAX_FUSED_DEGREE),
 * FEMTC_KERNEL_LINE or FEMTC_KERNEL_GENERAL.  The affine-geometry applies
 * (femtc_global_apply_affine) follow the same choice.  Assumes the 160 KB
 * of LDS per workgroup of gfx950; no device call. */
#define FEMTC_KERNEL_FUSED 0
#define FEMTC_KERNEL_GENERAL 1
#define FEMTC_KERNEL_LINE 2
int femtc_apply_kernel(int op, int dtype, int degree, int nq, int mode,
                       int* kernel);

/* Streaming probe (diagnostic; no counterpart in the reference): the
 * practical HBM bandwidth of a read-dominated stream like the operator
 * apply.  dst[i] = sum_r src[r*n + i] over nread (1..8) arrays of n (even)
 * doubles when nwrite = 1; with nwrite = 0 the sums land in sink (one per
 * workgroup, nsink >= 1 values).  16-byte non-temporal loads, grid-stride
 * over wg_per_cu (1..64; 0: 2) workgroups of 256 lanes per compute unit. */
int femtc_stream_probe(int nread, int nwrite, const double* src, double* dst,
                       int64_t n, double* sink, int nsink, int wg_per_cu,
                       void* stream);

/* ---------------------------------------------------------------------- */
/* Affine cells, geometry computed in the kernel (no counterpart in the    */
/* reference, which streams G per point): per cell 7 values cellgeo =      */
/* (|det J| J^-1 J^-T packed as G, |det J|), scaled in the kernel by the   */
/* tensor-product Gauss weights.  Equals the stored-geometry operator for  */
/* per-cell constant J while reading 7 values per cell instead of 7 nq^3.  */
/* tables here hold [phi | dphi | w]: 2*N*nq + nq values.                  */
/* ---------------------------------------------------------------------- */
/* J [num_cells][9] (FP64) -> cellgeo [num_cells][7] in dtype. */
int femtc_geometry_cell(int64_t num_cells, const double* J, void* cellgeo,
                        int dtype, void* stream);
/* [phi | dphi | w] (host, FP64), 2*N*nq + nq values. */
int femtc_basis_tables_weights(int degree, int nq, double* tables);
/* y = A x (accumulate = 0, y overwritten) or y += A x through the dofmap. */
int femtc_global_apply_affine(int op, int dtype, int degree, int nq,
                              const void* tables, const void* x, void* y,
                              int64_t ndofs, const int32_t* dofmap,
                              const void* cellgeo, int64_t num_cells,
                              int accumulate, void* stream);

/* ---------------------------------------------------------------------- */
/* The same operators with bfloat16 operands on the matrix cores (MFMA)    */
/* and FP32 accumulation: the operator of the BF16 inner solve.  No        */
/* counterpart in the reference, whose contraction chain                   */
/* (src/stiffness.hpp:68-97, src/mass.hpp:36-47) runs in T = float or      */
/* double.  Affine cells of a structured nx x ny x nz grid (cell (i,j,k)   */
/* is (i*ny + j)*nz + k, dofs numbered as femtc_dofmap_structured numbers  */
/* them, so no dofmap is read), nq = degree + 1, degree 1..7, FP32         */
/* vectors.  Every 1-D contraction is a bfloat16 MFMA product with FP32    */
/* accumulation.  The tables are split into two bfloat16 parts (hi + lo);  */
/* split = 1 splits the cell values too (products hi.hi + hi.lo + lo.hi:   */
/* about 1e-5 relative to the FP64 operator), split = 0 rounds them to     */
/* bfloat16 once per contraction (about 3e-3).                             */
/* ---------------------------------------------------------------------- */
/* Bytes of the operand set (the tables as MFMA operands, and the weights). */
int femtc_mfma_tables_bytes(int split, size_t* bytes);
/* The operand set built on the host into host memory of
 * femtc_mfma_tables_bytes bytes: for each matrix m, lane l, slot s the
 * bfloat16 A[l % 16][8 (l / 16) + s] at ((m * 64 + l) * 8 + s), then the nq
 * Gauss weights as 8 floats.  No device work. */
int femtc_mfma_tables_host(int degree, int nq, int split,
                           const double* tables, void* operands);
/* The same, from the host tables [phi | dphi] (FP64, 2*N*nq values) into
 * device memory of femtc_mfma_tables_bytes bytes.  Synchronous. */
int femtc_mfma_tables(int degree, int nq, int split, const double* tables,
                      void* operands);
/* y = A x (accumulate = 0) or y += A x, op a femtc_op; cellgeo [C][7] in
 * FP32 as femtc_geometry_cell writes it; lattice: a device workspace of
 * C*(degree+1)^3 floats that receives the cell values, which a second
 * kernel sums per dof (no atomics: the result is deterministic). */
int femtc_global_apply_mfma(int op, int degree, int nq, int split,
                            const void* operands, const float* x, float* y,
                            int64_t ndofs, int nx, int ny, int nz,
                            const float* cellgeo, float* lattice,
                            int accumulate, void* stream);
/* The same apply for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid it); no
 * counterpart in the reference.  It calls no allocating routine: the
 * lattice is a static device array of the library, 16 Mi floats (64 MiB)
 * per GPU it runs on, held until the process ends, so nx*ny*nz*(degree+1)^3
 * must not exceed 16777216 (FEMTC_ERR_INVALID_ARGUMENT otherwise).  Every
 * call on a GPU shares that GPU's array: calls that overlap on one GPU (other
 * streams, other threads) race.  Results equal femtc_global_apply_mfma's. */
int femtc_global_apply_mfma_noalloc(int op, int degree, int nq, int split,
                                    const void* operands, const float* x,
                                    float* y, int64_t ndofs, int nx, int ny,
                                    int nz, const float* cellgeo,
                                    int accumulate, void* stream);

/* ---------------------------------------------------------------------- */
/* Reference-shaped synchronous entry points on HOST buffers: the C side   */
/* of the operators::mass_operator<T> / stiffness_operator<T> shims        */
/* (src/mass.hpp:196-199, src/stiffness.hpp:245-248).  Q = degree+1 as in  */
/* the reference.  phi_len is the number of table values the caller owns:  */
/* (P+1)^2 for mass, 2 (P+1)^2 for stiffness ([phi | dphi]).  Copies in,   */
/* applies on the current device, copies W back.                           */
/* ---------------------------------------------------------------------- */
int femtc_mass_operator(int dtype, const void* phi, size_t phi_len,
                        const void* U, void* W, const void* detJ,
                        size_t num_cells, int degree, int order, int Mb,
                        int Nb, int mode);
int femtc_stiffness_operator(int dtype, const void* phi, size_t phi_len,
                             const void* U, void* W, const void* G,
                             int num_cells, int degree, int order, int Mb,
                             int Nb, int mode);

/* ---------------------------------------------------------------------- */
/* Mesh and geometry (device, stream-ordered)                              */
/* ---------------------------------------------------------------------- */
/* Structured nx*ny*nz hexahedral mesh of degree P: cell (ci,cj,ck) has id
 * (ci*ny+cj)*nz+ck; global dofs are the lexicographic (nx P+1)(ny P+1)
 * (nz P+1) grid, first axis slowest. */
int femtc_dofmap_structured(int nx, int ny, int nz, int degree,
                            int32_t* dofmap, void* stream);
/* Per-cell constant Jacobians J [num_cells][9] (FP64, row-major,
 * J[r][s] = dx_r/dxi_s) -> G [num_cells][nq^3][6] and/or detJw
 * [num_cells][nq^3] (either may be NULL), computed in FP64 and stored as
 * dtype (F64 or F32). */
int femtc_geometry_affine(int64_t num_cells, int nq, const double* J,
                          void* G, void* detJw, int dtype, void* stream);

/* ---------------------------------------------------------------------- */
/* Scatter plans (no counterpart in the reference, which has no dofmap).   */
/* A plan describes a dofmap whose cells are numbered (ci*ny + cj)*nz + ck */
/* on an nx*ny*nz grid (as femtc_dofmap_structured numbers them).  Applies */
/* through a plan process 3-D blocks of cells per workgroup, sum shared    */
/* dofs in LDS and store the dofs only one block touches without atomics.  */
/* Results are correct for any dofmap; the blocking pays off when the      */
/* numbering is the structured one.  The plan keeps a pointer to dofmap    */
/* (which must stay alive) and 1 byte per dof of touch counts.             */
/* Plan creation also checks whether the dofmap is exactly the structured  */
/* numbering (femtc_dofmap_structured); then the assembled scatter is      */
/* available: a cell pass stores every cell-dof value without atomics into */
/* a plan-owned buffer of C*(P+1)^3 values (allocated when the scatter is  */
/* first selected or timed), and a second pass sums the values of each     */
/* global dof into y.  Applies of one plan through it share that buffer:   */
/* they must not overlap in time (e.g. on two streams).                    */
/* ---------------------------------------------------------------------- */
typedef struct femtc_plan femtc_plan; /* opaque */
int femtc_plan_create(femtc_plan** plan, int nx, int ny, int nz, int degree,
                      const int32_t* dofmap, int64_t ndofs, void* stream);
int femtc_plan_destroy(femtc_plan* plan);
/* Block shape (cells per axis) for one (nq, dtype, op); 0,0,0 selects the
 * direct atomic scatter, -1,-1,-1 the assembled scatter, -2,-2,-2 the
 * structured direct scatter (the atomic scatter with the gather and scatter
 * indices in closed form: the dofmap is not read) and -3,-3,-3 the
 * structured z-run scatter (closed-form indices; each workgroup sums a run
 * of cells along z in LDS and adds their dofs to y with atomics on runs of
 * consecutive dofs; nq <= 8); the last three are refused with
 * FEMTC_ERR_INVALID_ARGUMENT when the dofmap is not the structured one.
 * For blocks, cbz must be a multiple of 64/nq^2 and the block at most 8
 * waves of 64/nq^2 cells; nq <= 8. */
int femtc_plan_set_block(femtc_plan* plan, int nq, int dtype, int op, int cbx,
                         int cby, int cbz);
int femtc_plan_get_block(const femtc_plan* plan, int nq, int dtype, int op,
                         int* cbx, int* cby, int* cbz);
/* y = A x (accumulate = 0: y overwritten) or y += A x (accumulate = 1).
 * tables may be NULL when femtc_plan_set_tables registered them for
 * (dtype, nq). */
int femtc_plan_apply(const femtc_plan* plan, int op, int dtype, int nq,
                     const void* tables, const void* x, void* y,
                     const void* G, const void* detJw, int accumulate,
                     void* stream);
/* Registers a copy of the tables [phi | dphi] ([2][degree+1][nq], dtype;
 * taken now, synchronously on `stream`) for the applies and autotuning of
 * (dtype, nq) that pass tables = NULL; a second call replaces it.  When
 * nq = degree + 1 and the tables are those of a symmetric node set and
 * rule (phi symmetric and dphi antisymmetric under b -> N-1-b,
 * q -> nq-1-q, to 64 rounding units, as femtc_basis_tables makes them),
 * those applies contract each line in even-odd form (half-size products;
 * results equal to rounding) on the direct and assembled scatters;
 * *even_odd (may be NULL) reports whether they do.  No counterpart in the
 * reference. */
int femtc_plan_set_tables(femtc_plan* plan, int dtype, int nq,
                          const void* tables, void* stream, int* even_odd);
/* Time every candidate block shape (and the direct and, for the structured
 * dofmap, the structured direct, z-run and assembled scatters) on these
 * buffers, reps overwriting applies
 * each (y = A x, clearing included), and keep the fastest for
 * (nq, dtype, op).  y is left cleared.  best_us (may be NULL) receives its
 * time per apply.  tables may be NULL as for femtc_plan_apply (the
 * registered copy, in even-odd form where it applies, is then timed). */
int femtc_plan_autotune(femtc_plan* plan, int op, int dtype, int nq,
                        const void* tables, const void* x, void* y,
                        const void* G, const void* detJw, int reps,
                        void* stream, double* best_us);

/* ---------------------------------------------------------------------- */
/* Batches (no counterpart in the reference): many independent problems    */
/* of uneven size, degree, operator and precision applied together.        */
/* Problems sharing (op, dtype, degree, nq, cell-local/continuous) form a  */
/* group (at most 64 problems) that runs either as one fused launch        */
/* (each workgroup finds its problem by block index) or as one launch per  */
/* problem, whichever femtc_batch_autotune measured faster (default: per   */
/* problem).  The launches are spread over up to four forked streams,     */
/* longest first (by the autotuner's times, else by the bytes); the        */
/* sequence can be captured once into a hipGraph and replayed.  Every      */
/* output is overwritten (continuous y is cleared first).  The batch keeps */
/* the item pointers: the buffers must stay alive and in place.            */
/* ---------------------------------------------------------------------- */
typedef struct femtc_batch_item
{
  int op;     /* femtc_op */
  int dtype;  /* FEMTC_F64 or FEMTC_F32: every array of the item */
  int degree;
  int nq;
  const void* tables; /* [phi | dphi] */
  const void* x;      /* continuous: [ndofs]; cell-local: [num_cells][N^3] */
  void* y;            /* same shape as x */
  const void* G;      /* [num_cells][nq^3][6] or NULL (mass) */
  const void* detJw;  /* [num_cells][nq^3] or NULL (stiffness) */
  const int32_t* dofmap; /* [num_cells][N^3], or NULL for cell-local */
  int64_t num_cells;
  int64_t ndofs; /* continuous items: length of x and y */
  /* Optional (all zero when unknown): the nx x ny x nz grid when dofmap is
   * the structured numbering (checked at batch creation); a group whose
   * items all have one may run the assembled scatter. */
  int grid[3];
} femtc_batch_item;

typedef struct femtc_batch femtc_batch; /* opaque */
int femtc_batch_create(femtc_batch** batch, const femtc_batch_item* items,
                       int nitems, void* stream);
/* Apply every item, stream-ordered on the batch's stream. */
int femtc_batch_run(femtc_batch* batch);
/* The same through a hipGraph: captured at the first call (and after an
 * autotune), replayed afterwards. */
int femtc_batch_graph_launch(femtc_batch* batch);
/* Time, for every group, one launch per problem, the fused launch with
 * 1, 2, 4 and 8 blocks per workgroup and, when every item of the group has
 * a verified grid, the assembled and (nq <= 8) z-run scatters per problem
 * and one fused z-run launch over the group (reps launches each, in two
 * rounds, after four warm-up runs of the whole batch) and keep the
 * fastest; synchronous; outputs are left undefined. */
int femtc_batch_autotune(femtc_batch* batch, int reps);
int femtc_batch_stats(const femtc_batch* batch, int* ngroups,
                      int64_t* total_cells, int64_t* total_dofs);
/* per_wg = 0: the group runs one launch per problem; per_wg = k > 0: one
 * fused launch, k blocks per workgroup; per_wg = -1: the assembled scatter
 * per problem (its lattice buffers are allocated then, C*(P+1)^3 values per
 * item); per_wg = -2: the z-run scatter per problem (nq <= 8); per_wg =
 * -3: one fused z-run launch over the group's problems (nq <= 8); -1..-3
 * give FEMTC_ERR_INVALID_ARGUMENT unless every item has a verified grid.
 * set_group drops a captured graph. */
int femtc_batch_set_group(femtc_batch* batch, int group, int per_wg);
int femtc_batch_group(const femtc_batch* batch, int group, int* op,
                      int* dtype, int* degree, int* nq, int* gather,
                      int* nitems, int* per_wg);
int femtc_batch_destroy(femtc_batch* batch);

/* ---------------------------------------------------------------------- */
/* Solver: Jacobi-preconditioned CG on the continuous operator, and FP64   */
/* defect correction around an FP32 inner PCG.  The reference has no       */
/* solver; these entry points have no counterpart there.                   */
/* ---------------------------------------------------------------------- */
/* diag = diagonal of the continuous operator op (overwritten, FP64),
 * summed cell by cell with sum factorization; inputs of dtype F64 or F32. */
int femtc_global_diagonal(int op, int dtype, int degree, int nq,
                          const void* tables, double* diag, int64_t ndofs,
                          const int32_t* dofmap, const void* G,
                          const void* detJw, int64_t num_cells, void* stream);

/* The operator of a solve: device buffers in FP64, layouts as above.  The
 * caller keeps them alive and unchanged while a solver refers to them. */
typedef struct femtc_problem
{
  int op; /* femtc_op */
  int degree;
  int nq;
  int64_t num_cells;
  int64_t ndofs;
  const int32_t* dofmap;
  const double* tables; /* [phi | dphi], 2*N*nq */
  const double* G;      /* [num_cells][nq^3][6], or NULL for the mass op */
  const double* detJw;  /* [num_cells][nq^3], or NULL for the stiffness op */
  /* Optional per-cell Jacobians [num_cells][9] (FP64, row-major) of affine
   * cells.  When given, the inner operator computes its geometry from 7
   * values per cell (femtc_global_apply_affine) instead of streaming G and
   * detJw; G/detJw may then be NULL, and the FP64 residual uses the
   * per-cell geometry too. */
  const double* J;
  /* Optional structured cell numbering: cell (i, j, k) of an
   * grid[0] x grid[1] x grid[2] grid is (i*grid[1] + j)*grid[2] + k; all
   * zero when unknown.  With a grid the solver checks at creation whether
   * the dofmap is the structured numbering (femtc_dofmap_structured); then
   * the CG iteration can run on the assembled scatter (see the plans
   * above): the operator's cell pass stores its cell values without
   * atomics into a workspace of C*(P+1)^3 values and the CG update kernel
   * sums them per dof itself.  `scatter` picks the inner operator's
   * scatter: */
  int grid[3];
  int scatter; /* femtc_scatter */
  /* FEMTC_EVEN_ODD_AUTO (0): when nq = degree + 1 and the tables pass the
   * symmetry check of femtc_plan_set_tables (they are read once, at
   * creation), the even-odd form of the operators is timed over one CG
   * iteration step against the general one, with each scatter the solver
   * may use, and the fastest pair is kept.  FEMTC_EVEN_ODD_OFF (1): the
   * general form only.  FEMTC_EVEN_ODD_ON (2): the even-odd form whenever
   * the tables pass the check (not timed). */
  int even_odd;
} femtc_problem;

enum
{
  FEMTC_EVEN_ODD_AUTO = 0,
  FEMTC_EVEN_ODD_OFF = 1,
  FEMTC_EVEN_ODD_ON = 2
};

typedef enum femtc_scatter
{
  FEMTC_SCATTER_AUTO = 0,     /* time one iteration step with each (when the
                                 dofmap allows it), keep the faster */
  FEMTC_SCATTER_DIRECT = 1,   /* atomic scatter */
  FEMTC_SCATTER_ASSEMBLED = 2, /* assembled scatter; creation fails with
                                  FEMTC_ERR_INVALID_ARGUMENT if the dofmap
                                  is not the structured one */
  FEMTC_SCATTER_ZRUN = 3      /* structured z-run scatter (cells summed
                                 along z in LDS, atomics on runs of
                                 consecutive dofs; nq <= 8); fails as the
                                 assembled one */
} femtc_scatter;

typedef struct femtc_solve_info
{
  int status;           /* femtc_status of the solve */
  int outer_iterations; /* defect-correction steps (1 for plain PCG) */
  int inner_iterations; /* CG iterations, summed over the steps */
  double relres;        /* ||f - A u||_2 / ||f||_2, evaluated in FP64 */
  double seconds;       /* host wall time of the call */
} femtc_solve_info;

typedef struct femtc_solver femtc_solver; /* opaque */

/* Allocate the workspace on the current device and build the Jacobi
 * preconditioner from the FP64 operator.  inner_dtype FEMTC_F64: plain PCG
 * in FP64.  FEMTC_F32: FP32 inner PCG (FP32 vectors, tables and geometry
 * copy; FP64 dot products) inside FP64 defect correction.  FEMTC_BF16 and
 * FEMTC_BF16_ROUNDED: the same with the inner operator on the matrix cores
 * (bfloat16 operands, FP32 accumulation); they need the cell Jacobians J,
 * the structured grid (and a dofmap that is its structured numbering),
 * nq = degree + 1 and degree <= 7.  The solver and every later solve run on
 * `stream`.  Fails with FEMTC_ERR_INVALID_ARGUMENT when a diagonal entry is
 * not positive (a dof no cell touches). */
int femtc_solver_create(femtc_solver** solver, const femtc_problem* problem,
                        int inner_dtype, void* stream);
/* Device bytes the solver holds (its workspace, tables and graphs aside). */
int femtc_solver_bytes(const femtc_solver* solver, size_t* bytes);
/* The estimate of a solver's device bytes that the request server reserves
 * while the solver is created (host only, no device work). */
int femtc_solver_workspace_bytes(const femtc_problem* problem,
                                 int inner_dtype, size_t* bytes);
/* Solve A u = f (f, u FP64 device vectors of ndofs; u overwritten) to
 * ||f - A u|| <= rtol ||f||.  Plain PCG stops on its recursive residual;
 * defect correction stops on the FP64 residual recomputed after each inner
 * solve (inner tolerance inner_rtol).  maxit caps the CG iterations (summed
 * over inner solves).  Synchronous; info may be NULL.  Returns
 * FEMTC_ERR_NOT_CONVERGED or FEMTC_ERR_BREAKDOWN (with info filled in) when
 * the tolerance is not met. */
int femtc_solver_solve(femtc_solver* solver, const double* f, double* u,
                       double rtol, double inner_rtol, int maxit,
                       femtc_solve_info* info);
/* Device pointer to the solver's FP64 operator diagonal. */
int femtc_solver_diagonal(const femtc_solver* solver, const double** diag);
/* The scatter the solver chose for its CG iterations
 * (FEMTC_SCATTER_DIRECT, FEMTC_SCATTER_ASSEMBLED or FEMTC_SCATTER_ZRUN). */
int femtc_solver_scatter(const femtc_solver* solver, int* scatter);
/* 1 when the solver's operators contract in even-odd form, else 0. */
int femtc_solver_even_odd(const femtc_solver* solver, int* even_odd);
/* CG iterations queued between two reads of the convergence state
 * (check_every, default 16; iterations past convergence cost a kernel
 * launch each and do no work), and whether that batch is replayed from a
 * hipGraph captured at the first solve (use_graph, default 1) or queued
 * kernel by kernel. */
int femtc_solver_set_batch(femtc_solver* solver, int check_every,
                           int use_graph);
int femtc_solver_destroy(femtc_solver* solver);

/* ---------------------------------------------------------------------- */
/* Request server (no counterpart in the reference): apply and solve       */
/* requests on resident problems, served by `slots` host threads with one  */
/* HIP stream each, so requests admitted while others run overlap on the   */
/* device.  Solvers stay resident per (problem, inner dtype, slot); their  */
/* workspaces are admitted against `hbm_budget` bytes, idle ones evicted   */
/* least-recently-used first; a request that does not fit yet waits while */
/* later ones that fit go ahead.  Times are seconds since server creation. */
/* ---------------------------------------------------------------------- */
enum
{
  FEMTC_REQUEST_APPLY = 0, /* output = A input (FP64, overwritten) */
  FEMTC_REQUEST_SOLVE = 1  /* output = A^-1 input to rtol */
};

typedef struct femtc_request
{
  int kind;
  /* a problem registered with femtc_server_register (its copy is used) */
  const femtc_problem* problem;
  const void* input;            /* FP64 device vector of ndofs */
  void* output;                 /* FP64 device vector of ndofs */
  double rtol;                  /* solve */
  double inner_rtol;            /* refined solves (inner_dtype not F64) */
  int maxit;                    /* solve */
  int inner_dtype;              /* solve: as femtc_solver_create's */
  int64_t tag;                  /* returned with the result */
} femtc_request;

typedef struct femtc_request_result
{
  int64_t tag;
  int kind;
  int status;
  int slot;
  int iterations;       /* solve: CG iterations */
  int outer_iterations; /* solve: refinement steps */
  double relres;        /* solve: FP64 relative residual */
  double t_arrival, t_start, t_end;
} femtc_request_result;

typedef struct femtc_server femtc_server; /* opaque */
int femtc_server_create(femtc_server** server, int slots, size_t hbm_budget);
/* Make a problem known to the server, which keeps a copy of the struct; its
 * device buffers must stay alive and unchanged until it is unregistered.
 * Requests name it by the pointer registered.  Registering a pointer that
 * is registered already fails. */
int femtc_server_register(femtc_server* server, const femtc_problem* problem);
/* Forget a problem: its queued requests complete with
 * FEMTC_ERR_INVALID_ARGUMENT, its idle resident solvers are destroyed now
 * and busy ones when their request ends, so the address may then be reused
 * for another problem without meeting a stale solver. */
int femtc_server_unregister(femtc_server* server,
                            const femtc_problem* problem);
/* Queue a request (non-blocking); FEMTC_ERR_INVALID_ARGUMENT for an
 * unregistered problem, FEMTC_ERR_OUT_OF_MEMORY if its solve workspace
 * alone exceeds the budget. */
int femtc_server_submit(femtc_server* server, const femtc_request* request);
/* Block until every submitted request has completed. */
int femtc_server_drain(femtc_server* server);
/* Move up to max completed results to out (completion order); *n = count. */
int femtc_server_results(femtc_server* server, femtc_request_result* out,
                         int max, int* n);
int femtc_server_memory(const femtc_server* server, size_t* used,
                        int* cached_solvers);
/* Finish the queued requests, then stop the slots and free the solvers. */
int femtc_server_destroy(femtc_server* server);

/* ---------------------------------------------------------------------- */
/* The server's admission, queueing and eviction code driven by a fake     */
/* clock and a fake allocator (host only, no device work; no counterpart   */
/* in the reference), for tests of the serving logic.                      */
/* ---------------------------------------------------------------------- */
typedef struct femtc_sim_request
{
  double t_arrival; /* seconds */
  int kind;         /* FEMTC_REQUEST_APPLY or FEMTC_REQUEST_SOLVE */
  int problem;      /* 0 .. nproblems-1 */
  int inner;        /* inner precision (part of the workspace key) */
  size_t bytes;     /* a solve's workspace estimate (charged at admission) */
  size_t actual;    /* its size once created (charged from then on) */
  double service;   /* run time with a resident workspace */
  double create;    /* extra run time when the workspace is created */
} femtc_sim_request;

typedef struct femtc_sim_result
{
  int status; /* FEMTC_OK; FEMTC_ERR_OUT_OF_MEMORY: refused at submission;
                 FEMTC_ERR_INVALID_ARGUMENT: its problem was removed first */
  int slot;
  int cached; /* 1 when a resident workspace served it */
  double t_start, t_end;
} femtc_sim_result;

typedef struct femtc_sim_stats
{
  size_t high_water;     /* largest charge against the budget */
  int64_t evictions;     /* idle workspaces evicted for room */
  int64_t created;       /* workspaces created */
  int64_t destroyed;     /* workspaces destroyed (evicted or retired) */
  double makespan;       /* time of the last event */
  int64_t left_in_queue; /* requests never admitted */
} femtc_sim_stats;

/* Run the requests (any order of arrival times) through `slots` slots and
 * a budget of `budget` bytes.  remove_at[p] >= 0 removes problem p at that
 * time (NULL: never).  out[i] is the result of reqs[i]. */
int femtc_admission_simulate(const femtc_sim_request* reqs, int n,
                             int nproblems, const double* remove_at, int slots,
                             size_t budget, femtc_sim_result* out,
                             femtc_sim_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* FEMTC_H */
