/* bosss_hip.h -- C ABI of the MI355X (gfx950) DG spatial-operator library.
 *
 * Drop-in boundary for the BoSSS DG operator hot path.  A C# host binds these
 * symbols by name through a DynLibLoader subclass (MPI.Wrappers/Utils/DynLibLoader.cs:30-450,
 * the same mechanism as BLAS.cs:1065-1090 / LAPACK.cs:28-66).  Conventions follow
 * that precedent:
 *   - every function returns int: 0 = ok, < 0 = error (message via
 *     bosss_hip_last_error(), or the block / XDG / peaks variant of the family
 *     that failed; each holds the last error of the calling host thread),
 *     > 0 = LAPACK-style INFO (1-based pivot index of the first singular block);
 *   - host pointers are only read/written during the call (managed arrays are
 *     pinned with `fixed` for one call only, DGField_Evaluation.cs:78-110);
 *   - a context is bound to one host thread and one GPU; calls on one context
 *     must not overlap (evaluators are not re-entrant, NECQuadratureVolume.cs:191-283).
 *
 * Coefficient vectors use the reference layout of UnsetteledCoordinateMapping
 * (UnsetteledCoordinateMapping.cs:696-706): entry (cell j, variable f, mode n)
 * is at j*V*N + f*N + n (all variables share one basis of N modes).
 */
#ifndef BOSSS_HIP_H
#define BOSSS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BOSSS_OK 0
#define BOSSS_ERR_ARG (-1)         /* invalid argument / size mismatch          */
#define BOSSS_ERR_HIP (-2)         /* HIP runtime error                          */
#define BOSSS_ERR_UNSUPPORTED (-3) /* form/degree/grid not covered by a kernel  */
#define BOSSS_ERR_NONFINITE (-4)   /* NaN/Inf detected (CheckForNanOrInf)        */

/* Equation components with hand-written kernels.  Anything else stays on the
 * reference CPU evaluator (the C# side keeps base.GetEvaluatorEx as fallback). */
enum bosss_form_kind {
  BOSSS_FORM_SIP = 0,        /* SIPLaplace, constant Nu, Dirichlet g=0      (ipLaplace.cs:30-269)          */
  BOSSS_FORM_ADVDIFF = 1,    /* upwind ScalarTransportFlux3D + SIPLaplace   (TransportFlux.cs:112-190)     */
  BOSSS_FORM_EULER_LF = 2,   /* Euler components + RusanovFlux              (RusanovFlux.cs:55-65)         */
  BOSSS_FORM_EULER_HLLC = 3  /* Euler components + HLLC{Density,Momentum,Energy}Flux (HLLCFlux.cs:48-86)   */
};

enum bosss_precision { BOSSS_FP64 = 0, BOSSS_FP32 = 1 };

typedef struct bosss_ctx bosss_ctx;
typedef struct bosss_grid bosss_grid;
typedef struct bosss_op bosss_op;

/* Affine cell and edge metrics as held by GridData (GridData_CellData.cs:293-387, 656-698;
 * GridData_EdgeData.cs:213-222, 2321, 2524-2580). */
typedef struct {
  int32_t D, J, E;
  const double* detJ;      /* [J]       JacobiDet                                         */
  const double* invJ;      /* [J][D][D] InverseTransformation: invJ[j][e][d] = dxi_e/dx_d  */
  const double* scaling;   /* [J]       ChefBasis.Scaling = 1/sqrt(detJ)                  */
  const double* lenscale;  /* [J]       Cells.CellLengthScale                             */
  const int32_t* cells;    /* [E][2]    CellIndices, OUT = -1 on boundary edges           */
  const int32_t* trafo;    /* [E][2]    Edge2CellTrafoIndex (face of IN / OUT cell)       */
  const double* normals;   /* [E][D]    NormalsForAffine (outward from IN)                */
  const double* sqrtgram;  /* [E]       SqrtGramian                                       */
} bosss_grid_desc;

/* Reference-element tables (RefElement quadrature + Basis evaluation, BasisData.cs:44-69). */
typedef struct {
  int32_t D, degree, N, Kv, Ke, ntrafo;
  const double* vol_weights;  /* [Kv]                                    */
  const double* phi;          /* [Kv][N]     phi_n(xi_k)                 */
  const double* dphi;         /* [Kv][N][D]  d phi_n / d xi_d            */
  const double* face_weights; /* [Ke]                                    */
  const double* trace;        /* [ntrafo][Ke][N]     EdgeEval            */
  const double* dtrace;       /* [ntrafo][Ke][N][D]  EdgeGradientEval    */
} bosss_tables_desc;

/* Form coefficients (SipControl.penalty_poisson, Nu, transport velocity, IdealGas). */
typedef struct {
  double penalty_base; /* SIPLaplace m_penalty_base                    */
  double nu;           /* SIPLaplace.Nu (signed)                       */
  double velocity[3];  /* constant transport velocity                  */
  double gamma;        /* HeatCapacityRatio                            */
  double mach;         /* MachNumber                                   */
} bosss_form_params;

const char* bosss_hip_last_error(void);

/* ---- device buffers (the "device-resident buffer handles" of a host without a GPU array
 *      library: a C# host keeps the returned pointer as an IntPtr and passes it, or an offset
 *      into it, to the *_dev entry points; BlockMsrMatrix- and DGField-sized arrays stay in HBM
 *      across calls, e.g. for the Newton loop, Newton.cs:189-365).  Copies run on the context's
 *      stream and return with the copy complete; create returns zero-filled memory. ---- */
int bosss_hip_buffer_create(bosss_ctx* ctx, int64_t bytes, void** dptr);
int bosss_hip_buffer_destroy(bosss_ctx* ctx, void* dptr);
int bosss_hip_buffer_upload(bosss_ctx* ctx, void* dptr, const void* host, int64_t bytes);
int bosss_hip_buffer_download(bosss_ctx* ctx, void* host, const void* dptr, int64_t bytes);

/* ---- context / grid / operator (replaces DifferentialOperator.GetEvaluatorEx setup,
 *      DifferentialOperator.cs:1013-1033, and the grid metrics it reads) ---- */
int bosss_hip_ctx_create(int device, bosss_ctx** ctx);
int bosss_hip_ctx_destroy(bosss_ctx* ctx);
int bosss_hip_sync(bosss_ctx* ctx);
/* The HIP stream the context launches on (hipStream_t). */
void* bosss_hip_ctx_stream(bosss_ctx* ctx);

int bosss_hip_grid_upload(bosss_ctx* ctx, const bosss_grid_desc* desc, bosss_grid** grid);
int bosss_hip_grid_destroy(bosss_grid* grid);

int bosss_hip_op_create(bosss_ctx* ctx, const bosss_grid* grid, const bosss_tables_desc* tables, int kind,
                        const bosss_form_params* params, int precision, bosss_op** op);
int bosss_hip_op_destroy(bosss_op* op);
/* ndof = J*V*N; nvars = V; nmodes = N. */
int bosss_hip_op_info(const bosss_op* op, int64_t* ndof, int32_t* nvars, int32_t* nmodes);
/* *tensor_tables = 1 if the operator runs the tensor-table kernels (1D factors, face tables;
 * built and checked against the uploaded tables at creation), 0 if it runs the generic
 * full-table kernels (BOSSS_HIP_GENERIC_TABLES=1, or tables that are not a tensor product).
 * No counterpart in the reference: lets a host or a test confirm which kernel family runs. */
int bosss_hip_op_kernel_family(const bosss_op* op, int32_t* tensor_tables);

/* ---- residual (replaces EvaluatorNonLin.Evaluate, DifferentialOperator.cs:1584-1661;
 *      IEvaluator.cs:81-97):  R <- beta*R + alpha*Op(U).  Host pointers, synchronous. ---- */
int bosss_hip_evaluate(bosss_op* op, double time, double alpha, double beta, const double* U, double* R);

/* ---- linearisation (replaces Newton.dirder FD-JVP, Newton.cs:1450-1500, and the
 *      transpose of FDJacobianBuilder, DifferentialOperator.cs:1905-2880) ---- */
int bosss_hip_jvp(bosss_op* op, const double* U, const double* dU, double* JdU);
int bosss_hip_vjp(bosss_op* op, const double* U, const double* W, double* JtW);

/* CheckForNanOrInf of the outputs (DifferentialOperator.cs:1602,1623,1658, there in DEBUG
 * builds), off by default.  When enabled, evaluate / jvp / vjp and their _dev forms scan every
 * output vector after the launch, wait for the stream (so the _dev forms become synchronous and
 * cannot be captured into a graph) and return BOSSS_ERR_NONFINITE, with the first NaN or Inf
 * entry (cell, variable, mode) in bosss_hip_last_error(), if there is one. */
int bosss_hip_op_set_check_finite(bosss_op* op, int32_t enable);

/* ---- device-resident variants: pointers are device memory, work is queued on
 *      `stream` (a hipStream_t; NULL is HIP's null stream, bosss_hip_ctx_stream()
 *      gives the context's own stream) and returns without synchronising.
 *      Each vector holds ndof entries of the operator's precision: double for
 *      BOSSS_FP64, float for BOSSS_FP32 (passed through the same double* parameter,
 *      which is then read as float*).  The host-pointer entry points above always
 *      take double and round to the operator's precision on the device. ---- */
int bosss_hip_evaluate_dev(bosss_op* op, double alpha, double beta, const double* U, double* R, void* stream);
int bosss_hip_jvp_dev(bosss_op* op, const double* U, const double* dU, double* JdU, void* stream);
int bosss_hip_vjp_dev(bosss_op* op, const double* U, const double* W, double* JtW, void* stream);
/* fused: R <- beta*R + alpha*Op(U) and JtW <- J(U)^T W in one pass over U */
int bosss_hip_evaluate_vjp_dev(bosss_op* op, double alpha, double beta, const double* U, double* R, const double* W,
                               double* JtW, void* stream);

/* ---- operator matrix (replaces EvaluatorLinear.ComputeMatrix, DifferentialOperator.cs:1721-1838,
 *      into a BlockMsrMatrix, BlockMsrMatrix.cs:186/1388, and the coloured
 *      FDJacobianBuilder, DifferentialOperator.cs:1905-2880, with exact JVPs).
 *      Layout "block-ELL": blocks[j][s] is the Nt x Nt (Nt = V*N) row-major block of
 *      cell row j and column cell cols[j][s]; s = 0 is the diagonal, s = 1 + f the
 *      neighbour across local face f (-1: no block -- a boundary face, or a neighbour an
 *      earlier slot of the row already holds, across a periodic direction of two cells;
 *      assembly sets such slots to beta * old, i.e. zero for beta = 0).
 *      entries = J*(1+2D)*Nt*Nt. ---- */
int bosss_hip_op_matrix_layout(bosss_op* op, int32_t* blocks_per_row, int64_t* entries, int32_t* ncolors, int32_t* cols);
/* blocks <- beta*blocks + alpha*dR/dU(U)  (device pointers; the Jacobian of a linear form is its matrix) */
int bosss_hip_op_assemble_dev(bosss_op* op, const double* U, double alpha, double beta, double* blocks, void* stream);
/* Diagonal cell blocks of the Jacobian, for cell-block preconditioners (the blocks BlockJacobi.InitImpl
 * reads from the assembled matrix, BlockJacobi.cs:63-120, without assembling the off-diagonal ones):
 * blocks[j] <- alpha * dR_j/dU_j (U) + shift * I, row-major Nt x Nt (Nt = V*N), J blocks of FP64;
 * with transposed != 0 each block is stored transposed (column-major), which is how the probes
 * produce it (one column per probe, written contiguously): inverting those blocks in place gives
 * the transposed inverses that bosss_hip_block_gemv_uniform_dev applies.
 * Nt probing JVPs with the neighbour tangents taken as zero, in the operator's precision (U is
 * double for BOSSS_FP64, float for BOSSS_FP32; an FP32 operator gives FP32-accurate blocks, widened).
 * Works for every form, linear or not. */
int bosss_hip_op_diag_blocks_dev(bosss_op* op, const void* U, double alpha, double shift, int32_t transposed, double* blocks,
                                 void* stream);
/* y <- alpha*A x + beta*y with A in the operator's block-ELL layout (BlockMsrMatrix.SpMV, BlockMsrMatrix.cs:2554) */
int bosss_hip_bell_spmv_dev(bosss_op* op, const double* blocks, double alpha, const double* x, double beta, double* y, void* stream);
/* Host-pointer forms of the two calls above (blocks, U, x, y host arrays of entries / ndof
 * doubles): BlockMsrMatrix-side callers that hold the matrix in host memory (AccBlock,
 * BlockMsrMatrix.cs:1388-1420; ReadBlock :1497).  Copies per call, synchronous. */
int bosss_hip_op_assemble(bosss_op* op, const double* U, double alpha, double beta, double* blocks);
int bosss_hip_bell_spmv(bosss_op* op, const double* blocks, double alpha, const double* x, double beta, double* y);
/* Tile shape of that apply: the panel size NL (64-element loads per lane and panel) of the
 * streaming kernel.  The first bosss_hip_bell_spmv_dev call on an operator times every candidate
 * on its own matrix and keeps the fastest (an autotuner; not while the stream is being
 * captured).  nl = 0 asks for a new tuning at the next apply; another value fixes the shape and
 * must be one of the candidates.  Results do not depend on the shape (same summation order). */
int bosss_hip_op_set_spmv_tile(bosss_op* op, int32_t nl);
/* The chosen NL (0 until tuned), the number of candidates, and up to max candidates with their
 * timed ms per apply (< 0: not timed). */
int bosss_hip_op_spmv_tiles(const bosss_op* op, int32_t* chosen, int32_t* ncand, int32_t* cand, float* ms, int32_t max);

/* ---- batched uneven cell-block factorisation (replaces BlockJacobi.InitImpl's
 *      ReadBlock -> InvertInPlace -> LAPACK DGETRF/DGETRI per block, BlockJacobi.cs:63-120,
 *      IMatrixExtensions.cs:2128-2230, and Schwarz's LAPACK block solver for blocks up to
 *      4096 DOF, Schwarz.cs:827-915; above that the reference switches to PARDISO).
 *      Blocks are dense row-major at element offsets off[b], sizes n[b] <= nmax <= 4096:
 *      nmax <= 128 holds each block in registers (up to 20 DOF several blocks per
 *      wavefront, above that one 256-thread workgroup per block); larger nmax
 *      factors in place in global memory, blocked: getf2 on 32-column panels (the same
 *      pivots), the trailing update as a rank-32 product on the matrix units (one
 *      256-thread workgroup per block, or, for fewer than 256 blocks with nmax >= 512,
 *      two launches per panel with the trailing update over many workgroups; that form
 *      synchronises the stream before it returns).
 *      ipiv (1-based, LAPACK convention) at pivoff[b]; info[b] as DGETRF's INFO
 *      (0, or j+1 for the first zero pivot; -1 if n[b] exceeds nmax).  The inverse
 *      (want_inverse) needs nmax <= 128.  In
 *      bosss_hip_block_lu_dev, ipiv (with pivoff) and info may be NULL.  ctx selects the
 *      device (NULL: the calling thread's current device); stream must belong to it.
 *      want_inverse != 0 overwrites each block with its inverse (DGETRI's result, above
 *      20 DOF by in-place Gauss-Jordan with the same pivot rows), n <= 128; a singular
 *      block keeps its LU factors instead, with INFO set.
 *      All arrays are device memory; one launch covers the whole batch. ---- */
int bosss_hip_block_lu_dev(bosss_ctx* ctx, int nblk, int32_t nmax, const int32_t* n, const int64_t* off, double* A,
                           int32_t* ipiv, const int64_t* pivoff, int32_t* info, int32_t want_inverse, void* stream);
/* The same factorisation without allocating: no counterpart in the reference; for callers that
 * cannot allocate device memory (code under HIP graph capture, tight solver loops, hosts that
 * forbid hipMalloc).  Blocks above 128 DOF always use the split form (two launches per panel,
 * same pivots and factors as above) with its per-block status words in a fixed device
 * array of the library (8 KiB per device, held until the process ends), so at most 1024 such
 * blocks per call (BOSSS_ERR_ARG beyond); up to 128 DOF it is the register kernels above,
 * which need no scratch.  The call only queues work on `stream` and never synchronises.  Every call on one
 * device shares that array: calls that overlap on one device race. */
int bosss_hip_block_lu_noalloc_dev(bosss_ctx* ctx, int nblk, int32_t nmax, const int32_t* n, const int64_t* off, double* A,
                                   int32_t* ipiv, const int64_t* pivoff, int32_t* info, int32_t want_inverse, void* stream);
/* x_b <- (LU_b)^-1 x_b  (DGETRS), x_b at element offset xoff[b] */
int bosss_hip_block_solve_dev(bosss_ctx* ctx, int nblk, const int32_t* n, const int64_t* off, const double* LU,
                              const int32_t* ipiv, const int64_t* pivoff, const int64_t* xoff, double* x, void* stream);
/* y_b <- alpha Ainv_b x_b + beta y_b  (block-diagonal apply of BlockJacobi.Solve, BlockJacobi.cs:148-197) */
int bosss_hip_block_gemv_dev(bosss_ctx* ctx, int nblk, const int32_t* n, const int64_t* off, const double* Ainv,
                             const int64_t* xoff, double alpha, const double* x, double beta, double* y, void* stream);
/* Host-pointer forms of the three batch calls (BlockJacobi.InitImpl holds the blocks it
 * ReadBlock's in host arrays, BlockJacobi.cs:63-120; InvertInPlace works in place on them,
 * IMatrixExtensions.cs:2128-2230): same arguments without the stream; every array is host memory,
 * copied to the device and back on ctx's stream (ctx is required); synchronous.  The array
 * extents are max over b of off[b] + n[b]^2 (A, LU, Ainv), pivoff[b] + n[b] (ipiv) and
 * xoff[b] + n[b] (x, y); entries between blocks are copied back unchanged. */
int bosss_hip_block_lu(bosss_ctx* ctx, int nblk, int32_t nmax, const int32_t* n, const int64_t* off, double* A, int32_t* ipiv,
                       const int64_t* pivoff, int32_t* info, int32_t want_inverse);
int bosss_hip_block_solve(bosss_ctx* ctx, int nblk, const int32_t* n, const int64_t* off, const double* LU, const int32_t* ipiv,
                          const int64_t* pivoff, const int64_t* xoff, double* x);
int bosss_hip_block_gemv(bosss_ctx* ctx, int nblk, const int32_t* n, const int64_t* off, const double* Ainv, const int64_t* xoff,
                         double alpha, const double* x, double beta, double* y);
const char* bosss_hip_block_last_error(void);

/* ---- two-species XDG Poisson (XdgPoisson3 with XLaplace_Bulk + XLaplace_Interface + RHSSource,
 *      XdgPoisson3Main.cs:148-178, XLaplace.cs:50-373; agglomeration threshold 0) on
 *      precomputed cut-cell quadrature items (the Saye rules of the level-set tracker,
 *      computed on the host like the reference does).  Replaces the XDG matrix assembly of
 *      XEvaluatorLinear (XSpatialOperatorMk2_XEvaluators.cs:48-600), the block-Jacobi
 *      smoother's block inversion (BlockJacobi.cs:63-197) and the Krylov solve around it.
 *
 *      Species: 0 = A (level set < 0), 1 = B.  Cell j holds nspec[j] (1 or 2) species blocks
 *      of N coefficients, species spec0[j] first (XDGBasis.cs:92-113).  Matrix: block-ELL
 *      with uneven blocks, S = 1 + 2D slots per cell row (0 diagonal, 1 + f the neighbour
 *      across local face f, cols[j][s] = -1: none); block (j, s) is row-major
 *      (nspec[j] N) x (nspec[cols[j][s]] N) at element offset boff[j][s].
 *      Quadrature nodes are in the reference coordinates of the cell(s) they belong to,
 *      weights are physical (volume / area), normals unit: face normals IN -> OUT, interface
 *      normals from A to B.  face_slot[i] = {slot of OUT in IN's row, slot of IN in OUT's row}. ---- */
typedef struct bosss_xdg bosss_xdg;
typedef struct {
  int32_t D, degree, N, J;
  const int32_t* mode_exponents;   /* [N][D] exponents of the Legendre factors of each mode (Basis order) */
  const int32_t* nspec;            /* [J] */
  const int32_t* spec0;            /* [J] */
  const double* invJ;              /* [J][D][D] */
  const double* scaling;           /* [J] */
  const double* lscale;            /* [J][2] species cell length scales (MultiphaseAgglomerator.cs:746-960) */
  double penalty_base, mu[2], source[2];
  int32_t S;
  const int32_t* cols;             /* [J][S] */
  const int64_t* boff;             /* [J][S] offset of the row-major n_j x n_col block; the
                                      blocks of a cell row follow each other in slot order */
  int32_t nvol;                    /* volume items: (cell, species) */
  const int32_t *vol_cell, *vol_spec;
  const int64_t* vol_off;          /* [nvol+1] node offsets */
  const double *vol_nodes, *vol_w; /* [K][D], [K] */
  int32_t nface;                   /* face items: (edge, species) */
  const int32_t *face_cells, *face_slot, *face_spec;   /* [nface][2], [nface][2], [nface] */
  const int64_t* face_off;
  const double *face_nodes_in, *face_nodes_out, *face_w, *face_normal, *face_gD;  /* [K][D] x2, [K], [nface][D], [K] (NULL: g_D = 0) */
  int32_t nifc;                    /* interface items: cut cells */
  const int32_t* ifc_cell;
  const int64_t* ifc_off;
  const double *ifc_nodes, *ifc_w, *ifc_normal;         /* [K][D], [K], [K][D] */
} bosss_xdg_desc;

int bosss_hip_xdg_create(bosss_ctx* ctx, const bosss_xdg_desc* desc, bosss_xdg** xdg);
int bosss_hip_xdg_destroy(bosss_xdg* xdg);
int bosss_hip_xdg_info(bosss_xdg* xdg, int64_t* ndof, int64_t* nentries);
/* blocks <- M, affine <- b with R(u) = M u + b (device pointers; ComputeMatrix + ComputeAffine) */
int bosss_hip_xdg_assemble_dev(bosss_xdg* xdg, double* blocks, double* affine, void* stream);
/* y <- alpha M x + beta y */
int bosss_hip_xdg_spmv_dev(bosss_xdg* xdg, const double* blocks, double alpha, const double* x, double beta, double* y, void* stream);
/* Solve M sol = b by block-Jacobi preconditioned CG (diagonal blocks inverted by the batched
 * uneven block LU), sol holding the initial guess, to ||b - M sol|| <= tol ||b||.
 * precision BOSSS_FP64: CG in FP64 preconditioned by the inverted diagonal blocks;
 * BOSSS_FP32: unpreconditioned CG in FP32 on the diagonal-block equilibrated matrix
 * L^-1 M L^-T (L the Cholesky factor of each diagonal block, so the equilibrated diagonal
 * blocks are the identity and block-Jacobi would be a no-op; the diagonal blocks must be
 * symmetric positive definite, else BOSSS_ERR_ARG), rounded to FP32 once, inside FP64
 * iterative refinement (at most 40 outer steps).  All scalars stay on the device; each
 * group of check_every iterations is one HIP graph launch when use_graph != 0.  The call
 * returns with its work complete (sol final, relres = ||b - M sol|| / ||b|| of that sol).
 * Returns > 0 (the 1-based cell) if a diagonal block is singular. */
int bosss_hip_xdg_solve_dev(bosss_xdg* xdg, const double* blocks, const double* b, double* sol, double tol, int32_t maxit,
                            int32_t precision, int32_t use_graph, int32_t check_every, int32_t* iterations, int32_t* outer,
                            double* relres, void* stream);
/* Tile shape of the solver's cell-row kernels (apply and update, one wavefront per cell row): T
 * groups of four rows per workgroup, T = 1, 2 or 4.  The first solve in each precision times the
 * three shapes on its own matrix (an autotuner; not while the stream is being captured) and keeps the
 * fastest; t = 0 asks for a new tuning, 1, 2 or 4 fixes the shape.  Every shape forms the same
 * dot-product partial sums per group of four rows in the same order, so the CG iterates do not
 * depend on it.  bosss_hip_xdg_tiles reports the chosen T (0: not tuned yet) and the timed ms per
 * apply + update of T = 1, 2, 4 (< 0: not timed). */
int bosss_hip_xdg_set_tile(bosss_xdg* xdg, int32_t precision, int32_t t);
int bosss_hip_xdg_tiles(const bosss_xdg* xdg, int32_t precision, int32_t* chosen, float* ms3);
/* Host-pointer forms (blocks: nentries doubles, vectors: ndof doubles, host memory; copies on the
 * operator's stream, synchronous).  bosss_hip_xdg_solve runs the solver above with use_graph = 1
 * and check_every = 16. */
int bosss_hip_xdg_assemble(bosss_xdg* xdg, double* blocks, double* affine);
int bosss_hip_xdg_spmv(bosss_xdg* xdg, const double* blocks, double alpha, const double* x, double beta, double* y);
int bosss_hip_xdg_solve(bosss_xdg* xdg, const double* blocks, const double* b, double* sol, double tol, int32_t maxit, int32_t precision,
                        int32_t* iterations, int32_t* outer, double* relres);
const char* bosss_hip_xdg_last_error(void);

/* ---- device-resident GMRES (Newton.MatrixFreeGMRES, Newton.cs:1128-1500, with the Arnoldi
 *      process, the Hessenberg matrix and the Givens rotations on the GPU; the reference keeps
 *      them on the host).  Restart length m <= 256; V holds m+1 basis vectors of n entries at a
 *      row stride ldv >= n, double (BOSSS_FP64) or float (BOSSS_FP32); H is the (m+1) x m
 *      column-major Hessenberg matrix, cs, sn the m rotations, g the m+1 rotated right-hand side
 *      (g[0] = |r0| before step 0, the rest zero), all double; work holds the doubles that
 *      bosss_hip_gmres_workspace reports.  Both calls only queue kernels on `stream` (no
 *      allocation, no memset, no synchronisation), so a run of them can be captured into a graph.
 *
 *      arnoldi step k (0 <= k < m): on entry V[0..k] and w = A M^-1 V[k] (right preconditioning;
 *      w = A V[k] without); classical Gram-Schmidt twice against V[0..k], column k of H, V[k+1] =
 *      w / |w| (zero after a breakdown), the previous rotations applied to column k, rotation k,
 *      g updated; res[k] (m doubles) receives |g[k+1]|, the residual estimate of the k+1-step
 *      solution (g[k+1] itself changes in later steps).  w is overwritten.
 *      update: y = R^-1 g[0..k1) (R the leading k1 x k1 triangle of H, y: k1 doubles) and
 *      x += sum_{i<k1} y_i Z[i] (Z = V, or the preconditioned vectors M^-1 V[i], stride ldz). ---- */
int bosss_hip_gmres_workspace(int32_t m, int64_t* doubles);
int bosss_hip_gmres_arnoldi_dev(bosss_ctx* ctx, int64_t n, int32_t k, int32_t m, int32_t precision, void* V, int64_t ldv, void* w,
                                double* H, double* cs, double* sn, double* g, double* res, double* work, void* stream);
int bosss_hip_gmres_update_dev(bosss_ctx* ctx, int64_t n, int32_t k1, int32_t m, int32_t precision, const void* Z, int64_t ldz,
                               const double* H, const double* g, double* y, void* x, void* stream);
/* Cell-block preconditioner apply: y_b <- alpha A_b x_b + beta y_b for nblk blocks of uniform size
 * n <= 128 (x_b, y_b at b*n), A_b stored TRANSPOSED (AinvT[b][k][i] = A_b[i][k]: each block column
 * is contiguous, read coalesced), double or float like x and y.  For the inverted diagonal blocks
 * of a Newton step (BlockJacobi.Solve, BlockJacobi.cs:148-197), in FP32 for the mixed solver. */
int bosss_hip_block_gemv_uniform_dev(bosss_ctx* ctx, int nblk, int32_t n, int32_t precision, const void* AinvT, double alpha,
                                     const void* x, double beta, void* y, void* stream);
const char* bosss_hip_krylov_last_error(void);

/* ---- measured arithmetic ceilings (no counterpart in the reference: the rates the roofline
 *      fractions are quoted against, measured on the context's device) ----
 * Dense FP64 MFMA (v_mfma_f64_16x16x4f64), FP32 MFMA (v_mfma_f32_16x16x4f32) and FP64 FMA
 * issue rates in TFLOP/s, best of three launches of independent accumulation chains over
 * the whole device (a few milliseconds each). */
int bosss_hip_measure_peaks(bosss_ctx* ctx, double* tflops_fp64_mfma, double* tflops_fp32_mfma, double* tflops_fp64_fma);
/* FP64 rate (TFLOP/s, MFMA and FMA flops together) of a stream that follows every
 * v_mfma_f64_16x16x4f64 with 16 FP64 FMAs (equal flops): near the sum of the two single rates
 * if the matrix core ran beside the FP64 VALU, near either single rate if they share one FP64
 * rate (the ceiling of a kernel that issues both, as the fused Euler residual + VJP does). */
int bosss_hip_measure_fp64_mix(bosss_ctx* ctx, double* tflops_fp64_mix);
const char* bosss_hip_peaks_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* BOSSS_HIP_H */
