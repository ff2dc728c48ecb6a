/* mpet_hip.h — C ABI of libmpet_hip.so, the MI355X (gfx950) backend of the
 * total-pressure MPET time step.
 *
 * The reference (meg-simula/waterscapes, package `mpet`) has no FFI: its
 * boundary is the Python class API of MPETTotalPressureSolver, whose numerics
 * run inside DOLFIN / PETSc / hypre.  Each entry point below replaces one of
 * those calls; the reference call site is cited on each declaration.  The
 * Python package waterscapes_amd binds these symbols with ctypes.
 *
 * Conventions
 *  - All array pointers are DEVICE pointers unless marked (host).  Arrays
 *    passed in stay owned by the caller; handles own what they allocate and
 *    free it in their *_destroy call.
 *  - Return value: 0 on success, otherwise an MpetStatus code; the message is
 *    available from mpet_last_error (thread-local).
 *  - Work is enqueued on the given stream; nothing synchronises unless the
 *    entry point says so.  One host thread per handle.
 *  - Canonical DOF numbering: u DOF = 3*a + j (P2 node a, component j);
 *    p_f DOF = 3*nP2 + (A+1)*c + f (P1 node c, field f = 0..A, f = 0 is the
 *    total pressure).
 */
#ifndef MPET_HIP_H
#define MPET_HIP_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPET_MAXF 8 /* at most 7 networks: fields p_0..p_7 */

typedef enum {
  MPET_OK = 0,
  MPET_ERR_INVALID = 1,
  MPET_ERR_HIP = 2,
  MPET_ERR_NOT_CONVERGED = 3,
  MPET_ERR_NAN = 4,
  MPET_ERR_OOM = 5,
  MPET_ERR_BUDGET = 6
} MpetStatus;

/* Mesh of one instance.  cells are the P1 nodes (vertices, ascending per
 * cell); cell_p2 the 10 P2 nodes per cell in UFC order.  Cells are listed
 * by colour (colour_cells[colour_offsets[k] .. colour_offsets[k+1]]) such
 * that cells of one colour share no vertex.  colour_offsets is a HOST array. */
typedef struct {
  int32_t nV, nC, nP2, nP1;
  const double* coords;          /* nV*3 */
  const int32_t* cells;          /* nC*4 */
  const int32_t* cell_p2;        /* nC*10 */
  const int32_t* colour_cells;   /* nC */
  const int64_t* colour_offsets; /* (host) ncolours+1 */
  int32_t ncolours;
} MpetMeshDesc;

/* Node-pair sparsity patterns (CSR, sorted columns): rows/cols are P2 or P1
 * nodes: 22 = P2xP2, 21 = P2xP1, 12 = P1xP2, 11 = P1xP1.  Operator entry
 * points reject a pattern with no nonzeros (MPET_ERR_INVALID); rows may be
 * empty. */
typedef struct {
  const int32_t *rp22, *ci22;
  const int32_t *rp21, *ci21;
  const int32_t *rp12, *ci12;
  const int32_t *rp11, *ci11;
  int32_t nnz22, nnz21, nnz12, nnz11;
} MpetPatterns;

/* Primitive operator values on the node patterns (FP64):
 *   Kuu  nnz22*9 : 2 mu eps(phi_b e_j):eps(phi_a e_i)  (row-major 3x3 blocks)
 *   Lap  nnz22   : grad phi_a . grad phi_b (unscaled)
 *   Dv   nnz21*3 : int psi_c d_i phi_a   (row P2 node a, col P1 node c)
 *   Dt   nnz12*3 : the same values on the transposed pattern
 *   M1   nnz11   : P1 mass
 *   St   A*nnz11 : int K_i grad psi . grad psi, network-major */
typedef struct {
  double *Kuu, *Lap, *Dv, *Dt, *M1, *St;
} MpetPrimitives;

/* (A+1)x(A+1) pressure coupling (row-major, stride MPET_MAXF) multiplying
 * the P1 mass, plus per-field stiffness weights (field f uses St[f-1]). */
typedef struct {
  double C[MPET_MAXF * MPET_MAXF];
  double stiff[MPET_MAXF];
  double mu;
} MpetCoef;

/* Structured total-pressure operator (the matrix of lhs(F), or its FP32
 * copy), values structure-of-arrays on the node patterns:
 *   Kuu[c*nnz22 + k] (c = 3i+j), Dv[c*nnz21 + k], Dt[c*nnz12 + k],
 *   M1[k], St[i*nnz11 + k].  dtype 0 = float64, 1 = float32.
 * Rows with dmask[dof] != 0 (Dirichlet DOFs) are returned as 0. */
typedef struct {
  int32_t nP2, nP1, A;
  MpetPatterns pat;
  int32_t dtype;
  const void *Kuu, *Dv, *Dt, *M1, *St;
  MpetCoef coef;
  const uint8_t* dmask;
  int32_t lanes_u, lanes_p;   /* lanes per P2 / P1 node row: 4/8/16/32, 0 = 16 */
  int32_t xcd_chunk;          /* >0: runs of this many consecutive row tiles per
                                 XCD (the 8 XCDs take workgroups round-robin),
                                 0: hardware order */
} MpetOpDesc;

/* One level of the displacement-block hierarchy (device arrays, int32
 * CSR).  Values / dinv in the precision of the owning MpetPrecDesc. */
typedef struct {
  int32_t n, nc;                 /* rows, coarse columns (0 on the coarsest) */
  const int32_t *A_rp, *A_ci;
  const void *A_v, *dinv;
  double rho;                    /* spectral radius estimate of D^-1 A */
  const int32_t *P_rp, *P_ci;    /* prolongator n x nc */
  const void* P_v;
  const int32_t *R_rp, *R_ci;    /* restriction nc x n (= P^T) */
  const void* R_v;
  int32_t lpr, lprP, lprR;       /* lanes per row of the three SpMMs */
  /* Optional compact copy of the level operator for FP32 hierarchies (NULL:
   * none): values as IEEE half of A/A_vscale, columns as signed 16-bit
   * offsets from the row index (same row pointers A_rp). */
  const void* A_v16;
  const int16_t* A_dci;
  double A_vscale;
  /* Batches (NULL otherwise): instance of each row of this level, so that
   * the V-cycle skips the rows of instances whose solve has stopped. */
  const int32_t* grp;
} MpetAmgLevelDesc;

#define MPET_MAXLEV 16

/* One AMG hierarchy (SA setup, Chebyshev smoothing, dense coarsest solve). */
typedef struct {
  int32_t nlev;
  MpetAmgLevelDesc lev[MPET_MAXLEV];
  const void* coarse_inv;        /* dense n_coarse^2, row-major */
  int32_t degree;                /* Chebyshev degree of the smoother */
  double eig_ratio, eig_safety;  /* interval [safety*rho/ratio, safety*rho] */
  double scale;                  /* V-cycle approximates (scale * A0)^-1 */
  const uint8_t* mask0;          /* level-0 rows that are Dirichlet (may be NULL) */
  /* Block-diagonal coarsest level (a batch of instances), or all NULL for one
   * dense block: row r of the coarsest level belongs to block coarse_blk[r],
   * which spans rows coarse_off[b] .. coarse_off[b+1]-1 and whose dense
   * inverse starts at coarse_inv + coarse_ioff[b]. */
  const int32_t* coarse_blk;
  const int32_t* coarse_off;
  const int64_t* coarse_ioff;
  int32_t level_degree[MPET_MAXLEV];  /* per-level Chebyshev degree (0: degree) */
} MpetHierDesc;

/* Block-diagonal preconditioner of prec (:279-283) with Dirichlet rows and
 * columns eliminated (apply_symmetric(bc, P), :445-446):
 *   u: hierarchy on the scalar P2 Laplacian, 3 components at once, scale mu;
 *   p: one hierarchy on the block-diagonal pressure matrix diag(P_0..P_A)
 *      in field-major order (F*nP1 unknowns). */
typedef struct {
  MpetHierDesc u, p;
  int32_t nP1, F;
  /* NULL: pressure DOF F*c + f (node-interleaved) sits at f*nP1 + c of the
   * p hierarchy.  Otherwise pressure DOF i sits at pmap[i] (a batch: set
   * nP1 to the number of pressure DOFs and F to 1). */
  const int32_t* pmap;
} MpetPrecDesc;

typedef struct {
  long long nrows;
  const long long* rp;
  const int32_t* ci;
  const double* v;
  int32_t lpr;
} MpetCsrDesc;

typedef struct {
  int32_t mode;         /* 0: FP64 MINRES; 1: FP32 MINRES in FP64 refinement;
                           2: FP64 MINRES with the FP32 preconditioner */
  int32_t use_csr;      /* mode 0: use the scalar CSR A instead of the structured one */
  double rtol;          /* modes 0, 2: ||r||_{M^-1} <= rtol ||b_I||_{M^-1} */
  int32_t maxit;        /* per MINRES solve */
  int32_t check_every;  /* host reads the device state every k iterations */
  int32_t use_graphs;   /* replay k-iteration chunks from a captured hipGraph */
  double outer_tol;     /* mode 1: ||r||_{M^-1} <= outer_tol ||b_I||_{M^-1}, r = Pi(b - A x) in FP64 */
  double inner_rtol;    /* mode 1: FP32 MINRES relative tolerance */
  int32_t max_outer;
} MpetSolverOptions;

/* residual_norm: final ||r||_{M^-1} / ||b_I||_{M^-1} (preconditioned norms;
 * b_I = Pi (b - A g), the lifted right-hand side of the free rows). */
typedef struct {
  int32_t inner_iterations, outer_iterations, converged;
  double residual_norm;
} MpetSolveStats;

/* Error message of the last failing call on this thread. */
int mpet_last_error(char* buf, size_t n);

/* Upload the reference-element tables (host, MPET_TAB_SIZE doubles) to the
 * device constant bank of the current device. */
int mpet_upload_tables(const double* tab_host, int n);

/* Element tensors of lhs(F) (mpet/mpettotalpressuresolver.py:267-274) and of
 * prec (:279-283), scattered into the primitive arrays in colour order:
 * replaces `A = assemble(a)` (:437) and `P = assemble(prec)` (:444).
 * Kcell: nC*A per-cell K_i (DG0, cell-major).  Deterministic. */
int mpet_assemble_primitives(const MpetMeshDesc* mesh, const MpetPatterns* pat,
                             const double* Kcell, int A, double mu,
                             const MpetPrimitives* out, hipStream_t stream);

/* mpet_assemble_primitives of ninst instances at once (host arrays of
 * ninst descriptors, per-instance device data): one launch per colour
 * index assembles that colour of every instance, with the lanes and order
 * of the single-instance assembly (each instance's values are bitwise those
 * of its own assembly).  Synchronises the stream.  The reference assembles
 * one problem at a time (:437, :444); no counterpart. */
int mpet_assemble_primitives_batched(int ninst, const MpetMeshDesc* meshes,
                                     const MpetPatterns* pats, const double* const* Kcell,
                                     const int32_t* A, const double* mu,
                                     const MpetPrimitives* outs, hipStream_t stream);

/* A-posteriori error indicators of the two-field MPET step
 * (AdaptiveMPETSolver, mpet/adaptivempetsolver.py:82-185).  Mesh arrays as
 * in MpetMeshDesc (device); x, xp the current and previous two-field states
 * (u then p_1..p_A, F = A pressure fields per P1 node); constant material
 * parameters (S row-major, stride MPET_MAXF).  Cell quadrature: nq
 * barycentric points qbary (nq x 4) with weights qw normalised to sum 1
 * (exact for the integrand degree); sources as values at the P_k nodes of
 * every cell (Expression(degree=k) semantics): f, fp (f at the previous
 * time, the reference's f_) nC x nkf x 3, g nC x nkg x A (null = 0), with
 * the P_k basis at the quadrature points Lf (nq x nkf), Lg (nq x nkg). */
typedef struct {
  int32_t nC, nP2, nP1, A;
  const double* coords;
  const int32_t *cells, *cell_p2;
  double mu, lmbda, dt;
  double alpha[MPET_MAXF], c[MPET_MAXF], K[MPET_MAXF], S[MPET_MAXF * MPET_MAXF];
  int32_t nq;
  const double *qbary, *qw;
  int32_t nkf, nkg;
  const double *Lf, *Lg;
  const double *f, *fp, *g;
} MpetIndDesc;

/* Per-cell indicators out (4 x nC, device): rows zeta_p, zeta_u, zeta_u_dt
 * (eta_p_K_m, eta_u_K_m, eta_u_dt_K_m of estimate_error_at_t, :168-171) and
 * the cell contributions to R4p (:146-153).  fac (nF x 4: cell a, local
 * face of a, cell b, local face of b) lists the interior facets, cell_f
 * (nC x 4) the interior facet of each local face (-1 on the boundary);
 * tmp holds 3 nF doubles.  Deterministic. */
int mpet_error_indicators(const MpetIndDesc* d, const double* x, const double* xp,
                          const int32_t* fac, int nF, const int32_t* cell_f, double* tmp,
                          double* out, hipStream_t stream);

/* mpet_error_indicators without the tmp workspace, allocating nothing: no
 * counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight adaptive loops, harnesses that
 * forbid hipMalloc).  The facet values live in a library-owned device array
 * of 3 * MPET_NOALLOC_MAX_FACETS doubles (48 MiB) per GPU, held until the
 * process ends; nF above the maximum returns MPET_ERR_INVALID.  Calls that
 * overlap on one GPU share that array and race.  Same outputs as
 * mpet_error_indicators. */
#define MPET_NOALLOC_MAX_FACETS 2097152
int mpet_error_indicators_noalloc(const MpetIndDesc* d, const double* x, const double* xp,
                                  const int32_t* fac, int nF, const int32_t* cell_f,
                                  double* out, hipStream_t stream);

/* P2 mass int phi_a phi_b on the P2 node pattern (nnz22 doubles, device),
 * the vector mass of the u block that the reference adds to the
 * preconditioner with a displacement nullspace (inner(u, v) dx,
 * mpet/mpettotalpressuresolver.py:290-293) and the source of the rigid-
 * motion constraint rows.  Deterministic (colour order). */
int mpet_assemble_p2_mass(const MpetMeshDesc* mesh, const MpetPatterns* pat, double* M2,
                          hipStream_t stream);

/* Scalar CSR (int64 rowptr computed by the caller) of the un-eliminated A
 * (which = 0) or P (which = 1) in canonical numbering — the PETSc AIJ matrix
 * that `assemble(a)` / `assemble(prec)` produce (:437, :444). */
int mpet_expand_csr(const MpetPatterns* pat, const MpetPrimitives* prim, int nP2, int nP1,
                    int A, const MpetCoef* coef, int which, const long long* rowptr,
                    int* cols, double* vals, hipStream_t stream);

/* Two-field system of MPETSolver (mpet/mpetsolver.py:196-201; fields u,
 * p_1..p_A, F = A pressure fields per P1 node): coefficients of the
 * pressure rows (C row-major, stride MPET_MAXF, times the P1 mass; stiff[f]
 * times St_f), the Biot-Willis coefficients alpha_f and the Lame constants. */
typedef struct {
  double C[MPET_MAXF * MPET_MAXF];
  double stiff[MPET_MAXF];
  double alpha[MPET_MAXF];
  double mu, lmbda;
} MpetTwoFieldCoef;

/* Scalar FP64 CSR (int64 rowptr computed by the caller) of the two-field
 * system from the assembled primitives (Kuu, Lap, Dv, Dt, M1, St):
 * which = 0 the matrix of lhs(F) + Robin (`assemble(a)` + `a_robin`,
 * :334-338, :411-415), 1 the preconditioner prec (:269-271), 2 the matrix of
 * the up_-linear part of rhs(F) (`assemble(L)`, :356, :433; pressure rows
 * only).  The u-u block 2 mu eps:eps + lambda div div is formed from Kuu and
 * Lap (no new element integrals). */
int mpet_tf_expand_csr(const MpetPatterns* pat, const MpetPrimitives* prim, int nP2, int nP1,
                       int A, const MpetTwoFieldCoef* coef, int which, const long long* rowptr,
                       int* cols, double* vals, hipStream_t stream);

/* Pressure rows of b = rhs(F) at t_n (linear in up_): replaces
 * `b = assemble(L)` (:461); zeroes the u rows.  coef.C = rhs coupling,
 * coef.stiff[f] = dt*(1-theta). */
int mpet_rhs_pressure(const MpetPatterns* pat, const MpetPrimitives* prim, int nP2, int nP1,
                      int A, const MpetCoef* coef, const double* up_prev, double* b,
                      hipStream_t stream);

/* Volume source b += scale * int f_h . v for f interpolated into P_k on each
 * cell (fval: nC*nk*ncomp nodal values, Q: ntest*nk reference weights):
 * replaces `assemble(L1_i)` volume part (dt g_i w_i, :322, :471) and
 * `assemble(L0)` volume part (f . v, :313, :482). */
int mpet_cell_source(const MpetMeshDesc* mesh, const int* test_nodes, int ntest,
                     const double* Q, int nk, const double* fval, int ncomp, double scale,
                     long long dof_base, int dof_stride, int dof_off, double* b,
                     hipStream_t stream);

/* Exterior-facet terms, step 1 (per facet): for facet f, test node a,
 * component c:
 *   loc[f][a][c] = scale |F_f| sum_m sum_n T[a][m][n] u[f][m][c] w[f][n]
 * with T (ntest x nm x nn) a reference-triangle table, u either an array
 * (u + f*u_fstride + m*ncomp + c; u_fstride 0 broadcasts one block) or
 * gathered from the vector xg at gdof[f*nm + m] (ncomp 1), and w optional
 * (nn values per facet; NULL = 1).  Replaces the facet integrals of
 * assemble(L0) ds_m(1) (:313, :482), assemble(L1_i) ds_i(1) (:322, :471),
 * rhs(L2_i) and lhs(L2_i) ds_i(2) (:324, :439-441, :474-476). */
int mpet_facet_local(int nF, int ntest, int nm, int nn, int ncomp, const double* T,
                     const double* area, const double* u, long long u_fstride,
                     const double* xg, const int* gdof, const double* w, double scale,
                     double* loc, hipStream_t stream);

/* Step 2 for vectors: b[row_dof[r] + c] += sum_{k in inc_rp[r]..} loc[inc[k]][c]
 * (inc[k] = f * ntest + a: the incident facets of boundary node r, fixed
 * order; one writer per entry, no atomics). */
int mpet_facet_scatter_vec(int nrows, int ntest, int ncomp, const long long* row_dof,
                           const int* inc_rp, const int* inc, const double* loc, double* b,
                           hipStream_t stream);

/* Step 2 for the Robin matrix: vals (P1 node pattern rp11/ci11) at
 * (row_node[r], fnode[3f + b]) += loc[f][a][b] for (f, a) incident to r. */
int mpet_facet_scatter_mat(int nrows, const int* row_node, const int* inc_rp, const int* inc,
                           const int* fnode, const int* rp11, const int* ci11,
                           const double* loc, double* vals, hipStream_t stream);

/* y = A x (b == NULL) or y = b - A x, Dirichlet rows zeroed: the MatMult
 * of PETSc's KSP on the symmetrically eliminated Acopy (:457, :489, :493)
 * for iterates with zero Dirichlet components; with b it is the lifting
 * b_I - A_ID g (apply_symmetric, bc_symmetric.py:11-18) and the outer
 * FP64 residual of iterative refinement.  x, b, y of the op's dtype. */
int mpet_op_apply(const MpetOpDesc* op, const void* x, const void* b, void* y,
                  hipStream_t stream);

/* Cell-wise (matrix-free) form of the same operator, in FP32 (fp64 = 0:
 * the MatMult of the inner FP32 MINRES, :493) or FP64 (fp64 = 1: the
 * refinement's FP64 residuals, b - A x): the u rows and the divergence
 * part of the p0 rows (2 mu eps(u):eps(v) + p0 div v + div(u) w0,
 * :267-269) are integrated cell by cell from the geometry, exactly (P2
 * gradients are linear); the P1 x P1 pressure block (C (x) M1 + stiff St,
 * with the Robin terms in St) is read from the structured copy of the same
 * precision.  Summation order is
 * fixed: results are reproducible.  Cells are taken in blocks of
 * cells_per_block (= 256) consecutive cells; a block sums, per node its
 * cells touch, their contributions in slot order, and writes one partial
 * per (node, block), the block's partials contiguous (block-major); a row
 * sums its node's partials in block order, reached through the node-sorted
 * entry list.
 * Cell tables are structure-of-arrays, [k*nC + e]:
 *   node2 / node1  P2 node of local node k (UFC order) / P1 node of vertex k
 *   geom           10*nC values from mpet_cell_geometry (FP32) or
 *                  mpet_cell_geometry64 (FP64)
 * Partials, for P2 nodes (suffix 2, local slot c*10 + k, c the cell within
 * its block) and P1 nodes (suffix 1, local slot c*4 + v):
 *   bq[b] .. bq[b+1]-1   the partials of block b (nblocks + 1 entries)
 *   lrp[q] .. lrp[q+1]-1 partial q's local slots in lsl (ascending)
 *   pe[k]                the partial at entry k of the node-sorted list
 *   inc2_rp / inc1_rp    node -> first entry of its partials (n + 1)
 * Work: 3*npart2 + npart1 values of the descriptor's precision.  lanes_u
 * in {1, 4, 8, 16} (1: one thread per P2 node), lanes_p in {8, 16}
 * (others: 8).
 * The reference always multiplies by the assembled PETSc matrix (:437). */
typedef struct {
  int32_t nC, nP2, nP1, A;
  const int32_t *node2, *node1;  /* 10*nC, 4*nC */
  const void* geom;               /* float, or double when fp64 */
  int32_t cells_per_block, npart2, npart1;
  const int32_t *bq2, *lrp2, *pe2;
  const int16_t* lsl2;
  const int32_t *bq1, *lrp1, *pe1;
  const int16_t* lsl1;
  const int32_t *inc2_rp, *inc1_rp;   /* nP2 + 1, nP1 + 1 */
  const int32_t *rp11, *ci11;
  int32_t nnz11;
  const void *M1, *St;           /* structured copy (nnz11, A*nnz11), float or double */
  MpetCoef coef;
  const uint8_t* dmask;
  int32_t lanes_u, lanes_p, xcd_chunk;
  int32_t fp64;                  /* 0: FP32 arrays and vectors, 1: FP64 */
} MpetCellOpDesc;

/* Per-cell geometry of the cell-wise operator, FP32 structure-of-arrays:
 * geom[d*nC + e] = d lambda_{1 + d/3} / d x_{d%3} (d < 9), geom[9*nC + e]
 * = |T_e|, computed in FP64 from mesh->coords and rounded. */
int mpet_cell_geometry(const MpetMeshDesc* mesh, float* geom, hipStream_t stream);

/* The same geometry kept in FP64 (for the FP64 cell-wise operator).  No
 * counterpart in the reference. */
int mpet_cell_geometry64(const MpetMeshDesc* mesh, double* geom, hipStream_t stream);

/* y = A x / y = b - A x with Dirichlet rows zeroed, as mpet_op_apply,
 * through the cell-wise operator; x, b, y and work (3*npart2 + npart1
 * values) in the descriptor's precision. */
int mpet_cell_op_apply(const MpetCellOpDesc* op, const void* x, const void* b, void* y,
                       void* work, hipStream_t stream);

/* Scalar CSR SpMV (PETSc MatMult on AIJ, :493), int64 rowptr, lpr lanes
 * per row in {4,8,16,32,64}; optional residual b and row mask as above. */
int mpet_csr_spmv(int dtype, long long nrows, const long long* rp, const int* ci,
                  const void* val, const void* x, const void* b, const uint8_t* dmask,
                  void* y, int lpr, hipStream_t stream);

/* FP64 AoS primitive (ncomp per entry) -> SoA copy in float64/float32. */
int mpet_soa_convert(int dtype, const double* src, int ncomp, long long n, void* dst,
                     hipStream_t stream);

/* Node-pair sparsity pattern (host arrays) of all (r, c) sharing a cell:
 * the sparsity DOLFIN builds before `assemble(a)` (:437).  Call once with
 * cols == NULL to fill rowptr (nrow + 1), then with cols sized rowptr[nrow]. */
int mpet_build_pattern(int ncell, int nr, const int32_t* rmap, int nc, const int32_t* cmap,
                       int nrow, int64_t* rowptr, int32_t* cols);

/* Smoothed-aggregation setup (host arrays, FP64) of the displacement-block
 * hierarchy: replaces hypre BoomerAMG's setup inside
 * PETScKrylovSolver("minres", "hypre_amg") (:448).  P0 optional (n1 > 0):
 * prescribed first prolongator.  The hierarchy is read back with
 * mpet_amg_level_sizes / mpet_amg_level_copy / mpet_amg_coarse_inverse. */
typedef struct {
  double strength;
  int max_coarse, max_levels, power_iters;
  double omega_scale;
} MpetAmgParams;
int mpet_amg_setup(int n, const int64_t* rp, const int* ci, const double* v,
                   const MpetAmgParams* prm, int n1, const int64_t* p0_rp,
                   const int* p0_ci, const double* p0_v, void** handle);
/* OpenMP threads of the setups the calling host thread runs next (no
 * counterpart in the reference): instances set up concurrently from a
 * thread pool split the host cores among them. */
int mpet_amg_set_threads(int n);
int mpet_amg_num_levels(void* h);
int mpet_amg_level_sizes(void* h, int l, int64_t* out);
int mpet_amg_level_copy(void* h, int l, int64_t* A_rp, int* A_ci, double* A_v, double* dinv,
                        double* rho, int64_t* P_rp, int* P_ci, double* P_v, int64_t* R_rp,
                        int* R_ci, double* R_v);
int mpet_amg_coarse_inverse(void* h, double* out);
void mpet_amg_destroy(void* h);

/* Launch shapes of a solver or batch handle, as picked by the package's
 * autotuner (waterscapes_amd/autotune.py): lanes per node row of the
 * structured operator's u and p rows (8, 16 or 32; used by the fused MINRES
 * step and by mpet_op_apply), the run length of row tiles per XCD
 * (xcd_chunk >= 0; < 0 keeps the current one), and the lanes per row of
 * each hierarchy level's three SpMMs (level operator, prolongator,
 * restriction; 1, 2, 4, .., 64).  Zero lanes keep the current value.  No
 * counterpart in the reference (hypre and PETSc pick their own launch
 * shapes, mpet/mpettotalpressuresolver.py:448). */
typedef struct {
  int32_t lanes_u, lanes_p, xcd_chunk;
  int32_t u_lpr[MPET_MAXLEV][3];
  int32_t p_lpr[MPET_MAXLEV][3];
} MpetTuning;

/* Krylov solver handle: MINRES + block preconditioner on the structured
 * (or CSR) operator — PETScKrylovSolver("minres","hypre_amg") (:448).
 * Borrows every device array of the descriptors (caller keeps them alive);
 * owns its workspaces.  op32/prec32 are needed for mode 1, prec64 for
 * mode 0, prec32 for mode 2. */
/* FP64 MINRES on a scalar CSR operator only (no structured operator): the
 * two-field MPETSolver system, F pressure fields per P1 node, replacing the
 * LU solve of step / solve_direct (mpet/mpetsolver.py:347-379, :422-456)
 * with the block preconditioner of prec.  opt->mode must be 0; dmask marks
 * the Dirichlet DOFs (n = 3 nP2 + F nP1 entries). */
int mpet_solver_create_csr(const MpetCsrDesc* csr64, const uint8_t* dmask, int nP2, int nP1,
                           int F, const MpetPrecDesc* prec64, const MpetSolverOptions* opt,
                           void** handle);

/* Mixed precision on a CSR operator alone (the two-field system): FP32
 * MINRES with the FP32 block preconditioner prec32 on an FP32 copy of the
 * CSR values (made here, owned by the handle), inside FP64 iterative
 * refinement with the FP64 CSR residual, to opt->outer_tol as in
 * mpet_solver_create's mode 1.  opt->mode must be 1.  No counterpart in the
 * reference (which factorises the two-field system with MUMPS,
 * mpet/mpetsolver.py:347-379); an alternative to mpet_solver_create_csr. */
int mpet_solver_create_csr_mixed(const MpetCsrDesc* csr64, const uint8_t* dmask, int nP2,
                                 int nP1, int F, const MpetPrecDesc* prec32,
                                 const MpetSolverOptions* opt, void** handle);

/* FP64 MINRES on the bordered system [A B^T; B 0] of the nullspace
 * multipliers (rigid motions of u, constants of chosen pressure fields:
 * mpet/mpettotalpressuresolver.py:179-202, 285-306; mpet/mpetsolver.py:
 * 107-128, 203-215), replacing the LU / MINRES solve of the augmented
 * mixed space.  A is the CSR operator of mpet_solver_create_csr (F pressure
 * fields); B (m x n, row-major, device, borrowed) the constraint rows of
 * mpet_constraint_rows, m <= 16; omega > 0 the multiplier block of the
 * preconditioner (omega I; the reference's int z r dx gives |Omega|).
 * Vectors b and x of mpet_solver_solve then hold n + m entries (fields,
 * then multipliers); the multiplier rows of b are the constraint values
 * (0 in the reference).  opt->mode must be 0. */
int mpet_solver_create_bordered(const MpetCsrDesc* csr64, const uint8_t* dmask, int nP2,
                                int nP1, int F, const MpetPrecDesc* prec64, const double* B,
                                int m, double omega, const MpetSolverOptions* opt,
                                void** handle);

/* Constraint rows B of the bordered system (device, (nz + nf) x n,
 * row-major): rows k < nz = int Z_k . v for the P2 vector test functions,
 * from the P2 mass M2 (nnz22, mpet_assemble_p2_mass) and Zt (nz x 3 nP2,
 * the rigid motions at the P2 nodes, node-interleaved; nz = 0 or 6);
 * rows nz + j = int q over the P1 test functions of pressure field
 * fields[j] (host array, field numbering of the DOF layout), from the P1
 * mass M1 (nnz11).  Restates the constraint terms of
 * mpet/mpettotalpressuresolver.py:285-303 and mpet/mpetsolver.py:203-215. */
int mpet_constraint_rows(const MpetPatterns* pat, const double* M2, const double* M1, int nP2,
                         int nP1, int F, const double* Zt, int nz, const int* fields, int nf,
                         double* B, hipStream_t stream);

int mpet_solver_create(const MpetOpDesc* op64, const MpetOpDesc* op32,
                       const MpetCsrDesc* csr64, const MpetPrecDesc* prec64,
                       const MpetPrecDesc* prec32, const MpetSolverOptions* opt,
                       void** handle);
int mpet_solver_destroy(void* h);

/* One time-step solve: solver.set_operators(Acopy, P); solver.solve(x, b)
 * (:492-493) with the symmetric Dirichlet elimination of :487-489 folded in.
 * b: full right-hand side (FP64, un-lifted); x: in = initial guess with the
 * Dirichlet values in place, out = solution.  Synchronises the stream.
 * Returns MPET_ERR_NOT_CONVERGED (stats filled) when maxit is reached. */
int mpet_solver_solve(void* h, const double* b, double* x, MpetSolveStats* stats,
                      hipStream_t stream);

/* One SpMM of an AMG level as the V-cycle launches it (the autotuner's
 * timing probe): NC = 1 or 3 node-interleaved columns, mode 0 y = A x,
 * 1 y = b - A x, 3 one Chebyshev step (d = cd d + cr dinv (b - A x),
 * y = x + d); lpr lanes per row; v16 / dci / vscale the compact FP32 copy
 * (dtype 1 only; v may then be NULL).  No counterpart in the reference. */
int mpet_level_spmm(int dtype, int nc, int mode, int nrows, const int32_t* rp,
                    const int32_t* ci, const void* v, const void* x, const void* b, void* y,
                    const void* dinv, void* d, double cd, double cr, const uint8_t* mask,
                    int lpr, const void* v16, const int16_t* dci, double vscale,
                    hipStream_t stream);

/* Apply launch shapes (MpetTuning) to the operator and preconditioner of
 * precision dtype (0 FP64, 1 FP32) of a solver handle; captured graphs are
 * dropped and captured again by the next solve.  Results change only by
 * the summation order of the reductions inside a row. */
int mpet_solver_set_tuning(void* h, int dtype, const MpetTuning* t);

/* Inner FP32 MINRES of a mixed-precision solver handle on the cell-wise
 * operator (op != NULL; the handle allocates its work floats and
 * keeps a copy of the descriptor, whose arrays must outlive it) or back on
 * the FP32 structured operator (op == NULL).  Captured graphs are dropped
 * and captured again by the next solve.  No counterpart in the reference. */
int mpet_solver_set_cell_operator(void* h, const MpetCellOpDesc* op);

/* FP64 residuals b - A x of a solver handle (the refinement's outer
 * residuals in mixed mode, the start residuals in FP64 mode) through the
 * FP64 cell-wise operator (op != NULL, fp64 = 1; the handle allocates its
 * work and keeps a copy of the descriptor, whose arrays must outlive it)
 * or back through the FP64 structured operator (op == NULL).  Not used
 * when the handle's FP64 operator is a CSR matrix.  No counterpart in the
 * reference (PETSc's MatMult on the assembled matrix, :493). */
int mpet_solver_set_cell_operator64(void* h, const MpetCellOpDesc* op);

/* z = P^-1 r with the FP64 (dtype 0) or FP32 (1) preconditioner. */
int mpet_precond_apply(void* h, int dtype, const void* r, void* z, hipStream_t stream);

/* ---------------------------------------------------------------------------
 * Batches: many independent instances (different meshes, network counts,
 * parameters and boundary conditions) stepped together.  No counterpart in
 * the reference, which steps one MPETTotalPressureSolver at a time (:456-502);
 * each instance's step is the one mpet_solver_solve performs.
 *
 * Packed numbering of a batch of K instances:
 *   u DOFs of all instances first: instance k's P2 nodes are the global P2
 *   nodes p2_off[k] .. p2_off[k+1]-1, u DOF = 3a + j;
 *   then the pressure DOFs: instance k's are 3*nP2 + pd_off[k] + F_k*c' + f
 *   (c' = local P1 node, node-interleaved as for one instance).
 * ------------------------------------------------------------------------- */

/* Per-instance constants of a packed operator (device array, K entries). */
typedef struct {
  double C[MPET_MAXF * MPET_MAXF];   /* pressure coupling, as MpetCoef */
  double stiff[MPET_MAXF];
  int64_t st_base;    /* offset of the instance's St arrays in the packed St */
  int64_t k11_base;   /* first P11 nonzero of the instance */
  int64_t nnz11;      /* its P11 nonzeros (stride of its St arrays) */
  int32_t F, pad;     /* fields A+1 */
} MpetBatchCoef;

/* Packed structured operator: the node patterns of all instances
 * concatenated (row pointers global), values structure-of-arrays with the
 * packed nonzero counts as strides (Kuu[c*nnz22 + k] etc., St per instance
 * at st_base).  Column indices: ci22, ci12 = global P2 node; ci21, ci11 =
 * packed DOF of the column node's p_0 (fields f follow at +f).  p_inst /
 * p_base: instance and packed p_0 DOF of every global P1 node. */
typedef struct {
  int32_t ninst, nP2, nP1;
  const int32_t *rp22, *ci22, *rp21, *ci21, *rp12, *ci12, *rp11, *ci11;
  int64_t nnz22, nnz21, nnz12, nnz11;
  int32_t dtype;                      /* 0 float64, 1 float32 */
  const void *Kuu, *Dv, *Dt, *M1, *St;
  const int32_t *p_inst, *p_base;
  const MpetBatchCoef* coef;
  const uint8_t* dmask;               /* packed Dirichlet mask (required) */
} MpetBatchOpDesc;

/* Instance ranges (host arrays of K+1 prefix sums). */
typedef struct {
  int32_t ninst;
  const int64_t* p2_off;   /* P2 nodes */
  const int64_t* p1_off;   /* P1 nodes */
  const int64_t* pd_off;   /* pressure DOFs, relative to the first one */
} MpetBatchLayout;

/* Batched solver handle over packed operators and preconditioners (the
 * hierarchies of the instances concatenated level by level, block-diagonal
 * coarsest level, pmap set).  opt->mode 0 (FP64) or 1 (mixed); one MINRES
 * state per instance.  Borrows the descriptors' arrays; owns workspaces. */
int mpet_batch_create(const MpetBatchOpDesc* op64, const MpetBatchOpDesc* op32,
                      const MpetBatchLayout* layout, const MpetPrecDesc* prec64,
                      const MpetPrecDesc* prec32, const MpetSolverOptions* opt,
                      void** handle);
int mpet_batch_destroy(void* h);

/* One time step of the instances with active[k] != 0 (NULL: all): for each,
 * the solve of mpet_solver_solve on its part of the packed b and x
 * (packed numbering).  stats: host array of K entries.  Synchronises. */
int mpet_batch_solve(void* h, const double* b, double* x, const int32_t* active,
                     MpetSolveStats* stats, hipStream_t stream);

/* y = A x or y = b - A x of every instance in the packed operator of the
 * handle (dtype 0 FP64, 1 FP32), Dirichlet rows 0 (as mpet_op_apply). */
int mpet_batch_apply(void* h, int dtype, const void* x, const void* b, void* y,
                     hipStream_t stream);

/* Lanes per row of the batch's hierarchy levels (the u_lpr / p_lpr part of
 * MpetTuning; the packed operator keeps its shapes); drops captured graphs. */
int mpet_batch_set_tuning(void* h, int dtype, const MpetTuning* t);

/* z = P^-1 r with the batch's block preconditioner (as mpet_precond_apply). */
int mpet_batch_precond_apply(void* h, int dtype, const void* r, void* z, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* MPET_HIP_H */
