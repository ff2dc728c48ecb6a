// Preconditioned MINRES with the block-diagonal MPET preconditioner, FP64
// or FP32 inner iterations inside FP64 iterative refinement.
//
// Replaces PETScKrylovSolver("minres", "hypre_amg") + solve
// (mpet/mpettotalpressuresolver.py:448, :492-494).
//
// Preconditioner (block diagonal, SPD, a fixed linear operator):
//   u : one V-cycle of the hierarchy built on the P2 Laplacian (mu scaled),
//       applied to the 3 displacement components at once (node-interleaved
//       vectors: every matrix value is read once for three columns);
//   p : one V-cycle of the hierarchy built on diag(P_0, .., P_A) (pressure
//       blocks in field-major order: aggregation never couples fields, so
//       all A+1 blocks share every launch).
//   Chebyshev smoothing (degree k on [rho/ratio, safety*rho] of D^-1 A),
//   dense inverse on the coarsest level.
//
// MINRES (Paige & Saunders 1975) keeps its scalar recurrences on the device.
// Dot products write one partial per block; a one-block finalize kernel
// sums them in fixed order (deterministic) and updates the state.  No block
// ever waits on another inside a kernel: ordering comes from kernel
// boundaries only.  Iteration i reads the flag inactive[i & 1]; the
// finalize kernel of iteration i writes inactive[(i + 1) & 1], so the
// iteration that converges still completes its update and every later one
// is a no-op.  Chunks of iterations are therefore enqueued (or replayed from
// a captured hipGraph) with one host check per chunk.
//
// Dirichlet invariant: Krylov vectors have zero Dirichlet components and
// operators zero their Dirichlet rows (see spmv.hip).
#include "amg_cycle.h"
#include "cellwise.h"
#include "op_rows.h"
#include <vector>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

namespace mpet {

// ---------------------------------------------------------------------------
// MINRES state
// ---------------------------------------------------------------------------
struct MinresState {
  double beta1, beta, oldb, dbar, epsln, phibar, cs, sn, phi, gamma, delta, oldeps, alfa;
  double tol;            // absolute threshold on phibar (||r||_{M^-1})
  int it, maxit, nan, done;
  int inactive[2];
};


// ---------------------------------------------------------------------------
// MINRES kernels
// ---------------------------------------------------------------------------
// Epilogue of the fused operator (K1), one output DOF:
//   y = s (A z) - c1 r1,  v = s z,  returns v . y  (partial of alfa).
template <typename T>
struct EpiK1 {
  T* y; T* v; const T* z; const T* r1; const uint8_t* mask;   // v may be null
  __device__ __forceinline__ double operator()(long long i, T ax, T s, T c1) const {
    T yy = (mask && mask[i]) ? (T)0 : s * ax - c1 * r1[i];
    T vv = s * z[i];
    y[i] = yy;
    if (v) v[i] = vv;
    return (double)vv * (double)yy;
  }
};

template <typename T>
__device__ __forceinline__ void k1_scalars(const MinresState* st, T& s, T& c1) {
  s = (T)(1.0 / st->beta);
  c1 = st->it > 0 ? (T)(st->beta / st->oldb) : (T)0;
}

// K1 epilogue for the cell-wise operator's row kernel (cellwise.h).
struct CwEpiK1 {
  EpiK1<float> e;
  const MinresState* st;
  const int* skip;
  float s = 0.f, c1 = 0.f;
  __device__ __forceinline__ bool skipped() const { return *skip != 0; }
  __device__ __forceinline__ void prep() { k1_scalars(st, s, c1); }
  __device__ __forceinline__ double operator()(long long i, float ax) const {
    return e(i, ax, s, c1);
  }
};

// K1 on the structured operator, grid-stride over row tiles: GU lanes per
// P2 node row (tiles of NT / GU rows), GP lanes per P1 node row (tiles of
// NT / GP rows); the u tiles come first.  The lane counts are the
// operator's lanes_u / lanes_p (8, 16 or 32; picked by the autotuner).
template <typename T, int F, int GU, int GP>
__global__ void __launch_bounds__(NT)
k1_op_kernel(MpetOpT<T> op, const T* __restrict__ z, EpiK1<T> epi, const MinresState* st,
             const int* __restrict__ skip, double* __restrict__ partial) {
  if (*skip) return;
  T s, c1;
  k1_scalars(st, s, c1);
  constexpr int RU = NT / GU, RP = NT / GP;
  const long long nu = 3ll * op.nP2;
  const int ntu = (op.nP2 + RU - 1) / RU, ntp = (op.nP1 + RP - 1) / RP;
  const long long n22 = op.nnz22, n21 = op.nnz21, n12 = op.nnz12, n11 = op.nnz11;
  double acc = 0.0;
  for (int t = blockIdx.x; t < ntu + ntp; t += gridDim.x) {
    const int tile = xcd_tile(t, ntu + ntp, op.xcd_chunk);
    if (tile < ntu) {
      const int lane = threadIdx.x % GU;
      const int a = tile * RU + threadIdx.x / GU;
      if (a < op.nP2) {
        T y0 = 0, y1 = 0, y2 = 0;
        urow_apply<T, GU>(op.rp22, op.ci22, op.Kuu, n22, op.rp21, op.ci21, op.Dv, n21, z, nu,
                          F, a, lane, y0, y1, y2);
        y0 = group_allsum<GU>(y0);
        y1 = group_allsum<GU>(y1);
        y2 = group_allsum<GU>(y2);
        if (lane < 3) {
          const T yc = lane == 0 ? y0 : (lane == 1 ? y1 : y2);
          acc += epi(3ll * a + lane, yc, s, c1);
        }
      }
    } else {
      const int lane = threadIdx.x % GP;
      const int c = (tile - ntu) * RP + threadIdx.x / GP;
      if (c < op.nP1) {
        T yd = 0, m[F], sv[F];
#pragma unroll
        for (int g = 0; g < F; ++g) { m[g] = 0; sv[g] = 0; }
        prow_apply<T, GP, F>(op.rp12, op.ci12, op.Dt, n12, op.rp11, op.ci11, op.M1, op.St, n11,
                             0, z, nu, F, F, c, lane, yd, m, sv);
        yd = group_allsum<GP>(yd);
        T mine = 0;
#pragma unroll
        for (int g = 0; g < F; ++g) {
          m[g] = group_allsum<GP>(m[g]);
          T svg = group_allsum<GP>(sv[g]);
          if (lane == g) mine = (g == 0) ? yd : (T)op.stiff[g] * svg;
        }
        if (lane < F) {
#pragma unroll
          for (int g = 0; g < F; ++g) mine += (T)op.C[lane * MPET_MAXF + g] * m[g];
          acc += epi(nu + (long long)F * c + lane, mine, s, c1);
        }
      }
    }
  }
  double bs = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = bs;
}

// Lane counts of the fused K1 (8, 16 or 32; anything else the default 16,
// and GP at least the number of fields).
static int k1_lanes(int g, int lo) {
  const int v = (g == 8 || g == 16 || g == 32) ? g : 16;
  return v < lo ? 16 : v;
}

// K1 on a scalar CSR matrix (plain FP64 path), LPR lanes per row.
template <typename T, int LPR>
__global__ void __launch_bounds__(NT)
k1_csr_kernel(long long nrows, const long long* __restrict__ rp, const int* __restrict__ ci,
              const T* __restrict__ val, const T* __restrict__ z, EpiK1<T> epi,
              const MinresState* st, const int* __restrict__ skip, double* __restrict__ partial) {
  if (*skip) return;
  T s, c1;
  k1_scalars(st, s, c1);
  const int lane = threadIdx.x % LPR;
  constexpr int ROWS = NT / LPR;
  const long long ntiles = (nrows + ROWS - 1) / ROWS;
  double acc = 0.0;
  for (long long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    long long r = tile * ROWS + threadIdx.x / LPR;
    if (r < nrows) {
      T sum = csr_row<T, LPR>(rp, ci, val, z, r, lane);
      sum = group_allsum<LPR>(sum);
      if (lane == 0) acc += epi(r, sum, s, c1);
    }
  }
  double bs = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = bs;
}

// alfa = sum of the K1 partials.
__global__ void __launch_bounds__(NT)
k1f_kernel(const double* __restrict__ partial, int np, MinresState* st,
           const int* __restrict__ skip) {
  if (*skip) return;
  double s = sum_partials(partial, np);
  if (threadIdx.x == 0) st->alfa = s;
}

// K2: r2_new = y - (alfa/beta) r2_old (written over r1).
template <typename T>
__global__ void __launch_bounds__(NT)
k2_kernel(long long n, const T* __restrict__ y, const T* __restrict__ r2o, T* __restrict__ r2n,
          const MinresState* st, const int* __restrict__ skip) {
  if (*skip) return;
  const T c = (T)(st->alfa / st->beta);
  long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i < n) r2n[i] = y[i] - c * r2o[i];
}

// Partials of a . b (double accumulation).
template <typename T>
__global__ void __launch_bounds__(NT)
dot_kernel(long long n, const T* __restrict__ a, const T* __restrict__ b,
           const int* __restrict__ skip, double* __restrict__ partial) {
  if (skip && *skip) return;
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * NT + threadIdx.x; i < n;
       i += (long long)gridDim.x * NT)
    acc += (double)a[i] * (double)b[i];
  double bs = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = bs;
}

// K4 finalize: beta = sqrt(r2 . z) and the Givens update; decides whether
// iteration i + 1 runs (writes inactive[next]).
__global__ void __launch_bounds__(NT)
k4f_kernel(const double* __restrict__ partial, int np, MinresState* st, int cur, int next) {
  if (st->inactive[cur]) {
    if (threadIdx.x == 0) st->inactive[next] = 1;
    return;
  }
  double tot = sum_partials(partial, np);
  if (threadIdx.x == 0) {
    MinresState s = *st;
    s.oldb = s.beta;
    // fmax(NaN, c) is c: test the raw sum and the rotation for NaN / inf
    // before clamping, so a breakdown is reported instead of read as 0
    const bool bad_in = !(tot == tot) || isinf(tot);
    s.beta = sqrt(fmax(tot, 0.0));
    s.oldeps = s.epsln;
    s.delta = s.cs * s.dbar + s.sn * s.alfa;
    double gbar = s.sn * s.dbar - s.cs * s.alfa;
    s.epsln = s.sn * s.beta;
    s.dbar = -s.cs * s.beta;
    const double hg = hypot(gbar, s.beta);
    s.gamma = fmax(hg, 1e-300);
    s.cs = gbar / s.gamma;
    s.sn = s.beta / s.gamma;
    s.phi = s.cs * s.phibar;
    s.phibar = s.sn * s.phibar;
    s.it += 1;
    bool bad = bad_in || !(hg == hg) || !(s.phibar == s.phibar) || isinf(s.phibar);
    s.nan = bad ? 1 : 0;
    int conv = (bad || s.phibar <= s.tol || s.beta == 0.0 || s.it >= s.maxit) ? 1 : 0;
    s.done = conv;
    s.inactive[next] = conv;
    *st = s;
  }
}

// K5: w = (v - oldeps w1 - delta w2) / gamma (written over w1); x += phi w.
template <typename T>
__global__ void __launch_bounds__(NT)
k5_kernel(long long n, const T* __restrict__ v, T* __restrict__ w1, const T* __restrict__ w2,
          T* __restrict__ x, const MinresState* st, const int* __restrict__ skip) {
  if (*skip) return;
  long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i >= n) return;
  const T oe = (T)st->oldeps, de = (T)st->delta, ig = (T)(1.0 / st->gamma), ph = (T)st->phi;
  T w = (v[i] - oe * w1[i] - de * w2[i]) * ig;
  w1[i] = w;
  x[i] += ph * w;
}

// K5 with v = s z formed here from this iteration's z (s = 1 / oldb: the
// 1 / beta that K1 scaled z with, bit for bit), so that K1 need not store v.
template <typename T>
__global__ void __launch_bounds__(NT)
k5z_kernel(long long n, const T* __restrict__ z, T* __restrict__ w1, const T* __restrict__ w2,
           T* __restrict__ x, const MinresState* st, const int* __restrict__ skip) {
  if (*skip) return;
  long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i >= n) return;
  const T sv = (T)(1.0 / st->oldb);
  const T oe = (T)st->oldeps, de = (T)st->delta, ig = (T)(1.0 / st->gamma), ph = (T)st->phi;
  T w = (sv * z[i] - oe * w1[i] - de * w2[i]) * ig;
  w1[i] = w;
  x[i] += ph * w;
}

// init: r1 = r2 = b, w1 = w2 = x = 0.
template <typename T>
__global__ void __launch_bounds__(NT)
minres_init_vec_kernel(long long n, const T* __restrict__ b, T* r1, T* r2, T* w1, T* w2, T* x) {
  long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i >= n) return;
  T v = b[i];
  r1[i] = v; r2[i] = v;
  w1[i] = 0; w2[i] = 0; x[i] = 0;
}

// rtol > 0: stop at phibar <= rtol * beta1; rtol < 0: stop at phibar <= -rtol
// (an absolute threshold computed by the caller).
__global__ void __launch_bounds__(NT)
minres_init_state_kernel(const double* __restrict__ partial, int np, MinresState* st,
                         double rtol, int maxit) {
  double tot = sum_partials(partial, np);
  if (threadIdx.x == 0) {
    MinresState s;
    memset(&s, 0, sizeof(s));
    const bool bad = !(tot == tot) || isinf(tot);    // before fmax, which drops NaN
    s.beta1 = bad ? tot : sqrt(fmax(tot, 0.0));
    s.beta = s.beta1;
    s.phibar = s.beta1;
    s.cs = -1.0;
    s.sn = 0.0;
    s.tol = rtol >= 0 ? rtol * s.beta1 : -rtol;
    s.maxit = maxit;
    s.nan = bad ? 1 : 0;
    s.done = (s.beta1 == 0.0 || s.beta1 <= s.tol || s.nan) ? 1 : 0;
    s.inactive[0] = s.done;
    s.inactive[1] = s.done;
    *st = s;
  }
}

// ---------------------------------------------------------------------------
// precision conversion and FP64 refinement helpers
// ---------------------------------------------------------------------------
__global__ void f64_to_f32_kernel(long long n, const double* __restrict__ a, float* __restrict__ b) {
  long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}
__global__ void add_f32_to_f64_kernel(long long n, const float* __restrict__ d, double* __restrict__ x) {
  long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i < n) x[i] += (double)d[i];
}
__global__ void add_f64_kernel(long long n, const double* __restrict__ d, double* __restrict__ x) {
  long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i < n) x[i] += d[i];
}
// g = Pi_D x (the Dirichlet values of x, zero elsewhere).
__global__ void dirichlet_part_kernel(long long n, const double* __restrict__ x,
                                      const uint8_t* __restrict__ mask, double* __restrict__ g) {
  long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i < n) g[i] = mask[i] ? x[i] : 0.0;
}
__global__ void __launch_bounds__(NT)
sum_to_kernel(const double* __restrict__ partial, int np, double* __restrict__ out) {
  double s = sum_partials(partial, np);
  if (threadIdx.x == 0) *out = s;
}
__global__ void cvt_down_kernel(long long n, const double* __restrict__ a, float* __restrict__ b,
                                const int* __restrict__ skip) {
  if (skip && *skip) return;
  long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}
__global__ void cvt_up_kernel(long long n, const float* __restrict__ a, double* __restrict__ b,
                              const int* __restrict__ skip) {
  if (skip && *skip) return;
  long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i < n) b[i] = (double)a[i];
}

// ---------------------------------------------------------------------------
// Bordered systems [A B^T; B 0] (nullspace multipliers): m constraint rows
// B (m x n, row-major, dense) after the n field rows; Krylov vectors carry
// the m multipliers after the fields.
// ---------------------------------------------------------------------------
constexpr int MAXB = 16;   // constraint rows

// K1 epilogue on the field rows: y = s (A z + B^T z_t) - c1 r1 (Dirichlet
// rows 0), v = s z; partials of v . y (row 0 of partial) and of B z (row
// 1 + k), stride gridDim.x.  az = A z.
__global__ void __launch_bounds__(NT)
k1b_kernel(long long n, int m, const double* __restrict__ B, const double* __restrict__ az,
           const double* __restrict__ z, const double* __restrict__ r1,
           const uint8_t* __restrict__ mask, double* __restrict__ y, double* __restrict__ v,
           const MinresState* st, const int* __restrict__ skip, double* __restrict__ partial) {
  if (*skip) return;
  double s, c1;
  k1_scalars(st, s, c1);
  double zt[MAXB], bz[MAXB];
#pragma unroll
  for (int k = 0; k < MAXB; ++k) {
    zt[k] = k < m ? z[n + k] : 0.0;
    bz[k] = 0.0;
  }
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * NT + threadIdx.x; i < n;
       i += (long long)gridDim.x * NT) {
    double ax = az[i];
    const double zi = z[i];
#pragma unroll
    for (int k = 0; k < MAXB; ++k) {
      if (k < m) {
        const double bk = B[k * n + i];
        ax += bk * zt[k];
        bz[k] += bk * zi;
      }
    }
    const double yy = (mask && mask[i]) ? 0.0 : s * ax - c1 * r1[i];
    const double vv = s * zi;
    y[i] = yy;
    v[i] = vv;
    acc += vv * yy;
  }
  double t = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
#pragma unroll
  for (int k = 0; k < MAXB; ++k) {
    if (k < m) {
      __syncthreads();
      t = block_sum(bz[k]);
      if (threadIdx.x == 0) partial[(long long)(k + 1) * gridDim.x + blockIdx.x] = t;
    }
  }
}

// K1 finalize of a bordered system: the multiplier rows y_t = s B z - c1
// r1_t, v_t = s z_t, and alfa = v . y over fields and multipliers.
__global__ void __launch_bounds__(NT)
k1bf_kernel(const double* __restrict__ partial, int np, int m, long long n,
            const double* __restrict__ z, const double* __restrict__ r1, double* __restrict__ y,
            double* __restrict__ v, MinresState* st, const int* __restrict__ skip) {
  if (*skip) return;
  double s, c1;
  k1_scalars(st, s, c1);
  double alfa = sum_partials(partial, np);
  for (int k = 0; k < m; ++k) {
    __syncthreads();
    const double bzk = sum_partials(partial + (long long)(k + 1) * np, np);
    if (threadIdx.x == 0) {
      const double yy = s * bzk - c1 * r1[n + k];
      const double vv = s * z[n + k];
      y[n + k] = yy;
      v[n + k] = vv;
      alfa += vv * yy;
    }
  }
  if (threadIdx.x == 0) st->alfa = alfa;
}

// z_t = sc r_t on the multiplier rows (their preconditioner block is
// omega I, sc = 1 / omega).
__global__ void tail_scale_kernel(int m, const double* __restrict__ r, double* __restrict__ z,
                                  double sc, const int* __restrict__ skip) {
  if (skip && *skip) return;
  if ((int)threadIdx.x < m) z[threadIdx.x] = sc * r[threadIdx.x];
}

// Residual of a bordered system, field rows: r -= B^T x_t (r holds
// Pi (b - A x); Dirichlet rows stay 0); partials of B x (row k, stride
// gridDim.x).
__global__ void __launch_bounds__(NT)
bres_kernel(long long n, int m, const double* __restrict__ B, const double* __restrict__ x,
            const uint8_t* __restrict__ mask, double* __restrict__ r,
            double* __restrict__ partial) {
  double xt[MAXB], bx[MAXB];
#pragma unroll
  for (int k = 0; k < MAXB; ++k) {
    xt[k] = k < m ? x[n + k] : 0.0;
    bx[k] = 0.0;
  }
  for (long long i = (long long)blockIdx.x * NT + threadIdx.x; i < n;
       i += (long long)gridDim.x * NT) {
    double ri = r[i];
    const double xi = x[i];
#pragma unroll
    for (int k = 0; k < MAXB; ++k) {
      if (k < m) {
        const double bk = B[k * n + i];
        ri -= bk * xt[k];
        bx[k] += bk * xi;
      }
    }
    if (!(mask && mask[i])) r[i] = ri;
  }
#pragma unroll
  for (int k = 0; k < MAXB; ++k) {
    if (k < m) {
      if (k) __syncthreads();
      const double t = block_sum(bx[k]);
      if (threadIdx.x == 0) partial[(long long)k * gridDim.x + blockIdx.x] = t;
    }
  }
}

// Multiplier rows of the residual: r_t = b_t - B x.
__global__ void __launch_bounds__(NT)
bresf_kernel(const double* __restrict__ partial, int np, int m, long long n,
             const double* __restrict__ b, double* __restrict__ r) {
  for (int k = 0; k < m; ++k) {
    if (k) __syncthreads();
    const double t = sum_partials(partial + (long long)k * np, np);
    if (threadIdx.x == 0) r[n + k] = b[n + k] - t;
  }
}

}  // namespace mpet

// ===========================================================================
// host side
// ===========================================================================
using namespace mpet;

// Fraction of the outer target that the last inner solve of a refinement
// aims for (FP32 rounding makes the FP64 residual differ slightly from the
// inner recurrence's estimate).
constexpr double INNER_MARGIN = 0.3;

// --- MINRES workspace in precision T -------------------------------------
template <typename T>
struct Minres {
  long long n = 0, nu = 0;
  T *R[2] = {nullptr, nullptr}, *W[2] = {nullptr, nullptr};
  // z alternates with z1 between iterations (the preconditioner writes the
  // next iteration's z while this iteration's is still read by K5)
  T *z = nullptr, *z1 = nullptr, *y = nullptr, *v = nullptr, *x = nullptr, *b = nullptr;
  MinresState* st = nullptr;
  MinresState* st_host = nullptr;   // pinned
  double* partial = nullptr;        // MAX_PARTIALS
};

struct MpetSolver {
  MpetSolverOptions opt;
  long long n = 0, nu = 0;
  int F = 0;
  MpetOpT<double> op64;
  MpetOpT<float> op32;
  bool has_op32 = false;
  // optional cell-wise FP32 operator of the inner MINRES (cellwise.h)
  bool has_cw = false;
  CellOp cw;
  float* cw_work = nullptr;
  long long cw_work_n = 0;
  // optional cell-wise FP64 operator of the FP64 residuals (op_residual64)
  bool has_cw64 = false;
  CellOpT<double> cw64;
  double* cw64_work = nullptr;
  long long cw64_work_n = 0;
  // optional scalar CSR (plain FP64 path)
  long long csr_rows = 0;
  const long long* csr_rp = nullptr;
  const int* csr_ci = nullptr;
  const double* csr_v = nullptr;
  int csr_lpr = 16;
  Precond<double> pc64;
  Precond<float> pc32;
  bool has_pc64 = false, has_pc32 = false;
  Minres<double> m64;
  Minres<float> m32;
  double* r64 = nullptr;     // FP64 residual
  std::vector<void*> allocs;
  std::vector<void*> host_allocs;
  hipGraphExec_t graph32 = nullptr;
  hipGraphExec_t graph64 = nullptr;
  hipStream_t graph_stream = nullptr;
  const uint8_t* opt_mask64 = nullptr;
  float *r32 = nullptr, *z32 = nullptr;   // mode 2: FP32 preconditioner buffers
  double* g64 = nullptr;                  // Dirichlet part of the initial guess
  double* scal = nullptr;                 // device scalar
  double* scal_host = nullptr;            // pinned
  double* partial = nullptr;
  // bordered system (nullspace multipliers): mb constraint rows B (mb x n),
  // multiplier block of the preconditioner omega I (tail_inv = 1 / omega)
  int mb = 0;
  const double* B = nullptr;
  double tail_inv = 0.0;
  long long nv = 0;                       // Krylov vector length n + mb
  double* yraw = nullptr;                 // A z of the field rows
  // CSR-only solvers (no structured operator): the operator of every mode
  // is the CSR matrix; mixed mode iterates on its FP32 copy
  bool csr_only = false;
  float* csr_v32 = nullptr;
};

// Grid of the bordered kernels: one partial per block for each of the m + 1
// reductions they produce.
static int bgrid(long long n, int m) { return grid_for(n, NT * 4, MAX_PARTIALS / (m + 1)); }

// z = M^-1 r.  Mode 2 applies the FP32 preconditioner to FP64 vectors
// (convert, V-cycles in FP32, convert back).
// Whether apply_M is the plain preconditioner apply and takes the K2
// update fused (not for the FP32 preconditioner of the FP64 solve or the
// bordered systems).
template <typename T>
static bool k2_direct(const MpetSolver* S, const Precond<T>& P) {
  if constexpr (sizeof(T) == 8) {
    if (S->opt.mode == 2 || S->mb) return false;
  }
  return k2_fusable(P);
}

template <typename T>
static void apply_M(MpetSolver* S, const Precond<T>& P, const T* r, T* z, const int* skip,
                    hipStream_t s) {
  if constexpr (sizeof(T) == 8) {
    if (S->opt.mode == 2) {
      long long n = S->n;
      int eb = (int)((n + NT - 1) / NT);
      hipLaunchKernelGGL(cvt_down_kernel, dim3(eb), dim3(NT), 0, s, n, r, S->r32, skip);
      precond_apply(S->pc32, (const float*)S->r32, S->z32, skip, s);
      hipLaunchKernelGGL(cvt_up_kernel, dim3(eb), dim3(NT), 0, s, n, (const float*)S->z32, z,
                         skip);
      return;
    }
    if (S->mb) {
      precond_apply(P, r, z, skip, s);
      hipLaunchKernelGGL(tail_scale_kernel, dim3(1), dim3(64), 0, s, S->mb, r + S->n, z + S->n,
                         S->tail_inv, skip);
      return;
    }
  }
  precond_apply(P, r, z, skip, s);
}


template <typename T>
static int build_minres(MpetSolver* S, Minres<T>& M) {
  M.n = S->nv;
  M.nu = S->nu;
  int rc;
  if ((rc = alloc_dev(S->allocs, &M.R[0], M.n)) || (rc = alloc_dev(S->allocs, &M.R[1], M.n)) ||
      (rc = alloc_dev(S->allocs, &M.W[0], M.n)) || (rc = alloc_dev(S->allocs, &M.W[1], M.n)) ||
      (rc = alloc_dev(S->allocs, &M.z, M.n)) || (rc = alloc_dev(S->allocs, &M.z1, M.n)) ||
      (rc = alloc_dev(S->allocs, &M.y, M.n)) ||
      (rc = alloc_dev(S->allocs, &M.v, M.n)) || (rc = alloc_dev(S->allocs, &M.x, M.n)) ||
      (rc = alloc_dev(S->allocs, &M.b, M.n)) || (rc = alloc_dev(S->allocs, &M.st, 1)) ||
      (rc = alloc_dev(S->allocs, &M.partial, MAX_PARTIALS)))
    return rc;
  HIPCHK(hipHostMalloc((void**)&M.st_host, sizeof(MinresState), hipHostMallocDefault));
  S->host_allocs.push_back(M.st_host);
  return OK;
}

static int dot_grid(long long n) { return grid_for(n, NT * 8, 2048); }

template <typename T, int F, int GU>
static void launch_k1_gp(int gp, dim3 grid, hipStream_t s, const MpetOpT<T>& op, const T* z,
                         EpiK1<T> epi, const MinresState* st, const int* skip,
                         double* partial) {
  switch (gp) {
    case 8: hipLaunchKernelGGL((k1_op_kernel<T, F, GU, 8>), grid, dim3(NT), 0, s, op, z, epi, st, skip, partial); break;
    case 32: hipLaunchKernelGGL((k1_op_kernel<T, F, GU, 32>), grid, dim3(NT), 0, s, op, z, epi, st, skip, partial); break;
    default: hipLaunchKernelGGL((k1_op_kernel<T, F, GU, 16>), grid, dim3(NT), 0, s, op, z, epi, st, skip, partial); break;
  }
}
template <typename T, int F>
static void launch_k1_f(int gu, int gp, dim3 grid, hipStream_t s, const MpetOpT<T>& op,
                        const T* z, EpiK1<T> epi, const MinresState* st, const int* skip,
                        double* partial) {
  switch (gu) {
    case 8: launch_k1_gp<T, F, 8>(gp, grid, s, op, z, epi, st, skip, partial); break;
    case 32: launch_k1_gp<T, F, 32>(gp, grid, s, op, z, epi, st, skip, partial); break;
    default: launch_k1_gp<T, F, 16>(gp, grid, s, op, z, epi, st, skip, partial); break;
  }
}
template <typename T>
static void launch_k1_op(int F, int gu, int gp, dim3 grid, hipStream_t s, const MpetOpT<T>& op,
                         const T* z, EpiK1<T> epi, const MinresState* st, const int* skip,
                         double* partial) {
  switch (F) {
    case 2: launch_k1_f<T, 2>(gu, gp, grid, s, op, z, epi, st, skip, partial); break;
    case 3: launch_k1_f<T, 3>(gu, gp, grid, s, op, z, epi, st, skip, partial); break;
    case 4: launch_k1_f<T, 4>(gu, gp, grid, s, op, z, epi, st, skip, partial); break;
    case 5: launch_k1_f<T, 5>(gu, gp, grid, s, op, z, epi, st, skip, partial); break;
    case 6: launch_k1_f<T, 6>(gu, gp, grid, s, op, z, epi, st, skip, partial); break;
    case 7: launch_k1_f<T, 7>(gu, gp, grid, s, op, z, epi, st, skip, partial); break;
    case 8: launch_k1_f<T, 8>(gu, gp, grid, s, op, z, epi, st, skip, partial); break;
    default: break;
  }
}

// One MINRES iteration i (i selects the buffer rotation and skip flag).
template <typename T>
static void minres_iter(MpetSolver* S, Minres<T>& M, const Precond<T>& P, bool csr, int i,
                        hipStream_t s) {
  T* Ra = M.R[i & 1];
  T* Rb = M.R[(i + 1) & 1];
  T* Wa = M.W[i & 1];
  T* Wb = M.W[(i + 1) & 1];
  const int cur = i & 1, next = (i + 1) & 1;
  const int* skip = &M.st->inactive[cur];
  // this iteration's z (read by K1 and K5) and the next one's (written by
  // the preconditioner)
  T* zc = (i & 1) ? M.z1 : M.z;
  T* zn = (i & 1) ? M.z : M.z1;
  // v = z / beta is stored for K5 only by the bordered path; otherwise K5
  // forms it from zc
  const bool keep_v = csr && S->mb;
  EpiK1<T> epi{M.y, keep_v ? M.v : nullptr, zc, Ra, nullptr};
  long long n = M.n;
  int eb = (int)((n + NT - 1) / NT);
  int np = 1;
  if constexpr (sizeof(T) == 8) {
    if (csr && S->mb) {
      launch_csr_spmv<double>(S->csr_rows, S->csr_rp, S->csr_ci, S->csr_v, zc,
                              (const double*)nullptr, (const uint8_t*)nullptr, S->yraw,
                              S->csr_lpr, s);
      const int bg = bgrid(S->n, S->mb);
      hipLaunchKernelGGL(k1b_kernel, dim3(bg), dim3(NT), 0, s, S->n, S->mb, S->B,
                         (const double*)S->yraw, (const double*)zc, (const double*)Ra,
                         S->opt_mask64, M.y, M.v, M.st, skip, M.partial);
      hipLaunchKernelGGL(k1bf_kernel, dim3(1), dim3(NT), 0, s, M.partial, bg, S->mb, S->n,
                         (const double*)zc, (const double*)Ra, M.y, M.v, M.st, skip);
    } else if (csr) {
      epi.mask = S->opt_mask64;
      int lpr = S->csr_lpr;
      int rows_per_block = NT / lpr;
      long long ntiles = (S->csr_rows + rows_per_block - 1) / rows_per_block;
      np = grid_for(ntiles, 1, MAX_PARTIALS);
#define K1CSR(LP)                                                                              \
  case LP:                                                                                     \
    hipLaunchKernelGGL((k1_csr_kernel<double, LP>), dim3(np), dim3(NT), 0, s, S->csr_rows,     \
                       S->csr_rp, S->csr_ci, S->csr_v, zc, epi, M.st, skip, M.partial);        \
    break;
      switch (lpr) { K1CSR(4) K1CSR(8) K1CSR(16) K1CSR(32) K1CSR(64) default: break; }
#undef K1CSR
    }
  } else {
    if (csr) {                      // FP32 copy of a CSR-only operator
      epi.mask = S->opt_mask64;
      int lpr = S->csr_lpr;
      int rows_per_block = NT / lpr;
      long long ntiles = (S->csr_rows + rows_per_block - 1) / rows_per_block;
      np = grid_for(ntiles, 1, MAX_PARTIALS);
#define K1CSR(LP)                                                                              \
  case LP:                                                                                     \
    hipLaunchKernelGGL((k1_csr_kernel<float, LP>), dim3(np), dim3(NT), 0, s, S->csr_rows,      \
                       S->csr_rp, S->csr_ci, (const float*)S->csr_v32, zc, epi, M.st, skip,     \
                       M.partial);                                                             \
    break;
      switch (lpr) { K1CSR(4) K1CSR(8) K1CSR(16) K1CSR(32) K1CSR(64) default: break; }
#undef K1CSR
    }
  }
  bool cw_done = false;
  if constexpr (sizeof(T) == 4) {
    if (!csr && S->has_cw) {
      CwEpiK1 cepi{EpiK1<float>{M.y, nullptr, zc, Ra, S->cw.dmask}, M.st, skip};
      np = cw_launch(S->cw, zc, cepi, M.partial, skip, s);
      cw_done = true;
    }
  }
  if (!csr && !cw_done) {
    const MpetOpT<T>* op;
    if constexpr (sizeof(T) == 8) op = &S->op64; else op = &S->op32;
    epi.mask = op->dmask;
    const int gu = k1_lanes(op->lanes_u, 3), gp = k1_lanes(op->lanes_p, S->F);
    long long ntiles = (op->nP2 + NT / gu - 1) / (NT / gu) + (op->nP1 + NT / gp - 1) / (NT / gp);
    np = grid_for(ntiles, 1, MAX_PARTIALS);
    launch_k1_op<T>(S->F, gu, gp, dim3(np), s, *op, zc, epi, M.st, skip, M.partial);
  }
  if (!(csr && S->mb))
    hipLaunchKernelGGL(k1f_kernel, dim3(1), dim3(NT), 0, s, M.partial, np, M.st, skip);
  if (k2_direct(S, P)) {
    // K2 made by the preconditioner's first kernels (one vector pass and
    // one launch fewer on the iteration's critical path)
    const K2In<T> k2{M.y, Rb, Ra, &M.st->alfa, &M.st->beta};
    Skip sk;
    sk.all = skip;
    precond_apply(P, Ra, zn, sk, s, &k2);
  } else {
    hipLaunchKernelGGL((k2_kernel<T>), dim3(eb), dim3(NT), 0, s, n, M.y, Rb, Ra, M.st, skip);
    apply_M(S, P, Ra, zn, skip, s);
  }
  int dg = dot_grid(n);
  hipLaunchKernelGGL((dot_kernel<T>), dim3(dg), dim3(NT), 0, s, n, Ra, zn, skip, M.partial);
  hipLaunchKernelGGL(k4f_kernel, dim3(1), dim3(NT), 0, s, M.partial, dg, M.st, cur, next);
  if (keep_v)
    hipLaunchKernelGGL((k5_kernel<T>), dim3(eb), dim3(NT), 0, s, n, M.v, Wa, Wb, M.x, M.st,
                       skip);
  else
    hipLaunchKernelGGL((k5z_kernel<T>), dim3(eb), dim3(NT), 0, s, n, (const T*)zc, Wa, Wb, M.x,
                       M.st, skip);
}

// MINRES start: r1 = r2 = b, z = M^-1 b, beta1 = sqrt(b . z).  Returns beta1
// (synchronises the stream).
template <typename T>
static int minres_start(MpetSolver* S, Minres<T>& M, const Precond<T>& P, double rtol,
                        int maxit, hipStream_t s, double* beta1) {
  long long n = M.n;
  int eb = (int)((n + NT - 1) / NT);
  hipLaunchKernelGGL((minres_init_vec_kernel<T>), dim3(eb), dim3(NT), 0, s, n, M.b, M.R[0],
                     M.R[1], M.W[0], M.W[1], M.x);
  apply_M(S, P, M.R[1], M.z, (const int*)nullptr, s);
  int dg = dot_grid(n);
  hipLaunchKernelGGL((dot_kernel<T>), dim3(dg), dim3(NT), 0, s, n, M.R[1], M.z,
                     (const int*)nullptr, M.partial);
  hipLaunchKernelGGL(minres_init_state_kernel, dim3(1), dim3(NT), 0, s, M.partial, dg, M.st,
                     rtol, maxit);
  HIPCHK(hipMemcpyAsync(M.st_host, M.st, sizeof(MinresState), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  *beta1 = M.st_host->beta1;
  return M.st_host->nan ? ERR_NAN : OK;
}

// MINRES iterations after minres_start, in chunks of `check` iterations
// (replayed from a captured hipGraph when enabled); the host reads the
// device state once per chunk.
template <typename T>
static int minres_continue(MpetSolver* S, Minres<T>& M, const Precond<T>& P, bool csr,
                           double rtol, int maxit, int check, hipGraphExec_t* graph,
                           hipStream_t s, int* iters, int* conv) {
  if (check < 1) check = 1;
  if (check & 1) check += 1;     // buffer rotation and skip flags have period 2
  int done_iters = 0;
  while (!M.st_host->done) {
    if (graph && *graph) {
      HIPCHK(hipGraphLaunch(*graph, s));
    } else if (graph && S->opt.use_graphs) {
      hipStream_t cs = S->graph_stream;
      hipGraph_t g;
      HIPCHK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
      for (int i = 0; i < check; ++i) minres_iter(S, M, P, csr, i, cs);
      HIPCHK(hipStreamEndCapture(cs, &g));
      HIPCHK(hipGraphInstantiate(graph, g, nullptr, nullptr, 0));
      HIPCHK(hipGraphDestroy(g));
      HIPCHK(hipGraphLaunch(*graph, s));
    } else {
      for (int i = 0; i < check; ++i) minres_iter(S, M, P, csr, i, s);
    }
    done_iters += check;
    HIPCHK(hipMemcpyAsync(M.st_host, M.st, sizeof(MinresState), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (done_iters >= maxit + check) break;
  }
  *iters = M.st_host->it;
  *conv = M.st_host->nan ? -1 : (M.st_host->phibar <= M.st_host->tol ||
                                  M.st_host->beta == 0.0);
  return M.st_host->nan ? ERR_NAN : OK;
}

template <typename T>
static int minres_solve(MpetSolver* S, Minres<T>& M, const Precond<T>& P, bool csr,
                        double rtol, int maxit, int check, hipGraphExec_t* graph,
                        hipStream_t s, int* iters, int* conv) {
  double beta1;
  int rc = minres_start(S, M, P, rtol, maxit, s, &beta1);
  if (rc) return rc;
  return minres_continue(S, M, P, csr, rtol, maxit, check, graph, s, iters, conv);
}

// r = Pi (b - K x) in FP64 (K = A, or the bordered [A B^T; B 0]).
static void op_residual64(MpetSolver* S, const double* b, const double* x, double* r,
                          hipStream_t s) {
  if (S->csr_rp && (S->csr_only || (S->opt.mode == 0 && S->opt.use_csr)))
    launch_csr_spmv<double>(S->csr_rows, S->csr_rp, S->csr_ci, S->csr_v, x, b, S->opt_mask64,
                            r, S->csr_lpr, s);
  else if (S->has_cw64)
    cw_launch(S->cw64, x, CwEpiApplyT<double>{r, b, S->cw64.dmask}, (double*)nullptr,
              (const int*)nullptr, s);
  else
    launch_op_apply<double>(S->op64, x, b, r, s);
  if (S->mb) {
    const int bg = bgrid(S->n, S->mb);
    hipLaunchKernelGGL(bres_kernel, dim3(bg), dim3(NT), 0, s, S->n, S->mb, S->B, x,
                       S->opt_mask64, r, S->partial);
    hipLaunchKernelGGL(bresf_kernel, dim3(1), dim3(NT), 0, s, (const double*)S->partial, bg,
                       S->mb, S->n, b, r);
  }
}

// ||Pi (b - A g)||_{M^-1}: the preconditioned norm of the right-hand side of
// the reduced system A_II x_I = b_I - A_ID g (g = the Dirichlet values, the
// rows of x0 on the Dirichlet set).  Relative residuals are measured against
// it: unlike the norm of the full eliminated vector (which also holds g on
// the Dirichlet rows and is dominated by it), it scales with the unknowns,
// and unlike ||r_0|| it rewards a good initial guess.
template <typename T>
static int bnorm_M(MpetSolver* S, Minres<T>& M, const Precond<T>& P, const double* b,
                   const double* x0, hipStream_t s, double* out) {
  long long n = S->n;
  int eb = (int)((n + NT - 1) / NT);
  hipLaunchKernelGGL(dirichlet_part_kernel, dim3(eb), dim3(NT), 0, s, n, x0, S->opt_mask64,
                     S->g64);
  op_residual64(S, b, S->g64, S->r64, s);              // Pi (b - A g)
  if constexpr (sizeof(T) == 4)
    hipLaunchKernelGGL(f64_to_f32_kernel, dim3(eb), dim3(NT), 0, s, n, S->r64, M.b);
  else
    HIPCHK(hipMemcpyAsync(M.b, S->r64, sizeof(double) * M.n, hipMemcpyDeviceToDevice, s));
  double beta;
  int rc = minres_start(S, M, P, 1.0, 1, s, &beta);   // synchronises
  if (rc) return rc;
  *out = beta;
  return OK;
}

template <typename T>
MpetOpT<T> mpet_make_op(const MpetOpDesc* d);

static int finish_create(MpetSolver* S, const MpetPrecDesc* prec64, const MpetPrecDesc* prec32,
                         void** handle);

static int create_csr(const char* who, const MpetCsrDesc* csr64, const uint8_t* dmask, int nP2,
                      int nP1, int F, const MpetPrecDesc* prec64, const MpetPrecDesc* prec32,
                      const double* B, int m, double omega, const MpetSolverOptions* opt,
                      void** handle) {
  char msg[160];
  if (!csr64 || !csr64->rp || !csr64->ci || !csr64->v || !dmask || !(prec64 || prec32) ||
      !opt || !handle || F < 1 || F > MPET_MAXF || nP2 < 1 || nP1 < 1 ||
      csr64->nrows != 3ll * nP2 + (long long)F * nP1 || m < 0 || m > MAXB ||
      (m > 0 && (!B || !(omega > 0.0)))) {
    snprintf(msg, sizeof(msg), "%s: invalid argument", who);
    mpet_set_error(msg);
    return ERR_INVALID;
  }
  const int want = prec32 ? 1 : 0;
  if (opt->mode != want) {
    snprintf(msg, sizeof(msg), "%s: mode %d does not match the entry point (mode %d)", who,
             opt->mode, want);
    mpet_set_error(msg);
    return ERR_INVALID;
  }
  MpetSolver* S = new MpetSolver();
  S->opt = *opt;
  S->opt.use_csr = 1;
  S->F = F;
  S->nu = 3ll * nP2;
  S->n = S->nu + (long long)F * nP1;
  S->mb = m;
  S->B = B;
  S->tail_inv = m > 0 ? 1.0 / omega : 0.0;
  S->nv = S->n + m;
  S->op64.nP2 = nP2;
  S->op64.nP1 = nP1;
  S->opt_mask64 = dmask;
  S->csr_rows = csr64->nrows;
  S->csr_rp = csr64->rp;
  S->csr_ci = csr64->ci;
  S->csr_v = csr64->v;
  S->csr_lpr = csr64->lpr ? csr64->lpr : 16;
  S->csr_only = true;
  if (prec32) {
    // FP32 copy of the CSR values for the inner iterations
    long long nnz = 0;
    if (hipMemcpy(&nnz, csr64->rp + csr64->nrows, sizeof(long long), hipMemcpyDeviceToHost) !=
            hipSuccess || nnz < 1) {
      snprintf(msg, sizeof(msg), "%s: cannot read the CSR row pointers", who);
      mpet_set_error(msg);
      delete S;
      return ERR_INVALID;
    }
    int rc = alloc_dev(S->allocs, &S->csr_v32, (size_t)nnz);
    if (rc) { mpet_solver_destroy(S); return rc; }
    const long long blocks = (nnz + NT - 1) / NT;
    hipLaunchKernelGGL(f64_to_f32_kernel, dim3((unsigned)blocks), dim3(NT), 0, 0, nnz,
                       csr64->v, S->csr_v32);
    if (hipDeviceSynchronize() != hipSuccess) {
      mpet_set_error("mpet_solver_create_csr_mixed: conversion of the CSR values failed");
      mpet_solver_destroy(S);
      return ERR_HIP;
    }
  }
  return finish_create(S, prec64, prec32, handle);
}

extern "C" int mpet_solver_create_csr_mixed(const MpetCsrDesc* csr64, const uint8_t* dmask,
                                            int nP2, int nP1, int F, const MpetPrecDesc* prec32,
                                            const MpetSolverOptions* opt, void** handle) {
  if (!prec32) {
    mpet_set_error("mpet_solver_create_csr_mixed: invalid argument");
    return ERR_INVALID;
  }
  return create_csr("mpet_solver_create_csr_mixed", csr64, dmask, nP2, nP1, F, nullptr, prec32,
                    nullptr, 0, 0.0, opt, handle);
}

extern "C" int mpet_solver_create_csr(const MpetCsrDesc* csr64, const uint8_t* dmask, int nP2,
                                      int nP1, int F, const MpetPrecDesc* prec64,
                                      const MpetSolverOptions* opt, void** handle) {
  if (!prec64) {
    mpet_set_error("mpet_solver_create_csr: invalid argument");
    return ERR_INVALID;
  }
  return create_csr("mpet_solver_create_csr", csr64, dmask, nP2, nP1, F, prec64, nullptr,
                    nullptr, 0, 0.0, opt, handle);
}

extern "C" int mpet_solver_create_bordered(const MpetCsrDesc* csr64, const uint8_t* dmask,
                                           int nP2, int nP1, int F, const MpetPrecDesc* prec64,
                                           const double* B, int m, double omega,
                                           const MpetSolverOptions* opt, void** handle) {
  if (m < 1) {
    mpet_set_error("mpet_solver_create_bordered: needs at least one constraint row");
    return ERR_INVALID;
  }
  if (!prec64) {
    mpet_set_error("mpet_solver_create_bordered: invalid argument");
    return ERR_INVALID;
  }
  return create_csr("mpet_solver_create_bordered", csr64, dmask, nP2, nP1, F, prec64, nullptr,
                    B, m, omega, opt, handle);
}

extern "C" int mpet_solver_create(const MpetOpDesc* op64, const MpetOpDesc* op32,
                                  const MpetCsrDesc* csr64, const MpetPrecDesc* prec64,
                                  const MpetPrecDesc* prec32, const MpetSolverOptions* opt,
                                  void** handle) {
  if (!op64 || !opt || !handle || (!prec64 && !prec32) || op64->A < 1 ||
      op64->A > MPET_MAXF - 1 || op64->pat.nnz22 < 1 || op64->pat.nnz21 < 1 ||
      op64->pat.nnz12 < 1 || op64->pat.nnz11 < 1) {
    mpet_set_error("mpet_solver_create: invalid argument");
    return ERR_INVALID;
  }
  if (opt->mode == 1 && (!op32 || !prec32)) {
    mpet_set_error("mpet_solver_create: mixed mode needs the FP32 operator and preconditioner");
    return ERR_INVALID;
  }
  if (opt->mode == 0 && !prec64) {
    mpet_set_error("mpet_solver_create: FP64 mode needs the FP64 preconditioner");
    return ERR_INVALID;
  }
  if (opt->mode == 2 && !prec32) {
    mpet_set_error("mpet_solver_create: mode 2 needs the FP32 preconditioner");
    return ERR_INVALID;
  }
  MpetSolver* S = new MpetSolver();
  S->opt = *opt;
  S->F = op64->A + 1;
  S->nu = 3ll * op64->nP2;
  S->n = S->nu + (long long)S->F * op64->nP1;
  S->nv = S->n;
  S->op64 = mpet_make_op<double>(op64);
  S->opt_mask64 = op64->dmask;
  if (op32) {
    S->op32 = mpet_make_op<float>(op32);
    S->has_op32 = true;
  }
  if (csr64) {
    S->csr_rows = csr64->nrows;
    S->csr_rp = csr64->rp;
    S->csr_ci = csr64->ci;
    S->csr_v = csr64->v;
    S->csr_lpr = csr64->lpr ? csr64->lpr : 16;
  }
  return finish_create(S, prec64, prec32, handle);
}

// Preconditioners, MINRES workspaces and streams of a solver whose sizes,
// operators and options are set; frees S on failure.
static int finish_create(MpetSolver* S, const MpetPrecDesc* prec64, const MpetPrecDesc* prec32,
                         void** handle) {
  int rc = OK;
  if (prec64 && (rc = build_precond(S->allocs, S->nu, prec64, S->pc64))) { mpet_solver_destroy(S); return rc; }
  if (prec32 && (rc = build_precond(S->allocs, S->nu, prec32, S->pc32))) { mpet_solver_destroy(S); return rc; }
  S->has_pc64 = prec64 != nullptr;
  S->has_pc32 = prec32 != nullptr;
  if ((prec64 && (rc = make_fork(S->pc64))) || (prec32 && (rc = make_fork(S->pc32)))) {
    mpet_solver_destroy(S);
    return rc;
  }
  const bool need64 = S->opt.mode != 1, need32 = S->opt.mode == 1;
  if (need64 && (rc = build_minres(S, S->m64))) { mpet_solver_destroy(S); return rc; }
  if (need32 && (rc = build_minres(S, S->m32))) { mpet_solver_destroy(S); return rc; }
  if (S->opt.mode == 2 &&
      ((rc = alloc_dev(S->allocs, &S->r32, S->n)) || (rc = alloc_dev(S->allocs, &S->z32, S->n)))) {
    mpet_solver_destroy(S);
    return rc;
  }
  if ((rc = alloc_dev(S->allocs, &S->r64, S->nv)) ||
      (rc = alloc_dev(S->allocs, &S->g64, S->nv)) || (rc = alloc_dev(S->allocs, &S->scal, 1)) ||
      (rc = alloc_dev(S->allocs, &S->partial, MAX_PARTIALS)) ||
      (S->mb && (rc = alloc_dev(S->allocs, &S->yraw, S->n)))) {
    mpet_solver_destroy(S);
    return rc;
  }
  // multiplier rows of the Dirichlet part g of an initial guess are zero
  if (S->mb && hipMemset(S->g64 + S->n, 0, sizeof(double) * S->mb) != hipSuccess) {
    mpet_set_error("mpet_solver_create: HIP setup failed");
    mpet_solver_destroy(S);
    return ERR_HIP;
  }
  if (hipHostMalloc((void**)&S->scal_host, sizeof(double), hipHostMallocDefault) != hipSuccess) {
    mpet_set_error("mpet_solver_create: pinned allocation failed");
    mpet_solver_destroy(S);
    return ERR_HIP;
  }
  S->host_allocs.push_back(S->scal_host);
  if (hipStreamCreateWithFlags(&S->graph_stream, hipStreamNonBlocking) != hipSuccess) {
    mpet_set_error("mpet_solver_create: HIP setup failed");
    mpet_solver_destroy(S);
    return ERR_HIP;
  }
  *handle = S;
  return OK;
}

extern "C" int mpet_solver_destroy(void* h) {
  MpetSolver* S = (MpetSolver*)h;
  if (!S) return OK;
  (void)hipDeviceSynchronize();
  if (S->graph32) (void)hipGraphExecDestroy(S->graph32);
  if (S->graph64) (void)hipGraphExecDestroy(S->graph64);
  if (S->graph_stream) (void)hipStreamDestroy(S->graph_stream);
  free_fork(S->pc64);
  free_fork(S->pc32);
  for (void* p : S->allocs)
    if (p) (void)hipFree(p);
  for (void* p : S->host_allocs)
    if (p) (void)hipHostFree(p);
  delete S;
  return OK;
}

// Solve A_elim x = b_elim for one time step.  On entry x holds the initial
// guess with the Dirichlet values already in place (x_D = g); b is the full
// un-lifted right-hand side.  On exit x is the solution.
extern "C" int mpet_solver_solve(void* h, const double* b, double* x, MpetSolveStats* stats,
                                 hipStream_t s) {
  MpetSolver* S = (MpetSolver*)h;
  if (!S || !b || !x) { mpet_set_error("mpet_solver_solve: invalid argument"); return ERR_INVALID; }
  MpetSolveStats st;
  memset(&st, 0, sizeof(st));
  long long n = S->n;
  int eb = (int)((n + NT - 1) / NT);
  int rc = OK;
  static const bool dbg = getenv("MPET_DEBUG") != nullptr;
  if (S->opt.mode == 0 || S->opt.mode == 2) {
    // FP64 MINRES (mode 2: FP32 preconditioner): r0 = Pi (b - A x0); MINRES
    // until ||r||_{M^-1} <= rtol ||b_elim||_{M^-1}
    double bn = 0.0;
    rc = bnorm_M(S, S->m64, S->pc64, b, x, s, &bn);
    if (rc) return rc;
    op_residual64(S, b, x, S->m64.b, s);
    int it = 0, conv = 0;
    double beta1;
    rc = minres_start(S, S->m64, S->pc64, -S->opt.rtol * bn, S->opt.maxit, s, &beta1);
    if (!rc)
      rc = minres_continue(S, S->m64, S->pc64, S->csr_rp != nullptr && S->opt.use_csr,
                           S->opt.rtol, S->opt.maxit, S->opt.check_every,
                           S->opt.use_graphs ? &S->graph64 : nullptr, s, &it, &conv);
    if (dbg) fprintf(stderr, "[mpet] fp64 ||b||_M^-1 %.6e  ||r0|| %.6e  its %d\n", bn, beta1, it);
    hipLaunchKernelGGL(add_f64_kernel, dim3((int)((S->nv + NT - 1) / NT)), dim3(NT), 0, s,
                       S->nv, S->m64.x, x);
    st.inner_iterations = it;
    st.outer_iterations = 1;
    st.converged = conv > 0;
    st.residual_norm = bn > 0 ? S->m64.st_host->phibar / bn : 0.0;
  } else {
    // mixed: FP32 MINRES corrections inside FP64 refinement.  Convergence is
    // measured like the FP64 path: beta_k = ||r_k||_{M^-1} (r_k = Pi (b - A
    // x_k) in FP64, M the FP32 preconditioner; each inner MINRES computes it
    // at its start) must fall below outer_tol * ||b_elim||_{M^-1}.
    double bn = 0.0;
    rc = bnorm_M(S, S->m32, S->pc32, b, x, s, &bn);
    if (rc) return rc;
    op_residual64(S, b, x, S->r64, s);
    int total = 0, outer = 0;
    double beta0 = -1.0, beta = 0.0;
    bool conv_outer = false;
    for (;;) {
      hipLaunchKernelGGL(f64_to_f32_kernel, dim3(eb), dim3(NT), 0, s, n, S->r64, S->m32.b);
      rc = minres_start(S, S->m32, S->pc32, S->opt.inner_rtol, S->opt.maxit, s, &beta);
      if (rc) break;
      if (beta0 < 0) beta0 = beta;
      if (dbg) fprintf(stderr, "[mpet] outer %d  ||r||_M^-1 = %.6e  (||b||_M^-1 %.6e)\n", outer, beta, bn);
      if (!(beta == beta)) { rc = ERR_NAN; break; }
      if (beta <= S->opt.outer_tol * bn || beta == 0.0) { conv_outer = true; break; }
      if (outer >= S->opt.max_outer) break;
      // Inner target: inner_rtol relative, but no deeper than needed to reach
      // the outer target (with margin): ||r||_{M^-1} <= INNER_MARGIN * outer_tol
      // * ||b_I||_{M^-1}.  Avoids overshooting the outer tolerance by up to
      // inner_rtol in the last refinement step.
      {
        const double tol = fmax(S->opt.inner_rtol * beta, INNER_MARGIN * S->opt.outer_tol * bn);
        if (tol != S->m32.st_host->tol) {
          S->m32.st_host->tol = tol;
          HIPCHK(hipMemcpyAsync(S->m32.st, S->m32.st_host, sizeof(MinresState),
                                hipMemcpyHostToDevice, s));
        }
      }
      int it = 0, conv = 0;
      rc = minres_continue(S, S->m32, S->pc32, S->csr_only, S->opt.inner_rtol, S->opt.maxit,
                           S->opt.check_every, S->opt.use_graphs ? &S->graph32 : nullptr, s,
                           &it, &conv);
      total += it;
      if (rc) break;
      hipLaunchKernelGGL(add_f32_to_f64_kernel, dim3(eb), dim3(NT), 0, s, n, S->m32.x, x);
      op_residual64(S, b, x, S->r64, s);
      ++outer;
    }
    st.inner_iterations = total;
    st.outer_iterations = outer;
    st.residual_norm = bn > 0 ? beta / bn : 0.0;
    st.converged = conv_outer;
  }
  if (hipGetLastError() != hipSuccess) { mpet_set_error("kernel launch failed"); return ERR_HIP; }
  if (stats) *stats = st;
  if (rc) return rc;
  if (!st.converged) {
    mpet_set_error("MINRES did not converge within maxit");
    return ERR_NOT_CONVERGED;
  }
  return OK;
}

// One level SpMM as the V-cycle launches it (the autotuner's probe).
template <typename T, int NC>
static int level_spmm_t(int mode, int nrows, const int* rp, const int* ci, const void* v,
                        const void* x, const void* b, void* y, const void* dinv, void* d,
                        double cd, double cr, const uint8_t* mask, int lpr, const void* v16,
                        const int16_t* dci, double vscale, hipStream_t s) {
  LevelDev<T> L;
  L.n = nrows;
  L.rp = rp;
  L.ci = ci;
  L.v = (const T*)v;
  L.dinv = (const T*)dinv;
  L.mask = mask;
  L.lpr = lpr;
  L.v16 = (const __half*)v16;
  L.dci = dci;
  L.vscale = vscale;
  Skip sk;
  switch (mode) {
    case 0: spmm<T, NC, 0>(L, 0, (const T*)x, nullptr, (T*)y, nullptr, 0, 0, sk, s); break;
    case 1: spmm<T, NC, 1>(L, 0, (const T*)x, (const T*)b, (T*)y, nullptr, 0, 0, sk, s); break;
    case 3: spmm<T, NC, 3>(L, 0, (const T*)x, (const T*)b, (T*)y, (T*)d, (T)cd, (T)cr, sk, s,
                           nullptr, (T*)d);
            break;
    default: return ERR_INVALID;
  }
  return OK;
}

extern "C" int mpet_level_spmm(int dtype, int nc, int mode, int nrows, const int* rp,
                               const int* ci, const void* v, const void* x, const void* b,
                               void* y, const void* dinv, void* d, double cd, double cr,
                               const uint8_t* mask, int lpr, const void* v16,
                               const int16_t* dci, double vscale, hipStream_t s) {
  if ((dtype != 0 && dtype != 1) || (nc != 1 && nc != 3) || nrows < 0 || !rp || !ci ||
      !(v || v16) || !x || !y || !lpr_ok(lpr) || lpr == 0 || (v16 && (dtype != 1 || !dci)) ||
      ((mode == 1 || mode == 3) && !b) || (mode == 3 && (!d || !dinv))) {
    mpet_set_error("mpet_level_spmm: invalid argument");
    return ERR_INVALID;
  }
  int rc;
  if (dtype == 0)
    rc = nc == 3 ? level_spmm_t<double, 3>(mode, nrows, rp, ci, v, x, b, y, dinv, d, cd, cr,
                                           mask, lpr, nullptr, nullptr, 1.0, s)
                 : level_spmm_t<double, 1>(mode, nrows, rp, ci, v, x, b, y, dinv, d, cd, cr,
                                           mask, lpr, nullptr, nullptr, 1.0, s);
  else
    rc = nc == 3 ? level_spmm_t<float, 3>(mode, nrows, rp, ci, v, x, b, y, dinv, d, cd, cr,
                                          mask, lpr, v16, dci, vscale, s)
                 : level_spmm_t<float, 1>(mode, nrows, rp, ci, v, x, b, y, dinv, d, cd, cr,
                                          mask, lpr, v16, dci, vscale, s);
  if (rc) {
    mpet_set_error("mpet_level_spmm: invalid mode");
    return rc;
  }
  return hipGetLastError() == hipSuccess ? OK : ERR_HIP;
}

// Launch shapes from the autotuner; the captured iteration graphs are
// dropped (they hold the old launch configurations).
extern "C" int mpet_solver_set_tuning(void* h, int dtype, const MpetTuning* t) {
  MpetSolver* S = (MpetSolver*)h;
  const bool lanes_ok = t && (t->lanes_u == 0 || t->lanes_u == 8 || t->lanes_u == 16 ||
                              t->lanes_u == 32) &&
                        (t->lanes_p == 0 || t->lanes_p == 8 || t->lanes_p == 16 ||
                         t->lanes_p == 32);
  if (!S || (dtype != 0 && dtype != 1) || !tuning_ok(t) || !lanes_ok) {
    mpet_set_error("mpet_solver_set_tuning: invalid argument");
    return ERR_INVALID;
  }
  HIPCHK(hipDeviceSynchronize());
  if (dtype == 1) {
    if (t->lanes_u) S->op32.lanes_u = t->lanes_u;
    if (t->lanes_p) S->op32.lanes_p = t->lanes_p;
    if (t->xcd_chunk >= 0) S->op32.xcd_chunk = t->xcd_chunk;
    if (S->has_pc32) set_precond_tuning(S->pc32, t);
  } else {
    if (t->lanes_u) S->op64.lanes_u = t->lanes_u;
    if (t->lanes_p) S->op64.lanes_p = t->lanes_p;
    if (t->xcd_chunk >= 0) S->op64.xcd_chunk = t->xcd_chunk;
    if (S->has_pc64) set_precond_tuning(S->pc64, t);
  }
  // both graphs: the FP64 one of mode 2 holds the FP32 preconditioner
  if (S->graph32) HIPCHK(hipGraphExecDestroy(S->graph32));
  if (S->graph64) HIPCHK(hipGraphExecDestroy(S->graph64));
  S->graph32 = S->graph64 = nullptr;
  return OK;
}

// Inner FP32 MINRES on the cell-wise operator (op != NULL) or back on the
// structured FP32 operator (op == NULL); the captured graphs are dropped.
extern "C" int mpet_solver_set_cell_operator(void* h, const MpetCellOpDesc* op) {
  MpetSolver* S = (MpetSolver*)h;
  if (!S || !S->has_op32 || S->csr_only || S->mb) {
    mpet_set_error("mpet_solver_set_cell_operator: needs a mixed-precision solver on the "
                   "structured operator");
    return ERR_INVALID;
  }
  HIPCHK(hipDeviceSynchronize());
  if (S->graph32) HIPCHK(hipGraphExecDestroy(S->graph32));
  if (S->graph64) HIPCHK(hipGraphExecDestroy(S->graph64));
  S->graph32 = S->graph64 = nullptr;
  if (!op) {
    S->has_cw = false;
    return OK;
  }
  if (op->nP2 != S->op32.nP2 || op->nP1 != S->op32.nP1 || op->A != S->op32.A ||
      op->nnz11 != S->op32.nnz11) {
    mpet_set_error("mpet_solver_set_cell_operator: sizes do not match the solver's operator");
    return ERR_INVALID;
  }
  if (op->npart2 <= 0 || op->npart1 <= 0) {
    mpet_set_error("invalid cell-wise operator description");
    return ERR_INVALID;
  }
  const long long need = cw_work_floats(op);
  if (S->cw_work_n < need) {
    int rc = alloc_dev(S->allocs, &S->cw_work, (size_t)need);
    if (rc) return rc;
    S->cw_work_n = need;
  }
  CellOp c;
  int rc = cw_from_desc<float>(op, S->cw_work, c);
  if (rc) return rc;
  S->cw = c;
  S->has_cw = true;
  return OK;
}

// FP64 residuals through the cell-wise FP64 operator (op != NULL) or back
// through the structured FP64 operator (op == NULL).  The residuals run
// outside the captured MINRES graphs, so no graph is dropped.
extern "C" int mpet_solver_set_cell_operator64(void* h, const MpetCellOpDesc* op) {
  MpetSolver* S = (MpetSolver*)h;
  if (!S || S->csr_only || S->mb) {
    mpet_set_error("mpet_solver_set_cell_operator64: needs a solver on the structured "
                   "operator");
    return ERR_INVALID;
  }
  HIPCHK(hipDeviceSynchronize());
  if (!op) {
    S->has_cw64 = false;
    return OK;
  }
  if (!op->fp64 || op->nP2 != S->op64.nP2 || op->nP1 != S->op64.nP1 || op->A != S->op64.A ||
      op->nnz11 != S->op64.nnz11 || op->npart2 <= 0 || op->npart1 <= 0) {
    mpet_set_error("mpet_solver_set_cell_operator64: not an FP64 description of the "
                   "solver's operator");
    return ERR_INVALID;
  }
  const long long need = cw_work_floats(op);
  if (S->cw64_work_n < need) {
    int rc = alloc_dev(S->allocs, &S->cw64_work, (size_t)need);
    if (rc) return rc;
    S->cw64_work_n = need;
  }
  CellOpT<double> c;
  int rc = cw_from_desc<double>(op, S->cw64_work, c);
  if (rc) return rc;
  S->cw64 = c;
  S->has_cw64 = true;
  return OK;
}

// z = P^-1 r with the FP64 (dtype 0) or FP32 (1) preconditioner.
extern "C" int mpet_precond_apply(void* h, int dtype, const void* r, void* z, hipStream_t s) {
  MpetSolver* S = (MpetSolver*)h;
  if (!S) return ERR_INVALID;
  if (dtype == 0 && S->has_pc64) {
    precond_apply(S->pc64, (const double*)r, (double*)z, (const int*)nullptr, s);
  } else if (dtype == 1 && S->has_pc32) {
    precond_apply(S->pc32, (const float*)r, (float*)z, (const int*)nullptr, s);
  } else {
    mpet_set_error("mpet_precond_apply: preconditioner of that dtype not built");
    return ERR_INVALID;
  }
  return hipGetLastError() == hipSuccess ? OK : ERR_HIP;
}
