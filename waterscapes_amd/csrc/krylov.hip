// Preconditioned MINRES with the block-diagonal MPET preconditioner, FP64
// or FP32 inner iterations inside FP64 iterative refinement.
//
// Replaces PETScKrylovSolver("minres", "hypre_amg") + solve
// (mpet/mpettotalpressuresolver.py:448, :492-494).
//
// Preconditioner (block diagonal, SPD, fixed linear operator):
//   u   : one V-cycle of the hierarchy built on the P2 Laplacian (mu scaled),
//         applied to the 3 displacement components at once (node-interleaved
//         vectors: every matrix value is read once for three columns);
//         Chebyshev smoothing (degree k, interval [rho/ratio, safety*rho] of
//         D^-1 A), dense inverse on the coarsest level;
//   p_0 : Chebyshev-Jacobi on the P1 mass;
//   p_f : Chebyshev-Jacobi on (Pd_f M1 + Ps_f St_f), Dirichlet rows masked.
//
// MINRES (Paige & Saunders 1975) keeps its scalar recurrences on the device:
// every dot product ends in a "last block" reduction (per-block partials,
// fixed summation order -> deterministic) whose final block updates the
// state; no host round trip per iteration.  Kernels return immediately once
// the state says the solve is done, so chunks of iterations can be enqueued
// (or replayed from a captured hipGraph) without a host check in between.
//
// Dirichlet invariant: all Krylov vectors have zero Dirichlet components;
// operators zero their Dirichlet rows (see spmv.hip).
#include "mpet_ops.h"
#include <vector>
#include <math.h>
#include <stdio.h>
#include <string.h>

namespace mpet {

constexpr int NT = 256;             // threads per block everywhere
constexpr int RED_GRID = 2048;      // max blocks of reducing kernels
constexpr int MAXLEV = 16;
constexpr int MAXDEG = 16;

// ---------------------------------------------------------------------------
// block / grid reductions
// ---------------------------------------------------------------------------
__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r += sh[i];
  return r;   // valid in thread 0
}

struct Reducer {
  double* partial;        // >= RED_GRID * 2 doubles
  unsigned int* counter;  // zero between kernels (reset by the last block)
};

// Sum v over the grid.  Returns true in every thread of the last block to
// finish, with tot[0..1] valid there.
template <int NV>
__device__ bool grid_reduce(const double* v, Reducer rd, double* tot) {
  __shared__ double sh[16];
  __shared__ int last;
  __shared__ double res[NV];
  double bs[NV];
  for (int j = 0; j < NV; ++j) bs[j] = block_sum(v[j], sh);
  if (threadIdx.x == 0) {
    for (int j = 0; j < NV; ++j) rd.partial[(size_t)j * RED_GRID + blockIdx.x] = bs[j];
    __threadfence();
    unsigned t = atomicAdd(rd.counter, 1u);
    last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
  for (int j = 0; j < NV; ++j) {
    double s = 0.0;
    for (int i = threadIdx.x; i < (int)gridDim.x; i += blockDim.x)
      s += __hip_atomic_load(&rd.partial[(size_t)j * RED_GRID + i], __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
    s = block_sum(s, sh);
    if (threadIdx.x == 0) res[j] = s;
  }
  if (threadIdx.x == 0) __hip_atomic_store(rd.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  for (int j = 0; j < NV; ++j) tot[j] = res[j];
  return true;
}

static int red_grid(long long work_items) {
  long long b = (work_items + NT - 1) / NT;
  if (b < 1) b = 1;
  return (int)(b < RED_GRID ? b : RED_GRID);
}

// ---------------------------------------------------------------------------
// MINRES state
// ---------------------------------------------------------------------------
struct MinresState {
  double beta1, beta, oldb, dbar, epsln, phibar, cs, sn, phi, gamma, delta, oldeps, alfa;
  double tol;
  int it, done, conv, nan, maxit, pad;
};

// ---------------------------------------------------------------------------
// 3-column scalar-matrix kernels (AMG levels)
// MODE 0: y = A x        MODE 1: y = b - A x       MODE 2: y += A x
// MODE 3: Chebyshev step d = cd d + cr Dinv (b - A x); y = x + d
// ---------------------------------------------------------------------------
template <typename T, int LPR, int MODE>
__global__ void __launch_bounds__(NT)
spmm3_kernel(int nrows, const int* __restrict__ rp, const int* __restrict__ ci,
             const T* __restrict__ val, const T* __restrict__ x, const T* __restrict__ b,
             T* __restrict__ y, const T* __restrict__ dinv, T* __restrict__ d, T cd, T cr,
             const uint8_t* __restrict__ mask, const int* __restrict__ skip) {
  if (skip && *skip) return;
  long long gid = (long long)blockIdx.x * NT + threadIdx.x;
  int r = (int)(gid / LPR);
  int lane = (int)(gid % LPR);
  if (r >= nrows) return;
  T s0 = 0, s1 = 0, s2 = 0;
  for (int k = rp[r] + lane; k < rp[r + 1]; k += LPR) {
    const T a = val[k];
    const long long c = 3ll * ci[k];
    s0 += a * x[c];
    s1 += a * x[c + 1];
    s2 += a * x[c + 2];
  }
  s0 = group_sum<LPR>(s0);
  s1 = group_sum<LPR>(s1);
  s2 = group_sum<LPR>(s2);
  if (lane != 0) return;
  const long long o = 3ll * r;
  const bool m = mask && mask[r];
  if (MODE == 0) {
    y[o] = m ? (T)0 : s0; y[o + 1] = m ? (T)0 : s1; y[o + 2] = m ? (T)0 : s2;
  } else if (MODE == 1) {
    y[o] = m ? (T)0 : b[o] - s0;
    y[o + 1] = m ? (T)0 : b[o + 1] - s1;
    y[o + 2] = m ? (T)0 : b[o + 2] - s2;
  } else if (MODE == 2) {
    if (!m) { y[o] += s0; y[o + 1] += s1; y[o + 2] += s2; }
  } else {
    const T di = dinv[r];
    T d0 = cd * d[o] + cr * di * (b[o] - s0);
    T d1 = cd * d[o + 1] + cr * di * (b[o + 1] - s1);
    T d2 = cd * d[o + 2] + cr * di * (b[o + 2] - s2);
    if (m) { d0 = d1 = d2 = 0; }
    d[o] = d0; d[o + 1] = d1; d[o + 2] = d2;
    y[o] = m ? (T)0 : x[o] + d0;
    y[o + 1] = m ? (T)0 : x[o + 1] + d1;
    y[o + 2] = m ? (T)0 : x[o + 2] + d2;
  }
}

// First Chebyshev step from x0 = 0: d = y = cr * Dinv * (scale * b).
template <typename T>
__global__ void __launch_bounds__(NT)
cheb0_kernel(int nrows, const T* __restrict__ b, T scale, const T* __restrict__ dinv, T cr,
             T* __restrict__ d, T* __restrict__ y, T* __restrict__ bcopy,
             const uint8_t* __restrict__ mask, const int* __restrict__ skip) {
  if (skip && *skip) return;
  long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i >= 3ll * nrows) return;
  int r = (int)(i / 3);
  T bv = scale * b[i];
  if (bcopy) bcopy[i] = bv;
  T v = (mask && mask[r]) ? (T)0 : cr * dinv[r] * bv;
  d[i] = v;
  y[i] = v;
}

// Dense coarsest solve x = Cinv b (3 columns).
template <typename T>
__global__ void __launch_bounds__(NT)
coarse_kernel(int n, const T* __restrict__ cinv, const T* __restrict__ b, T* __restrict__ x,
              const int* __restrict__ skip) {
  if (skip && *skip) return;
  int i = blockIdx.x * NT + threadIdx.x;
  if (i >= 3 * n) return;
  int r = i / 3, c = i % 3;
  T s = 0;
  for (int k = 0; k < n; ++k) s += cinv[(long long)r * n + k] * b[3 * k + c];
  x[i] = s;
}

// ---------------------------------------------------------------------------
// pressure-block Chebyshev (all F fields of a P1 node in one group)
// ---------------------------------------------------------------------------
struct PCoefStep {
  double cd[MPET_MAXF], cr[MPET_MAXF];
  double Pd[MPET_MAXF], Ps[MPET_MAXF];
};

template <typename T, int F>
__global__ void __launch_bounds__(NT)
pcheb0_kernel(int nP1, const T* __restrict__ b, const T* __restrict__ dinv, PCoefStep c,
              T* __restrict__ d, T* __restrict__ y, const uint8_t* __restrict__ mask,
              const int* __restrict__ skip) {
  if (skip && *skip) return;
  long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i >= (long long)F * nP1) return;
  int f = (int)(i % F);
  T v = (mask && mask[i]) ? (T)0 : (T)c.cr[f] * dinv[i] * b[i];
  d[i] = v;
  y[i] = v;
}

constexpr int PG = 8;   // lanes per P1 row in the pressure kernels

template <typename T, int F>
__global__ void __launch_bounds__(NT)
pcheb_kernel(int nP1, const int* __restrict__ rp, const int* __restrict__ ci,
             const T* __restrict__ M1, const T* __restrict__ St, long long nnz11,
             const T* __restrict__ b, const T* __restrict__ x, const T* __restrict__ dinv,
             PCoefStep c, T* __restrict__ d, T* __restrict__ y,
             const uint8_t* __restrict__ mask, const int* __restrict__ skip) {
  if (skip && *skip) return;
  long long gid = (long long)blockIdx.x * NT + threadIdx.x;
  int r = (int)(gid / PG);
  int lane = (int)(gid % PG);
  if (r >= nP1) return;
  T m[F], s[F];
#pragma unroll
  for (int f = 0; f < F; ++f) { m[f] = 0; s[f] = 0; }
  for (int k = rp[r] + lane; k < rp[r + 1]; k += PG) {
    const T* xp = x + (long long)F * ci[k];
    const T mv = M1[k];
#pragma unroll
    for (int f = 0; f < F; ++f) {
      m[f] += mv * xp[f];
      if (f >= 1) s[f] += St[(f - 1) * nnz11 + k] * xp[f];
    }
  }
#pragma unroll
  for (int f = 0; f < F; ++f) {
    m[f] = group_sum<PG>(m[f]);
    s[f] = group_sum<PG>(s[f]);
  }
  if (lane != 0) return;
#pragma unroll
  for (int f = 0; f < F; ++f) {
    const long long o = (long long)F * r + f;
    T ax = (T)c.Pd[f] * m[f] + (f >= 1 ? (T)c.Ps[f] * s[f] : (T)0);
    T dd = (T)c.cd[f] * d[o] + (T)c.cr[f] * dinv[o] * (b[o] - ax);
    if (mask && mask[o]) dd = 0;
    d[o] = dd;
    y[o] = (mask && mask[o]) ? (T)0 : x[o] + dd;
  }
}

// ---------------------------------------------------------------------------
// MINRES kernels
// ---------------------------------------------------------------------------
// Epilogue of the fused operator (K1): y = s (A z) - c1 r1, v = s z,
// partial alfa += v . y.
template <typename T>
struct EpiK1 {
  T* y; T* v; const T* z; const T* r1; const uint8_t* mask;
  __device__ __forceinline__ double operator()(long long i, T ax, T s, T c1) const {
    T yy = (mask && mask[i]) ? (T)0 : s * ax - c1 * r1[i];
    T vv = s * z[i];
    y[i] = yy;
    v[i] = vv;
    return (double)vv * (double)yy;
  }
};

template <typename T>
__device__ __forceinline__ void k1_scalars(const MinresState* st, T& s, T& c1) {
  s = (T)(1.0 / st->beta);
  c1 = st->it > 0 ? (T)(st->beta / st->oldb) : (T)0;
}

// K1 on the structured operator (grid-stride over 16-row tiles).
template <typename T, int F>
__global__ void __launch_bounds__(NT)
k1_op_kernel(MpetOpT<T> op, const T* __restrict__ z, EpiK1<T> epi, MinresState* st,
             Reducer rd) {
  if (st->done) return;
  constexpr int G = 16, ROWS = NT / G;
  T s, c1;
  k1_scalars(st, s, c1);
  const int lane = threadIdx.x % G;
  const long long nu = 3ll * op.nP2;
  const int ntu = (op.nP2 + ROWS - 1) / ROWS, ntp = (op.nP1 + ROWS - 1) / ROWS;
  double acc = 0.0;
  for (int tile = blockIdx.x; tile < ntu + ntp; tile += gridDim.x) {
    if (tile < ntu) {
      int a = tile * ROWS + threadIdx.x / G;
      if (a < op.nP2) {
        T y0 = 0, y1 = 0, y2 = 0;
        const long long n22 = op.nnz22;
        for (int k = op.rp22[a] + lane; k < op.rp22[a + 1]; k += G) {
          const int cb = 3 * op.ci22[k];
          const T x0 = z[cb], x1 = z[cb + 1], x2 = z[cb + 2];
          y0 += op.Kuu[k] * x0 + op.Kuu[n22 + k] * x1 + op.Kuu[2 * n22 + k] * x2;
          y1 += op.Kuu[3 * n22 + k] * x0 + op.Kuu[4 * n22 + k] * x1 + op.Kuu[5 * n22 + k] * x2;
          y2 += op.Kuu[6 * n22 + k] * x0 + op.Kuu[7 * n22 + k] * x1 + op.Kuu[8 * n22 + k] * x2;
        }
        const long long n21 = op.nnz21;
        for (int k = op.rp21[a] + lane; k < op.rp21[a + 1]; k += G) {
          const T p0 = z[nu + (long long)F * op.ci21[k]];
          y0 += op.Dv[k] * p0;
          y1 += op.Dv[n21 + k] * p0;
          y2 += op.Dv[2 * n21 + k] * p0;
        }
        y0 = group_sum<G>(y0);
        y1 = group_sum<G>(y1);
        y2 = group_sum<G>(y2);
        if (lane == 0) {
          const long long o = 3ll * a;
          acc += epi(o, y0, s, c1) + epi(o + 1, y1, s, c1) + epi(o + 2, y2, s, c1);
        }
      }
    } else {
      int c = (tile - ntu) * ROWS + threadIdx.x / G;
      if (c < op.nP1) {
        T yd = 0;
        const long long n12 = op.nnz12;
        for (int k = op.rp12[c] + lane; k < op.rp12[c + 1]; k += G) {
          const int cb = 3 * op.ci12[k];
          yd += op.Dt[k] * z[cb] + op.Dt[n12 + k] * z[cb + 1] + op.Dt[2 * n12 + k] * z[cb + 2];
        }
        T m[F], sv[F];
#pragma unroll
        for (int g = 0; g < F; ++g) { m[g] = 0; sv[g] = 0; }
        const long long n11 = op.nnz11;
        for (int k = op.rp11[c] + lane; k < op.rp11[c + 1]; k += G) {
          const T* xp = z + nu + (long long)F * op.ci11[k];
          const T mv = op.M1[k];
#pragma unroll
          for (int g = 0; g < F; ++g) {
            m[g] += mv * xp[g];
            if (g >= 1) sv[g] += op.St[(g - 1) * n11 + k] * xp[g];
          }
        }
        yd = group_sum<G>(yd);
#pragma unroll
        for (int g = 0; g < F; ++g) {
          m[g] = group_sum<G>(m[g]);
          sv[g] = group_sum<G>(sv[g]);
        }
        if (lane == 0) {
          const long long o = nu + (long long)F * c;
#pragma unroll
          for (int f = 0; f < F; ++f) {
            T v = (f == 0) ? yd : (T)op.stiff[f] * sv[f];
#pragma unroll
            for (int g = 0; g < F; ++g) v += (T)op.C[f * MPET_MAXF + g] * m[g];
            acc += epi(o + f, v, s, c1);
          }
        }
      }
    }
  }
  double tot;
  if (grid_reduce<1>(&acc, rd, &tot) && threadIdx.x == 0) st->alfa = tot;
}

// K1 on a scalar CSR matrix (plain FP64 path), LPR lanes per row.
template <typename T, int LPR>
__global__ void __launch_bounds__(NT)
k1_csr_kernel(long long nrows, const long long* __restrict__ rp, const int* __restrict__ ci,
              const T* __restrict__ val, const T* __restrict__ z, EpiK1<T> epi,
              MinresState* st, Reducer rd) {
  if (st->done) return;
  T s, c1;
  k1_scalars(st, s, c1);
  const int lane = threadIdx.x % LPR;
  constexpr int ROWS = NT / LPR;
  const long long ntiles = (nrows + ROWS - 1) / ROWS;
  double acc = 0.0;
  for (long long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    long long r = tile * ROWS + threadIdx.x / LPR;
    if (r < nrows) {
      T sum = 0;
      for (long long k = rp[r] + lane; k < rp[r + 1]; k += LPR) sum += val[k] * z[ci[k]];
      sum = group_sum<LPR>(sum);
      if (lane == 0) acc += epi(r, sum, s, c1);
    }
  }
  double tot;
  if (grid_reduce<1>(&acc, rd, &tot) && threadIdx.x == 0) st->alfa = tot;
}

// K2: r2_new = y - (alfa/beta) r2_old (written over r1); the u part scaled
// by 1/mu is also the right-hand side of the V-cycle.
template <typename T>
__global__ void __launch_bounds__(NT)
k2_kernel(long long n, long long nu, const T* __restrict__ y, const T* __restrict__ r2o,
          T* __restrict__ r2n, const MinresState* st) {
  if (st->done) return;
  const T c = (T)(st->alfa / st->beta);
  long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i < n) r2n[i] = y[i] - c * r2o[i];
}

// K4: beta = sqrt(r2 . z) and the Givens update of the MINRES recurrences.
template <typename T>
__global__ void __launch_bounds__(NT)
k4_kernel(long long n, const T* __restrict__ r2, const T* __restrict__ z, MinresState* st,
          Reducer rd) {
  if (st->done) return;
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * NT + threadIdx.x; i < n; i += (long long)gridDim.x * NT)
    acc += (double)r2[i] * (double)z[i];
  double tot;
  if (grid_reduce<1>(&acc, rd, &tot) && threadIdx.x == 0) {
    MinresState s = *st;
    s.oldb = s.beta;
    s.beta = sqrt(fmax(tot, 0.0));
    s.oldeps = s.epsln;
    s.delta = s.cs * s.dbar + s.sn * s.alfa;
    double gbar = s.sn * s.dbar - s.cs * s.alfa;
    s.epsln = s.sn * s.beta;
    s.dbar = -s.cs * s.beta;
    s.gamma = fmax(hypot(gbar, s.beta), 1e-300);
    s.cs = gbar / s.gamma;
    s.sn = s.beta / s.gamma;
    s.phi = s.cs * s.phibar;
    s.phibar = s.sn * s.phibar;
    s.it += 1;
    bool bad = !(s.phibar == s.phibar) || !(s.gamma == s.gamma) || isinf(s.phibar);
    s.nan = bad ? 1 : 0;
    s.conv = (bad || s.phibar <= s.tol * s.beta1 || s.beta == 0.0 || s.it >= s.maxit) ? 1 : 0;
    *st = s;
  }
}

// K5: w = (v - oldeps w1 - delta w2) / gamma (written over w1); x += phi w.
template <typename T>
__global__ void __launch_bounds__(NT)
k5_kernel(long long n, const T* __restrict__ v, T* __restrict__ w1, const T* __restrict__ w2,
          T* __restrict__ x, MinresState* st, Reducer rd) {
  if (st->done) return;
  const T oe = (T)st->oldeps, de = (T)st->delta, ig = (T)(1.0 / st->gamma), ph = (T)st->phi;
  for (long long i = (long long)blockIdx.x * NT + threadIdx.x; i < n; i += (long long)gridDim.x * NT) {
    T w = (v[i] - oe * w1[i] - de * w2[i]) * ig;
    w1[i] = w;
    x[i] += ph * w;
  }
  double zero = 0.0, tot;
  if (grid_reduce<1>(&zero, rd, &tot) && threadIdx.x == 0) st->done = st->conv;
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

template <typename T>
__global__ void __launch_bounds__(NT)
minres_init_state_kernel(long long n, const T* __restrict__ b, const T* __restrict__ z,
                         MinresState* st, Reducer rd, double tol, int maxit) {
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * NT + threadIdx.x; i < n; i += (long long)gridDim.x * NT)
    acc += (double)b[i] * (double)z[i];
  double tot;
  if (grid_reduce<1>(&acc, rd, &tot) && threadIdx.x == 0) {
    MinresState s;
    memset(&s, 0, sizeof(s));
    s.beta1 = sqrt(fmax(tot, 0.0));
    s.beta = s.beta1;
    s.phibar = s.beta1;
    s.cs = -1.0;
    s.sn = 0.0;
    s.tol = tol;
    s.maxit = maxit;
    s.nan = !(s.beta1 == s.beta1) ? 1 : 0;
    s.done = (s.beta1 == 0.0 || s.nan) ? 1 : 0;
    s.conv = s.done;
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
// out[0] = sum a^2 (+ out[1] = sum b^2 if b): FP64 norms on the device.
__global__ void __launch_bounds__(NT)
norm2_kernel(long long n, const double* __restrict__ a, const double* __restrict__ b,
             double* out, Reducer rd) {
  double acc[2] = {0.0, 0.0};
  for (long long i = (long long)blockIdx.x * NT + threadIdx.x; i < n; i += (long long)gridDim.x * NT) {
    acc[0] += a[i] * a[i];
    if (b) acc[1] += b[i] * b[i];
  }
  double tot[2];
  if (grid_reduce<2>(acc, rd, tot) && threadIdx.x == 0) { out[0] = tot[0]; out[1] = tot[1]; }
}

}  // namespace mpet

// ===========================================================================
// host side
// ===========================================================================
using namespace mpet;

#define HIPCHK(expr)                                                        \
  do {                                                                      \
    hipError_t _e = (expr);                                                 \
    if (_e != hipSuccess) {                                                 \
      char _m[400];                                                         \
      snprintf(_m, sizeof(_m), "%s: %s", #expr, hipGetErrorString(_e));     \
      mpet_set_error(_m);                                                   \
      return ERR_HIP;                                                       \
    }                                                                       \
  } while (0)

static void chebyshev(double lmin, double lmax, int deg, double* cd, double* cr) {
  double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
  double sigma = theta / delta, rho = 1.0 / sigma;
  cd[0] = 0.0;
  cr[0] = 1.0 / theta;
  for (int i = 1; i < deg; ++i) {
    double rn = 1.0 / (2.0 * sigma - rho);
    cd[i] = rn * rho;
    cr[i] = 2.0 * rn / delta;
    rho = rn;
  }
}

template <typename T>
struct LevelDev {
  int n = 0, nc = 0;
  const int *rp = nullptr, *ci = nullptr;
  const T *v = nullptr, *dinv = nullptr;
  const int *Prp = nullptr, *Pci = nullptr, *Rrp = nullptr, *Rci = nullptr;
  const T *Pv = nullptr, *Rv = nullptr;
  const uint8_t* mask = nullptr;
  double cd[MAXDEG], cr[MAXDEG];
  int lpr = 16, lprP = 8, lprR = 16;
  T *B = nullptr, *X = nullptr, *R = nullptr, *E = nullptr, *S1 = nullptr, *S2 = nullptr,
    *D = nullptr;
};

template <typename T>
struct Precond {
  int nlev = 0, deg = 2, nc = 0;
  LevelDev<T> lev[MAXLEV];
  const T* cinv = nullptr;
  double inv_mu = 1.0;
  // pressure blocks
  int nP1 = 0, F = 0, pdeg = 0;
  long long nnz11 = 0;
  const int *rp11 = nullptr, *ci11 = nullptr;
  const T *M1 = nullptr, *St = nullptr, *dinvp = nullptr;
  const uint8_t* pmask = nullptr;
  PCoefStep pstep[MAXDEG];
  T *pS1 = nullptr, *pS2 = nullptr, *pD = nullptr;
  long long nu = 0;
  const int* skip = nullptr;
};

template <typename T, int MODE>
static void spmm3(const LevelDev<T>& L, int which, const T* x, const T* b, T* y, T* d, T cd,
                  T cr, const int* skip, hipStream_t s) {
  // which: 0 level operator, 1 prolongator (rows n), 2 restriction (rows nc)
  int nrows = which == 2 ? L.nc : L.n;
  const int* rp = which == 0 ? L.rp : (which == 1 ? L.Prp : L.Rrp);
  const int* ci = which == 0 ? L.ci : (which == 1 ? L.Pci : L.Rci);
  const T* v = which == 0 ? L.v : (which == 1 ? L.Pv : L.Rv);
  int lpr = which == 0 ? L.lpr : (which == 1 ? L.lprP : L.lprR);
  const uint8_t* mask = which == 2 ? nullptr : L.mask;
  long long threads = (long long)nrows * lpr;
  int blocks = (int)((threads + NT - 1) / NT);
  if (blocks == 0) return;
#define SPMM3_CASE(LP)                                                                      \
  case LP:                                                                                  \
    hipLaunchKernelGGL((spmm3_kernel<T, LP, MODE>), dim3(blocks), dim3(NT), 0, s, nrows, rp, \
                       ci, v, x, b, y, L.dinv, d, cd, cr, mask, skip);                      \
    break;
  switch (lpr) {
    SPMM3_CASE(4) SPMM3_CASE(8) SPMM3_CASE(16) SPMM3_CASE(32) SPMM3_CASE(64)
    default: break;
  }
#undef SPMM3_CASE
}

// k Chebyshev steps on level l; from zero (x0 == nullptr, b scaled by
// bscale and optionally copied to bcopy) or from x0; result in out.
template <typename T>
static void smooth(const Precond<T>& P, int l, const T* b, T bscale, T* bcopy, const T* x0,
                   T* out, hipStream_t s) {
  const LevelDev<T>& L = P.lev[l];
  const int k = P.deg;
  const T* cur = x0;
  int i = 0;
  T* bufs[2] = {L.S1, L.S2};
  int flip = 0;
  const T* bb = b;
  if (!x0) {
    T* dst = (k == 1) ? out : bufs[flip];
    int blocks = (int)((3ll * L.n + NT - 1) / NT);
    hipLaunchKernelGGL((cheb0_kernel<T>), dim3(blocks), dim3(NT), 0, s, L.n, b, bscale, L.dinv,
                       (T)L.cr[0], L.D, dst, bcopy, L.mask, P.skip);
    if (bcopy) bb = bcopy;
    cur = dst;
    i = 1;
    flip ^= 1;
  }
  for (; i < k; ++i) {
    T* dst = (i == k - 1) ? out : bufs[flip];
    if (dst == cur) dst = bufs[flip ^ 1];
    spmm3<T, 3>(L, 0, cur, bb, dst, L.D, (T)L.cd[i], (T)L.cr[i], P.skip, s);
    cur = dst;
    flip ^= 1;
  }
}

template <typename T>
static void vcycle(const Precond<T>& P, int l, const T* b, T bscale, T* out, hipStream_t s) {
  const LevelDev<T>& L = P.lev[l];
  if (l == P.nlev - 1) {
    int blocks = (3 * L.n + NT - 1) / NT;
    hipLaunchKernelGGL((coarse_kernel<T>), dim3(blocks), dim3(NT), 0, s, L.n, P.cinv, b, out,
                       P.skip);
    return;
  }
  // level 0 receives the unscaled MINRES vector: scale it once into L.B
  T* bcopy = (bscale != (T)1) ? L.B : nullptr;
  smooth<T>(P, l, b, bscale, bcopy, nullptr, L.X, s);
  const T* bb = bcopy ? bcopy : b;
  spmm3<T, 1>(L, 0, L.X, bb, L.R, nullptr, 0, 0, P.skip, s);
  const LevelDev<T>& C = P.lev[l + 1];
  spmm3<T, 0>(L, 2, L.R, nullptr, C.B, nullptr, 0, 0, P.skip, s);
  vcycle(P, l + 1, C.B, (T)1, C.E, s);
  spmm3<T, 2>(L, 1, C.E, nullptr, L.X, nullptr, 0, 0, P.skip, s);
  smooth<T>(P, l, bb, (T)1, nullptr, L.X, out, s);
}

template <typename T, int F>
static void pblock(const Precond<T>& P, const T* b, T* out, hipStream_t s) {
  const int k = P.pdeg;
  long long n = (long long)F * P.nP1;
  int eb = (int)((n + NT - 1) / NT);
  T* bufs[2] = {P.pS1, P.pS2};
  int flip = 0;
  T* dst = (k == 1) ? out : bufs[flip];
  hipLaunchKernelGGL((pcheb0_kernel<T, F>), dim3(eb), dim3(NT), 0, s, P.nP1, b, P.dinvp,
                     P.pstep[0], P.pD, dst, P.pmask, P.skip);
  const T* cur = dst;
  flip ^= 1;
  int sb = (int)(((long long)P.nP1 * PG + NT - 1) / NT);
  for (int i = 1; i < k; ++i) {
    dst = (i == k - 1) ? out : bufs[flip];
    hipLaunchKernelGGL((pcheb_kernel<T, F>), dim3(sb), dim3(NT), 0, s, P.nP1, P.rp11, P.ci11,
                       P.M1, P.St, P.nnz11, b, cur, P.dinvp, P.pstep[i], P.pD, dst, P.pmask,
                       P.skip);
    cur = dst;
    flip ^= 1;
  }
}

template <typename T>
static void precond_apply(const Precond<T>& P, const T* r, T* z, hipStream_t s) {
  vcycle(P, 0, r, (T)P.inv_mu, z, s);
  const T* rp_ = r + P.nu;
  T* zp = z + P.nu;
  switch (P.F) {
    case 2: pblock<T, 2>(P, rp_, zp, s); break;
    case 3: pblock<T, 3>(P, rp_, zp, s); break;
    case 4: pblock<T, 4>(P, rp_, zp, s); break;
    case 5: pblock<T, 5>(P, rp_, zp, s); break;
    case 6: pblock<T, 6>(P, rp_, zp, s); break;
    case 7: pblock<T, 7>(P, rp_, zp, s); break;
    case 8: pblock<T, 8>(P, rp_, zp, s); break;
    default: break;
  }
}

// --- MINRES workspace in precision T -------------------------------------
template <typename T>
struct Minres {
  long long n = 0, nu = 0;
  T *R[2] = {nullptr, nullptr}, *W[2] = {nullptr, nullptr};
  T *z = nullptr, *y = nullptr, *v = nullptr, *x = nullptr, *b = nullptr;
  MinresState* st = nullptr;
  MinresState* st_host = nullptr;   // pinned
  Reducer rd;
};

struct MpetSolver {
  MpetSolverOptions opt;
  long long n = 0, nu = 0;
  int F = 0;
  MpetOpT<double> op64;
  MpetOpT<float> op32;
  bool has_op32 = false;
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
  double* norms = nullptr;   // device [2]
  double* norms_host = nullptr;
  Reducer rd64;
  std::vector<void*> allocs;
  hipGraphExec_t graph32 = nullptr;
  hipGraphExec_t graph64 = nullptr;
  hipStream_t graph_stream = nullptr;
  const uint8_t* opt_mask64 = nullptr;
  double bnorm = 0.0;
};

template <typename T>
static int alloc_dev(MpetSolver* S, T** p, size_t count) {
  void* q = nullptr;
  if (hipMalloc(&q, sizeof(T) * (count ? count : 1)) != hipSuccess) {
    mpet_set_error("hipMalloc failed (out of device memory)");
    return ERR_OOM;
  }
  S->allocs.push_back(q);
  *p = (T*)q;
  return OK;
}

template <typename T>
static int build_precond(MpetSolver* S, const MpetPrecDesc* d, Precond<T>& P) {
  P.nlev = d->nlev;
  P.deg = d->degree;
  if (P.nlev < 1 || P.nlev > MAXLEV || P.deg < 1 || P.deg > MAXDEG || d->pdeg < 1 ||
      d->pdeg > MAXDEG) {
    mpet_set_error("invalid preconditioner description");
    return ERR_INVALID;
  }
  P.inv_mu = d->inv_mu;
  for (int l = 0; l < P.nlev; ++l) {
    const MpetAmgLevelDesc& ld = d->lev[l];
    LevelDev<T>& L = P.lev[l];
    L.n = ld.n;
    L.nc = ld.nc;
    L.rp = ld.A_rp; L.ci = ld.A_ci; L.v = (const T*)ld.A_v; L.dinv = (const T*)ld.dinv;
    L.Prp = ld.P_rp; L.Pci = ld.P_ci; L.Pv = (const T*)ld.P_v;
    L.Rrp = ld.R_rp; L.Rci = ld.R_ci; L.Rv = (const T*)ld.R_v;
    L.mask = l == 0 ? d->mask0 : nullptr;
    L.lpr = ld.lpr ? ld.lpr : 16;
    L.lprP = ld.lprP ? ld.lprP : 4;
    L.lprR = ld.lprR ? ld.lprR : 16;
    double lmax = ld.rho * d->eig_safety;
    chebyshev(lmax / d->eig_ratio, lmax, P.deg, L.cd, L.cr);
    size_t m = 3ull * L.n;
    int rc;
    if ((rc = alloc_dev(S, &L.B, m)) || (rc = alloc_dev(S, &L.X, m)) ||
        (rc = alloc_dev(S, &L.R, m)) || (rc = alloc_dev(S, &L.E, m)) ||
        (rc = alloc_dev(S, &L.S1, m)) || (rc = alloc_dev(S, &L.S2, m)) ||
        (rc = alloc_dev(S, &L.D, m)))
      return rc;
  }
  P.cinv = (const T*)d->coarse_inv;
  P.nP1 = d->nP1;
  P.F = d->F;
  P.nu = S->nu;
  P.rp11 = d->rp11; P.ci11 = d->ci11; P.nnz11 = d->nnz11;
  P.M1 = (const T*)d->M1; P.St = (const T*)d->St; P.dinvp = (const T*)d->dinvp;
  P.pmask = d->pmask;
  P.pdeg = d->pdeg;
  for (int i = 0; i < MAXDEG; ++i) memset(&P.pstep[i], 0, sizeof(PCoefStep));
  for (int f = 0; f < P.F; ++f) {
    double cd[MAXDEG], cr[MAXDEG];
    chebyshev(d->plmin[f], d->plmax[f], P.pdeg, cd, cr);
    for (int i = 0; i < P.pdeg; ++i) {
      P.pstep[i].cd[f] = cd[i];
      P.pstep[i].cr[f] = cr[i];
      P.pstep[i].Pd[f] = d->Pd[f];
      P.pstep[i].Ps[f] = d->Ps[f];
    }
  }
  size_t np = (size_t)P.F * P.nP1;
  int rc;
  if ((rc = alloc_dev(S, &P.pS1, np)) || (rc = alloc_dev(S, &P.pS2, np)) ||
      (rc = alloc_dev(S, &P.pD, np)))
    return rc;
  return OK;
}

template <typename T>
static int build_minres(MpetSolver* S, Minres<T>& M) {
  M.n = S->n;
  M.nu = S->nu;
  int rc;
  if ((rc = alloc_dev(S, &M.R[0], M.n)) || (rc = alloc_dev(S, &M.R[1], M.n)) ||
      (rc = alloc_dev(S, &M.W[0], M.n)) || (rc = alloc_dev(S, &M.W[1], M.n)) ||
      (rc = alloc_dev(S, &M.z, M.n)) || (rc = alloc_dev(S, &M.y, M.n)) ||
      (rc = alloc_dev(S, &M.v, M.n)) || (rc = alloc_dev(S, &M.x, M.n)) ||
      (rc = alloc_dev(S, &M.b, M.n)) || (rc = alloc_dev(S, &M.st, 1)) ||
      (rc = alloc_dev(S, &M.rd.partial, 2 * RED_GRID)) ||
      (rc = alloc_dev(S, &M.rd.counter, 1)))
    return rc;
  HIPCHK(hipMemset(M.rd.counter, 0, sizeof(unsigned)));
  HIPCHK(hipHostMalloc((void**)&M.st_host, sizeof(MinresState), hipHostMallocDefault));
  S->allocs.push_back(nullptr);   // marker: host allocs freed separately
  return OK;
}

// One MINRES iteration (index i selects the buffer rotation).
template <typename T>
static void minres_iter(MpetSolver* S, Minres<T>& M, const Precond<T>& P, bool csr, int i,
                        hipStream_t s) {
  T* Ra = M.R[i & 1];
  T* Rb = M.R[(i + 1) & 1];
  T* Wa = M.W[i & 1];
  T* Wb = M.W[(i + 1) & 1];
  EpiK1<T> epi{M.y, M.v, M.z, Ra, nullptr};
  long long n = M.n;
  int eb = (int)((n + NT - 1) / NT);
  if constexpr (sizeof(T) == 8) {
    if (csr) {
      epi.mask = S->opt_mask64;
      int lpr = S->csr_lpr;
      long long rows_per_block = NT / lpr;
      int g = (int)((S->csr_rows + rows_per_block - 1) / rows_per_block);
      if (g > RED_GRID) g = RED_GRID;
#define K1CSR(LP)                                                                              \
  case LP:                                                                                     \
    hipLaunchKernelGGL((k1_csr_kernel<double, LP>), dim3(g), dim3(NT), 0, s, S->csr_rows,      \
                       S->csr_rp, S->csr_ci, S->csr_v, M.z, epi, M.st, M.rd);                  \
    break;
      switch (lpr) { K1CSR(4) K1CSR(8) K1CSR(16) K1CSR(32) K1CSR(64) default: break; }
#undef K1CSR
    }
  }
  if (!csr) {
    const MpetOpT<T>* op;
    if constexpr (sizeof(T) == 8) op = &S->op64; else op = &S->op32;
    epi.mask = op->dmask;
    int ntiles = (op->nP2 + 15) / 16 + (op->nP1 + 15) / 16;
    int g = ntiles < RED_GRID ? ntiles : RED_GRID;
#define K1OP(FF)                                                                                 \
  case FF:                                                                                       \
    hipLaunchKernelGGL((k1_op_kernel<T, FF>), dim3(g), dim3(NT), 0, s, *op, M.z, epi, M.st, M.rd); \
    break;
    switch (S->F) { K1OP(2) K1OP(3) K1OP(4) K1OP(5) K1OP(6) K1OP(7) K1OP(8) default: break; }
#undef K1OP
  }
  hipLaunchKernelGGL((k2_kernel<T>), dim3(eb), dim3(NT), 0, s, n, M.nu, M.y, Rb, Ra, M.st);
  precond_apply(P, Ra, M.z, s);
  int rg = red_grid(n);
  hipLaunchKernelGGL((k4_kernel<T>), dim3(rg), dim3(NT), 0, s, n, Ra, M.z, M.st, M.rd);
  hipLaunchKernelGGL((k5_kernel<T>), dim3(rg), dim3(NT), 0, s, n, M.v, Wa, Wb, M.x, M.st, M.rd);
}

// MINRES on rhs M.b (zero initial guess) -> M.x.  Host checks the state
// every `check` iterations (chunks may be replayed from a captured graph).
template <typename T>
static int minres_solve(MpetSolver* S, Minres<T>& M, const Precond<T>& P, bool csr,
                        double rtol, int maxit, int check, hipGraphExec_t* graph,
                        hipStream_t s, int* iters, int* conv) {
  long long n = M.n;
  int eb = (int)((n + NT - 1) / NT);
  hipLaunchKernelGGL((minres_init_vec_kernel<T>), dim3(eb), dim3(NT), 0, s, n, M.b, M.R[0],
                     M.R[1], M.W[0], M.W[1], M.x);
  // the preconditioner must run before the state exists: skip is null here
  Precond<T> P0 = P;
  P0.skip = nullptr;
  precond_apply(P0, M.R[1], M.z, s);
  hipLaunchKernelGGL((minres_init_state_kernel<T>), dim3(red_grid(n)), dim3(NT), 0, s, n,
                     M.R[1], M.z, M.st, M.rd, rtol, maxit);
  if (check < 1) check = 1;
  if (check & 1) check += 1;     // buffer rotation has period 2
  int done_iters = 0;
  for (;;) {
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
    if (M.st_host->done || done_iters >= maxit + check) break;
  }
  *iters = M.st_host->it;
  *conv = M.st_host->nan ? -1 : (M.st_host->phibar <= rtol * M.st_host->beta1 ||
                                  M.st_host->beta == 0.0);
  return M.st_host->nan ? ERR_NAN : OK;
}

static void op_residual64(MpetSolver* S, const double* b, const double* x, double* r,
                          hipStream_t s) {
  if (S->opt.mode == 0 && S->csr_rp)
    launch_csr_spmv<double>(S->csr_rows, S->csr_rp, S->csr_ci, S->csr_v, x, b, S->opt_mask64,
                            r, S->csr_lpr, s);
  else
    launch_op_apply<double>(S->op64, x, b, r, s);
}

static int norms64(MpetSolver* S, const double* a, const double* b, double* out,
                   hipStream_t s) {
  hipLaunchKernelGGL(norm2_kernel, dim3(red_grid(S->n)), dim3(NT), 0, s, S->n, a, b, S->norms,
                     S->rd64);
  HIPCHK(hipMemcpyAsync(S->norms_host, S->norms, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  out[0] = sqrt(S->norms_host[0]);
  out[1] = sqrt(S->norms_host[1]);
  return OK;
}

template <typename T>
MpetOpT<T> mpet_make_op(const MpetOpDesc* d);

extern "C" int mpet_solver_create(const MpetOpDesc* op64, const MpetOpDesc* op32,
                                  const MpetCsrDesc* csr64, const MpetPrecDesc* prec64,
                                  const MpetPrecDesc* prec32, const MpetSolverOptions* opt,
                                  void** handle) {
  if (!op64 || !opt || !handle || (!prec64 && !prec32) || op64->A < 1 ||
      op64->A > MPET_MAXF - 1) {
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
  MpetSolver* S = new MpetSolver();
  S->opt = *opt;
  S->F = op64->A + 1;
  S->nu = 3ll * op64->nP2;
  S->n = S->nu + (long long)S->F * op64->nP1;
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
  int rc = OK;
  if (prec64 && (rc = build_precond(S, prec64, S->pc64))) { mpet_solver_destroy(S); return rc; }
  if (prec32 && (rc = build_precond(S, prec32, S->pc32))) { mpet_solver_destroy(S); return rc; }
  S->has_pc64 = prec64 != nullptr;
  S->has_pc32 = prec32 != nullptr;
  if (S->has_pc64 && (rc = build_minres(S, S->m64))) { mpet_solver_destroy(S); return rc; }
  if (S->has_pc32 && (rc = build_minres(S, S->m32))) { mpet_solver_destroy(S); return rc; }
  S->pc64.skip = S->has_pc64 ? &S->m64.st->done : nullptr;
  S->pc32.skip = S->has_pc32 ? &S->m32.st->done : nullptr;
  if ((rc = alloc_dev(S, &S->r64, S->n)) || (rc = alloc_dev(S, &S->norms, 2)) ||
      (rc = alloc_dev(S, &S->rd64.partial, 2 * RED_GRID)) ||
      (rc = alloc_dev(S, &S->rd64.counter, 1))) {
    mpet_solver_destroy(S);
    return rc;
  }
  if (hipMemset(S->rd64.counter, 0, sizeof(unsigned)) != hipSuccess ||
      hipHostMalloc((void**)&S->norms_host, 2 * sizeof(double), hipHostMallocDefault) !=
          hipSuccess ||
      hipStreamCreateWithFlags(&S->graph_stream, hipStreamNonBlocking) != hipSuccess) {
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
  for (void* p : S->allocs)
    if (p) (void)hipFree(p);
  if (S->m64.st_host) (void)hipHostFree(S->m64.st_host);
  if (S->m32.st_host) (void)hipHostFree(S->m32.st_host);
  if (S->norms_host) (void)hipHostFree(S->norms_host);
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
  if (S->opt.mode == 0) {
    // plain FP64: r0 = Pi (b - A x0); MINRES to rtol on the preconditioned norm
    op_residual64(S, b, x, S->m64.b, s);
    int it = 0, conv = 0;
    rc = minres_solve(S, S->m64, S->pc64, S->csr_rp != nullptr && S->opt.use_csr, S->opt.rtol,
                      S->opt.maxit, S->opt.check_every, S->opt.use_graphs ? &S->graph64 : nullptr,
                      s, &it, &conv);
    hipLaunchKernelGGL(add_f64_kernel, dim3(eb), dim3(NT), 0, s, n, S->m64.x, x);
    st.inner_iterations = it;
    st.outer_iterations = 1;
    st.converged = conv > 0;
    // true residual for reporting
    op_residual64(S, b, x, S->r64, s);
    double nr[2];
    norms64(S, S->r64, nullptr, nr, s);
    st.residual_norm = nr[0];
  } else {
    // mixed: FP32 MINRES corrections inside FP64 refinement on the true
    // residual r = Pi (b - A x); stop when ||r|| <= tol ||r_0|| ... relative
    // to ||b_elim|| which the caller passes in opt.bnorm or computes here.
    op_residual64(S, b, x, S->r64, s);
    double nr[2];
    norms64(S, S->r64, nullptr, nr, s);
    double bn = S->bnorm > 0 ? S->bnorm : nr[0];
    double target = S->opt.outer_tol * bn;
    int total = 0, outer = 0;
    while (nr[0] > target && outer < S->opt.max_outer) {
      hipLaunchKernelGGL(f64_to_f32_kernel, dim3(eb), dim3(NT), 0, s, n, S->r64, S->m32.b);
      int it = 0, conv = 0;
      rc = minres_solve(S, S->m32, S->pc32, false, S->opt.inner_rtol, S->opt.maxit,
                        S->opt.check_every, S->opt.use_graphs ? &S->graph32 : nullptr, s, &it,
                        &conv);
      total += it;
      if (rc) break;
      hipLaunchKernelGGL(add_f32_to_f64_kernel, dim3(eb), dim3(NT), 0, s, n, S->m32.x, x);
      op_residual64(S, b, x, S->r64, s);
      norms64(S, S->r64, nullptr, nr, s);
      ++outer;
      if (!(nr[0] == nr[0])) { rc = ERR_NAN; break; }
    }
    st.inner_iterations = total;
    st.outer_iterations = outer;
    st.residual_norm = nr[0];
    st.converged = nr[0] <= target;
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

extern "C" int mpet_solver_set_bnorm(void* h, double bnorm) {
  ((MpetSolver*)h)->bnorm = bnorm;
  return OK;
}

// z = P^-1 r with the FP64 (dtype 0) or FP32 (1) preconditioner.
extern "C" int mpet_precond_apply(void* h, int dtype, const void* r, void* z, hipStream_t s) {
  MpetSolver* S = (MpetSolver*)h;
  if (!S) return ERR_INVALID;
  if (dtype == 0 && S->has_pc64) {
    Precond<double> P = S->pc64;
    P.skip = nullptr;
    precond_apply(P, (const double*)r, (double*)z, s);
  } else if (dtype == 1 && S->has_pc32) {
    Precond<float> P = S->pc32;
    P.skip = nullptr;
    precond_apply(P, (const float*)r, (float*)z, s);
  } else {
    mpet_set_error("mpet_precond_apply: preconditioner of that dtype not built");
    return ERR_INVALID;
  }
  return hipGetLastError() == hipSuccess ? OK : ERR_HIP;
}
