// AMG V-cycle machinery shared by the single-instance solver (krylov.hip)
// and the batched solver (batch.hip): level SpMM kernels on node-interleaved
// NC-column vectors, Chebyshev smoothing, dense coarsest solves, and the
// block-diagonal MPET preconditioner built on them.
//
// Replaces the hypre BoomerAMG V-cycle behind PETSc's "hypre_amg"
// preconditioner of the MINRES solve (mpet/mpettotalpressuresolver.py:448).
#pragma once
#include "mpet_ops.h"
#include <hip/hip_fp16.h>
#include <vector>
#include <stdio.h>

namespace mpet {

constexpr int NT = 256;            // threads per block everywhere
constexpr int MAX_PARTIALS = 8192; // per-block partial sums of one reduction
constexpr int MAXLEV = 16;
constexpr int MAXDEG = 16;

// Work skipped by the level kernels: everything when *all is set; the rows
// of instance k when inst[k] is set (grp: row -> instance of the level).
struct Skip {
  const int* all = nullptr;
  const int* inst = nullptr;
};
__device__ __forceinline__ bool row_skipped(const Skip& sk, const int* __restrict__ grp, int r) {
  return sk.inst && grp && sk.inst[grp[r]];
}

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
// All lanes of an aligned group of G lanes get the group sum.
template <int G, typename T>
__device__ __forceinline__ T group_allsum(T v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block sum of v, valid in thread 0.
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double sh[NT / 64];
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < NT / 64; ++i) r += sh[i];
  return r;
}

// Fixed-order sum of p[lo, hi) by the threads of a block into four
// accumulators per thread (four loads in flight); per-thread value.
__device__ __forceinline__ double thread_sum(const double* __restrict__ p, int lo, int hi) {
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int i = lo + threadIdx.x;
  for (; i + 3 * NT < hi; i += 4 * NT) {
    a0 += p[i];
    a1 += p[i + NT];
    a2 += p[i + 2 * NT];
    a3 += p[i + 3 * NT];
  }
  for (; i < hi; i += NT) a0 += p[i];
  return (a0 + a1) + (a2 + a3);
}

// Fixed-order sum of n partials by one block; valid in thread 0.
__device__ __forceinline__ double sum_partials(const double* __restrict__ p, int n) {
  return block_sum(thread_sum(p, 0, n));
}

static int grid_for(long long items, int per_block, int cap) {
  long long b = (items + per_block - 1) / per_block;
  if (b < 1) b = 1;
  return (int)(b < cap ? b : cap);
}

// ---------------------------------------------------------------------------
// AMG level kernels on NC-column node-interleaved vectors (NC = 3 for the
// displacement block: one scalar matrix, three components; NC = 1 for the
// pressure blocks).  LPR lanes per row; after an xor all-reduce, lane c of
// the row group finishes component c.
// MODE 0: y = A x        MODE 1: y = b - A x       MODE 2: y += A x
// MODE 3: Chebyshev step d' = cd d + cr Dinv (b - A x); y = x + d'; d is
//         read only when cd != 0 (not by the first step from x), d' is
//         written to y2 when y2 is given (not by the last step of a
//         smoothing sequence; y2 may be d: each entry is read, then written,
//         by one thread)
// MODE 4: restriction entering the coarse level's smoother: y = A x, then
//         its first Chebyshev step from zero d = y2 = cr Dinv y (dinv and
//         cr of the coarse level)
// ---------------------------------------------------------------------------
template <typename T, int NC, int LPR, int MODE, bool COMPACT = false>
__global__ void __launch_bounds__(NT)
spmm_kernel(int nrows, const int* __restrict__ rp, const int* __restrict__ ci,
            const T* __restrict__ val, const T* __restrict__ x, const T* __restrict__ b,
            T* __restrict__ y, const T* __restrict__ dinv, T* d, T cd, T cr,
            const uint8_t* __restrict__ mask, Skip sk, const int* __restrict__ grp,
            T* y2, const __half* __restrict__ val16 = nullptr,
            const int16_t* __restrict__ dci = nullptr, T vscale = (T)1) {
  // COMPACT: values stored as half precision times 1/vscale, columns as
  // signed 16-bit offsets from the row (half the bytes per nonzero)
  static_assert(NC <= LPR, "one lane per component in the epilogue");
  if (sk.all && *sk.all) return;
  long long gid = (long long)blockIdx.x * NT + threadIdx.x;
  int r = (int)(gid / LPR);
  int lane = (int)(gid % LPR);
  if (r >= nrows || row_skipped(sk, grp, r)) return;   // whole row groups leave
  T acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = 0;
  const int kb = rp[r], ke = rp[r + 1];
  auto value = [&](int k) -> T {
    if constexpr (COMPACT) return (T)__half2float(val16[k]);
    else return val[k];
  };
  auto column = [&](int k) -> long long {
    if constexpr (COMPACT) return (long long)NC * (r + (int)dci[k]);
    else return (long long)NC * ci[k];
  };
  // the epilogue operands are loaded first (by every lane, at clamped
  // indices, so that no branch splits the loads) and their latency overlaps
  // the row loop
  const long long o = (long long)NC * r + (lane < NC ? lane : NC - 1);
  T pb = 0, px = 0, pd = 0, pdinv = 0;
  if constexpr (MODE == 1 || MODE == 3) pb = b[o];
  if constexpr (MODE == 2) px = y[o];
  if constexpr (MODE == 3) {
    px = x[o];
    if (cd != (T)0) pd = d[o];
    pdinv = dinv[r];
  }
  if constexpr (MODE == 4) pdinv = dinv[r];
  const int mrow = mask ? (int)mask[r] : 0;
  // U nonzeros per lane per trip: every index, value and gather load of a
  // trip is issued before the first use (a missing nonzero reloads the
  // row's last one and contributes 0).  A lane's nonzeros are summed in the
  // order k, k + LPR, k + 2 LPR, ... (the same order for any unrolling).
  constexpr int U = (NC == 3 && LPR <= 4) ? 8 : 4;
  for (int k = kb + lane; k < ke; k += U * LPR) {
    T a[U];
    long long cc[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const int kk = k + j * LPR;
      const int kc = kk < ke ? kk : ke - 1;
      const T v = value(kc);
      cc[j] = column(kc);
      a[j] = kk < ke ? v : (T)0;
    }
    T xv[U][NC];
#pragma unroll
    for (int j = 0; j < U; ++j) load_nc<T, NC>(x + cc[j], xv[j]);
#pragma unroll
    for (int j = 0; j < U; ++j)
#pragma unroll
      for (int c = 0; c < NC; ++c) acc[c] += a[j] * xv[j][c];
  }
  if constexpr (COMPACT) {
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] *= vscale;
  }
  T mine = 0;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    T v = group_allsum<LPR>(acc[c]);
    if (lane == c) mine = v;
  }
  if (lane >= NC) return;
  const bool m = mrow != 0;
  if (MODE == 0) {
    y[o] = m ? (T)0 : mine;
  } else if (MODE == 1) {
    y[o] = m ? (T)0 : pb - mine;
  } else if (MODE == 2) {
    if (!m) y[o] = px + mine;
  } else if (MODE == 4) {
    y[o] = mine;
    const T v = cr * pdinv * mine;
    d[o] = v;
    y2[o] = v;
  } else {
    T dd = m ? (T)0 : cd * pd + cr * pdinv * (pb - mine);
    if (y2) y2[o] = dd;
    y[o] = m ? (T)0 : px + dd;
  }
}

// The same SpMM modes with one thread per row doing all NC columns (rows of
// a few nonzeros, such as the prolongator from the P1 level: no lanes idle
// on short rows).  Nonzeros are summed in row order.
template <typename T, int NC, int MODE>
__global__ void __launch_bounds__(NT)
spmm_row_kernel(int nrows, const int* __restrict__ rp, const int* __restrict__ ci,
                const T* __restrict__ val, const T* __restrict__ x, const T* __restrict__ b,
                T* __restrict__ y, const T* __restrict__ dinv, T* d, T cd, T cr,
                const uint8_t* __restrict__ mask, Skip sk, const int* __restrict__ grp,
                T* y2) {
  if (sk.all && *sk.all) return;
  const int r = blockIdx.x * NT + threadIdx.x;
  if (r >= nrows || row_skipped(sk, grp, r)) return;
  T acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = 0;
  const int kb = rp[r], ke = rp[r + 1];
  int k = kb;
  for (; k + 1 < ke; k += 2) {
    const T va = val[k], vb = val[k + 1];
    const long long ca = (long long)NC * ci[k], cb = (long long)NC * ci[k + 1];
    T xa[NC], xb[NC];
    load_nc<T, NC>(x + ca, xa);
    load_nc<T, NC>(x + cb, xb);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      acc[c] += va * xa[c];
      acc[c] += vb * xb[c];
    }
  }
  if (k < ke) {
    const T va = val[k];
    T xa[NC];
    load_nc<T, NC>(x + (long long)NC * ci[k], xa);
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] += va * xa[c];
  }
  const bool m = mask ? mask[r] != 0 : false;
  const T di = (MODE == 3 || MODE == 4) ? dinv[r] : (T)0;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const long long o = (long long)NC * r + c;
    const T mine = acc[c];
    if (MODE == 0) {
      y[o] = m ? (T)0 : mine;
    } else if (MODE == 1) {
      y[o] = m ? (T)0 : b[o] - mine;
    } else if (MODE == 2) {
      if (!m) y[o] = y[o] + mine;
    } else if (MODE == 4) {
      y[o] = mine;
      const T v = cr * di * mine;
      d[o] = v;
      y2[o] = v;
    } else {
      const T dp = cd != (T)0 ? cd * d[o] : (T)0;
      const T dd = m ? (T)0 : dp + cr * di * (b[o] - mine);
      if (y2) y2[o] = dd;
      y[o] = m ? (T)0 : x[o] + dd;
    }
  }
}

// A MINRES vector update fused into the first kernels of a preconditioner
// apply that read its input (off when y is null): the input entry i is
// r[i] = y[i] - (alfa / beta) r2o[i], computed where it is first read and
// stored to r[i] for the later readers.
template <typename T>
struct K2In {
  const T* y = nullptr;
  const T* r2o = nullptr;
  T* r = nullptr;
  const double* alfa = nullptr;
  const double* beta = nullptr;
  __device__ __forceinline__ T make(long long i) const {
    const T c = (T)(*alfa / *beta);
    const T v = y[i] - c * r2o[i];
    r[i] = v;
    return v;
  }
};

// First Chebyshev step from x0 = 0: d = y = cr * Dinv * (scale * b) (d
// only when given); the scaled right-hand side is kept in bcopy when given;
// with k2 on, b is made by k2 (and b itself is not read).
template <typename T, int NC>
__global__ void __launch_bounds__(NT)
cheb0_kernel(int nrows, const T* b, T scale, const T* __restrict__ dinv, T cr,
             T* __restrict__ d, T* __restrict__ y, T* __restrict__ bcopy,
             const uint8_t* __restrict__ mask, Skip sk, const int* __restrict__ grp,
             K2In<T> k2 = K2In<T>()) {
  if (sk.all && *sk.all) return;
  long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i >= (long long)NC * nrows) return;
  int r = (int)(i / NC);
  if (row_skipped(sk, grp, r)) return;
  T bv = scale * (k2.y ? k2.make(i) : b[i]);
  if (bcopy) bcopy[i] = bv;
  T v = (mask && mask[r]) ? (T)0 : cr * dinv[r] * bv;
  if (d) d[i] = v;
  y[i] = v;
}

// Dense coarsest solve x = Cinv b (NC columns), one wave per row; each lane
// keeps four row entries (and their b values) in flight per trip.
template <typename T, int NC>
__global__ void __launch_bounds__(NT)
coarse_kernel(int n, const T* __restrict__ cinv, const T* __restrict__ b, T* __restrict__ x,
              const int* __restrict__ skip) {
  if (skip && *skip) return;
  const int r = blockIdx.x * (NT / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (r >= n) return;
  T acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = 0;
  const T* __restrict__ row = cinv + (long long)r * n;
  for (int k0 = lane; k0 < n; k0 += 256) {
    T a[4], bv[4][NC];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = k0 + 64 * j;
      const int kc = k < n ? k : n - 1;
      a[j] = k < n ? row[kc] : (T)0;
      load_nc<T, NC>(b + NC * kc, bv[j]);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int c = 0; c < NC; ++c) acc[c] += a[j] * bv[j][c];
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    T v = group_allsum<64>(acc[c]);
    if (lane == c) x[NC * r + c] = v;
  }
}

// Pressure vectors: node-interleaved [c][F] <-> field-major [f][c].
template <typename T>
__global__ void __launch_bounds__(NT)
p_gather_kernel(int nP1, int F, const T* src, T* __restrict__ dst,
                const int* __restrict__ skip, K2In<T> k2 = K2In<T>()) {
  if (skip && *skip) return;
  long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i >= (long long)F * nP1) return;
  int c = (int)(i / F), f = (int)(i % F);
  dst[(long long)f * nP1 + c] = k2.y ? k2.make(i) : src[i];
}
template <typename T>
__global__ void __launch_bounds__(NT)
p_scatter_kernel(int nP1, int F, const T* __restrict__ src, T* __restrict__ dst,
                 const int* __restrict__ skip) {
  if (skip && *skip) return;
  long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i >= (long long)F * nP1) return;
  int c = (int)(i / F), f = (int)(i % F);
  dst[i] = src[(long long)f * nP1 + c];
}

// Block-diagonal dense coarsest solve (many independent instances): row r
// lies in block blk[r] spanning rows off[b] .. off[b+1]-1, whose dense
// inverse is stored row-major at inv + ioff[b].
template <typename T, int NC>
__global__ void __launch_bounds__(NT)
coarse_blockdiag_kernel(int n, const int* __restrict__ blk, const int* __restrict__ off,
                        const int64_t* __restrict__ ioff, const T* __restrict__ inv,
                        const T* __restrict__ b, T* __restrict__ x, Skip sk,
                        const int* __restrict__ grp) {
  if (sk.all && *sk.all) return;
  const int r = blockIdx.x * (NT / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (r >= n || row_skipped(sk, grp, r)) return;
  const int bk = blk[r];
  const int o = off[bk], sz = off[bk + 1] - o;
  const T* row = inv + ioff[bk] + (long long)(r - o) * sz;
  T acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = 0;
  for (int k = lane; k < sz; k += 64) {
    const T a = row[k];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] += a * b[NC * (o + k) + c];
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    T v = group_allsum<64>(acc[c]);
    if (lane == c) x[NC * r + c] = v;
  }
}

// Pressure vectors through an explicit map: dst[map[i]] = src[i] (gather
// into hierarchy order) and dst[i] = src[map[i]] (scatter back).
template <typename T>
__global__ void __launch_bounds__(NT)
p_gather_map_kernel(long long np, const int* __restrict__ map, const T* __restrict__ src,
                    T* __restrict__ dst, const int* __restrict__ skip) {
  if (skip && *skip) return;
  long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i < np) dst[map[i]] = src[i];
}
template <typename T>
__global__ void __launch_bounds__(NT)
p_scatter_map_kernel(long long np, const int* __restrict__ map, const T* __restrict__ src,
                     T* __restrict__ dst, const int* __restrict__ skip) {
  if (skip && *skip) return;
  long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i < np) dst[i] = src[map[i]];
}

}  // namespace mpet

// --- host side ---
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
  // row -> instance of this level and of the next coarser one (batches)
  const int *grp = nullptr, *grpc = nullptr;
  // optional compact copy of the level operator (half values times 1/vscale,
  // 16-bit column offsets from the row), used by the level SpMMs when set
  const __half* v16 = nullptr;
  const int16_t* dci = nullptr;
  double vscale = 1.0;
  int deg = 2;                     // Chebyshev degree of this level's smoother
  double cd[MAXDEG], cr[MAXDEG];
  int lpr = 8, lprP = 4, lprR = 8;
  T *B = nullptr, *X = nullptr, *R = nullptr, *E = nullptr, *S1 = nullptr, *S2 = nullptr,
    *D = nullptr;
};

// One AMG hierarchy applied to NC-column vectors.
template <typename T>
struct Hier {
  int nlev = 0, deg = 2, NC = 1;
  LevelDev<T> lev[MAXLEV];
  const T* cinv = nullptr;
  // block-diagonal coarsest level (batched instances); null: one dense block
  const int *cblk = nullptr, *coff = nullptr;
  const int64_t* cioff = nullptr;
  double scale = 1.0;      // the V-cycle approximates (scale * A0)^-1
};

template <typename T>
struct Precond {
  Hier<T> u, p;
  int nP1 = 0, F = 0;
  long long nu = 0;
  const int* pmap = nullptr;           // explicit pressure order (batches)
  T *pin = nullptr, *pout = nullptr;   // field-major pressure buffers
  // optional second stream: the pressure V-cycle runs beside the
  // displacement one (fork / join through the two events; inside a captured
  // graph these become parallel branches)
  hipStream_t ps = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
};

// Create the fork stream and events of P (default priority, no timing).
template <typename T>
static int make_fork(Precond<T>& P) {
  if (hipStreamCreateWithFlags(&P.ps, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&P.ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&P.ev_join, hipEventDisableTiming) != hipSuccess) {
    mpet_set_error("preconditioner stream / event creation failed");
    return ERR_HIP;
  }
  return OK;
}
template <typename T>
static void free_fork(Precond<T>& P) {
  if (P.ev_fork) (void)hipEventDestroy(P.ev_fork);
  if (P.ev_join) (void)hipEventDestroy(P.ev_join);
  if (P.ps) (void)hipStreamDestroy(P.ps);
  P.ev_fork = P.ev_join = nullptr;
  P.ps = nullptr;
}

template <typename T, int NC, int MODE>
static void spmm(const LevelDev<T>& L, int which, const T* x, const T* b, T* y, T* d, T cd,
                 T cr, Skip sk, hipStream_t s, const T* dinv = nullptr, T* y2 = nullptr) {
  // which: 0 level operator, 1 prolongator (rows n), 2 restriction (rows nc);
  // dinv: the diagonal inverse used by the epilogue (default the level's)
  int nrows = which == 2 ? L.nc : L.n;
  const int* rp = which == 0 ? L.rp : (which == 1 ? L.Prp : L.Rrp);
  const int* ci = which == 0 ? L.ci : (which == 1 ? L.Pci : L.Rci);
  const T* v = which == 0 ? L.v : (which == 1 ? L.Pv : L.Rv);
  int lpr = which == 0 ? L.lpr : (which == 1 ? L.lprP : L.lprR);
  const bool compact = which == 0 && L.v16;
  if (NC > 1 && lpr == 1 && !compact) {     // one thread per row, all columns
    const int blocks1 = (nrows + NT - 1) / NT;
    if (blocks1 == 0) return;
    hipLaunchKernelGGL((spmm_row_kernel<T, NC, MODE>), dim3(blocks1), dim3(NT), 0, s, nrows,
                       rp, ci, v, x, b, y, dinv ? dinv : L.dinv, d, cd, cr,
                       which == 2 ? nullptr : L.mask, sk, which == 2 ? L.grpc : L.grp, y2);
    return;
  }
  if (lpr < NC) lpr = NC == 3 ? 4 : NC;
  const uint8_t* mask = which == 2 ? nullptr : L.mask;
  const int* grp = which == 2 ? L.grpc : L.grp;
  if (!dinv) dinv = L.dinv;
  long long threads = (long long)nrows * lpr;
  int blocks = (int)((threads + NT - 1) / NT);
  if (blocks == 0) return;
  if (which == 0 && L.v16) {
    if constexpr (sizeof(T) == 4 && MODE != 4) {
#define SPMM_CASE(LP)                                                                          \
  case LP:                                                                                     \
    hipLaunchKernelGGL((spmm_kernel<T, NC, (LP < NC ? 4 : LP), MODE, true>), dim3(blocks),     \
                       dim3(NT), 0, s, nrows, rp, ci, v, x, b, y, dinv, d, cd, cr, mask, sk, grp, \
                       y2, L.v16, L.dci, (T)L.vscale);                                         \
    break;
      switch (lpr) {
        SPMM_CASE(4) SPMM_CASE(8) SPMM_CASE(16) SPMM_CASE(32) SPMM_CASE(64)
        default: hipLaunchKernelGGL((spmm_kernel<T, NC, (NC > 4 ? 8 : 4), MODE, true>),
                                    dim3((unsigned)(((long long)nrows * (NC > 4 ? 8 : 4) + NT - 1) / NT)),
                                    dim3(NT), 0, s, nrows, rp, ci, v, x, b, y, dinv, d, cd, cr,
                                    mask, sk, grp, y2, L.v16, L.dci, (T)L.vscale);
      }
#undef SPMM_CASE
      return;
    }
  }
#define SPMM_CASE(LP)                                                                          \
  case LP:                                                                                     \
    hipLaunchKernelGGL((spmm_kernel<T, NC, (LP < NC ? 4 : LP), MODE>), dim3(blocks), dim3(NT), \
                       0, s, nrows, rp, ci, v, x, b, y, dinv, d, cd, cr, mask, sk, grp, y2);  \
    break;
  switch (lpr) {
    SPMM_CASE(1) SPMM_CASE(2) SPMM_CASE(4) SPMM_CASE(8) SPMM_CASE(16) SPMM_CASE(32)
    SPMM_CASE(64)
    default: break;
  }
#undef SPMM_CASE
}

// Buffer of the first Chebyshev step from zero on a level whose smoothing
// result goes to out.
template <typename T>
static T* first_dst(const LevelDev<T>& L, T* out) {
  return L.deg == 1 ? out : L.S1;
}

// k Chebyshev steps on level l; from zero (x0 == nullptr, b scaled by
// bscale and copied to bcopy when given; entered: the first step is already
// in first_dst, made by the restriction) or from x0; result in out.
template <typename T, int NC>
static void smooth(const Hier<T>& H, int l, const T* b, T bscale, T* bcopy, const T* x0,
                   T* out, Skip sk, hipStream_t s, bool entered = false,
                   K2In<T> k2 = K2In<T>()) {
  const LevelDev<T>& L = H.lev[l];
  const int k = L.deg;
  const T* cur = x0;
  int i = 0;
  T* bufs[2] = {L.S1, L.S2};
  int flip = 0;
  const T* bb = b;
  if (!x0) {
    T* dst = first_dst(L, out);
    int blocks = (int)(((long long)NC * L.n + NT - 1) / NT);
    if (!entered)
      hipLaunchKernelGGL((cheb0_kernel<T, NC>), dim3(blocks), dim3(NT), 0, s, L.n, b, bscale,
                         L.dinv, (T)L.cr[0], k == 1 ? nullptr : L.D, dst, bcopy, L.mask, sk,
                         L.grp, k2);
    if (bcopy) bb = bcopy;
    cur = dst;
    i = 1;
    flip ^= 1;
  }
  for (; i < k; ++i) {
    T* dst = (i == k - 1) ? out : bufs[flip];
    if (dst == cur) dst = bufs[flip ^ 1];
    // the last step's direction is not needed again
    spmm<T, NC, 3>(L, 0, cur, bb, dst, L.D, (T)L.cd[i], (T)L.cr[i], sk, s, nullptr,
                   i == k - 1 ? nullptr : L.D);
    cur = dst;
    flip ^= 1;
  }
}

template <typename T, int NC>
static void vcycle(const Hier<T>& H, int l, const T* b, T bscale, T* out, Skip sk,
                   hipStream_t s, bool entered = false, K2In<T> k2 = K2In<T>()) {
  // k2 (a fused input update) only on a level with a smoother
  const LevelDev<T>& L = H.lev[l];
  if (l == H.nlev - 1) {
    const T* bb = b;
    if (bscale != (T)1) {   // single-level hierarchy: scale into L.B first
      int blocks = (int)(((long long)NC * L.n + NT - 1) / NT);
      hipLaunchKernelGGL((cheb0_kernel<T, NC>), dim3(blocks), dim3(NT), 0, s, L.n, b, bscale,
                         L.dinv, (T)0, L.D, L.X, L.B, L.mask, sk, L.grp);
      bb = L.B;
    }
    int blocks = (L.n + NT / 64 - 1) / (NT / 64);
    if (H.cblk)
      hipLaunchKernelGGL((coarse_blockdiag_kernel<T, NC>), dim3(blocks), dim3(NT), 0, s, L.n,
                         H.cblk, H.coff, H.cioff, H.cinv, bb, out, sk, L.grp);
    else
      hipLaunchKernelGGL((coarse_kernel<T, NC>), dim3(blocks), dim3(NT), 0, s, L.n, H.cinv, bb,
                         out, sk.all);
    return;
  }
  T* bcopy = (bscale != (T)1) ? L.B : nullptr;
  smooth<T, NC>(H, l, b, bscale, bcopy, nullptr, L.X, sk, s, entered, k2);
  const T* bb = bcopy ? bcopy : b;
  spmm<T, NC, 1>(L, 0, L.X, bb, L.R, nullptr, 0, 0, sk, s);
  const LevelDev<T>& C = H.lev[l + 1];
  if (l + 1 < H.nlev - 1) {
    // restriction fused with the coarse level's first smoothing step (its
    // pre-smoothing writes to C.X)
    spmm<T, NC, 4>(L, 2, L.R, nullptr, C.B, C.D, 0, (T)C.cr[0], sk, s, C.dinv,
                   first_dst(C, C.X));
    vcycle<T, NC>(H, l + 1, C.B, (T)1, C.E, sk, s, true);
  } else {
    spmm<T, NC, 0>(L, 2, L.R, nullptr, C.B, nullptr, 0, 0, sk, s);
    vcycle<T, NC>(H, l + 1, C.B, (T)1, C.E, sk, s);
  }
  spmm<T, NC, 2>(L, 1, C.E, nullptr, L.X, nullptr, 0, 0, sk, s);
  smooth<T, NC>(H, l, bb, (T)1, nullptr, L.X, out, sk, s);
}

// Whether precond_apply can take a fused input update (k2): the first
// kernels reading the input are the displacement level-0 smoother and the
// pressure gather.
template <typename T>
static bool k2_fusable(const Precond<T>& P) {
  return P.u.nlev > 1 && !P.pmap;
}

// z = M^-1 r; with k2 (k2_fusable only) r is first made as k2 describes.
template <typename T>
static void precond_apply(const Precond<T>& P, const T* r, T* z, Skip sk, hipStream_t s,
                          const K2In<T>* k2 = nullptr) {
  const int* skip = sk.all;
  hipStream_t sp = s;
  if (P.ps) {                      // pressure blocks on the second stream
    (void)hipEventRecord(P.ev_fork, s);
    (void)hipStreamWaitEvent(P.ps, P.ev_fork, 0);
    sp = P.ps;
  }
  K2In<T> ku, kp;
  if (k2 && k2->y) {
    ku = *k2;
    kp = K2In<T>{k2->y + P.nu, k2->r2o + P.nu, k2->r + P.nu, k2->alfa, k2->beta};
  }
  vcycle<T, 3>(P.u, 0, r, (T)(1.0 / P.u.scale), z, sk, s, false, ku);
  long long np = (long long)P.F * P.nP1;
  int eb = (int)((np + NT - 1) / NT);
  if (P.pmap)
    hipLaunchKernelGGL((p_gather_map_kernel<T>), dim3(eb), dim3(NT), 0, sp, np, P.pmap, r + P.nu,
                       P.pin, skip);
  else
    hipLaunchKernelGGL((p_gather_kernel<T>), dim3(eb), dim3(NT), 0, sp, P.nP1, P.F, r + P.nu,
                       P.pin, skip, kp);
  vcycle<T, 1>(P.p, 0, P.pin, (T)1, P.pout, sk, sp);
  if (P.pmap)
    hipLaunchKernelGGL((p_scatter_map_kernel<T>), dim3(eb), dim3(NT), 0, sp, np, P.pmap, P.pout,
                       z + P.nu, skip);
  else
    hipLaunchKernelGGL((p_scatter_kernel<T>), dim3(eb), dim3(NT), 0, sp, P.nP1, P.F, P.pout,
                       z + P.nu, skip);
  if (P.ps) {
    (void)hipEventRecord(P.ev_join, P.ps);
    (void)hipStreamWaitEvent(s, P.ev_join, 0);
  }
}
template <typename T>
static void precond_apply(const Precond<T>& P, const T* r, T* z, const int* skip,
                          hipStream_t s) {
  Skip sk;
  sk.all = skip;
  precond_apply(P, r, z, sk, s);
}

// Lanes per row of the levels' SpMMs from an MpetTuning table (0 keeps).
static bool lpr_ok(int v) {
  return v == 0 || v == 1 || v == 2 || v == 4 || v == 8 || v == 16 || v == 32 || v == 64;
}
static bool tuning_ok(const MpetTuning* t) {
  if (!t) return false;
  for (int l = 0; l < MPET_MAXLEV; ++l)
    for (int k = 0; k < 3; ++k)
      if (!lpr_ok(t->u_lpr[l][k]) || !lpr_ok(t->p_lpr[l][k])) return false;
  return true;
}
template <typename T>
static void set_hier_lpr(Hier<T>& H, const int32_t (*lpr)[3]) {
  for (int l = 0; l < H.nlev && l < MPET_MAXLEV; ++l) {
    if (lpr[l][0] > 0) H.lev[l].lpr = lpr[l][0];
    if (lpr[l][1] > 0) H.lev[l].lprP = lpr[l][1];
    if (lpr[l][2] > 0) H.lev[l].lprR = lpr[l][2];
  }
}
template <typename T>
static void set_precond_tuning(Precond<T>& P, const MpetTuning* t) {
  set_hier_lpr(P.u, t->u_lpr);
  set_hier_lpr(P.p, t->p_lpr);
}

template <typename T>
static int alloc_dev(std::vector<void*>& allocs, T** p, size_t count) {
  void* q = nullptr;
  if (hipMalloc(&q, sizeof(T) * (count ? count : 1)) != hipSuccess) {
    mpet_set_error("hipMalloc failed (out of device memory)");
    return ERR_OOM;
  }
  allocs.push_back(q);
  *p = (T*)q;
  return OK;
}

template <typename T>
static int build_hier(std::vector<void*>& allocs, const MpetHierDesc* d, int NC,
                      Hier<T>& H) {
  H.nlev = d->nlev;
  H.deg = d->degree;
  H.NC = NC;
  H.scale = d->scale != 0.0 ? d->scale : 1.0;
  if (H.nlev < 1 || H.nlev > MAXLEV || H.deg < 1 || H.deg > MAXDEG) {
    mpet_set_error("invalid hierarchy description");
    return ERR_INVALID;
  }
  for (int l = 0; l < H.nlev; ++l) {
    const MpetAmgLevelDesc& ld = d->lev[l];
    LevelDev<T>& L = H.lev[l];
    L.n = ld.n;
    L.nc = ld.nc;
    L.rp = ld.A_rp; L.ci = ld.A_ci; L.v = (const T*)ld.A_v; L.dinv = (const T*)ld.dinv;
    L.Prp = ld.P_rp; L.Pci = ld.P_ci; L.Pv = (const T*)ld.P_v;
    L.Rrp = ld.R_rp; L.Rci = ld.R_ci; L.Rv = (const T*)ld.R_v;
    L.mask = l == 0 ? d->mask0 : nullptr;
    L.grp = ld.grp;
    L.v16 = (const __half*)ld.A_v16;
    L.dci = ld.A_dci;
    L.vscale = ld.A_vscale != 0.0 ? ld.A_vscale : 1.0;
    L.lpr = ld.lpr ? ld.lpr : 8;
    L.lprP = ld.lprP ? ld.lprP : 4;
    L.lprR = ld.lprR ? ld.lprR : 8;
    L.deg = d->level_degree[l] > 0 ? d->level_degree[l] : H.deg;
    if (L.deg > MAXDEG) {
      mpet_set_error("invalid hierarchy description");
      return ERR_INVALID;
    }
    double lmax = ld.rho * d->eig_safety;
    chebyshev(lmax / d->eig_ratio, lmax, L.deg, L.cd, L.cr);
    size_t m = (size_t)NC * L.n;
    int rc;
    if ((rc = alloc_dev(allocs, &L.B, m)) || (rc = alloc_dev(allocs, &L.X, m)) ||
        (rc = alloc_dev(allocs, &L.R, m)) || (rc = alloc_dev(allocs, &L.E, m)) ||
        (rc = alloc_dev(allocs, &L.S1, m)) || (rc = alloc_dev(allocs, &L.S2, m)) ||
        (rc = alloc_dev(allocs, &L.D, m)))
      return rc;
  }
  for (int l = 0; l + 1 < H.nlev; ++l) H.lev[l].grpc = H.lev[l + 1].grp;
  H.cinv = (const T*)d->coarse_inv;
  H.cblk = d->coarse_blk;
  H.coff = d->coarse_off;
  H.cioff = d->coarse_ioff;
  return OK;
}

template <typename T>
static int build_precond(std::vector<void*>& allocs, long long nu, const MpetPrecDesc* d,
                         Precond<T>& P) {
  int rc;
  if ((rc = build_hier(allocs, &d->u, 3, P.u)) || (rc = build_hier(allocs, &d->p, 1, P.p)))
    return rc;
  P.nP1 = d->nP1;
  P.F = d->F;
  P.nu = nu;
  P.pmap = d->pmap;
  if (P.u.lev[0].n * 3ll != nu || (long long)P.p.lev[0].n != (long long)P.F * P.nP1) {
    mpet_set_error("preconditioner sizes do not match the operator");
    return ERR_INVALID;
  }
  size_t np = (size_t)P.F * P.nP1;
  if ((rc = alloc_dev(allocs, &P.pin, np)) || (rc = alloc_dev(allocs, &P.pout, np))) return rc;
  return OK;
}
