// Row computations of the structured total-pressure operator, shared by the
// plain apply (spmv.hip), the fused MINRES step (krylov.hip) and the packed
// batch (batch.hip).  G lanes cooperate on one node row; every lane takes
// two nonzeros per trip so that both sets of index, value and gather loads
// are in flight together (the summation order per lane is that of one
// nonzero at a time, so results do not depend on the unrolling).
#pragma once
#include "mpet_common.h"

namespace mpet {

// u rows of P2 node a:  (y0, y1, y2) += Kuu(a, :) x_u + Dv(a, :) p_0, with
// p_0 of P1 column node c at x[p0base + p0stride * c].
template <typename T, int G>
__device__ __forceinline__ void urow_apply(const int* __restrict__ rp22,
                                           const int* __restrict__ ci22,
                                           const T* __restrict__ Kuu, long long n22,
                                           const int* __restrict__ rp21,
                                           const int* __restrict__ ci21,
                                           const T* __restrict__ Dv, long long n21,
                                           const T* __restrict__ x, long long p0base,
                                           int p0stride, int a, int lane, T& y0, T& y1, T& y2) {
  int k = rp22[a] + lane;
  const int ke = rp22[a + 1];
  for (; k + G < ke; k += 2 * G) {
    const long long ca = 3ll * ci22[k], cb = 3ll * ci22[k + G];
    T ka[9], kb[9];
#pragma unroll
    for (int m = 0; m < 9; ++m) {
      ka[m] = Kuu[m * n22 + k];
      kb[m] = Kuu[m * n22 + k + G];
    }
    const T a0 = x[ca], a1 = x[ca + 1], a2 = x[ca + 2];
    const T b0 = x[cb], b1 = x[cb + 1], b2 = x[cb + 2];
    y0 += ka[0] * a0 + ka[1] * a1 + ka[2] * a2;
    y1 += ka[3] * a0 + ka[4] * a1 + ka[5] * a2;
    y2 += ka[6] * a0 + ka[7] * a1 + ka[8] * a2;
    y0 += kb[0] * b0 + kb[1] * b1 + kb[2] * b2;
    y1 += kb[3] * b0 + kb[4] * b1 + kb[5] * b2;
    y2 += kb[6] * b0 + kb[7] * b1 + kb[8] * b2;
  }
  if (k < ke) {
    const long long ca = 3ll * ci22[k];
    const T a0 = x[ca], a1 = x[ca + 1], a2 = x[ca + 2];
    y0 += Kuu[k] * a0 + Kuu[n22 + k] * a1 + Kuu[2 * n22 + k] * a2;
    y1 += Kuu[3 * n22 + k] * a0 + Kuu[4 * n22 + k] * a1 + Kuu[5 * n22 + k] * a2;
    y2 += Kuu[6 * n22 + k] * a0 + Kuu[7 * n22 + k] * a1 + Kuu[8 * n22 + k] * a2;
  }
  k = rp21[a] + lane;
  const int ke1 = rp21[a + 1];
  for (; k + G < ke1; k += 2 * G) {
    const T pa = x[p0base + (long long)p0stride * ci21[k]];
    const T pb = x[p0base + (long long)p0stride * ci21[k + G]];
    const T da0 = Dv[k], da1 = Dv[n21 + k], da2 = Dv[2 * n21 + k];
    const T db0 = Dv[k + G], db1 = Dv[n21 + k + G], db2 = Dv[2 * n21 + k + G];
    y0 += da0 * pa;
    y1 += da1 * pa;
    y2 += da2 * pa;
    y0 += db0 * pb;
    y1 += db1 * pb;
    y2 += db2 * pb;
  }
  if (k < ke1) {
    const T p0 = x[p0base + (long long)p0stride * ci21[k]];
    y0 += Dv[k] * p0;
    y1 += Dv[n21 + k] * p0;
    y2 += Dv[2 * n21 + k] * p0;
  }
}

// p rows of P1 node c, divergence part: sum Dt(c, :) x_u.
template <typename T, int G>
__device__ __forceinline__ T prow_div(const int* __restrict__ rp12, const int* __restrict__ ci12,
                                      const T* __restrict__ Dt, long long n12,
                                      const T* __restrict__ x, int c, int lane) {
  T yd = 0;
  int k = rp12[c] + lane;
  const int ke = rp12[c + 1];
  for (; k + G < ke; k += 2 * G) {
    const long long ca = 3ll * ci12[k], cb = 3ll * ci12[k + G];
    const T da0 = Dt[k], da1 = Dt[n12 + k], da2 = Dt[2 * n12 + k];
    const T db0 = Dt[k + G], db1 = Dt[n12 + k + G], db2 = Dt[2 * n12 + k + G];
    const T a0 = x[ca], a1 = x[ca + 1], a2 = x[ca + 2];
    const T b0 = x[cb], b1 = x[cb + 1], b2 = x[cb + 2];
    yd += da0 * a0 + da1 * a1 + da2 * a2;
    yd += db0 * b0 + db1 * b1 + db2 * b2;
  }
  if (k < ke) {
    const long long ca = 3ll * ci12[k];
    yd += Dt[k] * x[ca] + Dt[n12 + k] * x[ca + 1] + Dt[2 * n12 + k] * x[ca + 2];
  }
  return yd;
}

// p rows of P1 node c, mass and stiffness parts for the F fields (F <= FM):
// m[g] += M1(c, :) x_g, sv[g] += St_g(c, :) x_g (g >= 1).  Field g of
// column node c' is at x[xbase + xstride * c' + g]; St_g's entry of nonzero
// k is St[(g - 1) * n11 + k - kbase].
template <typename T, int G, int FM>
__device__ __forceinline__ void prow_mass(const int* __restrict__ rp11,
                                          const int* __restrict__ ci11,
                                          const T* __restrict__ M1, const T* __restrict__ St,
                                          long long n11, long long kbase,
                                          const T* __restrict__ x, long long xbase, int xstride,
                                          int F, int c, int lane, T (&m)[FM], T (&sv)[FM]) {
  int k = rp11[c] + lane;
  const int ke = rp11[c + 1];
  for (; k + G < ke; k += 2 * G) {
    const T* xa = x + xbase + (long long)xstride * ci11[k];
    const T* xb = x + xbase + (long long)xstride * ci11[k + G];
    const T ma = M1[k], mb = M1[k + G];
    T va[FM], vb[FM], sa[FM], sb[FM];
#pragma unroll
    for (int g = 0; g < FM; ++g) {
      if (g < F) {
        va[g] = xa[g];
        vb[g] = xb[g];
        sa[g] = g >= 1 ? St[(g - 1) * n11 + k - kbase] : (T)0;
        sb[g] = g >= 1 ? St[(g - 1) * n11 + k + G - kbase] : (T)0;
      }
    }
#pragma unroll
    for (int g = 0; g < FM; ++g) {
      if (g < F) {
        m[g] += ma * va[g];
        if (g >= 1) sv[g] += sa[g] * va[g];
        m[g] += mb * vb[g];
        if (g >= 1) sv[g] += sb[g] * vb[g];
      }
    }
  }
  if (k < ke) {
    const T* xa = x + xbase + (long long)xstride * ci11[k];
    const T ma = M1[k];
#pragma unroll
    for (int g = 0; g < FM; ++g) {
      if (g < F) {
        const T v = xa[g];
        m[g] += ma * v;
        if (g >= 1) sv[g] += St[(g - 1) * n11 + k - kbase] * v;
      }
    }
  }
}

// Scalar CSR row r: sum over lane's nonzeros k = kb + lane + j LPR, two per
// trip (same per-lane order as one at a time).
template <typename T, int LPR>
__device__ __forceinline__ T csr_row(const long long* __restrict__ rp, const int* __restrict__ ci,
                                     const T* __restrict__ val, const T* __restrict__ x,
                                     long long r, int lane) {
  T s = 0;
  long long k = rp[r] + lane;
  const long long ke = rp[r + 1];
  for (; k + LPR < ke; k += 2 * LPR) {
    const T va = val[k], vb = val[k + LPR];
    const T xa = x[ci[k]], xb = x[ci[k + LPR]];
    s += va * xa;
    s += vb * xb;
  }
  if (k < ke) s += val[k] * x[ci[k]];
  return s;
}

}  // namespace mpet
