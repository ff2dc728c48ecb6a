// Row computations of the structured total-pressure operator, shared by the
// plain apply (spmv.hip), the fused MINRES step (krylov.hip) and the packed
// batch (batch.hip).  G lanes cooperate on one node row.
//
// Latency: a row is walked in trips, and every index, value and gather load
// of a trip is issued before the first of them is used.  A trip takes a
// fixed number of nonzeros per lane from each block of the row (the u row's
// P2 and P1 parts, the p row's P2 and P1 parts, together), predicated: a
// slot past the row end reads a valid index of the same array (slot_index)
// and multiplies a zero.  A lane's nonzeros of one block are summed in the
// order k, k + G, k + 2 G, ..., so results do not depend on the unrolling.
//
// The operator arrays must be non-empty (every pattern has at least one
// nonzero); rows may be empty.
#pragma once
#include "mpet_common.h"

namespace mpet {

// Index for slot k of a row ending at ke: k itself inside the row, else the
// row's last nonzero, and lo (the first index the caller may read) for an
// empty row at the start of the readable range.
__device__ __forceinline__ int slot_index(int k, int ke, int lo) {
  const int i = k < ke ? k : ke - 1;
  return i > lo ? i : lo;
}

// u rows of P2 node a:  (y0, y1, y2) += Kuu(a, :) x_u + Dv(a, :) p_0, with
// p_0 of P1 column node c at x[p0base + p0stride * c].  One trip: two P2
// nonzeros and one P1 nonzero per lane.
template <typename T, int G>
__device__ __forceinline__ void urow_apply(const int* __restrict__ rp22,
                                           const int* __restrict__ ci22,
                                           const T* __restrict__ Kuu, long long n22,
                                           const int* __restrict__ rp21,
                                           const int* __restrict__ ci21,
                                           const T* __restrict__ Dv, long long n21,
                                           const T* __restrict__ x, long long p0base,
                                           int p0stride, int a, int lane, T& y0, T& y1, T& y2) {
  const int ke = rp22[a + 1], ke1 = rp21[a + 1];
  int k = rp22[a] + lane, k1 = rp21[a] + lane;
  while (k < ke || k1 < ke1) {
    const int ia = slot_index(k, ke, 0), ib = slot_index(k + G, ke, 0);
    const int ip = slot_index(k1, ke1, 0);
    const bool oa = k < ke, ob = k + G < ke, op = k1 < ke1;
    const long long ca = 3ll * ci22[ia], cb = 3ll * ci22[ib];
    const long long cp = p0base + (long long)p0stride * ci21[ip];
    T ka[9], kb[9];
#pragma unroll
    for (int m = 0; m < 9; ++m) {
      ka[m] = Kuu[m * n22 + ia];
      kb[m] = Kuu[m * n22 + ib];
    }
    const T d0 = Dv[ip], d1 = Dv[n21 + ip], d2 = Dv[2 * n21 + ip];
    T xa[3], xb[3];
    load_nc<T, 3>(x + ca, xa);
    load_nc<T, 3>(x + cb, xb);
    T a0 = xa[0], a1 = xa[1], a2 = xa[2];
    T b0 = xb[0], b1 = xb[1], b2 = xb[2];
    T p = x[cp];
    if (!oa) a0 = a1 = a2 = 0;
    if (!ob) b0 = b1 = b2 = 0;
    if (!op) p = 0;
    y0 += ka[0] * a0 + ka[1] * a1 + ka[2] * a2;
    y1 += ka[3] * a0 + ka[4] * a1 + ka[5] * a2;
    y2 += ka[6] * a0 + ka[7] * a1 + ka[8] * a2;
    y0 += kb[0] * b0 + kb[1] * b1 + kb[2] * b2;
    y1 += kb[3] * b0 + kb[4] * b1 + kb[5] * b2;
    y2 += kb[6] * b0 + kb[7] * b1 + kb[8] * b2;
    y0 += d0 * p;
    y1 += d1 * p;
    y2 += d2 * p;
    k += 2 * G;
    k1 += G;
  }
}

// p rows of P1 node c:
//   yd += Dt(c, :) x_u                       (divergence part)
//   m[g] += M1(c, :) x_g, sv[g] += St_g(c, :) x_g (g >= 1), g < F <= FM
// Field g of column node c' is at x[xbase + xstride * c' + g]; St_g's entry
// of nonzero k is St[(g - 1) * n11 + k - kbase] (the caller's St holds the
// nonzeros kbase .. kbase + n11 - 1).  One trip: four P2 nonzeros and one P1
// nonzero per lane.
template <typename T, int G, int FM>
__device__ __forceinline__ void prow_apply(const int* __restrict__ rp12,
                                           const int* __restrict__ ci12,
                                           const T* __restrict__ Dt, long long n12,
                                           const int* __restrict__ rp11,
                                           const int* __restrict__ ci11,
                                           const T* __restrict__ M1, const T* __restrict__ St,
                                           long long n11, long long kbase, const T* __restrict__ x,
                                           long long xbase, int xstride, int F, int c, int lane,
                                           T& yd, T (&m)[FM], T (&sv)[FM]) {
  constexpr int U = 4;
  const int ke = rp12[c + 1], ke1 = rp11[c + 1];
  int k = rp12[c] + lane, k1 = rp11[c] + lane;
  while (k < ke || k1 < ke1) {
    int iu[U];
    long long cu[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
      iu[j] = slot_index(k + j * G, ke, 0);
      cu[j] = 3ll * ci12[iu[j]];
    }
    const int ip = slot_index(k1, ke1, (int)kbase);
    const bool op = k1 < ke1;
    const T* xp = x + xbase + (long long)xstride * ci11[ip];
    T dt[U][3], xu[U][3];
#pragma unroll
    for (int j = 0; j < U; ++j) {
#pragma unroll
      for (int e = 0; e < 3; ++e) dt[j][e] = Dt[e * n12 + iu[j]];
      load_nc<T, 3>(x + cu[j], xu[j]);
    }
    const T mm = M1[ip];
    T xv[FM], st[FM];
#pragma unroll
    for (int g = 0; g < FM; ++g) {
      if (g < F) {
        xv[g] = xp[g];
        st[g] = g >= 1 ? St[(g - 1) * n11 + ip - kbase] : (T)0;
      }
    }
#pragma unroll
    for (int j = 0; j < U; ++j) {
      if (!(k + j * G < ke)) xu[j][0] = xu[j][1] = xu[j][2] = 0;
      yd += dt[j][0] * xu[j][0] + dt[j][1] * xu[j][1] + dt[j][2] * xu[j][2];
    }
#pragma unroll
    for (int g = 0; g < FM; ++g) {
      if (g < F) {
        const T v = op ? xv[g] : (T)0;
        m[g] += mm * v;
        if (g >= 1) sv[g] += st[g] * v;
      }
    }
    k += U * G;
    k1 += G;
  }
}

// Scalar CSR row r: sum over the lane's nonzeros k = kb + lane + j LPR, two
// per trip and a remainder (measured faster than four predicated slots for
// the ~100-nonzero rows of the assembled FP64 matrix).
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
