// Two-field MPET system (MPETSolver, mpet/mpetsolver.py:196-201; stress of
// mpetproblem.py:18-24) as a scalar FP64 CSR built on the device from the
// primitives the total-pressure assembly already produces:
//
//   u-u    2 mu eps:eps + lambda div div.  With Kuu[k][i][j] = mu (delta_ij
//          Lap_k + int d_j phi_a d_i phi_b) (assemble.hip), the div-div entry
//          int d_i phi_a d_j phi_b is Kuu[k][j][i] / mu - delta_ij Lap_k, so no
//          further element integration is needed;
//   u-p_f  -alpha_f int p_f div v    = -alpha_f Dv        (:197)
//   p_f-u  -alpha_f int div u q_f    = -alpha_f Dt        (:199)
//   p_f-p_g  C[f][g] M1 + delta_fg stiff[f] St_f          (:198, 200-201)
//
// which = 0: the matrix of lhs(F) (Robin folded into St as for the total-
// pressure system); 1: the block-diagonal prec (:269-271); 2: the matrix of
// the up_-linear part of rhs(F) at t_n (pressure rows only).  F = A
// pressure fields per P1 node (p_1..p_A).  One thread per row; columns are
// emitted in ascending order.
#include "mpet_ops.h"
#include <stdio.h>

namespace mpet {

__global__ void __launch_bounds__(256)
tf_expand_kernel(MpetPatterns pat, MpetPrimitives prim, int nP2, int nP1, int A,
                 MpetTwoFieldCoef coef, int which, const long long* __restrict__ rowptr,
                 int* __restrict__ cols, double* __restrict__ vals) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nu = 3ll * nP2;
  const int F = A;
  const long long n = nu + (long long)F * nP1;
  if (r >= n) return;
  long long o = rowptr[r];
  const long long nnz11 = pat.nnz11;
  if (r < nu) {
    if (which == 2) return;                       // rhs(F) has no up_ term in u rows
    const int a = (int)(r / 3), i = (int)(r % 3);
    for (int k = pat.rp22[a]; k < pat.rp22[a + 1]; ++k) {
      const int b = pat.ci22[k];
      for (int j = 0; j < 3; ++j) {
        double v;
        if (which == 0) {
          const double dd = prim.Kuu[9ll * k + 3 * j + i] / coef.mu - (i == j ? prim.Lap[k] : 0.0);
          v = prim.Kuu[9ll * k + 3 * i + j] + coef.lmbda * dd;
        } else {
          v = i == j ? coef.mu * prim.Lap[k] : 0.0;
        }
        cols[o] = 3 * b + j;
        vals[o] = v;
        ++o;
      }
    }
    if (which == 0) {
      for (int k = pat.rp21[a]; k < pat.rp21[a + 1]; ++k) {
        const long long cb = nu + (long long)F * pat.ci21[k];
        const double d = prim.Dv[3ll * k + i];
        for (int f = 0; f < F; ++f) {
          cols[o] = (int)(cb + f);
          vals[o] = -coef.alpha[f] * d;
          ++o;
        }
      }
    }
  } else {
    const long long q = r - nu;
    const int c = (int)(q / F), f = (int)(q % F);
    if (which != 1) {
      for (int k = pat.rp12[c]; k < pat.rp12[c + 1]; ++k) {
        const int b = pat.ci12[k];
        for (int j = 0; j < 3; ++j) {
          cols[o] = 3 * b + j;
          vals[o] = -coef.alpha[f] * prim.Dt[3ll * k + j];
          ++o;
        }
      }
    }
    for (int k = pat.rp11[c]; k < pat.rp11[c + 1]; ++k) {
      const long long cb = nu + (long long)F * pat.ci11[k];
      const double m = prim.M1[k];
      const double st = prim.St[f * nnz11 + k];
      if (which != 1) {
        for (int g = 0; g < F; ++g) {
          double v = coef.C[f * MPET_MAXF + g] * m;
          if (g == f) v += coef.stiff[f] * st;
          cols[o] = (int)(cb + g);
          vals[o] = v;
          ++o;
        }
      } else {
        cols[o] = (int)(cb + f);
        vals[o] = coef.C[f * MPET_MAXF + f] * m + coef.stiff[f] * st;
        ++o;
      }
    }
  }
}

}  // namespace mpet

using namespace mpet;

extern "C" int mpet_tf_expand_csr(const MpetPatterns* pat, const MpetPrimitives* prim, int nP2,
                                  int nP1, int A, const MpetTwoFieldCoef* coef, int which,
                                  const long long* rowptr, int* cols, double* vals,
                                  hipStream_t stream) {
  if (!pat || !prim || !coef || !rowptr || !cols || !vals || A < 1 || A > MPET_MAXF ||
      which < 0 || which > 2 || nP2 < 1 || nP1 < 1 || !(coef->mu > 0.0)) {
    mpet_set_error("mpet_tf_expand_csr: invalid argument");
    return ERR_INVALID;
  }
  const long long n = 3ll * nP2 + (long long)A * nP1;
  const int blocks = (int)((n + 255) / 256);
  hipLaunchKernelGGL(tf_expand_kernel, dim3(blocks), dim3(256), 0, stream, *pat, *prim, nP2,
                     nP1, A, *coef, which, rowptr, cols, vals);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    char m[300];
    snprintf(m, sizeof(m), "mpet_tf_expand_csr: %s", hipGetErrorString(e));
    mpet_set_error(m);
    return ERR_HIP;
  }
  return OK;
}
