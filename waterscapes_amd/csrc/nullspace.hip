// Constraint rows of the nullspace multipliers (bordered systems).
//
// With u_has_nullspace the reference adds, for the L2-orthonormal rigid
// motions Z_0..Z_5 of mpet/rm_basis_L2.py:10-74, the terms
//   sum_k r_k int Z_k . u + z_k int Z_k . v
// (mpet/mpettotalpressuresolver.py:285-293, mpet/mpetsolver.py:203-206), and
// for every pressure field k with p_has_nullspace[k]
//   int p_k q_null + p_null int q_k
// (mpettotalpressuresolver.py:295-303, mpetsolver.py:208-215): the matrix
// becomes [A B^T; B 0] with one dense row of B per multiplier.
//
// Row k < nz:      B[k][3a + i] = int Z_k,i phi_a = sum_b M2(a, b) Z_k,i(x_b)
//                  (Z_k is linear, so its P2 interpolant Zt is exact);
// row nz + j:      B[nz + j][nu + F c + fields[j]] = int psi_c = sum_d M1(c, d).
// All other entries are zero.  One thread per DOF; deterministic.
#include "mpet_ops.h"
#include <stdio.h>

namespace mpet {

struct FieldList {
  int f[MPET_MAXF];
};

__global__ void __launch_bounds__(256)
constraint_rows_kernel(MpetPatterns pat, const double* __restrict__ M2,
                       const double* __restrict__ M1, int nP2, int nP1, int F,
                       const double* __restrict__ Zt, int nz, FieldList fl, int nf,
                       double* __restrict__ B) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nu = 3ll * nP2;
  const long long n = nu + (long long)F * nP1;
  if (r >= n) return;
  if (r < nu) {
    const int a = (int)(r / 3), i = (int)(r % 3);
    for (int k = 0; k < nz; ++k) {
      double s = 0.0;
      for (int q = pat.rp22[a]; q < pat.rp22[a + 1]; ++q)
        s += M2[q] * Zt[k * nu + 3ll * pat.ci22[q] + i];
      B[k * n + r] = s;
    }
    for (int j = 0; j < nf; ++j) B[(nz + j) * n + r] = 0.0;
  } else {
    const long long q = r - nu;
    const int c = (int)(q / F), f = (int)(q % F);
    for (int k = 0; k < nz; ++k) B[k * n + r] = 0.0;
    double m = 0.0;
    for (int k = pat.rp11[c]; k < pat.rp11[c + 1]; ++k) m += M1[k];
    for (int j = 0; j < nf; ++j) B[(nz + j) * n + r] = f == fl.f[j] ? m : 0.0;
  }
}

}  // namespace mpet

using namespace mpet;

extern "C" int mpet_constraint_rows(const MpetPatterns* pat, const double* M2, const double* M1,
                                    int nP2, int nP1, int F, const double* Zt, int nz,
                                    const int* fields, int nf, double* B, hipStream_t stream) {
  if (!pat || !M1 || !B || nP2 < 1 || nP1 < 1 || F < 1 || F > MPET_MAXF || nz < 0 ||
      nf < 0 || nf > F || nz + nf < 1 || (nz > 0 && (!M2 || !Zt)) || (nf > 0 && !fields)) {
    mpet_set_error("mpet_constraint_rows: invalid argument");
    return ERR_INVALID;
  }
  FieldList fl;
  for (int j = 0; j < MPET_MAXF; ++j) fl.f[j] = -1;
  for (int j = 0; j < nf; ++j) {
    if (fields[j] < 0 || fields[j] >= F) {
      mpet_set_error("mpet_constraint_rows: field index out of range");
      return ERR_INVALID;
    }
    fl.f[j] = fields[j];
  }
  const long long n = 3ll * nP2 + (long long)F * nP1;
  const int blocks = (int)((n + 255) / 256);
  hipLaunchKernelGGL(constraint_rows_kernel, dim3(blocks), dim3(256), 0, stream, *pat, M2, M1,
                     nP2, nP1, F, Zt, nz, fl, nf, B);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    char m[300];
    snprintf(m, sizeof(m), "mpet_constraint_rows: %s", hipGetErrorString(e));
    mpet_set_error(m);
    return ERR_HIP;
  }
  return OK;
}
