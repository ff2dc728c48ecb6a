// Exterior-facet terms of the total-pressure forms (FP64, deterministic):
//   traction  int_{ds_m(1)} s . v                     (L0, :313)
//   flux      dt int_{ds_i(1)} I_i w_i                 (L1_i, :322)
//   Robin     dt int_{ds_i(2)} beta_i ((1-theta) p_i^- - p_robin,i) w_i  (rhs(L2_i), :324)
//             -dt theta int_{ds_i(2)} beta_i p_i w_i   (lhs(L2_i), :324, into A)
// (mpet/mpettotalpressuresolver.py).  Coefficients are interpolated into
// P_k on each facet (Expression semantics) and products are integrated
// exactly with reference-triangle tables.
//
// Two steps, both without atomics:
//   facet_local_kernel   one thread per (facet, test node, component):
//       loc[f][a][c] = scale |F_f| sum_m sum_n T[a][m][n] u[f][m][c] w[f][n]
//     (u from an array, or gathered from a vector through per-facet DOFs;
//      w optional);
//   node-major scatter   one thread per (boundary node, component) sums its
//     incident (facet, local index) pairs in a fixed order, so every output
//     entry has a single writer and a fixed summation order.
#include "mpet_ops.h"

namespace mpet {

constexpr int FNT = 256;

__global__ void __launch_bounds__(FNT)
facet_local_kernel(int nF, int ntest, int nm, int nn, int ncomp, const double* __restrict__ T,
                   const double* __restrict__ area, const double* __restrict__ u,
                   long long u_fstride, const double* __restrict__ xg,
                   const int* __restrict__ gdof, const double* __restrict__ w, double scale,
                   double* __restrict__ loc) {
  const long long t = (long long)blockIdx.x * FNT + threadIdx.x;
  if (t >= (long long)nF * ntest * ncomp) return;
  const int c = (int)(t % ncomp);
  const int a = (int)((t / ncomp) % ntest);
  const int f = (int)(t / ((long long)ncomp * ntest));
  double acc = 0.0;
  for (int m = 0; m < nm; ++m) {
    const double um = xg ? xg[gdof[(long long)f * nm + m]]
                         : u[f * u_fstride + (long long)m * ncomp + c];
    double s = 0.0;
    for (int n = 0; n < nn; ++n) {
      const double wn = w ? w[(long long)f * nn + n] : 1.0;
      s += T[(a * nm + m) * nn + n] * wn;
    }
    acc += um * s;
  }
  loc[t] = scale * area[f] * acc;
}

// b[row_dof[r] + c] += sum over (f, a) in inc(r) of loc[f][a][c]
__global__ void __launch_bounds__(FNT)
facet_scatter_vec_kernel(int nrows, int ntest, int ncomp, const long long* __restrict__ row_dof,
                         const int* __restrict__ inc_rp, const int* __restrict__ inc,
                         const double* __restrict__ loc, double* __restrict__ b) {
  const long long t = (long long)blockIdx.x * FNT + threadIdx.x;
  if (t >= (long long)nrows * ncomp) return;
  const int r = (int)(t / ncomp), c = (int)(t % ncomp);
  double s = 0.0;
  for (int k = inc_rp[r]; k < inc_rp[r + 1]; ++k) {
    const int fa = inc[k];                       // f * ntest + a
    s += loc[(long long)fa * ncomp + c];
  }
  b[row_dof[r] + c] += s;
}

// vals[pos(r, node(f, b))] += loc[f][a][b] over (f, a) in inc(r): a
// per-facet 3x3 matrix scattered into the P1 node pattern (row = node r).
__global__ void __launch_bounds__(FNT)
facet_scatter_mat_kernel(int nrows, const int* __restrict__ row_node,
                         const int* __restrict__ inc_rp, const int* __restrict__ inc,
                         const int* __restrict__ fnode, const int* __restrict__ rp11,
                         const int* __restrict__ ci11, const double* __restrict__ loc,
                         double* __restrict__ vals) {
  const int r = blockIdx.x * FNT + threadIdx.x;
  if (r >= nrows) return;
  const int node = row_node[r];
  const int lo = rp11[node], hi = rp11[node + 1];
  for (int k = inc_rp[r]; k < inc_rp[r + 1]; ++k) {
    const int fa = inc[k];
    const int f = fa / 3;
    for (int bcol = 0; bcol < 3; ++bcol) {
      const int col = fnode[3 * f + bcol];
      const int pos = lower_bound_i32(ci11, lo, hi, col);
      if (pos < hi && ci11[pos] == col) vals[pos] += loc[(long long)fa * 3 + bcol];
    }
  }
}

}  // namespace mpet

using namespace mpet;

extern "C" int mpet_facet_local(int nF, int ntest, int nm, int nn, int ncomp, const double* T,
                                const double* area, const double* u, long long u_fstride,
                                const double* xg, const int* gdof, const double* w,
                                double scale, double* loc, hipStream_t s) {
  if (nF < 0 || ntest < 1 || nm < 1 || nn < 1 || ncomp < 1 || !T || !area || !loc ||
      (!u && !(xg && gdof)) || (xg && ncomp != 1)) {
    mpet_set_error("mpet_facet_local: invalid argument");
    return ERR_INVALID;
  }
  if (nF == 0) return OK;
  const long long items = (long long)nF * ntest * ncomp;
  hipLaunchKernelGGL(facet_local_kernel, dim3((unsigned)((items + FNT - 1) / FNT)), dim3(FNT),
                     0, s, nF, ntest, nm, nn, ncomp, T, area, u, u_fstride, xg, gdof, w, scale,
                     loc);
  return hipGetLastError() == hipSuccess ? OK : ERR_HIP;
}

extern "C" int mpet_facet_scatter_vec(int nrows, int ntest, int ncomp, const long long* row_dof,
                                      const int* inc_rp, const int* inc, const double* loc,
                                      double* b, hipStream_t s) {
  if (nrows < 0 || ntest < 1 || ncomp < 1 || (nrows && (!row_dof || !inc_rp || !inc || !loc || !b))) {
    mpet_set_error("mpet_facet_scatter_vec: invalid argument");
    return ERR_INVALID;
  }
  if (nrows == 0) return OK;
  const long long items = (long long)nrows * ncomp;
  hipLaunchKernelGGL(facet_scatter_vec_kernel, dim3((unsigned)((items + FNT - 1) / FNT)),
                     dim3(FNT), 0, s, nrows, ntest, ncomp, row_dof, inc_rp, inc, loc, b);
  return hipGetLastError() == hipSuccess ? OK : ERR_HIP;
}

extern "C" int mpet_facet_scatter_mat(int nrows, const int* row_node, const int* inc_rp,
                                      const int* inc, const int* fnode, const int* rp11,
                                      const int* ci11, const double* loc, double* vals,
                                      hipStream_t s) {
  if (nrows < 0 || (nrows && (!row_node || !inc_rp || !inc || !fnode || !rp11 || !ci11 ||
                              !loc || !vals))) {
    mpet_set_error("mpet_facet_scatter_mat: invalid argument");
    return ERR_INVALID;
  }
  if (nrows == 0) return OK;
  hipLaunchKernelGGL(facet_scatter_mat_kernel, dim3((nrows + FNT - 1) / FNT), dim3(FNT), 0, s,
                     nrows, row_node, inc_rp, inc, fnode, rp11, ci11, loc, vals);
  return hipGetLastError() == hipSuccess ? OK : ERR_HIP;
}
