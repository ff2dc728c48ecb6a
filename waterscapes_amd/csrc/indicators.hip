// A-posteriori error indicators of the two-field MPET step
// (AdaptiveMPETSolver.__init_error_estimators / estimate_error_at_t,
// mpet/adaptivempetsolver.py:82-185), per cell (DG0):
//
//   zeta_p(K)    = sum_j  h_K^2 ||RK_p,j||_K^2 + sum_F avg(h) avg(w) ||RE_p,j||_F^2
//   zeta_u(K)    =        h_K^2 ||RK_u||_K^2   + sum_F avg(h) avg(w) ||RE_u||_F^2
//   zeta_udt(K)  = the same with (RK_u - RK_u^-)/dt and (RE_u - RE_u^-)/dt
//   r4(K)        = the cell part of R4p (:146-153)
//
//   RK_u   = f + div sigma(u) - sum_i alpha_i grad p_i        (:121)
//   RE_u   = sum_i jump(alpha_i p_i, n) - jump(sigma(u), n)     (:122)
//   RK_p,j = -g_j - c_j (p_j - p_j^-)/dt - alpha_j div(u - u^-)/dt
//            + K_j div grad p_j - sum_i S_ij (p_j - p_i)        (:130-132)
//   RE_p,j = -K_j jump(grad p_j, n)                             (:133)
//   sigma(u) = 2 mu eps(u) + lambda div u I, h = cell diameter (longest
//   edge), the facet sums over interior facets (dS) with avg(w) = 1/2 for
//   each of the two cells.
//
// u is P2 (div sigma(u) is constant per cell), p_j P1 (div grad p_j = 0).
// Cell integrals use a Grundmann-Moeller rule exact for the integrand
// degree, facet integrals the 3-point edge-midpoint rule (exact for the
// quadratic |RE_u|^2).  One thread per cell / per interior facet; the facet
// values are summed into the cells by a gather kernel (deterministic).
#include "mpet_ops.h"
#include <mutex>
#include <stdio.h>

namespace mpet {

struct IGeom {
  double g[4][3];   // grad lambda_k
  double v0[3];
  double vol, h;
};

__device__ __forceinline__ void igeom(const double* __restrict__ X, const int* __restrict__ cells,
                                      int e, IGeom& G) {
  double v[4][3];
  for (int k = 0; k < 4; ++k)
    for (int d = 0; d < 3; ++d) v[k][d] = X[3ll * cells[4 * e + k] + d];
  double J[3][3];
  for (int d = 0; d < 3; ++d)
    for (int k = 0; k < 3; ++k) J[d][k] = v[k + 1][d] - v[0][d];
  const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  const double id = 1.0 / det;
  // rows of J^-1 = grad lambda_1..3
  double Ji[3][3];
  Ji[0][0] = c00 * id;
  Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
  Ji[1][0] = c01 * id;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
  Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
  Ji[2][0] = c02 * id;
  Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
  for (int m = 0; m < 3; ++m) {
    G.g[1][m] = Ji[0][m];
    G.g[2][m] = Ji[1][m];
    G.g[3][m] = Ji[2][m];
    G.g[0][m] = -(Ji[0][m] + Ji[1][m] + Ji[2][m]);
    G.v0[m] = v[0][m];
  }
  G.vol = fabs(det) / 6.0;
  double h2 = 0.0;
  for (int i = 0; i < 4; ++i)
    for (int j = i + 1; j < 4; ++j) {
      double s = 0.0;
      for (int d = 0; d < 3; ++d) s += (v[i][d] - v[j][d]) * (v[i][d] - v[j][d]);
      h2 = fmax(h2, s);
    }
  G.h = sqrt(h2);
}

// P2 nodes: vertices 0..3, edges (2,3), (1,3), (1,2), (0,3), (0,2), (0,1).
// The node loops below are unrolled, so every index into the geometry and
// the barycentric point is a compile-time constant (no private-memory
// arrays).
__device__ constexpr int kEdge[6][2] = {{2, 3}, {1, 3}, {1, 2}, {0, 3}, {0, 2}, {0, 1}};

// grad phi_a at barycentric point l: sum_k dphi_a/dlambda_k grad lambda_k.
__device__ __forceinline__ void p2_grad(const IGeom& G, const double l[4], int a, double out[3]) {
  if (a < 4) {
    const double s = 4.0 * l[a] - 1.0;
    for (int m = 0; m < 3; ++m) out[m] = s * G.g[a][m];
  } else {
    const int i = kEdge[a - 4][0], j = kEdge[a - 4][1];
    for (int m = 0; m < 3; ++m) out[m] = 4.0 * (l[j] * G.g[i][m] + l[i] * G.g[j][m]);
  }
}

// Hessian of phi_a (constant): 4 g_a g_a^T (vertex), 4 (g_i g_j^T + g_j g_i^T) (edge).
__device__ __forceinline__ void p2_hess(const IGeom& G, int a, double H[3][3]) {
  for (int m = 0; m < 3; ++m)
    for (int q = 0; q < 3; ++q) {
      if (a < 4) {
        H[m][q] = 4.0 * G.g[a][m] * G.g[a][q];
      } else {
        const int i = kEdge[a - 4][0], j = kEdge[a - 4][1];
        H[m][q] = 4.0 * (G.g[i][m] * G.g[j][q] + G.g[j][m] * G.g[i][q]);
      }
    }
}

// grad u (G[i][j] = d_j u_i) at barycentric point l of cell geometry G.
__device__ __forceinline__ void grad_u(const IGeom& Gm, const double l[4],
                                       const double* __restrict__ x, const int* p2,
                                       double Gu[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Gu[i][j] = 0.0;
#pragma unroll
  for (int a = 0; a < 10; ++a) {
    double gp[3];
    p2_grad(Gm, l, a, gp);
    const long long b = 3ll * p2[a];
    for (int i = 0; i < 3; ++i) {
      const double ui = x[b + i];
      for (int j = 0; j < 3; ++j) Gu[i][j] += ui * gp[j];
    }
  }
}

// div sigma(u) = mu Lap u + (mu + lambda) grad div u (constant on the cell).
__device__ __forceinline__ void div_sigma(const IGeom& Gm, const double* __restrict__ x,
                                          const int* p2, double mu, double lm, double out[3]) {
  out[0] = out[1] = out[2] = 0.0;
#pragma unroll
  for (int a = 0; a < 10; ++a) {
    double H[3][3];
    p2_hess(Gm, a, H);
    const double tr = H[0][0] + H[1][1] + H[2][2];
    const long long b = 3ll * p2[a];
    for (int i = 0; i < 3; ++i) {
      double s = mu * tr * x[b + i];
      for (int j = 0; j < 3; ++j) s += (mu + lm) * H[i][j] * x[b + j];
      out[i] += s;
    }
  }
}

__global__ void __launch_bounds__(128)
ind_cell_kernel(MpetIndDesc d, const double* __restrict__ x, const double* __restrict__ xp,
                double* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= d.nC) return;
  const int A = d.A;
  const long long nu = 3ll * d.nP2;
  IGeom G;
  igeom(d.coords, d.cells, e, G);
  int p2[10], v[4];
  for (int a = 0; a < 10; ++a) p2[a] = d.cell_p2[10 * e + a];
  for (int k = 0; k < 4; ++k) v[k] = d.cells[4 * e + k];
  // constant parts of RK_u, RK_u^-: div sigma(u) - sum_i alpha_i grad p_i
  double cu[3], cup[3];
  div_sigma(G, x, p2, d.mu, d.lmbda, cu);
  div_sigma(G, xp, p2, d.mu, d.lmbda, cup);
  double r4 = 0.0;
#pragma unroll
  for (int j = 0; j < MPET_MAXF; ++j) {
    if (j >= A) continue;
    double gp[3] = {0, 0, 0}, gpp[3] = {0, 0, 0};
    for (int k = 0; k < 4; ++k) {
      const long long q = nu + (long long)A * v[k] + j;
      for (int m = 0; m < 3; ++m) {
        gp[m] += x[q] * G.g[k][m];
        gpp[m] += xp[q] * G.g[k][m];
      }
    }
    for (int m = 0; m < 3; ++m) {
      cu[m] -= d.alpha[j] * gp[m];
      cup[m] -= d.alpha[j] * gpp[m];
      r4 += d.K[j] * (gp[m] - gpp[m]) * (gp[m] - gpp[m]) * G.vol;
    }
  }
  const double idt = 1.0 / d.dt;
  // div(u - u^-) is linear on the cell: its values at the four vertices
  double dv[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double lk[4] = {k == 0 ? 1.0 : 0.0, k == 1 ? 1.0 : 0.0, k == 2 ? 1.0 : 0.0,
                          k == 3 ? 1.0 : 0.0};
    double sdv = 0.0;
#pragma unroll
    for (int a = 0; a < 10; ++a) {
      double gphi[3];
      p2_grad(G, lk, a, gphi);
      const long long b = 3ll * p2[a];
      for (int i = 0; i < 3; ++i) sdv += (x[b + i] - xp[b + i]) * gphi[i];
    }
    dv[k] = sdv;
  }
  double zu = 0.0, zudt = 0.0, zp = 0.0, r4s = 0.0;
  for (int q = 0; q < d.nq; ++q) {
    const double* l = d.qbary + 4 * q;
    const double w = d.qw[q];
    double fq[3] = {0, 0, 0}, fpq[3] = {0, 0, 0};
    if (d.f)
      for (int m = 0; m < d.nkf; ++m) {
        const double L = d.Lf[q * d.nkf + m];
        for (int i = 0; i < 3; ++i) fq[i] += L * d.f[(3ll * (long long)e * d.nkf) + 3 * m + i];
      }
    if (d.fp)
      for (int m = 0; m < d.nkf; ++m) {
        const double L = d.Lf[q * d.nkf + m];
        for (int i = 0; i < 3; ++i) fpq[i] += L * d.fp[(3ll * (long long)e * d.nkf) + 3 * m + i];
      }
    double ru2 = 0.0, rd2 = 0.0;
    for (int i = 0; i < 3; ++i) {
      const double r = fq[i] + cu[i], rp = fpq[i] + cup[i];
      ru2 += r * r;
      rd2 += (r - rp) * idt * (r - rp) * idt;
    }
    zu += w * ru2;
    zudt += w * rd2;
    // div(u - u^-) at the point (linear on the cell)
    const double du = l[0] * dv[0] + l[1] * dv[1] + l[2] * dv[2] + l[3] * dv[3];
    double pj[MPET_MAXF], ppj[MPET_MAXF];
#pragma unroll
    for (int j = 0; j < MPET_MAXF; ++j) {
      pj[j] = ppj[j] = 0.0;
      if (j >= A) continue;
      for (int k = 0; k < 4; ++k) {
        const long long qq = nu + (long long)A * v[k] + j;
        pj[j] += l[k] * x[qq];
        ppj[j] += l[k] * xp[qq];
      }
    }
#pragma unroll
    for (int j = 0; j < MPET_MAXF; ++j) {
      if (j >= A) continue;
      double gj = 0.0;
      if (d.g)
        for (int m = 0; m < d.nkg; ++m)
          gj += d.Lg[q * d.nkg + m] * d.g[((long long)e * d.nkg + m) * A + j];
      double r = -gj - d.c[j] * (pj[j] - ppj[j]) * idt - d.alpha[j] * du * idt;
#pragma unroll
      for (int i = 0; i < MPET_MAXF; ++i)
        if (i < A) r -= d.S[i * MPET_MAXF + j] * (pj[j] - pj[i]);
      zp += w * r * r;
#pragma unroll
      for (int i = 0; i < MPET_MAXF; ++i)
        if (i < A && d.S[i * MPET_MAXF + j] > 3.0e-16)
          r4s += w * d.S[i * MPET_MAXF + j] * ((pj[j] - ppj[j]) - (pj[i] - ppj[i])) *
                 (pj[j] - ppj[j]);
    }
  }
  const double h2v = G.h * G.h * G.vol;
  out[e] = zp * h2v;
  out[(long long)d.nC + e] = zu * h2v;
  out[2ll * d.nC + e] = zudt * h2v;
  out[3ll * d.nC + e] = r4 + r4s * G.vol;
}

// Barycentric coordinates of point P in cell geometry G.
__device__ __forceinline__ void bary_of(const IGeom& G, const double P[3], double l[4]) {
  double s = 0.0;
  for (int k = 1; k < 4; ++k) {
    l[k] = G.g[k][0] * (P[0] - G.v0[0]) + G.g[k][1] * (P[1] - G.v0[1]) +
           G.g[k][2] * (P[2] - G.v0[2]);
    s += l[k];
  }
  l[0] = 1.0 - s;
}

__constant__ int c_face[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};

// sigma(u) n at barycentric point l of a cell (geometry G, P2 nodes p2).
__device__ __forceinline__ void traction(const IGeom& G, const double l[4],
                                         const double* __restrict__ x,
                                         const int* __restrict__ p2, double mu, double lm,
                                         const double n[3], double t[3]) {
  double Gu[3][3];
  grad_u(G, l, x, p2, Gu);
  const double tr = Gu[0][0] + Gu[1][1] + Gu[2][2];
  for (int i = 0; i < 3; ++i) {
    double s = lm * tr * n[i];
    for (int j = 0; j < 3; ++j) s += mu * (Gu[i][j] + Gu[j][i]) * n[j];
    t[i] = s;
  }
}

// sum_j alpha_j p_j at barycentric point l of a cell with vertices v.
__device__ __forceinline__ double alpha_p(const MpetIndDesc& d, const double l[4],
                                          const double* __restrict__ x,
                                          const int* __restrict__ v, long long nu) {
  double s = 0.0;
#pragma unroll
  for (int j = 0; j < MPET_MAXF; ++j) {
    if (j >= d.A) continue;
    double p = 0.0;
    for (int k = 0; k < 4; ++k) p += l[k] * x[nu + (long long)d.A * v[k] + j];
    s += d.alpha[j] * p;
  }
  return s;
}

// One interior facet: fac = (cell a, local face a, cell b, local face b).
// out_f[0..2][t] = avg(h) int_F |RE_p|^2 (sum over j), |RE_u|^2, |RE_u,dt|^2.
// The tractions of the two cells are formed one after the other (one
// displacement gradient live at a time).
__global__ void __launch_bounds__(128)
ind_facet_kernel(MpetIndDesc d, const double* __restrict__ x, const double* __restrict__ xp,
                 const int* __restrict__ fac, int nF, double* __restrict__ out_f) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nF) return;
  const int ca = fac[4 * t], la = fac[4 * t + 1], cb = fac[4 * t + 2];
  const int A = d.A;
  const long long nu = 3ll * d.nP2;
  IGeom Ga, Gb;
  igeom(d.coords, d.cells, ca, Ga);
  igeom(d.coords, d.cells, cb, Gb);
  const int* pa = d.cell_p2 + 10ll * ca;
  const int* pb = d.cell_p2 + 10ll * cb;
  const int* va = d.cells + 4ll * ca;
  const int* vb = d.cells + 4ll * cb;
  // facet vertices (those of cell a's face la), outward normal of cell a
  double P[3][3];
  for (int k = 0; k < 3; ++k)
    for (int m = 0; m < 3; ++m) P[k][m] = d.coords[3ll * va[c_face[la][k]] + m];
  double nrm[3];
  {
    double e1[3], e2[3];
    for (int m = 0; m < 3; ++m) {
      e1[m] = P[1][m] - P[0][m];
      e2[m] = P[2][m] - P[0][m];
    }
    nrm[0] = e1[1] * e2[2] - e1[2] * e2[1];
    nrm[1] = e1[2] * e2[0] - e1[0] * e2[2];
    nrm[2] = e1[0] * e2[1] - e1[1] * e2[0];
  }
  const double len = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
  const double area = 0.5 * len;
  double opp = 0.0;   // (opposite vertex of a - P0) . n  < 0 for the outward normal
  for (int m = 0; m < 3; ++m) opp += (d.coords[3ll * va[la] + m] - P[0][m]) * nrm[m];
  const double sg = opp > 0.0 ? -1.0 / len : 1.0 / len;
  for (int m = 0; m < 3; ++m) nrm[m] *= sg;
  const double havg = 0.5 * (Ga.h + Gb.h);
  const double idt = 1.0 / d.dt;
  // RE_p,j: constant on the facet
  double fp = 0.0;
#pragma unroll
  for (int j = 0; j < MPET_MAXF; ++j) {
    if (j >= A) continue;
    double jn = 0.0;
    for (int k = 0; k < 4; ++k) {
      const double xa = x[nu + (long long)A * va[k] + j], xb = x[nu + (long long)A * vb[k] + j];
      for (int m = 0; m < 3; ++m) jn += (xa * Ga.g[k][m] - xb * Gb.g[k][m]) * nrm[m];
    }
    const double r = -d.K[j] * jn;
    fp += r * r;
  }
  fp *= area;
  // RE_u = sum_j alpha_j (p_j^a - p_j^b) n - (sigma_a - sigma_b) n, and the
  // same for the previous state, at the three edge midpoints.  The four
  // (cell, state) tractions are formed one at a time, each cell's geometry
  // recomputed, so that only one displacement gradient is live.
  double fu = 0.0, fud = 0.0;
#pragma unroll 1
  for (int q = 0; q < 3; ++q) {
    double X[3];
    for (int m = 0; m < 3; ++m)   // midpoint of edge q (static indices into P)
      X[m] = 0.5 * (q == 0 ? P[0][m] + P[1][m] : (q == 1 ? P[1][m] + P[2][m] : P[2][m] + P[0][m]));
    double re[3] = {0.0, 0.0, 0.0}, rp[3] = {0.0, 0.0, 0.0};
#pragma unroll 1
    for (int c4 = 0; c4 < 4; ++c4) {
      const int cell = (c4 & 1) ? cb : ca;
      const double* xs = (c4 & 2) ? xp : x;
      IGeom G;
      igeom(d.coords, d.cells, cell, G);
      double l[4];
      bary_of(G, X, l);
      double tt[3];
      traction(G, l, xs, d.cell_p2 + 10ll * cell, d.mu, d.lmbda, nrm, tt);
      const double ap = alpha_p(d, l, xs, d.cells + 4ll * cell, nu);
      const double sg = (c4 & 1) ? 1.0 : -1.0;     // cell a: + p_a n - t_a; cell b: - p_b n + t_b
      for (int i = 0; i < 3; ++i) {
        const double c = sg * (tt[i] - ap * nrm[i]);
        if (c4 & 2) rp[i] += c; else re[i] += c;
      }
    }
    double a2 = 0.0, d2 = 0.0;
    for (int i = 0; i < 3; ++i) {
      a2 += re[i] * re[i];
      const double r = (re[i] - rp[i]) * idt;
      d2 += r * r;
    }
    fu += a2 * (area / 3.0);
    fud += d2 * (area / 3.0);
  }
  out_f[t] = havg * fp;
  out_f[(long long)nF + t] = havg * fu;
  out_f[2ll * nF + t] = havg * fud;
}

// out[k][e] += 1/2 sum over the interior facets of cell e of out_f[k].
__global__ void __launch_bounds__(256)
ind_gather_kernel(int nC, const int* __restrict__ cell_f, const double* __restrict__ out_f,
                  int nF, double* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nC) return;
  for (int k = 0; k < 3; ++k) {
    double s = 0.0;
    for (int l = 0; l < 4; ++l) {
      const int f = cell_f[4 * e + l];
      if (f >= 0) s += out_f[(long long)k * nF + f];
    }
    out[(long long)k * nC + e] += 0.5 * s;
  }
}

}  // namespace mpet

using namespace mpet;

static int launch_indicators(const MpetIndDesc* d, const double* x, const double* xp,
                             const int* fac, int nF, const int* cell_f, double* tmp,
                             double* out, hipStream_t stream, const char* who) {
  hipLaunchKernelGGL(ind_cell_kernel, dim3((d->nC + 127) / 128), dim3(128), 0, stream, *d, x, xp,
                     out);
  if (nF > 0) {
    hipLaunchKernelGGL(ind_facet_kernel, dim3((nF + 127) / 128), dim3(128), 0, stream, *d, x,
                       xp, fac, nF, tmp);
    hipLaunchKernelGGL(ind_gather_kernel, dim3((d->nC + 255) / 256), dim3(256), 0, stream,
                       d->nC, cell_f, (const double*)tmp, nF, out);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    char m[300];
    snprintf(m, sizeof(m), "%s: %s", who, hipGetErrorString(e));
    mpet_set_error(m);
    return ERR_HIP;
  }
  return OK;
}

static bool indicators_args_ok(const MpetIndDesc* d, const double* x, const double* xp,
                               const double* out, const int* fac, int nF, const int* cell_f) {
  return d && x && xp && out && d->coords && d->cells && d->cell_p2 && d->nC >= 1 &&
         d->A >= 1 && d->A <= MPET_MAXF && d->nq >= 1 && d->qbary && d->qw && d->dt > 0.0 &&
         !((d->f || d->fp) && (!d->Lf || d->nkf < 1)) && !(d->g && (!d->Lg || d->nkg < 1)) &&
         nF >= 0 && !(nF > 0 && (!fac || !cell_f));
}

extern "C" int mpet_error_indicators(const MpetIndDesc* d, const double* x, const double* xp,
                                     const int* fac, int nF, const int* cell_f, double* tmp,
                                     double* out, hipStream_t stream) {
  if (!indicators_args_ok(d, x, xp, out, fac, nF, cell_f) || (nF > 0 && !tmp)) {
    mpet_set_error("mpet_error_indicators: invalid argument");
    return ERR_INVALID;
  }
  return launch_indicators(d, x, xp, fac, nF, cell_f, tmp, out, stream,
                           "mpet_error_indicators");
}

// ---------------------------------------------------------------------------
// Allocation-free variant: the facet values (3 doubles per interior facet,
// the tmp workspace above) live in a file-scope device array, one copy per
// GPU for the life of the process.  Every facet value is written by the facet
// kernel before the gather reads it, so nothing needs clearing between
// calls; calls that overlap on one GPU share the array and race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[3ll * MPET_NOALLOC_MAX_FACETS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};
}  // namespace

// Device address of this GPU's copy of scratch, looked up once per device
// (with that device current) and kept.  The array's name is a host-side
// stand-in; only this address may be passed to kernels.
static int scratch_address(double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    mpet_set_error("mpet_error_indicators_noalloc: no current device");
    return ERR_HIP;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      char m[300];
      snprintf(m, sizeof(m), "mpet_error_indicators_noalloc: hipGetSymbolAddress: %s",
               hipGetErrorString(e));
      mpet_set_error(m);
      return ERR_HIP;
    }
    g_scratch[dev] = (double*)p;
  }
  *out = g_scratch[dev];
  return OK;
}

extern "C" int mpet_error_indicators_noalloc(const MpetIndDesc* d, const double* x,
                                             const double* xp, const int* fac, int nF,
                                             const int* cell_f, double* out,
                                             hipStream_t stream) {
  if (!indicators_args_ok(d, x, xp, out, fac, nF, cell_f)) {
    mpet_set_error("mpet_error_indicators_noalloc: invalid argument");
    return ERR_INVALID;
  }
  if (nF > MPET_NOALLOC_MAX_FACETS) {
    char m[200];
    snprintf(m, sizeof(m),
             "mpet_error_indicators_noalloc: %d interior facets exceed the maximum %d "
             "(use mpet_error_indicators with a workspace)",
             nF, (int)MPET_NOALLOC_MAX_FACETS);
    mpet_set_error(m);
    return ERR_INVALID;
  }
  double* tmp = nullptr;
  if (nF > 0) {
    const int rc = scratch_address(&tmp);
    if (rc) return rc;
  }
  return launch_indicators(d, x, xp, fac, nF, cell_f, tmp, out, stream,
                           "mpet_error_indicators_noalloc");
}
