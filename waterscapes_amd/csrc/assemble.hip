// Element assembly of the total-pressure MPET operator (FP64, deterministic).
//
// Forms (mpet/mpettotalpressuresolver.py:267-283):
//   u-u   2 mu eps(u):eps(v)            -> Kuu  (3x3 block per P2 node pair)
//   prec  grad u : grad v (unscaled)    -> Lap  (scalar per P2 node pair)
//   p0-v  p0 div v / u-w0  div u w0     -> Dv (P2 row, P1 col, 3 comps), Dt
//   p-p   P1 mass, K_i-weighted P1 stiffness -> M1, St[i]
// Every other entry of A and P is a linear combination of these primitives
// with (A+1)x(A+1) scalar coefficients (see expand_csr_kernel).
//
// Cells are processed one colour at a time; cells of one colour share no
// node, so each matrix row is written by at most one cell per launch and the
// summation order of every entry is fixed: assembly is bitwise reproducible
// without atomics.
#include "mpet_common.h"
#include <vector>
#include "../../include/mpet_hip.h"

namespace mpet {

__constant__ double c_tab[MPET_TAB_SIZE];

struct CellGeom {
  double g[4][3];   // gradients of the barycentric coordinates
  double vol;       // |K|
};

__device__ __forceinline__ void cell_geometry(const double* __restrict__ X,
                                              const int* __restrict__ cells, int e,
                                              CellGeom& G) {
  int v0 = cells[4 * e + 0], v1 = cells[4 * e + 1], v2 = cells[4 * e + 2],
      v3 = cells[4 * e + 3];
  double x0 = X[3 * v0], y0 = X[3 * v0 + 1], z0 = X[3 * v0 + 2];
  // J columns = v_k - v_0
  double J[3][3];
  J[0][0] = X[3 * v1] - x0; J[1][0] = X[3 * v1 + 1] - y0; J[2][0] = X[3 * v1 + 2] - z0;
  J[0][1] = X[3 * v2] - x0; J[1][1] = X[3 * v2 + 1] - y0; J[2][1] = X[3 * v2 + 2] - z0;
  J[0][2] = X[3 * v3] - x0; J[1][2] = X[3 * v3 + 1] - y0; J[2][2] = X[3 * v3 + 2] - z0;
  double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  double id = 1.0 / det;
  // rows of J^{-1} are grad lambda_1..3
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
  }
  G.vol = fabs(det) / 6.0;
}

// M[m][n] = int d_m phi_a d_n phi_b
__device__ __forceinline__ void p2_grad_gram(const CellGeom& G, int a, int b, double M[3][3]) {
  for (int m = 0; m < 3; ++m)
    for (int n = 0; n < 3; ++n) M[m][n] = 0.0;
  const double* W = c_tab + MPET_TAB_W + (a * 10 + b) * 16;
  for (int k = 0; k < 4; ++k) {
    for (int l = 0; l < 4; ++l) {
      double w = W[k * 4 + l];
      if (w == 0.0) continue;
      w *= G.vol;
      for (int m = 0; m < 3; ++m) {
        double gm = w * G.g[k][m];
        for (int n = 0; n < 3; ++n) M[m][n] += gm * G.g[l][n];
      }
    }
  }
}

// int psi_c d_m phi_b, m = 0..2
__device__ __forceinline__ void p2p1_div(const CellGeom& G, int b, int c, double d[3]) {
  const double* V = c_tab + MPET_TAB_V + (b * 4 + c) * 4;
  d[0] = d[1] = d[2] = 0.0;
  for (int k = 0; k < 4; ++k) {
    double w = V[k] * G.vol;
    d[0] += w * G.g[k][0];
    d[1] += w * G.g[k][1];
    d[2] += w * G.g[k][2];
  }
}

// Lane `task` of cell e: tasks 0..9 own the P2 rows of the cell, 10..13
// its P1 rows (each row written by one lane of one cell of the colour).
__device__ __forceinline__ void assemble_cell(const double* __restrict__ X,
                                              const int* __restrict__ cells,
                                              const int* __restrict__ cell_p2, int e, int task,
                                              const double* __restrict__ Kcell, int A,
                                              double mu, const MpetPatterns& pat,
                                              const MpetPrimitives& out) {
  CellGeom G;
  cell_geometry(X, cells, e, G);
  if (task < 10) {
    const int a = task;
    const int row = cell_p2[10 * e + a];
    const int rb = pat.rp22[row], re = pat.rp22[row + 1];
    for (int b = 0; b < 10; ++b) {
      int col = cell_p2[10 * e + b];
      int pos = lower_bound_i32(pat.ci22, rb, re, col);
      double M[3][3];
      p2_grad_gram(G, a, b, M);
      double tr = M[0][0] + M[1][1] + M[2][2];
      double* K = out.Kuu + 9ll * pos;
      // row (a,i), col (b,j): mu (delta_ij tr M + M[j][i])
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) K[3 * i + j] += mu * ((i == j ? tr : 0.0) + M[j][i]);
      out.Lap[pos] += tr;
    }
    const int qb = pat.rp21[row], qe = pat.rp21[row + 1];
    for (int c = 0; c < 4; ++c) {
      int col = cells[4 * e + c];
      int pos = lower_bound_i32(pat.ci21, qb, qe, col);
      double d[3];
      p2p1_div(G, a, c, d);
      double* D = out.Dv + 3ll * pos;
      D[0] += d[0]; D[1] += d[1]; D[2] += d[2];
    }
  } else {
    const int c = task - 10;
    const int row = cells[4 * e + c];
    const int qb = pat.rp12[row], qe = pat.rp12[row + 1];
    for (int b = 0; b < 10; ++b) {
      int col = cell_p2[10 * e + b];
      int pos = lower_bound_i32(pat.ci12, qb, qe, col);
      double d[3];
      p2p1_div(G, b, c, d);
      double* D = out.Dt + 3ll * pos;
      D[0] += d[0]; D[1] += d[1]; D[2] += d[2];
    }
    const int sb = pat.rp11[row], se = pat.rp11[row + 1];
    const long long nnz11 = pat.nnz11;
    for (int d = 0; d < 4; ++d) {
      int col = cells[4 * e + d];
      int pos = lower_bound_i32(pat.ci11, sb, se, col);
      out.M1[pos] += G.vol * (c == d ? 0.1 : 0.05);
      double s = G.vol * (G.g[c][0] * G.g[d][0] + G.g[c][1] * G.g[d][1] + G.g[c][2] * G.g[d][2]);
      for (int i = 0; i < A; ++i) out.St[i * nnz11 + pos] += Kcell[(long long)e * A + i] * s;
    }
  }
}

// One lane per (cell, local row node): lanes 0..9 own the P2 rows of the
// cell, lanes 10..13 its P1 rows; 16 lanes per cell, 4 cells per wave.
__global__ void __launch_bounds__(256)
assemble_colour_kernel(const double* __restrict__ X, const int* __restrict__ cells,
                       const int* __restrict__ cell_p2, const int* __restrict__ colour_cells,
                       int cbeg, int cend, const double* __restrict__ Kcell, int A, double mu,
                       MpetPatterns pat, MpetPrimitives out) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int ci = cbeg + t / 16;
  int task = t % 16;
  if (ci >= cend || task >= 14) return;
  assemble_cell(X, cells, cell_p2, colour_cells[ci], task, Kcell, A, mu, pat, out);
}

// One instance of a batched assembly (device copies of its descriptors).
struct AsmInst {
  const double* X;
  const int* cells;
  const int* cell_p2;
  const int* colour_cells;
  const double* Kcell;
  int A;
  double mu;
  MpetPatterns pat;
  MpetPrimitives out;
};

// Colour k of every instance in one launch: cells g of the colour, counted
// over the instances in order (off[i] .. off[i+1]-1 belong to instance i,
// its colour starting at colour_cells[cb[i]]); the same lanes and order as
// assemble_colour_kernel, so each instance's sums are those of its own
// assembly.
__global__ void __launch_bounds__(256)
assemble_colour_batched_kernel(const AsmInst* __restrict__ inst, int ninst,
                               const long long* __restrict__ off,
                               const long long* __restrict__ cb) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long g = t / 16;
  const int task = (int)(t % 16);
  if (g >= off[ninst] || task >= 14) return;
  int lo = 0, hi = ninst;                 // off[lo] <= g < off[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= g) lo = mid; else hi = mid;
  }
  const AsmInst& I = inst[lo];
  const long long ci = cb[lo] + (g - off[lo]);
  assemble_cell(I.X, I.cells, I.cell_p2, I.colour_cells[ci], task, I.Kcell, I.A, I.mu, I.pat,
                I.out);
}

// Zeroes segments seg[i] = (pointer, count) of doubles, one block row per
// segment.
struct ZeroSeg {
  double* p;
  long long n;
};
__global__ void __launch_bounds__(256)
zero_segments_kernel(const ZeroSeg* __restrict__ seg, int nseg) {
  const int s = blockIdx.y;
  if (s >= nseg) return;
  double* __restrict__ p = seg[s].p;
  const long long n = seg[s].n;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

// P2 mass int phi_a phi_b on the P2 pattern, one colour of cells: lanes
// 0..9 of each 16 own the cell's P2 rows (same ownership and order as
// assemble_colour_kernel, so the sums are reproducible).
__global__ void __launch_bounds__(256)
p2_mass_colour_kernel(const double* __restrict__ X, const int* __restrict__ cells,
                      const int* __restrict__ cell_p2, const int* __restrict__ colour_cells,
                      int cbeg, int cend, const int* __restrict__ rp22,
                      const int* __restrict__ ci22, double* __restrict__ M2) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int ci = cbeg + t / 16;
  int a = t % 16;
  if (ci >= cend || a >= 10) return;
  int e = colour_cells[ci];
  CellGeom G;
  cell_geometry(X, cells, e, G);
  const int row = cell_p2[10 * e + a];
  const int rb = rp22[row], re = rp22[row + 1];
  for (int b = 0; b < 10; ++b) {
    int pos = lower_bound_i32(ci22, rb, re, cell_p2[10 * e + b]);
    M2[pos] += G.vol * c_tab[MPET_TAB_M2 + a * 10 + b];
  }
}

// Scalar CSR of A (un-eliminated) or P (block diagonal, un-eliminated) in
// the canonical numbering; one thread per row, columns emitted in ascending
// order.  rowptr is computed on the host from the node-pattern row lengths.
__global__ void __launch_bounds__(256)
expand_csr_kernel(MpetPatterns pat, MpetPrimitives prim, int nP2, int nP1, int A,
                  MpetCoef coef, int which, const long long* __restrict__ rowptr,
                  int* __restrict__ cols, double* __restrict__ vals) {
  long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nu = 3ll * nP2;
  const int F = A + 1;
  const long long n = nu + (long long)F * nP1;
  if (r >= n) return;
  long long o = rowptr[r];
  const long long nnz11 = pat.nnz11;
  if (r < nu) {
    int a = (int)(r / 3), i = (int)(r % 3);
    for (int k = pat.rp22[a]; k < pat.rp22[a + 1]; ++k) {
      int b = pat.ci22[k];
      for (int j = 0; j < 3; ++j) {
        cols[o] = 3 * b + j;
        vals[o] = which == 0 ? prim.Kuu[9ll * k + 3 * i + j]
                             : (i == j ? coef.mu * prim.Lap[k] : 0.0);
        ++o;
      }
    }
    if (which == 0) {
      for (int k = pat.rp21[a]; k < pat.rp21[a + 1]; ++k) {
        cols[o] = (int)(nu + (long long)F * pat.ci21[k]);
        vals[o] = prim.Dv[3ll * k + i];
        ++o;
      }
    }
  } else {
    long long q = r - nu;
    int c = (int)(q / F), f = (int)(q % F);
    if (which == 0 && f == 0) {
      for (int k = pat.rp12[c]; k < pat.rp12[c + 1]; ++k) {
        int b = pat.ci12[k];
        for (int j = 0; j < 3; ++j) {
          cols[o] = 3 * b + j;
          vals[o] = prim.Dt[3ll * k + j];
          ++o;
        }
      }
    }
    for (int k = pat.rp11[c]; k < pat.rp11[c + 1]; ++k) {
      int d = pat.ci11[k];
      double m = prim.M1[k];
      if (which == 0) {
        for (int g = 0; g < F; ++g) {
          double v = coef.C[f * MPET_MAXF + g] * m;
          if (g == f && f >= 1) v += coef.stiff[f] * prim.St[(f - 1) * nnz11 + k];
          cols[o] = (int)(nu + (long long)F * d + g);
          vals[o] = v;
          ++o;
        }
      } else {
        double v = coef.C[f * MPET_MAXF + f] * m;
        if (f >= 1) v += coef.stiff[f] * prim.St[(f - 1) * nnz11 + k];
        cols[o] = (int)(nu + (long long)F * d + f);
        vals[o] = v;
        ++o;
      }
    }
  }
}

// Per-step right-hand side, pressure rows (rhs of F, linear in up_ at t_n):
//   b_pf = sum_g B[f][g] (M1 p_g^-) + stiff[f] (St_f p_f^-),  f >= 1;  b_p0 = 0.
// Volume / facet sources are added by the source kernels.  u rows are zeroed.
template <int F>
__global__ void __launch_bounds__(256)
rhs_pressure_kernel(MpetPatterns pat, const double* __restrict__ M1,
                    const double* __restrict__ St, MpetCoef coef, int nP2, int nP1,
                    const double* __restrict__ up_prev, double* __restrict__ b) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  const long long nu = 3ll * nP2;
  if (c >= nP1) return;
  double m[F], s[F];
  for (int g = 0; g < F; ++g) { m[g] = 0.0; s[g] = 0.0; }
  const double* xp = up_prev + nu;
  const long long nnz11 = pat.nnz11;
  for (int k = pat.rp11[c]; k < pat.rp11[c + 1]; ++k) {
    int d = pat.ci11[k];
    double mv = M1[k];
    for (int g = 0; g < F; ++g) {
      double xv = xp[(long long)F * d + g];
      m[g] += mv * xv;
      if (g >= 1) s[g] += St[(g - 1) * nnz11 + k] * xv;
    }
  }
  double* bp = b + nu + (long long)F * c;
  bp[0] = 0.0;
  for (int f = 1; f < F; ++f) {
    double v = coef.stiff[f] * s[f];
    for (int g = 0; g < F; ++g) v += coef.C[f * MPET_MAXF + g] * m[g];
    bp[f] = v;
  }
}

// Volume source  b[test dofs] += scale * |K| sum_m Q[i][m] fval[e][m][comp]
// for P2-vector (ncomp 3, test nodes cell_p2) or P1 (field f) tests; cells of
// one colour at a time (no atomics).
__global__ void __launch_bounds__(256)
cell_source_kernel(const double* __restrict__ X, const int* __restrict__ cells,
                   const int* __restrict__ test_nodes, int ntest,
                   const int* __restrict__ colour_cells, int cbeg, int cend,
                   const double* __restrict__ Q, int nk, const double* __restrict__ fval,
                   int ncomp, double scale, long long dof_base, int dof_stride, int dof_off,
                   double* __restrict__ b) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int ci = cbeg + t / 16;
  int i = t % 16;
  if (ci >= cend || i >= ntest) return;
  int e = colour_cells[ci];
  CellGeom G;
  cell_geometry(X, cells, e, G);
  int node = test_nodes[ntest * e + i];
  for (int comp = 0; comp < ncomp; ++comp) {
    double s = 0.0;
    for (int m = 0; m < nk; ++m) s += Q[i * nk + m] * fval[((long long)e * nk + m) * ncomp + comp];
    b[dof_base + (long long)dof_stride * node + dof_off + comp] += scale * G.vol * s;
  }
}

}  // namespace mpet

using namespace mpet;

static thread_local char g_err[512];
extern "C" void mpet_set_error(const char* msg) {
  snprintf(g_err, sizeof(g_err), "%s", msg);
}
extern "C" int mpet_last_error(char* buf, size_t n) {
  if (buf && n) snprintf(buf, n, "%s", g_err);
  return 0;
}

#define HIP_TRY(expr)                                                   \
  do {                                                                  \
    hipError_t _e = (expr);                                             \
    if (_e != hipSuccess) {                                             \
      char _m[400];                                                     \
      snprintf(_m, sizeof(_m), "%s: %s", #expr, hipGetErrorString(_e)); \
      mpet_set_error(_m);                                               \
      return ERR_HIP;                                                   \
    }                                                                   \
  } while (0)

extern "C" int mpet_upload_tables(const double* tab_host, int n) {
  if (n != MPET_TAB_SIZE) { mpet_set_error("table size mismatch"); return ERR_INVALID; }
  HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_tab), tab_host, sizeof(double) * MPET_TAB_SIZE));
  return OK;
}

extern "C" int mpet_assemble_primitives(const MpetMeshDesc* mesh, const MpetPatterns* pat,
                                        const double* Kcell, int A, double mu,
                                        const MpetPrimitives* out, hipStream_t stream) {
  if (!mesh || !pat || !out || A < 1 || A > MPET_MAXF - 1) {
    mpet_set_error("mpet_assemble_primitives: invalid argument");
    return ERR_INVALID;
  }
  HIP_TRY(hipMemsetAsync(out->Kuu, 0, sizeof(double) * 9ll * pat->nnz22, stream));
  HIP_TRY(hipMemsetAsync(out->Lap, 0, sizeof(double) * (long long)pat->nnz22, stream));
  HIP_TRY(hipMemsetAsync(out->Dv, 0, sizeof(double) * 3ll * pat->nnz21, stream));
  HIP_TRY(hipMemsetAsync(out->Dt, 0, sizeof(double) * 3ll * pat->nnz12, stream));
  HIP_TRY(hipMemsetAsync(out->M1, 0, sizeof(double) * (long long)pat->nnz11, stream));
  HIP_TRY(hipMemsetAsync(out->St, 0, sizeof(double) * (long long)A * pat->nnz11, stream));
  for (int k = 0; k < mesh->ncolours; ++k) {
    int cb = (int)mesh->colour_offsets[k], ce = (int)mesh->colour_offsets[k + 1];
    if (ce <= cb) continue;
    long long threads = 16ll * (ce - cb);
    int blocks = (int)((threads + 255) / 256);
    hipLaunchKernelGGL(assemble_colour_kernel, dim3(blocks), dim3(256), 0, stream,
                       mesh->coords, mesh->cells, mesh->cell_p2, mesh->colour_cells, cb, ce,
                       Kcell, A, mu, *pat, *out);
  }
  HIP_TRY(hipGetLastError());
  return OK;
}

extern "C" int mpet_assemble_primitives_batched(int ninst, const MpetMeshDesc* meshes,
                                                const MpetPatterns* pats,
                                                const double* const* Kcell, const int32_t* A,
                                                const double* mu, const MpetPrimitives* outs,
                                                hipStream_t stream) {
  if (ninst < 1 || !meshes || !pats || !Kcell || !A || !mu || !outs) {
    mpet_set_error("mpet_assemble_primitives_batched: invalid argument");
    return ERR_INVALID;
  }
  int ncol = 0;
  std::vector<AsmInst> inst(ninst);
  std::vector<ZeroSeg> segs;
  for (int i = 0; i < ninst; ++i) {
    const MpetMeshDesc& m = meshes[i];
    if (A[i] < 1 || A[i] > MPET_MAXF - 1 || m.ncolours < 0 || !m.colour_offsets) {
      mpet_set_error("mpet_assemble_primitives_batched: invalid instance");
      return ERR_INVALID;
    }
    ncol = m.ncolours > ncol ? m.ncolours : ncol;
    inst[i] = AsmInst{m.coords, m.cells, m.cell_p2, m.colour_cells, Kcell[i], A[i], mu[i],
                      pats[i], outs[i]};
    const MpetPatterns& pt = pats[i];
    const MpetPrimitives& o = outs[i];
    segs.push_back({o.Kuu, 9ll * pt.nnz22});
    segs.push_back({o.Lap, (long long)pt.nnz22});
    segs.push_back({o.Dv, 3ll * pt.nnz21});
    segs.push_back({o.Dt, 3ll * pt.nnz12});
    segs.push_back({o.M1, (long long)pt.nnz11});
    segs.push_back({o.St, (long long)A[i] * pt.nnz11});
  }
  // per colour: cell offsets over the instances (ninst + 1) and each
  // instance's first cell of the colour (ninst)
  std::vector<long long> tab((size_t)ncol * (2 * ninst + 1), 0);
  std::vector<long long> total(ncol, 0);
  for (int k = 0; k < ncol; ++k) {
    long long* off = tab.data() + (size_t)k * (2 * ninst + 1);
    long long* cb = off + ninst + 1;
    off[0] = 0;
    for (int i = 0; i < ninst; ++i) {
      const MpetMeshDesc& m = meshes[i];
      long long n = 0, b = 0;
      if (k < m.ncolours) {
        b = m.colour_offsets[k];
        n = m.colour_offsets[k + 1] - b;
      }
      off[i + 1] = off[i] + n;
      cb[i] = b;
    }
    total[k] = off[ninst];
  }
  AsmInst* d_inst = nullptr;
  long long* d_tab = nullptr;
  ZeroSeg* d_seg = nullptr;
  int rc = OK;
  if (hipMalloc(&d_inst, sizeof(AsmInst) * ninst) != hipSuccess ||
      hipMalloc(&d_tab, sizeof(long long) * (tab.size() ? tab.size() : 1)) != hipSuccess ||
      hipMalloc(&d_seg, sizeof(ZeroSeg) * segs.size()) != hipSuccess) {
    mpet_set_error("mpet_assemble_primitives_batched: out of device memory");
    rc = ERR_OOM;
  }
  if (!rc && (hipMemcpyAsync(d_inst, inst.data(), sizeof(AsmInst) * ninst,
                             hipMemcpyHostToDevice, stream) != hipSuccess ||
              (tab.size() && hipMemcpyAsync(d_tab, tab.data(), sizeof(long long) * tab.size(),
                                            hipMemcpyHostToDevice, stream) != hipSuccess) ||
              hipMemcpyAsync(d_seg, segs.data(), sizeof(ZeroSeg) * segs.size(),
                             hipMemcpyHostToDevice, stream) != hipSuccess)) {
    mpet_set_error("mpet_assemble_primitives_batched: upload failed");
    rc = ERR_HIP;
  }
  if (!rc) {
    hipLaunchKernelGGL(zero_segments_kernel, dim3(64, (unsigned)segs.size()), dim3(256), 0,
                       stream, d_seg, (int)segs.size());
    for (int k = 0; k < ncol; ++k) {
      if (total[k] <= 0) continue;
      const long long threads = 16ll * total[k];
      const int blocks = (int)((threads + 255) / 256);
      const long long* off = d_tab + (size_t)k * (2 * ninst + 1);
      hipLaunchKernelGGL(assemble_colour_batched_kernel, dim3(blocks), dim3(256), 0, stream,
                         d_inst, ninst, off, off + ninst + 1);
    }
    if (hipGetLastError() != hipSuccess) {
      mpet_set_error("mpet_assemble_primitives_batched: launch failed");
      rc = ERR_HIP;
    }
  }
  // the tables are freed once the launches that read them have completed
  if (hipStreamSynchronize(stream) != hipSuccess && !rc) {
    mpet_set_error("mpet_assemble_primitives_batched: stream synchronisation failed");
    rc = ERR_HIP;
  }
  if (d_inst) (void)hipFree(d_inst);
  if (d_tab) (void)hipFree(d_tab);
  if (d_seg) (void)hipFree(d_seg);
  return rc;
}

extern "C" int mpet_assemble_p2_mass(const MpetMeshDesc* mesh, const MpetPatterns* pat,
                                     double* M2, hipStream_t stream) {
  if (!mesh || !pat || !M2 || pat->nnz22 < 1) {
    mpet_set_error("mpet_assemble_p2_mass: invalid argument");
    return ERR_INVALID;
  }
  HIP_TRY(hipMemsetAsync(M2, 0, sizeof(double) * (long long)pat->nnz22, stream));
  for (int k = 0; k < mesh->ncolours; ++k) {
    int cb = (int)mesh->colour_offsets[k], ce = (int)mesh->colour_offsets[k + 1];
    if (ce <= cb) continue;
    long long threads = 16ll * (ce - cb);
    int blocks = (int)((threads + 255) / 256);
    hipLaunchKernelGGL(p2_mass_colour_kernel, dim3(blocks), dim3(256), 0, stream, mesh->coords,
                       mesh->cells, mesh->cell_p2, mesh->colour_cells, cb, ce, pat->rp22,
                       pat->ci22, M2);
  }
  HIP_TRY(hipGetLastError());
  return OK;
}

extern "C" int mpet_expand_csr(const MpetPatterns* pat, const MpetPrimitives* prim, int nP2,
                               int nP1, int A, const MpetCoef* coef, int which,
                               const long long* rowptr, int* cols, double* vals,
                               hipStream_t stream) {
  if (!pat || !prim || !coef || A < 1 || A > MPET_MAXF - 1 || (which != 0 && which != 1)) {
    mpet_set_error("mpet_expand_csr: invalid argument");
    return ERR_INVALID;
  }
  long long n = 3ll * nP2 + (long long)(A + 1) * nP1;
  int blocks = (int)((n + 255) / 256);
  hipLaunchKernelGGL(expand_csr_kernel, dim3(blocks), dim3(256), 0, stream, *pat, *prim, nP2,
                     nP1, A, *coef, which, rowptr, cols, vals);
  HIP_TRY(hipGetLastError());
  return OK;
}

template <int F>
static void launch_rhs(const MpetPatterns* pat, const double* M1, const double* St,
                       const MpetCoef* coef, int nP2, int nP1, const double* up_prev, double* b,
                       hipStream_t stream) {
  int blocks = (nP1 + 255) / 256;
  hipLaunchKernelGGL(rhs_pressure_kernel<F>, dim3(blocks), dim3(256), 0, stream, *pat, M1, St,
                     *coef, nP2, nP1, up_prev, b);
}

extern "C" int mpet_rhs_pressure(const MpetPatterns* pat, const MpetPrimitives* prim, int nP2,
                                 int nP1, int A, const MpetCoef* coef, const double* up_prev,
                                 double* b, hipStream_t stream) {
  if (!pat || !prim || !coef || A < 1 || A > MPET_MAXF - 1) {
    mpet_set_error("mpet_rhs_pressure: invalid argument");
    return ERR_INVALID;
  }
  HIP_TRY(hipMemsetAsync(b, 0, sizeof(double) * 3ll * nP2, stream));
  switch (A + 1) {
    case 2: launch_rhs<2>(pat, prim->M1, prim->St, coef, nP2, nP1, up_prev, b, stream); break;
    case 3: launch_rhs<3>(pat, prim->M1, prim->St, coef, nP2, nP1, up_prev, b, stream); break;
    case 4: launch_rhs<4>(pat, prim->M1, prim->St, coef, nP2, nP1, up_prev, b, stream); break;
    case 5: launch_rhs<5>(pat, prim->M1, prim->St, coef, nP2, nP1, up_prev, b, stream); break;
    case 6: launch_rhs<6>(pat, prim->M1, prim->St, coef, nP2, nP1, up_prev, b, stream); break;
    case 7: launch_rhs<7>(pat, prim->M1, prim->St, coef, nP2, nP1, up_prev, b, stream); break;
    case 8: launch_rhs<8>(pat, prim->M1, prim->St, coef, nP2, nP1, up_prev, b, stream); break;
    default: mpet_set_error("unsupported network count"); return ERR_INVALID;
  }
  HIP_TRY(hipGetLastError());
  return OK;
}

extern "C" int mpet_cell_source(const MpetMeshDesc* mesh, const int* test_nodes, int ntest,
                                const double* Q, int nk, const double* fval, int ncomp,
                                double scale, long long dof_base, int dof_stride, int dof_off,
                                double* b, hipStream_t stream) {
  if (!mesh || (ntest != 10 && ntest != 4) || ncomp < 1 || ncomp > 3) {
    mpet_set_error("mpet_cell_source: invalid argument");
    return ERR_INVALID;
  }
  for (int k = 0; k < mesh->ncolours; ++k) {
    int cb = (int)mesh->colour_offsets[k], ce = (int)mesh->colour_offsets[k + 1];
    if (ce <= cb) continue;
    long long threads = 16ll * (ce - cb);
    int blocks = (int)((threads + 255) / 256);
    hipLaunchKernelGGL(cell_source_kernel, dim3(blocks), dim3(256), 0, stream, mesh->coords,
                       mesh->cells, test_nodes, ntest, mesh->colour_cells, cb, ce, Q, nk, fval,
                       ncomp, scale, dof_base, dof_stride, dof_off, b);
  }
  HIP_TRY(hipGetLastError());
  return OK;
}
