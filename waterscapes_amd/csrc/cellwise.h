// Cell-wise (matrix-free) application of the total-pressure operator, in
// FP32 (the inner MINRES of the mixed path) or FP64 (the refinement's
// residuals: the same formulas in FP64 with FP64 geometry and partials).
//
// The u rows and the divergence part of the p0 rows, i.e. the bilinear form
//     2 mu eps(u):eps(v) + p0 div v + div(u) w0
// (mpet/mpettotalpressuresolver.py:267-269), are computed cell by cell from
// the geometry instead of being read from the assembled Kuu / Dv / Dt
// values, which are 9 + 3 + 3 FP32 numbers per node pair (1.1 GB per apply
// at 48^3 x 6 tets).  The P1 x P1 pressure block C (x) M1 + stiff St stays
// assembled (small, and it carries the per-cell permeabilities and the Robin
// terms).
//
// Exact integration without quadrature: the gradient of a P2 field is
// linear on a tetrahedron, so with g_n = grad u at vertex n and the linear
// stress s = mu (g + g^T) + p0 I (exact: p0 is P1),
//     int_T s : grad phi_i = |T|/20 sum_n (S + s_n) : (grad phi_i)(vertex n),
// S = sum_n s_n, from int_T lambda_m lambda_n = |T| (1 + delta_mn) / 20.
// With G_j = grad lambda_j and the P2 basis gradients (4 lambda_j - 1) G_j
// (vertex j) and 4 (lambda_k G_j + lambda_j G_k) (edge jk):
//     g_n          = 4 u_n G_n - sum_j u_j G_j + 4 sum_{k != n} u_{nk} G_k
//     y(vertex j)  = |T|/20 (4 s_j - S) G_j
//     y(edge jk)   = |T|/5 ((S + s_k) G_j + (S + s_j) G_k)
//     y_p0(vertex v) = |T|/20 (sum_n tr g_n + tr g_v)
// (u_j, g_n, s_n in 3-component / 3x3 form; y(.) are the 3 rows of a P2
// node's u DOFs.)
//
// Two launches.  cw_cell_kernel (one thread per cell, CW_NT cells per
// block) computes each cell's 30 + 4 local results into LDS, then sums, per
// node touched by the block's cells, that node's contributions from these
// cells (in slot order) and writes one partial per (node, block), a block's
// partials stored contiguously (~3x fewer stores than one per (cell,
// node)).  cw_rows_kernel gives every row the sum of its node's partials,
// listed contiguously in block order by a node-sorted entry list (fixed
// order, so results are reproducible), plus the assembled P1 x P1 part, and
// hands the row value to an epilogue (plain apply / masked residual, or the
// fused MINRES step).  Cell tables are structure-of-arrays ([k * nC + e]).
#pragma once
#include "amg_cycle.h"

namespace mpet {

// Device view of MpetCellOpDesc plus the two work arrays, in precision T
// (FP32 for the inner MINRES, FP64 for the refinement's residuals).
template <typename T>
struct CellOpT {
  int nC, nP2, nP1, A;
  const int* node2;       // [k * nC + e]: P2 node of local node k (UFC order)
  const int* node1;       // [v * nC + e]: P1 node of vertex v
  // partials of block b: q in [bq[b], bq[b+1]) (block-major, so a block's
  // stores are contiguous); partial q sums the local slots
  // lsl[lrp[q] .. lrp[q+1]) (c * 10 + k for P2, c * 4 + v for P1, c the
  // cell within the block); entry k of the node-sorted list is partial pe[k]
  const int *bq2, *lrp2, *pe2;
  const int16_t* lsl2;
  const int *bq1, *lrp1, *pe1;
  const int16_t* lsl1;
  const T* geom;          // SoA: [d * nC + e] = G_{1 + d/3}[d % 3] (d < 9); |T| at d = 9
  const int *inc2_rp, *inc1_rp;     // node -> first entry (nP2 + 1, nP1 + 1)
  const int *rp11, *ci11;
  int nnz11;
  const T *M1, *St;                 // St: A x nnz11, network-major
  T C[MPET_MAXF * MPET_MAXF];
  T stiff[MPET_MAXF];
  T mu;
  const uint8_t* dmask;
  int lanes_u, lanes_p, xcd_chunk;
  int npart2, npart1;     // partials of all blocks (P2, P1)
  T* Yu;                  // 3 x npart2, block-major
  T* Yp;                  // npart1, block-major
};
using CellOp = CellOpT<float>;

#ifndef MPET_CW_CELLS
#define MPET_CW_CELLS 256
#endif
constexpr int CW_NT = MPET_CW_CELLS;   // cells per block of the cell kernel

// Index of the UFC edge joining local vertices j != k.
__device__ __forceinline__ int cw_edge(int j, int k) {
  const int lo = j < k ? j : k, hi = j < k ? k : j;
  // (2,3)->0 (1,3)->1 (1,2)->2 (0,3)->3 (0,2)->4 (0,1)->5
  return lo == 0 ? 6 - hi : (lo == 1 ? 4 - hi : 0);
}

// Phase 1: one thread per cell.
// (A template, so that the translation units including this header share
// one instance.)
template <int NTB, typename T>
__global__ void __launch_bounds__(NTB)
cw_cell_kernel(CellOpT<T> op, const T* __restrict__ x, const int* __restrict__ skip) {
  static_assert(NTB == CW_NT, "one cell per thread of a CW_NT block");
  if (skip && *skip) return;
  __shared__ T su[CW_NT * 30];
  __shared__ T sp[CW_NT * 4];
  const int t = threadIdx.x;
  const long long e = (long long)blockIdx.x * CW_NT + t;
  if (e < op.nC) {
    const long long nC = op.nC;
    int n2[10], n1[4];
#pragma unroll
    for (int k = 0; k < 10; ++k) n2[k] = op.node2[k * nC + e];
#pragma unroll
    for (int v = 0; v < 4; ++v) n1[v] = op.node1[v * nC + e];
    T G[4][3];
#pragma unroll
    for (int d = 0; d < 9; ++d) G[1 + d / 3][d % 3] = op.geom[d * nC + e];
    const T vol = op.geom[9 * nC + e];
#pragma unroll
    for (int m = 0; m < 3; ++m) G[0][m] = -(G[1][m] + G[2][m] + G[3][m]);
    const int F = op.A + 1;
    const long long nu = 3ll * op.nP2;
    T u[10][3], p0[4];
#pragma unroll
    for (int k = 0; k < 10; ++k) load_nc<T, 3>(x + 3ll * n2[k], u[k]);
#pragma unroll
    for (int v = 0; v < 4; ++v) p0[v] = x[nu + (long long)F * n1[v]];
    // U = sum_j u_j (x) G_j
    T U[3][3];
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int d = 0; d < 3; ++d)
        U[b][d] = u[0][b] * G[0][d] + u[1][b] * G[1][d] + u[2][b] * G[2][d] + u[3][b] * G[3][d];
    // vertex stresses (symmetric: xx yy zz xy xz yz) and divergences
    T s[4][6], tr[4];
    const T mu = op.mu;
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      T g[3][3];
#pragma unroll
      for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int d = 0; d < 3; ++d) g[b][d] = (T)4 * u[n][b] * G[n][d] - U[b][d];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (k == n) continue;
        const int m = 4 + cw_edge(n, k);
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
          for (int d = 0; d < 3; ++d) g[b][d] += (T)4 * u[m][b] * G[k][d];
      }
      tr[n] = g[0][0] + g[1][1] + g[2][2];
      s[n][0] = (T)2 * mu * g[0][0] + p0[n];
      s[n][1] = (T)2 * mu * g[1][1] + p0[n];
      s[n][2] = (T)2 * mu * g[2][2] + p0[n];
      s[n][3] = mu * (g[0][1] + g[1][0]);
      s[n][4] = mu * (g[0][2] + g[2][0]);
      s[n][5] = mu * (g[1][2] + g[2][1]);
    }
    T S[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) S[q] = (s[0][q] + s[1][q]) + (s[2][q] + s[3][q]);
    const T c = vol * ((T)1 / (T)20);
    auto put = [&](int k, const T* o) {
      su[t * 30 + 3 * k] = o[0];
      su[t * 30 + 3 * k + 1] = o[1];
      su[t * 30 + 3 * k + 2] = o[2];
    };
    // symmetric 3x3 (xx yy zz xy xz yz) times a vector
    auto symv = [](const T* M, const T* w, T* out) {
      out[0] = M[0] * w[0] + M[3] * w[1] + M[4] * w[2];
      out[1] = M[3] * w[0] + M[1] * w[1] + M[5] * w[2];
      out[2] = M[4] * w[0] + M[5] * w[1] + M[2] * w[2];
    };
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      T M[6], o[3];
#pragma unroll
      for (int q = 0; q < 6; ++q) M[q] = c * ((T)4 * s[j][q] - S[q]);
      symv(M, G[j], o);
      put(j, o);
    }
    constexpr int EA[6] = {2, 1, 1, 0, 0, 0}, EB[6] = {3, 3, 2, 3, 2, 1};
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      const int j = EA[m], k = EB[m];
      T Mk[6], Mj[6], o1[3], o2[3];
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        Mk[q] = (T)4 * c * (S[q] + s[k][q]);
        Mj[q] = (T)4 * c * (S[q] + s[j][q]);
      }
      symv(Mk, G[j], o1);
      symv(Mj, G[k], o2);
      const T o[3] = {o1[0] + o2[0], o1[1] + o2[1], o1[2] + o2[2]};
      put(4 + m, o);
    }
    const T D = (tr[0] + tr[1]) + (tr[2] + tr[3]);
#pragma unroll
    for (int v = 0; v < 4; ++v) sp[t * 4 + v] = c * (D + tr[v]);
  }
  __syncthreads();
  // one partial per (node, block): the node's slots of this block, summed in
  // slot order, stored at the block's contiguous run of partials (scattered
  // 12-byte stores to node-sorted lists measured 3.7x their bytes in memory
  // writes: lines shared by blocks on different XCDs)
  const int b = blockIdx.x;
  T* __restrict__ Yu = op.Yu;
  for (int q = op.bq2[b] + t; q < op.bq2[b + 1]; q += CW_NT) {
    const int lb = op.lrp2[q], le = op.lrp2[q + 1];
    T a[3] = {(T)0, (T)0, (T)0};
    for (int i = lb; i < le; ++i) {
      const int sl = op.lsl2[i];
      a[0] += su[3 * sl];
      a[1] += su[3 * sl + 1];
      a[2] += su[3 * sl + 2];
    }
    __builtin_memcpy(Yu + 3ll * q, a, 3 * sizeof(T));
  }
  for (int q = op.bq1[b] + t; q < op.bq1[b + 1]; q += CW_NT) {
    const int lb = op.lrp1[q], le = op.lrp1[q + 1];
    T a = (T)0;
    for (int i = lb; i < le; ++i) a += sp[op.lsl1[i]];
    op.Yp[q] = a;
  }
}

// Phase-2 epilogue of a plain application: y = A x, or y = b - A x with b;
// rows with dmask set give 0.
template <typename T>
struct CwEpiApplyT {
  T* y;
  const T* b;
  const uint8_t* mask;
  __device__ __forceinline__ bool skipped() const { return false; }
  __device__ __forceinline__ void prep() {}
  __device__ __forceinline__ double operator()(long long i, T ax) const {
    T v = b ? b[i] - ax : ax;
    if (mask && mask[i]) v = (T)0;
    y[i] = v;
    return 0.0;
  }
};
using CwEpiApply = CwEpiApplyT<float>;

// Phase 2: GU lanes per P2 node row (GU = 1: one thread per node, which
// writes all three rows), GP lanes per P1 node row, grid-stride
// over row tiles (u tiles first), XCD-aware tile order.  The epilogue's
// per-row results are summed per block into partial[blockIdx.x] (when given).
template <int F, int GU, int GP, typename Epi, typename T>
__global__ void __launch_bounds__(256)
cw_rows_kernel(CellOpT<T> op, const T* __restrict__ x, Epi epi, double* __restrict__ partial) {
  if (epi.skipped()) return;
  epi.prep();
  constexpr int RU = 256 / GU, RP = 256 / GP;
  const long long nu = 3ll * op.nP2;
  const int ntu = (op.nP2 + RU - 1) / RU, ntp = (op.nP1 + RP - 1) / RP;
  const T* __restrict__ Yu = op.Yu;
  const T* __restrict__ Yp = op.Yp;
  double acc = 0.0;
  for (int tt = blockIdx.x; tt < ntu + ntp; tt += gridDim.x) {
    const int tile = xcd_tile(tt, ntu + ntp, op.xcd_chunk);
    if (tile < ntu) {
      const int lane = threadIdx.x % GU;
      const int a = tile * RU + threadIdx.x / GU;
      if (a < op.nP2) {
        const int kb = op.inc2_rp[a], ke = op.inc2_rp[a + 1];
        T y0 = (T)0, y1 = (T)0, y2 = (T)0;
        // the node's entries are contiguous: two per lane per trip, loads
        // (entry -> partial -> value) issued before use
        for (int k = kb + lane; k < ke; k += 2 * GU) {
          const int k2 = k + GU < ke ? k + GU : k;
          T av[3], bv[3];
          load_nc<T, 3>(Yu + 3ll * op.pe2[k], av);
          load_nc<T, 3>(Yu + 3ll * op.pe2[k2], bv);
          if (!(k + GU < ke)) bv[0] = bv[1] = bv[2] = (T)0;
          y0 += av[0];
          y1 += av[1];
          y2 += av[2];
          y0 += bv[0];
          y1 += bv[1];
          y2 += bv[2];
        }
#pragma unroll
        for (int off = GU / 2; off > 0; off >>= 1) {
          y0 += __shfl_xor(y0, off, 64);
          y1 += __shfl_xor(y1, off, 64);
          y2 += __shfl_xor(y2, off, 64);
        }
        if constexpr (GU == 1) {           // one thread per row: all three rows
          acc += epi(3ll * a, y0);
          acc += epi(3ll * a + 1, y1);
          acc += epi(3ll * a + 2, y2);
        } else if (lane < 3) {
          const T yc = lane == 0 ? y0 : (lane == 1 ? y1 : y2);
          acc += epi(3ll * a + lane, yc);
        }
      }
    } else {
      const int lane = threadIdx.x % GP;
      const int c = (tile - ntu) * RP + threadIdx.x / GP;
      if (c < op.nP1) {
        T yd = (T)0;
        const int kb = op.inc1_rp[c], ke = op.inc1_rp[c + 1];
        for (int k = kb + lane; k < ke; k += GP) yd += Yp[op.pe1[k]];
        T m[F], sv[F];
#pragma unroll
        for (int g = 0; g < F; ++g) { m[g] = (T)0; sv[g] = (T)0; }
        const int qb = op.rp11[c], qe = op.rp11[c + 1];
        const long long n11 = op.nnz11;
        for (int k = qb + lane; k < qe; k += GP) {
          const T* xp = x + nu + (long long)F * op.ci11[k];
          const T mm = op.M1[k];
          T xv[F], st[F];
#pragma unroll
          for (int g = 0; g < F; ++g) {
            xv[g] = xp[g];
            st[g] = g >= 1 ? op.St[(g - 1) * n11 + k] : (T)0;
          }
#pragma unroll
          for (int g = 0; g < F; ++g) {
            m[g] += mm * xv[g];
            if (g >= 1) sv[g] += st[g] * xv[g];
          }
        }
#pragma unroll
        for (int off = GP / 2; off > 0; off >>= 1) yd += __shfl_xor(yd, off, 64);
        T mine = (T)0;
#pragma unroll
        for (int g = 0; g < F; ++g) {
          T mg = m[g], sg = sv[g];
#pragma unroll
          for (int off = GP / 2; off > 0; off >>= 1) {
            mg += __shfl_xor(mg, off, 64);
            sg += __shfl_xor(sg, off, 64);
          }
          m[g] = mg;
          if (lane == g) mine = (g == 0) ? yd : op.stiff[g] * sg;
        }
        if (lane < F) {
#pragma unroll
          for (int g = 0; g < F; ++g) mine += op.C[lane * MPET_MAXF + g] * m[g];
          acc += epi(nu + (long long)F * c + lane, mine);
        }
      }
    }
  }
  if (partial) {
    double bs = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = bs;
  }
}

// Launch of phase 2 for F fields and the lane counts of the operator (GU in
// {4, 8, 16}, GP in {8, 16}; others fall back to 8).  Returns the grid
// (number of partials written when partial is given).
template <typename Epi, int F, int GU, typename T>
static void cw_rows_gp(const CellOpT<T>& op, int gp, dim3 grid, hipStream_t s, const T* x,
                       Epi epi, double* partial) {
  if (gp == 16)
    hipLaunchKernelGGL((cw_rows_kernel<F, GU, 16, Epi, T>), grid, dim3(256), 0, s, op, x, epi,
                       partial);
  else
    hipLaunchKernelGGL((cw_rows_kernel<F, GU, 8, Epi, T>), grid, dim3(256), 0, s, op, x, epi,
                       partial);
}
template <typename Epi, int F, typename T>
static void cw_rows_f(const CellOpT<T>& op, dim3 grid, hipStream_t s, const T* x, Epi epi,
                      double* partial) {
  switch (op.lanes_u) {
    case 1: cw_rows_gp<Epi, F, 1>(op, op.lanes_p, grid, s, x, epi, partial); break;
    case 4: cw_rows_gp<Epi, F, 4>(op, op.lanes_p, grid, s, x, epi, partial); break;
    case 16: cw_rows_gp<Epi, F, 16>(op, op.lanes_p, grid, s, x, epi, partial); break;
    default: cw_rows_gp<Epi, F, 8>(op, op.lanes_p, grid, s, x, epi, partial); break;
  }
}
static inline int cw_lanes_u(int g) { return (g == 1 || g == 4 || g == 8 || g == 16) ? g : 8; }
static inline int cw_lanes_p(int g) { return (g == 8 || g == 16) ? g : 8; }
template <typename T>
static inline int cw_row_tiles(const CellOpT<T>& op) {
  const int gu = cw_lanes_u(op.lanes_u), gp = cw_lanes_p(op.lanes_p);
  return (op.nP2 + 256 / gu - 1) / (256 / gu) + (op.nP1 + 256 / gp - 1) / (256 / gp);
}
// Both phases: cells (x -> Yu, Yp), then rows through epi.  Returns the
// number of blocks of phase 2 (one partial each when partial is given).
template <typename Epi, typename T>
static int cw_launch(const CellOpT<T>& op, const T* x, Epi epi, double* partial,
                     const int* skip, hipStream_t s) {
  const int cb = (op.nC + CW_NT - 1) / CW_NT;
  hipLaunchKernelGGL((cw_cell_kernel<CW_NT, T>), dim3(cb), dim3(CW_NT), 0, s, op, x, skip);
  const int np = grid_for(cw_row_tiles(op), 1, MAX_PARTIALS);
  const dim3 grid(np);
  switch (op.A + 1) {
    case 2: cw_rows_f<Epi, 2>(op, grid, s, x, epi, partial); break;
    case 3: cw_rows_f<Epi, 3>(op, grid, s, x, epi, partial); break;
    case 4: cw_rows_f<Epi, 4>(op, grid, s, x, epi, partial); break;
    case 5: cw_rows_f<Epi, 5>(op, grid, s, x, epi, partial); break;
    case 6: cw_rows_f<Epi, 6>(op, grid, s, x, epi, partial); break;
    case 7: cw_rows_f<Epi, 7>(op, grid, s, x, epi, partial); break;
    case 8: cw_rows_f<Epi, 8>(op, grid, s, x, epi, partial); break;
    default: break;
  }
  return np;
}

// Work values (of the descriptor's precision) of a descriptor: the P2 and
// P1 partials.
static inline long long cw_work_floats(const MpetCellOpDesc* d) {
  return 3ll * d->npart2 + d->npart1;
}

// Host: device view of a descriptor (coefficients rounded to T; the
// descriptor's fp64 flag must match T) with the work arrays Yu = work,
// Yp = work + 3 npart2; checks the descriptor.
template <typename T>
static int cw_from_desc(const MpetCellOpDesc* d, T* work, CellOpT<T>& op) {
  if (!d || (d->fp64 != 0) != (sizeof(T) == 8) || d->nC <= 0 || d->nP2 <= 0 || d->nP1 <= 0 || d->A < 1 || d->A + 1 > MPET_MAXF ||
      d->nnz11 <= 0 || !d->node2 || !d->node1 || !d->geom || d->cells_per_block != CW_NT ||
      d->npart2 <= 0 || d->npart1 <= 0 || !d->bq2 || !d->lrp2 || !d->pe2 || !d->lsl2 ||
      !d->bq1 || !d->lrp1 || !d->pe1 || !d->lsl1 ||
      !d->inc2_rp || !d->inc1_rp || !d->rp11 || !d->ci11 || !d->M1 ||
      !d->St || !work) {
    mpet_set_error("invalid cell-wise operator description");
    return ERR_INVALID;
  }
  op.nC = d->nC;
  op.nP2 = d->nP2;
  op.nP1 = d->nP1;
  op.A = d->A;
  op.node2 = d->node2;
  op.node1 = d->node1;
  op.bq2 = d->bq2;
  op.lrp2 = d->lrp2;
  op.pe2 = d->pe2;
  op.lsl2 = d->lsl2;
  op.bq1 = d->bq1;
  op.lrp1 = d->lrp1;
  op.pe1 = d->pe1;
  op.lsl1 = d->lsl1;
  op.npart2 = d->npart2;
  op.npart1 = d->npart1;
  op.geom = (const T*)d->geom;
  op.inc2_rp = d->inc2_rp;
  op.inc1_rp = d->inc1_rp;
  op.rp11 = d->rp11;
  op.ci11 = d->ci11;
  op.nnz11 = d->nnz11;
  op.M1 = (const T*)d->M1;
  op.St = (const T*)d->St;
  for (int i = 0; i < MPET_MAXF * MPET_MAXF; ++i) op.C[i] = (T)d->coef.C[i];
  for (int i = 0; i < MPET_MAXF; ++i) op.stiff[i] = (T)d->coef.stiff[i];
  op.mu = (T)d->coef.mu;
  op.dmask = d->dmask;
  op.lanes_u = cw_lanes_u(d->lanes_u);
  op.lanes_p = cw_lanes_p(d->lanes_p);
  op.xcd_chunk = d->xcd_chunk;
  op.Yu = work;
  op.Yp = work + 3ll * d->npart2;
  return OK;
}

}  // namespace mpet
