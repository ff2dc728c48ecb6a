// Batched MPET step: many instances of uneven size, network count and
// parameters packed into one block-diagonal system and solved together, so
// every kernel launch covers all instances.  Each instance is one
// MPETTotalPressureSolver time step (mpet/mpettotalpressuresolver.py:456-502);
// the reference steps one problem at a time and has no batched form.
//
// Packed layout (all instances, see MpetBatchOpDesc in mpet_hip.h)
//   vectors : [u of all instances: 3 * sum nP2] [p of all instances:
//             sum F_k nP1_k, node-interleaved within an instance]
//   P2 rows : global P2 node a; its u DOFs are 3a .. 3a+2
//   P21/P11 : column = absolute DOF of the column node's p_0 (fields follow)
//   P12     : column = global P2 node
//   per P1 row: instance id and absolute DOF of its p_0
//   St      : instance k's A_k arrays at St + st_base[k], each nnz11_k long,
//             indexed by (k - k11_base[k])
// The AMG hierarchies of the instances are concatenated level by level
// (block diagonal, amg_cycle.h) with a block-diagonal dense coarsest solve.
//
// MINRES keeps one state per instance.  Row tiles and vector segments never
// straddle instances, so every partial sum belongs to one instance and one
// block per instance finalises its scalars in a fixed order (deterministic).
// Instance k in iteration i is skipped when inactive[k][i & 1] is set: its
// operator rows, vector segments and preconditioner rows (through the
// per-level row -> instance maps) are left alone, and the preconditioner
// launches stop once every instance is inactive.
#include "amg_cycle.h"
#include "op_rows.h"
#include <algorithm>
#include <vector>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

namespace mpet {
namespace batch {

constexpr int G = 16;            // lanes per node row of the operator kernels
constexpr int ROWS = NT / G;     // node rows per operator tile
constexpr int SEG = 2048;        // DOFs per elementwise segment
constexpr double INNER_MARGIN = 0.3;   // as in the single-instance solver

struct State {                   // MINRES state of one instance
  double beta1, beta, oldb, dbar, epsln, phibar, cs, sn, phi, gamma, delta, oldeps, alfa;
  double tol;
  int it, maxit, nan, done;
  int inactive[2];
};

struct Tile {                    // a row or DOF range of one instance
  int begin, end, inst, pad;
};

template <typename T>
struct BOp {                     // packed structured operator
  int nP2, nP1;
  const int *rp22, *ci22, *rp21, *ci21, *rp12, *ci12, *rp11, *ci11;
  long long nnz22, nnz21, nnz12, nnz11;
  const T *Kuu, *Dv, *Dt, *M1, *St;
  const int* p_inst;
  const int* p_base;
  const MpetBatchCoef* coef;
  const uint8_t* dmask;
};

// MINRES K1 epilogue arguments (st == nullptr: plain apply / residual).
template <typename T>
struct K1Args {
  const State* st;
  T* v;
  const T* r1;
  double* partial;
  int parity;
};

template <typename T>
__device__ __forceinline__ bool k1_setup(const K1Args<T>& k1, int inst, T& s, T& c1) {
  s = 1;
  c1 = 0;
  if (!k1.st) return true;
  const State& S = k1.st[inst];
  if (S.inactive[k1.parity]) return false;
  s = (T)(1.0 / S.beta);
  c1 = S.it > 0 ? (T)(S.beta / S.oldb) : (T)0;
  return true;
}

// u rows: y = A x (b == null) or y = b - A x, Dirichlet rows 0; in K1 mode
// (k1.st set, x = z): y = s A z - c1 r1, v = s z, partial[tile] = v . y.
template <typename T>
__global__ void __launch_bounds__(NT)
bop_u_kernel(BOp<T> op, const Tile* __restrict__ tiles, const T* __restrict__ x,
             const T* __restrict__ b, T* __restrict__ y, K1Args<T> k1) {
  const int t = blockIdx.x;
  const Tile tl = tiles[t];
  T s, c1;
  const bool active = k1_setup(k1, tl.inst, s, c1);
  double acc = 0.0;
  const int lane = threadIdx.x % G;
  const int a = tl.begin + threadIdx.x / G;
  if (active && a < tl.end) {
    T y0 = 0, y1 = 0, y2 = 0;
    urow_apply<T, G>(op.rp22, op.ci22, op.Kuu, op.nnz22, op.rp21, op.ci21, op.Dv, op.nnz21, x,
                     0, 1, a, lane, y0, y1, y2);
    y0 = group_allsum<G>(y0);
    y1 = group_allsum<G>(y1);
    y2 = group_allsum<G>(y2);
    if (lane < 3) {
      const long long o = 3ll * a + lane;
      T v = lane == 0 ? y0 : (lane == 1 ? y1 : y2);
      const bool m = op.dmask && op.dmask[o];
      if (k1.st) {
        const T yy = m ? (T)0 : s * v - c1 * k1.r1[o];
        const T vv = s * x[o];
        y[o] = yy;
        k1.v[o] = vv;
        acc = (double)vv * (double)yy;
      } else {
        if (b) v = b[o] - v;
        y[o] = m ? (T)0 : v;
      }
    }
  }
  if (k1.st) {
    const double bs = block_sum(acc);
    if (threadIdx.x == 0) k1.partial[t] = bs;
  }
}

// p rows (all fields of a P1 node in one row group), same modes; partials
// go to partial[part_off + tile].
template <typename T>
__global__ void __launch_bounds__(NT)
bop_p_kernel(BOp<T> op, const Tile* __restrict__ tiles, int part_off, const T* __restrict__ x,
             const T* __restrict__ b, T* __restrict__ y, K1Args<T> k1) {
  const int t = blockIdx.x;
  const Tile tl = tiles[t];
  T s, c1;
  const bool active = k1_setup(k1, tl.inst, s, c1);
  double acc = 0.0;
  const int lane = threadIdx.x % G;
  const int c = tl.begin + threadIdx.x / G;
  if (active && c < tl.end) {
    const MpetBatchCoef& ic = op.coef[tl.inst];
    const int F = ic.F;
    T yd = 0, m[MPET_MAXF], sv[MPET_MAXF];
#pragma unroll
    for (int g = 0; g < MPET_MAXF; ++g) { m[g] = 0; sv[g] = 0; }
    prow_apply<T, G, MPET_MAXF>(op.rp12, op.ci12, op.Dt, op.nnz12, op.rp11, op.ci11, op.M1,
                                op.St + ic.st_base, ic.nnz11, ic.k11_base, x, 0, 1, F, c, lane,
                                yd, m, sv);
    yd = group_allsum<G>(yd);
    T mine = 0;
#pragma unroll
    for (int g = 0; g < MPET_MAXF; ++g) {
      if (g < F) {
        m[g] = group_allsum<G>(m[g]);
        const T sg = group_allsum<G>(sv[g]);
        if (lane == g) mine = (g == 0) ? yd : (T)ic.stiff[g] * sg;
      }
    }
    if (lane < F) {
#pragma unroll
      for (int g = 0; g < MPET_MAXF; ++g)
        if (g < F) mine += (T)ic.C[lane * MPET_MAXF + g] * m[g];
      const long long o = (long long)op.p_base[c] + lane;
      const bool msk = op.dmask && op.dmask[o];
      if (k1.st) {
        const T yy = msk ? (T)0 : s * mine - c1 * k1.r1[o];
        const T vv = s * x[o];
        y[o] = yy;
        k1.v[o] = vv;
        acc = (double)vv * (double)yy;
      } else {
        if (b) mine = b[o] - mine;
        y[o] = msk ? (T)0 : mine;
      }
    }
  }
  if (k1.st) {
    const double bs = block_sum(acc);
    if (threadIdx.x == 0) k1.partial[part_off + t] = bs;
  }
}

// Sum of instance k's partials: ranges [r0[k], r0[k+1]) and [r1[k], r1[k+1])
// (its u part and its p part), fixed order; valid in thread 0.
__device__ __forceinline__ double inst_sum(const double* __restrict__ partial,
                                           const int* __restrict__ r0, const int* __restrict__ r1,
                                           int k) {
  return block_sum(thread_sum(partial, r0[k], r0[k + 1]) +
                   thread_sum(partial, r1[k], r1[k + 1]));
}

__global__ void __launch_bounds__(NT)
bk1f_kernel(const double* __restrict__ partial, const int* __restrict__ r0,
            const int* __restrict__ r1, State* st, int parity) {
  const int k = blockIdx.x;
  if (st[k].inactive[parity]) return;
  const double s = inst_sum(partial, r0, r1, k);
  if (threadIdx.x == 0) st[k].alfa = s;
}

// r2n = y - (alfa / beta) r2o on the segments of active instances.
template <typename T>
__global__ void __launch_bounds__(NT)
bk2_kernel(const Tile* __restrict__ segs, const T* __restrict__ y, const T* __restrict__ r2o,
           T* __restrict__ r2n, const State* st, int parity) {
  const Tile sg = segs[blockIdx.x];
  const State& S = st[sg.inst];
  if (S.inactive[parity]) return;
  const T c = (T)(S.alfa / S.beta);
  for (int i = sg.begin + threadIdx.x; i < sg.end; i += NT) r2n[i] = y[i] - c * r2o[i];
}

// partial[segment] = a . b over the segment (0 for inactive instances when
// st is given).
template <typename T>
__global__ void __launch_bounds__(NT)
bdot_kernel(const Tile* __restrict__ segs, const T* __restrict__ a, const T* __restrict__ b,
            const State* st, int parity, double* __restrict__ partial) {
  const Tile sg = segs[blockIdx.x];
  double acc = 0.0;
  if (!st || !st[sg.inst].inactive[parity])
    for (int i = sg.begin + threadIdx.x; i < sg.end; i += NT) acc += (double)a[i] * (double)b[i];
  const double bs = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = bs;
}

// Givens update of one iteration (Paige & Saunders), convergence test, and
// the skip flag of the next iteration.
__global__ void __launch_bounds__(NT)
bk4f_kernel(const double* __restrict__ partial, const int* __restrict__ r0,
            const int* __restrict__ r1, State* st, int cur, int next) {
  const int k = blockIdx.x;
  if (st[k].inactive[cur]) {
    if (threadIdx.x == 0) st[k].inactive[next] = 1;
    return;
  }
  const double tot = inst_sum(partial, r0, r1, k);
  if (threadIdx.x == 0) {
    State s = st[k];
    s.oldb = s.beta;
    // fmax(NaN, c) is c: test the raw sum and the rotation for NaN / inf
    // before clamping, so a breakdown is reported instead of read as 0
    const bool bad_in = !(tot == tot) || isinf(tot);
    s.beta = sqrt(fmax(tot, 0.0));
    s.oldeps = s.epsln;
    s.delta = s.cs * s.dbar + s.sn * s.alfa;
    const double gbar = s.sn * s.dbar - s.cs * s.alfa;
    s.epsln = s.sn * s.beta;
    s.dbar = -s.cs * s.beta;
    const double hg = hypot(gbar, s.beta);
    s.gamma = fmax(hg, 1e-300);
    s.cs = gbar / s.gamma;
    s.sn = s.beta / s.gamma;
    s.phi = s.cs * s.phibar;
    s.phibar = s.sn * s.phibar;
    s.it += 1;
    const bool bad = bad_in || !(hg == hg) || !(s.phibar == s.phibar) || isinf(s.phibar);
    s.nan = bad ? 1 : 0;
    const int stop = (bad || s.phibar <= s.tol || s.beta == 0.0 || s.it >= s.maxit) ? 1 : 0;
    s.done = stop;
    s.inactive[next] = stop;
    st[k] = s;
  }
}

// all[p] = 1 when every instance is inactive for parity p; iflag[p][k] =
// instance k inactive for parity p (read by the preconditioner kernels).
__global__ void __launch_bounds__(NT)
ball_kernel(const State* st, int ninst, int p, int* all, int* iflag) {
  __shared__ int any;
  if (threadIdx.x == 0) any = 0;
  __syncthreads();
  for (int k = threadIdx.x; k < ninst; k += NT) {
    const int ia = st[k].inactive[p];
    iflag[(long long)p * ninst + k] = ia;
    if (!ia) any = 1;
  }
  __syncthreads();
  if (threadIdx.x == 0) all[p] = any ? 0 : 1;
}

// w_new = (v - oldeps w1 - delta w2) / gamma; x += phi w_new  (w1 <- w_new).
template <typename T>
__global__ void __launch_bounds__(NT)
bk5_kernel(const Tile* __restrict__ segs, const T* __restrict__ v, T* __restrict__ w1,
           const T* __restrict__ w2, T* __restrict__ x, const State* st, int parity) {
  const Tile sg = segs[blockIdx.x];
  const State& S = st[sg.inst];
  if (S.inactive[parity]) return;
  const T oe = (T)S.oldeps, de = (T)S.delta, ig = (T)(1.0 / S.gamma), ph = (T)S.phi;
  for (int i = sg.begin + threadIdx.x; i < sg.end; i += NT) {
    const T w = (v[i] - oe * w1[i] - de * w2[i]) * ig;
    w1[i] = w;
    x[i] += ph * w;
  }
}

template <typename T>
__global__ void __launch_bounds__(NT)
binit_vec_kernel(long long n, const T* __restrict__ b, T* r1, T* r2, T* w1, T* w2, T* x) {
  const long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i >= n) return;
  const T v = b[i];
  r1[i] = v;
  r2[i] = v;
  w1[i] = 0;
  w2[i] = 0;
  x[i] = 0;
}

// Start of the MINRES solves: beta1 = sqrt(b . M^-1 b) per instance;
// tol = rtol * beta1 (rtol >= 0) or -rtol * ref[k] (rtol < 0); instances
// with keep[k] == 0 start inactive.
__global__ void __launch_bounds__(NT)
binit_state_kernel(const double* __restrict__ partial, const int* __restrict__ r0,
                   const int* __restrict__ r1, State* st, double rtol,
                   const double* __restrict__ ref, int maxit, const int* __restrict__ keep) {
  const int k = blockIdx.x;
  const double tot = inst_sum(partial, r0, r1, k);
  if (threadIdx.x == 0) {
    State s;
    memset(&s, 0, sizeof(s));
    const bool bad = !(tot == tot) || isinf(tot);    // before fmax, which drops NaN
    s.beta1 = bad ? tot : sqrt(fmax(tot, 0.0));
    s.beta = s.beta1;
    s.phibar = s.beta1;
    s.cs = -1.0;
    s.tol = rtol >= 0 ? rtol * s.beta1 : -rtol * ref[k];
    s.maxit = maxit;
    s.nan = bad ? 1 : 0;
    s.done = (s.beta1 == 0.0 || s.beta1 <= s.tol || s.nan || !keep[k]) ? 1 : 0;
    s.inactive[0] = s.done;
    s.inactive[1] = s.done;
    st[k] = s;
  }
}

// x += d on the segments of instances with upd[k] set (FP32 or FP64 d).
template <typename T>
__global__ void __launch_bounds__(NT)
badd_kernel(const Tile* __restrict__ segs, const T* __restrict__ d, double* __restrict__ x,
            const int* __restrict__ upd) {
  const Tile sg = segs[blockIdx.x];
  if (!upd[sg.inst]) return;
  for (int i = sg.begin + threadIdx.x; i < sg.end; i += NT) x[i] += (double)d[i];
}

__global__ void bf64_to_f32_kernel(long long n, const double* __restrict__ a,
                                   float* __restrict__ b) {
  const long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}
// g = Pi_D x (Dirichlet values of x, zero elsewhere).
__global__ void bdirichlet_part_kernel(long long n, const double* __restrict__ x,
                                       const uint8_t* __restrict__ mask, double* __restrict__ g) {
  const long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i < n) g[i] = mask[i] ? x[i] : 0.0;
}

}  // namespace batch
}  // namespace mpet

// ===========================================================================
// host side
// ===========================================================================
using namespace mpet::batch;

namespace {

template <typename T>
struct BMinres {
  T *R[2] = {nullptr, nullptr}, *W[2] = {nullptr, nullptr};
  T *z = nullptr, *y = nullptr, *v = nullptr, *x = nullptr, *b = nullptr;
  State* st = nullptr;
  int* all = nullptr;               // all[2]
  int* iflag = nullptr;             // iflag[2][ninst]
};

struct Batch {
  MpetSolverOptions opt;
  int ninst = 0;
  long long n = 0, nu = 0;
  BOp<double> op64;
  BOp<float> op32;
  bool has_op32 = false;
  const uint8_t* mask64 = nullptr;
  Tile *utiles = nullptr, *ptiles = nullptr, *segs = nullptr;
  int nut = 0, npt = 0, nseg = 0;
  int *kr0 = nullptr, *kr1 = nullptr;   // K1 partial ranges per instance
  int *sr0 = nullptr, *sr1 = nullptr;   // segment partial ranges per instance
  double* partial = nullptr;            // max(nut + npt, nseg)
  Precond<double> pc64;
  Precond<float> pc32;
  bool has_pc64 = false, has_pc32 = false;
  BMinres<double> m64;
  BMinres<float> m32;
  double *r64 = nullptr, *g64 = nullptr;
  double* ref = nullptr;                // per-instance reference norms (device)
  int* keep = nullptr;                  // per-instance flags (device)
  State* st_host = nullptr;             // pinned, ninst
  double* ref_host = nullptr;           // pinned
  int* keep_host = nullptr;             // pinned
  std::vector<void*> allocs, host_allocs;
  hipGraphExec_t graph[2] = {nullptr, nullptr};   // FP64 / FP32 iteration chunks
  hipStream_t graph_stream = nullptr;
};

template <typename T>
BOp<T> make_bop(const MpetBatchOpDesc* d) {
  BOp<T> o;
  o.nP2 = d->nP2;
  o.nP1 = d->nP1;
  o.rp22 = d->rp22; o.ci22 = d->ci22; o.rp21 = d->rp21; o.ci21 = d->ci21;
  o.rp12 = d->rp12; o.ci12 = d->ci12; o.rp11 = d->rp11; o.ci11 = d->ci11;
  o.nnz22 = d->nnz22; o.nnz21 = d->nnz21; o.nnz12 = d->nnz12; o.nnz11 = d->nnz11;
  o.Kuu = (const T*)d->Kuu; o.Dv = (const T*)d->Dv; o.Dt = (const T*)d->Dt;
  o.M1 = (const T*)d->M1; o.St = (const T*)d->St;
  o.p_inst = d->p_inst;
  o.p_base = d->p_base;
  o.coef = d->coef;
  o.dmask = d->dmask;
  return o;
}

template <typename T>
void op_launch(const Batch* B, const BOp<T>& op, const T* x, const T* b, T* y, K1Args<T> k1,
               hipStream_t s) {
  hipLaunchKernelGGL((bop_u_kernel<T>), dim3(B->nut), dim3(NT), 0, s, op, B->utiles, x, b, y,
                     k1);
  hipLaunchKernelGGL((bop_p_kernel<T>), dim3(B->npt), dim3(NT), 0, s, op, B->ptiles, B->nut, x,
                     b, y, k1);
}

template <typename T>
const BOp<T>& op_of(const Batch* B) {
  if constexpr (sizeof(T) == 8) return B->op64; else return B->op32;
}
template <typename T>
BMinres<T>& minres_of(Batch* B) {
  if constexpr (sizeof(T) == 8) return B->m64; else return B->m32;
}
template <typename T>
const Precond<T>& pc_of(const Batch* B) {
  if constexpr (sizeof(T) == 8) return B->pc64; else return B->pc32;
}

template <typename T>
void minres_iter(Batch* B, int i, hipStream_t s) {
  BMinres<T>& M = minres_of<T>(B);
  const Precond<T>& P = pc_of<T>(B);
  T* Ra = M.R[i & 1];
  T* Rb = M.R[(i + 1) & 1];
  T* Wa = M.W[i & 1];
  T* Wb = M.W[(i + 1) & 1];
  const int cur = i & 1, next = (i + 1) & 1;
  K1Args<T> k1{M.st, M.v, Ra, B->partial, cur};
  op_launch<T>(B, op_of<T>(B), M.z, nullptr, M.y, k1, s);
  hipLaunchKernelGGL(bk1f_kernel, dim3(B->ninst), dim3(NT), 0, s, B->partial, B->kr0, B->kr1,
                     M.st, cur);
  hipLaunchKernelGGL((bk2_kernel<T>), dim3(B->nseg), dim3(NT), 0, s, B->segs, M.y, Rb, Ra, M.st,
                     cur);
  Skip sk;
  sk.all = &M.all[cur];
  sk.inst = M.iflag + (long long)cur * B->ninst;
  precond_apply(P, Ra, M.z, sk, s);
  hipLaunchKernelGGL((bdot_kernel<T>), dim3(B->nseg), dim3(NT), 0, s, B->segs, Ra, M.z,
                     (const State*)M.st, cur, B->partial);
  hipLaunchKernelGGL(bk4f_kernel, dim3(B->ninst), dim3(NT), 0, s, B->partial, B->sr0, B->sr1,
                     M.st, cur, next);
  hipLaunchKernelGGL(ball_kernel, dim3(1), dim3(NT), 0, s, M.st, B->ninst, next, M.all,
                     M.iflag);
  hipLaunchKernelGGL((bk5_kernel<T>), dim3(B->nseg), dim3(NT), 0, s, B->segs, M.v, Wa, Wb, M.x,
                     M.st, cur);
}

// r1 = r2 = b, z = M^-1 b, beta1 per instance; instances with keep == 0
// start inactive.  Reads the states back (synchronises).
template <typename T>
int minres_start(Batch* B, double rtol, int maxit, hipStream_t s) {
  BMinres<T>& M = minres_of<T>(B);
  const int eb = (int)((B->n + NT - 1) / NT);
  hipLaunchKernelGGL((binit_vec_kernel<T>), dim3(eb), dim3(NT), 0, s, B->n, M.b, M.R[0], M.R[1],
                     M.W[0], M.W[1], M.x);
  precond_apply(pc_of<T>(B), M.R[1], M.z, (const int*)nullptr, s);
  hipLaunchKernelGGL((bdot_kernel<T>), dim3(B->nseg), dim3(NT), 0, s, B->segs, M.R[1], M.z,
                     (const State*)nullptr, 0, B->partial);
  hipLaunchKernelGGL(binit_state_kernel, dim3(B->ninst), dim3(NT), 0, s, B->partial, B->sr0,
                     B->sr1, M.st, rtol, B->ref, maxit, B->keep);
  hipLaunchKernelGGL(ball_kernel, dim3(1), dim3(NT), 0, s, M.st, B->ninst, 0, M.all, M.iflag);
  hipLaunchKernelGGL(ball_kernel, dim3(1), dim3(NT), 0, s, M.st, B->ninst, 1, M.all, M.iflag);
  HIPCHK(hipMemcpyAsync(B->st_host, M.st, sizeof(State) * B->ninst, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return OK;
}

bool all_done(const Batch* B) {
  for (int k = 0; k < B->ninst; ++k)
    if (!B->st_host[k].done) return false;
  return true;
}

// Iterations in chunks of `check` (replayed from a captured graph when
// enabled) until every instance has stopped.
template <typename T>
int minres_continue(Batch* B, int maxit, hipStream_t s) {
  BMinres<T>& M = minres_of<T>(B);
  int check = B->opt.check_every < 1 ? 1 : B->opt.check_every;
  if (check & 1) check += 1;       // buffer rotation and skip flags have period 2
  hipGraphExec_t* graph = &B->graph[sizeof(T) == 8 ? 0 : 1];
  int done_iters = 0;
  while (!all_done(B)) {
    if (B->opt.use_graphs) {
      if (!*graph) {
        hipGraph_t g;
        HIPCHK(hipStreamBeginCapture(B->graph_stream, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < check; ++i) minres_iter<T>(B, i, B->graph_stream);
        HIPCHK(hipStreamEndCapture(B->graph_stream, &g));
        HIPCHK(hipGraphInstantiate(graph, g, nullptr, nullptr, 0));
        HIPCHK(hipGraphDestroy(g));
      }
      HIPCHK(hipGraphLaunch(*graph, s));
    } else {
      for (int i = 0; i < check; ++i) minres_iter<T>(B, i, s);
    }
    done_iters += check;
    HIPCHK(hipMemcpyAsync(B->st_host, M.st, sizeof(State) * B->ninst, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (done_iters >= maxit + check) break;
  }
  return OK;
}

template <typename T>
int alloc_n(Batch* B, T** p, size_t count) {
  return alloc_dev(B->allocs, p, count);
}

template <typename T>
int build_minres(Batch* B, BMinres<T>& M) {
  int rc;
  const size_t n = (size_t)B->n;
  if ((rc = alloc_n(B, &M.R[0], n)) || (rc = alloc_n(B, &M.R[1], n)) ||
      (rc = alloc_n(B, &M.W[0], n)) || (rc = alloc_n(B, &M.W[1], n)) ||
      (rc = alloc_n(B, &M.z, n)) || (rc = alloc_n(B, &M.y, n)) || (rc = alloc_n(B, &M.v, n)) ||
      (rc = alloc_n(B, &M.x, n)) || (rc = alloc_n(B, &M.b, n)) ||
      (rc = alloc_n(B, &M.st, (size_t)B->ninst)) || (rc = alloc_n(B, &M.all, 2)) ||
      (rc = alloc_n(B, &M.iflag, 2 * (size_t)B->ninst)))
    return rc;
  return OK;
}

template <typename T>
int upload(Batch* B, const std::vector<T>& h, T** d) {
  int rc = alloc_n(B, d, h.size());
  if (rc) return rc;
  HIPCHK(hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return OK;
}

// Tiles, segments and the per-instance partial ranges.
int build_layout(Batch* B, const MpetBatchLayout* lay) {
  const int K = lay->ninst;
  std::vector<Tile> ut, pt, sg_u, sg_p;
  std::vector<int> kr0(K + 1), kr1(K + 1), sr0(K + 1), sr1(K + 1);
  for (int k = 0; k < K; ++k) {
    kr0[k] = (int)ut.size();
    const long long a1 = lay->p2_off[k + 1];
    for (long long a = lay->p2_off[k]; a < a1; a += ROWS)
      ut.push_back(Tile{(int)a, (int)std::min(a + ROWS, a1), k, 0});
    sr0[k] = (int)sg_u.size();
    for (long long i = 3 * lay->p2_off[k]; i < 3 * a1; i += SEG)
      sg_u.push_back(Tile{(int)i, (int)std::min(i + SEG, 3 * a1), k, 0});
  }
  kr0[K] = (int)ut.size();
  sr0[K] = (int)sg_u.size();
  const long long nu = 3 * lay->p2_off[K];
  for (int k = 0; k < K; ++k) {
    kr1[k] = (int)(ut.size() + pt.size());
    const long long c1 = lay->p1_off[k + 1];
    for (long long c = lay->p1_off[k]; c < c1; c += ROWS)
      pt.push_back(Tile{(int)c, (int)std::min(c + ROWS, c1), k, 0});
    sr1[k] = (int)(sg_u.size() + sg_p.size());
    const long long i1 = nu + lay->pd_off[k + 1];
    for (long long i = nu + lay->pd_off[k]; i < i1; i += SEG)
      sg_p.push_back(Tile{(int)i, (int)std::min(i + SEG, i1), k, 0});
  }
  kr1[K] = (int)(ut.size() + pt.size());
  sr1[K] = (int)(sg_u.size() + sg_p.size());
  std::vector<Tile> segs = sg_u;
  segs.insert(segs.end(), sg_p.begin(), sg_p.end());
  B->nut = (int)ut.size();
  B->npt = (int)pt.size();
  B->nseg = (int)segs.size();
  int rc;
  if ((rc = upload(B, ut, &B->utiles)) || (rc = upload(B, pt, &B->ptiles)) ||
      (rc = upload(B, segs, &B->segs)) || (rc = upload(B, kr0, &B->kr0)) ||
      (rc = upload(B, kr1, &B->kr1)) || (rc = upload(B, sr0, &B->sr0)) ||
      (rc = upload(B, sr1, &B->sr1)))
    return rc;
  return alloc_n(B, &B->partial, (size_t)std::max(B->nut + B->npt, B->nseg));
}

// FP64 residual r = Pi (b - A x) of every instance.
void residual64(Batch* B, const double* b, const double* x, double* r, hipStream_t s) {
  op_launch<double>(B, B->op64, x, b, r, K1Args<double>{nullptr, nullptr, nullptr, nullptr, 0},
                    s);
}

// ||Pi (b - A g)||_{M^-1} per instance (g = Dirichlet part of x) into
// ref_host / ref (the measure of every relative tolerance), via a MINRES
// start in precision T.
template <typename T>
int bnorm_M(Batch* B, const double* b, const double* x, hipStream_t s) {
  const int eb = (int)((B->n + NT - 1) / NT);
  hipLaunchKernelGGL(bdirichlet_part_kernel, dim3(eb), dim3(NT), 0, s, B->n, x, B->mask64,
                     B->g64);
  residual64(B, b, B->g64, B->r64, s);
  BMinres<T>& M = minres_of<T>(B);
  if constexpr (sizeof(T) == 4)
    hipLaunchKernelGGL(bf64_to_f32_kernel, dim3(eb), dim3(NT), 0, s, B->n, B->r64, M.b);
  else
    HIPCHK(hipMemcpyAsync(M.b, B->r64, sizeof(double) * B->n, hipMemcpyDeviceToDevice, s));
  int rc = minres_start<T>(B, 1.0, 1, s);
  if (rc) return rc;
  for (int k = 0; k < B->ninst; ++k) B->ref_host[k] = B->st_host[k].beta1;
  HIPCHK(hipMemcpyAsync(B->ref, B->ref_host, sizeof(double) * B->ninst, hipMemcpyHostToDevice,
                        s));
  return OK;
}

int set_keep(Batch* B, hipStream_t s) {
  HIPCHK(hipMemcpyAsync(B->keep, B->keep_host, sizeof(int) * B->ninst, hipMemcpyHostToDevice,
                        s));
  return OK;
}

int host_alloc(Batch* B, void** p, size_t bytes) {
  if (hipHostMalloc(p, bytes, hipHostMallocDefault) != hipSuccess) {
    mpet_set_error("pinned host allocation failed");
    return ERR_HIP;
  }
  B->host_allocs.push_back(*p);
  return OK;
}

}  // namespace

extern "C" int mpet_batch_destroy(void* h) {
  Batch* B = (Batch*)h;
  if (!B) return OK;
  (void)hipDeviceSynchronize();
  for (auto& g : B->graph)
    if (g) (void)hipGraphExecDestroy(g);
  if (B->graph_stream) (void)hipStreamDestroy(B->graph_stream);
  free_fork(B->pc64);
  free_fork(B->pc32);
  for (void* p : B->allocs)
    if (p) (void)hipFree(p);
  for (void* p : B->host_allocs)
    if (p) (void)hipHostFree(p);
  delete B;
  return OK;
}

extern "C" int mpet_batch_create(const MpetBatchOpDesc* op64, const MpetBatchOpDesc* op32,
                                 const MpetBatchLayout* lay, const MpetPrecDesc* prec64,
                                 const MpetPrecDesc* prec32, const MpetSolverOptions* opt,
                                 void** handle) {
  if (!op64 || !lay || !opt || !handle || lay->ninst < 1 || op64->ninst != lay->ninst ||
      !lay->p2_off || !lay->p1_off || !lay->pd_off || !op64->dmask || op64->nnz22 < 1 ||
      op64->nnz21 < 1 || op64->nnz12 < 1 || op64->nnz11 < 1) {
    mpet_set_error("mpet_batch_create: invalid argument");
    return ERR_INVALID;
  }
  if (opt->mode != 0 && opt->mode != 1) {
    mpet_set_error("mpet_batch_create: mode must be 0 (FP64) or 1 (mixed)");
    return ERR_INVALID;
  }
  if (opt->mode == 1 && (!op32 || !prec32)) {
    mpet_set_error("mpet_batch_create: mixed mode needs the FP32 operator and preconditioner");
    return ERR_INVALID;
  }
  if (opt->mode == 0 && !prec64) {
    mpet_set_error("mpet_batch_create: FP64 mode needs the FP64 preconditioner");
    return ERR_INVALID;
  }
  const int K = lay->ninst;
  const long long nP2 = lay->p2_off[K], nP1 = lay->p1_off[K], np = lay->pd_off[K];
  if (nP2 != op64->nP2 || nP1 != op64->nP1 || 3 * nP2 + np >= (1ll << 31) ||
      (op32 && (op32->nP2 != op64->nP2 || op32->nP1 != op64->nP1))) {
    mpet_set_error("mpet_batch_create: layout does not match the operator (or exceeds 2^31 DOFs)");
    return ERR_INVALID;
  }
  Batch* B = new Batch();
  B->opt = *opt;
  B->ninst = K;
  B->nu = 3 * nP2;
  B->n = B->nu + np;
  B->op64 = make_bop<double>(op64);
  B->mask64 = op64->dmask;
  if (op32) {
    B->op32 = make_bop<float>(op32);
    B->has_op32 = true;
  }
  int rc = build_layout(B, lay);
  if (!rc && prec64) rc = build_precond(B->allocs, B->nu, prec64, B->pc64);
  if (!rc && prec32) rc = build_precond(B->allocs, B->nu, prec32, B->pc32);
  B->has_pc64 = prec64 != nullptr;
  B->has_pc32 = prec32 != nullptr;
  if (!rc && prec64) rc = make_fork(B->pc64);
  if (!rc && prec32) rc = make_fork(B->pc32);
  if (!rc && opt->mode == 0) rc = build_minres(B, B->m64);
  if (!rc && opt->mode == 1) rc = build_minres(B, B->m32);
  if (!rc && ((rc = alloc_n(B, &B->r64, (size_t)B->n)) ||
              (rc = alloc_n(B, &B->g64, (size_t)B->n)) || (rc = alloc_n(B, &B->ref, (size_t)K)) ||
              (rc = alloc_n(B, &B->keep, (size_t)K)))) {
  }
  if (!rc && ((rc = host_alloc(B, (void**)&B->st_host, sizeof(State) * K)) ||
              (rc = host_alloc(B, (void**)&B->ref_host, sizeof(double) * K)) ||
              (rc = host_alloc(B, (void**)&B->keep_host, sizeof(int) * K)))) {
  }
  if (!rc && hipStreamCreateWithFlags(&B->graph_stream, hipStreamNonBlocking) != hipSuccess) {
    mpet_set_error("mpet_batch_create: stream creation failed");
    rc = ERR_HIP;
  }
  if (rc) {
    mpet_batch_destroy(B);
    return rc;
  }
  *handle = B;
  return OK;
}

// y = A x (b == NULL) or y = b - A x for every instance of the batch, in the
// handle's FP64 (dtype 0) or FP32 (1) packed operator; Dirichlet rows 0.
extern "C" int mpet_batch_apply(void* h, int dtype, const void* x, const void* b, void* y,
                                hipStream_t s) {
  Batch* B = (Batch*)h;
  if (!B || !x || !y || (dtype == 1 && !B->has_op32) || (dtype != 0 && dtype != 1)) {
    mpet_set_error("mpet_batch_apply: invalid argument");
    return ERR_INVALID;
  }
  if (dtype == 0)
    op_launch<double>(B, B->op64, (const double*)x, (const double*)b, (double*)y,
                      K1Args<double>{nullptr, nullptr, nullptr, nullptr, 0}, s);
  else
    op_launch<float>(B, B->op32, (const float*)x, (const float*)b, (float*)y,
                     K1Args<float>{nullptr, nullptr, nullptr, nullptr, 0}, s);
  return hipGetLastError() == hipSuccess ? OK : ERR_HIP;
}

// z = P^-1 r with the batch's FP64 (dtype 0) or FP32 (1) preconditioner.
extern "C" int mpet_batch_precond_apply(void* h, int dtype, const void* r, void* z,
                                        hipStream_t s) {
  Batch* B = (Batch*)h;
  if (!B || !r || !z) {
    mpet_set_error("mpet_batch_precond_apply: invalid argument");
    return ERR_INVALID;
  }
  if (dtype == 0 && B->has_pc64)
    precond_apply(B->pc64, (const double*)r, (double*)z, (const int*)nullptr, s);
  else if (dtype == 1 && B->has_pc32)
    precond_apply(B->pc32, (const float*)r, (float*)z, (const int*)nullptr, s);
  else {
    mpet_set_error("mpet_batch_precond_apply: preconditioner of that dtype not built");
    return ERR_INVALID;
  }
  return hipGetLastError() == hipSuccess ? OK : ERR_HIP;
}

// Lanes per row of the packed hierarchies' level SpMMs (the operator keeps
// its shapes); the captured iteration graphs are dropped.
extern "C" int mpet_batch_set_tuning(void* h, int dtype, const MpetTuning* t) {
  Batch* B = (Batch*)h;
  if (!B || (dtype != 0 && dtype != 1) || !tuning_ok(t)) {
    mpet_set_error("mpet_batch_set_tuning: invalid argument");
    return ERR_INVALID;
  }
  HIPCHK(hipDeviceSynchronize());
  if (dtype == 1 && B->has_pc32) set_precond_tuning(B->pc32, t);
  if (dtype == 0 && B->has_pc64) set_precond_tuning(B->pc64, t);
  hipGraphExec_t& g = B->graph[dtype == 0 ? 0 : 1];
  if (g) HIPCHK(hipGraphExecDestroy(g));
  g = nullptr;
  return OK;
}

// One time step of every instance with active[k] != 0 (NULL: all), same
// contract per instance as mpet_solver_solve.  stats[k] is filled for every
// instance (zeros for the inactive ones).  Returns MPET_ERR_NAN on a
// breakdown, MPET_ERR_NOT_CONVERGED when some active instance did not
// converge; the other instances' solutions are valid either way.
extern "C" int mpet_batch_solve(void* h, const double* b, double* x, const int32_t* active,
                                MpetSolveStats* stats, hipStream_t s) {
  Batch* B = (Batch*)h;
  if (!B || !b || !x || !stats) {
    mpet_set_error("mpet_batch_solve: invalid argument");
    return ERR_INVALID;
  }
  const int K = B->ninst;
  const int eb = (int)((B->n + NT - 1) / NT);
  std::vector<int> want(K), outer(K, 0), inner(K, 0), conv(K, 0), bad(K, 0);
  std::vector<double> resid(K, 0.0);
  for (int k = 0; k < K; ++k) {
    want[k] = active ? (active[k] != 0) : 1;
    B->keep_host[k] = want[k];
  }
  int rc = set_keep(B, s);
  if (rc) return rc;
  if (B->opt.mode == 0) {
    if ((rc = bnorm_M<double>(B, b, x, s))) return rc;
    residual64(B, b, x, B->m64.b, s);
    if ((rc = minres_start<double>(B, -B->opt.rtol, B->opt.maxit, s))) return rc;
    if ((rc = minres_continue<double>(B, B->opt.maxit, s))) return rc;
    hipLaunchKernelGGL((badd_kernel<double>), dim3(B->nseg), dim3(NT), 0, s, B->segs,
                       (const double*)B->m64.x, x, B->keep);
    for (int k = 0; k < K; ++k) {
      if (!want[k]) continue;
      const State& st = B->st_host[k];
      inner[k] = st.it;
      outer[k] = 1;
      bad[k] = st.nan;
      conv[k] = !st.nan && (st.phibar <= st.tol || st.beta == 0.0);
      resid[k] = B->ref_host[k] > 0 ? st.phibar / B->ref_host[k] : 0.0;
    }
  } else {
    if ((rc = bnorm_M<float>(B, b, x, s))) return rc;
    residual64(B, b, x, B->r64, s);
    for (;;) {
      // beta_k = ||r_k||_{M^-1} of every instance still refining
      hipLaunchKernelGGL(bf64_to_f32_kernel, dim3(eb), dim3(NT), 0, s, B->n, B->r64, B->m32.b);
      if ((rc = minres_start<float>(B, B->opt.inner_rtol, B->opt.maxit, s))) return rc;
      int left = 0, dropped = 0;
      for (int k = 0; k < K; ++k) {
        if (!B->keep_host[k]) continue;
        const State& st = B->st_host[k];
        resid[k] = B->ref_host[k] > 0 ? st.beta1 / B->ref_host[k] : 0.0;
        bool stop = true;
        if (st.nan) bad[k] = 1;
        else if (st.beta1 <= B->opt.outer_tol * B->ref_host[k] || st.beta1 == 0.0) conv[k] = 1;
        else if (outer[k] < B->opt.max_outer) stop = false;
        if (stop) {
          B->keep_host[k] = 0;
          ++dropped;
        } else {
          ++left;
        }
      }
      if (!left) break;
      if (dropped) {
        // restart without the instances that stopped refining
        if ((rc = set_keep(B, s))) return rc;
        if ((rc = minres_start<float>(B, B->opt.inner_rtol, B->opt.maxit, s))) return rc;
      }
      // per-instance inner target: inner_rtol relative, but no deeper than
      // needed for the outer target (INNER_MARGIN of it)
      {
        bool upd = false;
        for (int k = 0; k < K; ++k) {
          State& st = B->st_host[k];
          if (!B->keep_host[k] || st.done) continue;
          const double tol = fmax(B->opt.inner_rtol * st.beta1,
                                  INNER_MARGIN * B->opt.outer_tol * B->ref_host[k]);
          if (tol != st.tol) {
            st.tol = tol;
            upd = true;
          }
        }
        if (upd)
          HIPCHK(hipMemcpyAsync(B->m32.st, B->st_host, sizeof(State) * K,
                                hipMemcpyHostToDevice, s));
      }
      if ((rc = minres_continue<float>(B, B->opt.maxit, s))) return rc;
      int broke = 0;
      for (int k = 0; k < K; ++k)
        if (B->keep_host[k]) {
          inner[k] += B->st_host[k].it;
          outer[k] += 1;
          if (B->st_host[k].nan) {   // keep x of a broken-down instance finite
            bad[k] = 1;
            B->keep_host[k] = 0;
            ++broke;
          }
        }
      if (broke && (rc = set_keep(B, s))) return rc;
      hipLaunchKernelGGL((badd_kernel<float>), dim3(B->nseg), dim3(NT), 0, s, B->segs,
                         (const float*)B->m32.x, x, B->keep);
      residual64(B, b, x, B->r64, s);
    }
  }
  if (hipGetLastError() != hipSuccess) {
    mpet_set_error("mpet_batch_solve: kernel launch failed");
    return ERR_HIP;
  }
  HIPCHK(hipStreamSynchronize(s));
  int status = OK;
  for (int k = 0; k < K; ++k) {
    stats[k].inner_iterations = inner[k];
    stats[k].outer_iterations = outer[k];
    stats[k].converged = want[k] ? conv[k] : 0;
    stats[k].residual_norm = resid[k];
    if (want[k] && bad[k]) status = ERR_NAN;
    else if (want[k] && !conv[k] && status == OK) status = ERR_NOT_CONVERGED;
  }
  if (status == ERR_NAN)
    mpet_set_error("mpet_batch_solve: MINRES breakdown (NaN) in an instance");
  else if (status == ERR_NOT_CONVERGED)
    mpet_set_error("mpet_batch_solve: an instance did not converge within maxit / max_outer");
  return status;
}
