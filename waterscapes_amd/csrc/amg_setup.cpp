// Smoothed-aggregation AMG setup (host, C++/OpenMP) for the u-block of the
// MPET preconditioner: mu * grad u : grad v (mpet/mpettotalpressuresolver.py:279),
// a component-decoupled vector Laplacian, so ONE scalar hierarchy built on
// the P2 Laplacian serves all three displacement components.
//
// Algorithm (Vanek, Mandel, Brezina 1996):
//   strength   |a_ij| >= eps * sqrt(|a_ii a_jj|)
//   aggregates three passes, deterministic in node order:
//              1) seed a new aggregate at a node whose strong neighbourhood
//                 is entirely unaggregated;
//              2) attach leftovers to the aggregate of the strongest
//                 aggregated neighbour;
//              3) group what remains with its unaggregated strong neighbours.
//              Nodes without strong off-diagonal couplings (Dirichlet
//              identity rows) stay unaggregated: their prolongator row is 0.
//   tentative  T[i, agg(i)] = 1/sqrt(|agg|)   (constant near-nullspace)
//   smoothing  P = (I - omega D^-1 A) T,  omega = 4 / (3 rho(D^-1 A))
//   Galerkin   A_c = P^T A P  (row-wise Gustavson SpGEMM with dense markers)
//   coarsest   dense inverse of A_c (Cholesky) once n_c <= max_coarse.
// rho(D^-1 A) is estimated per level by power iteration (fixed start vector,
// deterministic) and stored for the Chebyshev smoother.
#include <stdint.h>
#include <string.h>
#include <math.h>
#include <stdio.h>
#include <vector>
#include <algorithm>
#include <omp.h>
#include <chrono>
#include <stdlib.h>
#include "../../include/mpet_hip.h"

extern "C" void mpet_set_error(const char* msg);

namespace {

struct Csr {
  int n = 0, m = 0;                 // rows, cols
  std::vector<int64_t> rp;
  std::vector<int> ci;
  std::vector<double> v;
  int64_t nnz() const { return rp.empty() ? 0 : rp[n]; }
};

struct Level {
  Csr A;            // level operator
  Csr P, R;         // prolongator (n x nc) and restriction = P^T (unused on coarsest)
  std::vector<double> dinv;
  double rho = 0;   // spectral radius estimate of D^-1 A
};

struct Hier {
  std::vector<Level> lev;
  std::vector<double> coarse_inv;   // dense inverse of the coarsest operator
  int nc = 0;
};

Csr transpose(const Csr& A) {
  Csr T;
  T.n = A.m; T.m = A.n;
  T.rp.assign(T.n + 1, 0);
  for (int64_t k = 0; k < A.nnz(); ++k) T.rp[A.ci[k] + 1]++;
  for (int i = 0; i < T.n; ++i) T.rp[i + 1] += T.rp[i];
  T.ci.resize(A.nnz());
  T.v.resize(A.nnz());
  std::vector<int64_t> pos(T.rp.begin(), T.rp.end() - 1);
  for (int i = 0; i < A.n; ++i)
    for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
      int64_t p = pos[A.ci[k]]++;
      T.ci[p] = i;
      T.v[p] = A.v[k];
    }
  return T;   // rows of A visited in order -> columns sorted
}

// C = A * B, row-parallel Gustavson with per-thread dense accumulators.
Csr spgemm(const Csr& A, const Csr& B) {
  Csr C;
  C.n = A.n; C.m = B.m;
  std::vector<std::vector<int>> rcols(A.n);
  std::vector<std::vector<double>> rvals(A.n);
#pragma omp parallel
  {
    std::vector<int> marker(B.m, -1);
    std::vector<double> acc(B.m, 0.0);
    std::vector<int> list;
#pragma omp for schedule(dynamic, 256)
    for (int i = 0; i < A.n; ++i) {
      list.clear();
      for (int64_t ka = A.rp[i]; ka < A.rp[i + 1]; ++ka) {
        int j = A.ci[ka];
        double a = A.v[ka];
        for (int64_t kb = B.rp[j]; kb < B.rp[j + 1]; ++kb) {
          int c = B.ci[kb];
          if (marker[c] != i) { marker[c] = i; acc[c] = 0.0; list.push_back(c); }
          acc[c] += a * B.v[kb];
        }
      }
      std::sort(list.begin(), list.end());
      rcols[i] = list;
      rvals[i].resize(list.size());
      for (size_t t = 0; t < list.size(); ++t) rvals[i][t] = acc[list[t]];
    }
  }
  C.rp.assign(C.n + 1, 0);
  for (int i = 0; i < C.n; ++i) C.rp[i + 1] = C.rp[i] + (int64_t)rcols[i].size();
  C.ci.resize(C.rp[C.n]);
  C.v.resize(C.rp[C.n]);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < C.n; ++i) {
    std::copy(rcols[i].begin(), rcols[i].end(), C.ci.begin() + C.rp[i]);
    std::copy(rvals[i].begin(), rvals[i].end(), C.v.begin() + C.rp[i]);
  }
  return C;
}

std::vector<double> diag_inv(const Csr& A) {
  std::vector<double> d(A.n, 1.0);
#pragma omp parallel for
  for (int i = 0; i < A.n; ++i) {
    for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
      if (A.ci[k] == i) { d[i] = A.v[k] != 0.0 ? 1.0 / A.v[k] : 1.0; break; }
  }
  return d;
}

// Largest eigenvalue of D^-1 A (similar to the SPD D^-1/2 A D^-1/2) by
// power iteration from a fixed pseudo-random start.
// Sum of f(i) over 0 <= i < n in a fixed order whatever the thread count:
// fixed blocks of rows summed in parallel, then the block sums in order.
template <typename Fn>
double fixed_sum(int n, Fn f) {
  constexpr int B = 4096;
  const int nb = (n + B - 1) / B;
  std::vector<double> part(nb, 0.0);
#pragma omp parallel for schedule(static)
  for (int b = 0; b < nb; ++b) {
    double t = 0.0;
    const int hi = std::min(n, (b + 1) * B);
    for (int i = b * B; i < hi; ++i) t += f(i);
    part[b] = t;
  }
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += part[b];
  return s;
}

// rho(D^-1 A) by power iteration in the D-inner product from a fixed start
// vector; every reduction in a fixed order, so the estimate (and the
// smoothed prolongators built from it) does not depend on the thread count.
double spectral_radius(const Csr& A, const std::vector<double>& dinv, int iters) {
  int n = A.n;
  std::vector<double> x(n), y(n), t(n);
  uint64_t s = 0x9E3779B97F4A7C15ull;
  for (int i = 0; i < n; ++i) {
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    x[i] = 0.5 + (double)(s % 1000003) / 1000003.0;
  }
  double lam = 0;
  for (int it = 0; it < iters; ++it) {
    const double nrm = sqrt(fixed_sum(n, [&](int i) { return x[i] * x[i] / dinv[i]; }));
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      double a = 0;
      for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k) a += A.v[k] * x[A.ci[k]];
      t[i] = a / nrm;
      y[i] = dinv[i] * t[i];
    }
    const double num = fixed_sum(n, [&](int i) { return x[i] * t[i]; }) / nrm;
    lam = num;         // Rayleigh quotient in the D-inner product
    const double ny = sqrt(fixed_sum(n, [&](int i) { return y[i] * y[i] / dinv[i]; }));
    for (int i = 0; i < n; ++i) x[i] = y[i] / ny;
  }
  return lam;
}

// Returns aggregate id per node (-1 = unaggregated) and the aggregate count.
int aggregate(const Csr& A, double eps, std::vector<int>& agg) {
  int n = A.n;
  std::vector<double> d(n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
      if (A.ci[k] == i) d[i] = fabs(A.v[k]);
  // strength pattern (excluding diagonal)
  std::vector<int64_t> srp(n + 1, 0);
  std::vector<int> sci;
  std::vector<double> sval;
  sci.reserve(A.nnz());
  sval.reserve(A.nnz());
  for (int i = 0; i < n; ++i) {
    for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
      int j = A.ci[k];
      if (j == i) continue;
      double a = fabs(A.v[k]);
      if (a > 0 && a >= eps * sqrt(d[i] * d[j])) {
        sci.push_back(j);
        sval.push_back(a / sqrt(d[i] * d[j]));
      }
    }
    srp[i + 1] = (int64_t)sci.size();
  }
  agg.assign(n, -1);
  int na = 0;
  // pass 1
  for (int i = 0; i < n; ++i) {
    if (agg[i] != -1 || srp[i + 1] == srp[i]) continue;
    bool free = true;
    for (int64_t k = srp[i]; k < srp[i + 1]; ++k)
      if (agg[sci[k]] != -1) { free = false; break; }
    if (!free) continue;
    agg[i] = na;
    for (int64_t k = srp[i]; k < srp[i + 1]; ++k) agg[sci[k]] = na;
    ++na;
  }
  // pass 2
  std::vector<int> agg1 = agg;
  for (int i = 0; i < n; ++i) {
    if (agg1[i] != -1 || srp[i + 1] == srp[i]) continue;
    double best = -1;
    int ba = -1;
    for (int64_t k = srp[i]; k < srp[i + 1]; ++k) {
      int j = sci[k];
      if (agg1[j] != -1 && sval[k] > best) { best = sval[k]; ba = agg1[j]; }
    }
    if (ba >= 0) agg[i] = ba;
  }
  // pass 3
  for (int i = 0; i < n; ++i) {
    if (agg[i] != -1 || srp[i + 1] == srp[i]) continue;
    agg[i] = na;
    for (int64_t k = srp[i]; k < srp[i + 1]; ++k)
      if (agg[sci[k]] == -1) agg[sci[k]] = na;
    ++na;
  }
  return na;
}

// Dense inverse of the SPD coarsest operator: Cholesky A = L L^T, W = L^-1
// (stored transposed: row c of Wt is column c of W), A^-1 = W^T W.  Every
// entry is one fixed-order dot product of contiguous rows, so the result
// does not depend on the thread count; rows / columns are shared over the
// OpenMP threads (n^3 / 3 multiply-adds for each of the three parts).
bool cholesky_inverse(const Csr& A, std::vector<double>& inv) {
  const int n = A.n;
  std::vector<double> L((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k) L[(size_t)i * n + A.ci[k]] = A.v[k];
  for (int j = 0; j < n; ++j) {
    const double* Lj = &L[(size_t)j * n];
    double s = Lj[j];
    for (int k = 0; k < j; ++k) s -= Lj[k] * Lj[k];
    if (!(s > 0)) return false;
    const double ljj = sqrt(s);
    L[(size_t)j * n + j] = ljj;
#pragma omp parallel for schedule(static)
    for (int i = j + 1; i < n; ++i) {
      double* Li = &L[(size_t)i * n];
      double t = Li[j];
      for (int k = 0; k < j; ++k) t -= Li[k] * Lj[k];
      Li[j] = t / ljj;
    }
  }
  // Wt[c][i] = (L^-1)[i][c], i >= c
  std::vector<double> Wt((size_t)n * n, 0.0);
#pragma omp parallel for schedule(dynamic, 8)
  for (int c = 0; c < n; ++c) {
    double* w = &Wt[(size_t)c * n];
    for (int i = c; i < n; ++i) {
      const double* Li = &L[(size_t)i * n];
      double t = (i == c) ? 1.0 : 0.0;
      for (int k = c; k < i; ++k) t -= Li[k] * w[k];
      w[i] = t / Li[i];
    }
  }
  // A^-1 [i][j] = sum_{k >= max(i, j)} W[k][i] W[k][j]
  inv.assign((size_t)n * n, 0.0);
#pragma omp parallel for schedule(dynamic, 8)
  for (int i = 0; i < n; ++i) {
    const double* wi = &Wt[(size_t)i * n];
    for (int j = i; j < n; ++j) {
      const double* wj = &Wt[(size_t)j * n];
      double t = 0.0;
      for (int k = j; k < n; ++k) t += wi[k] * wj[k];
      inv[(size_t)i * n + j] = t;
      inv[(size_t)j * n + i] = t;
    }
  }
  return true;
}

}  // namespace

extern "C" {

// P0 (optional, n x n1 CSR): a prescribed first prolongator; level 1 is then
// the Galerkin product P0^T A P0 and aggregation starts from there.  The
// displacement block passes the exact P1 -> P2 embedding (nested spaces).
int mpet_amg_setup(int n, const int64_t* rp, const int* ci, const double* v,
                   const MpetAmgParams* prm, int n1, const int64_t* p0_rp,
                   const int* p0_ci, const double* p0_v, void** handle) {
  if (n <= 0 || !rp || !ci || !v || !prm || !handle) {
    mpet_set_error("mpet_amg_setup: invalid argument");
    return 1;
  }
  Hier* H = new Hier();
  Csr A0;
  A0.n = A0.m = n;
  A0.rp.assign(rp, rp + n + 1);
  A0.ci.assign(ci, ci + rp[n]);
  A0.v.assign(v, v + rp[n]);
  H->lev.emplace_back();
  H->lev.back().A = std::move(A0);
  for (;;) {
    Level& L = H->lev.back();
    L.dinv = diag_inv(L.A);
    L.rho = spectral_radius(L.A, L.dinv, prm->power_iters);
    if (H->lev.size() == 1 && p0_rp) {
      Csr P;
      P.n = n; P.m = n1;
      P.rp.assign(p0_rp, p0_rp + n + 1);
      P.ci.assign(p0_ci, p0_ci + p0_rp[n]);
      P.v.assign(p0_v, p0_v + p0_rp[n]);
      Csr R = transpose(P);
      Csr AP = spgemm(L.A, P);
      Csr Ac = spgemm(R, AP);
      L.P = std::move(P);
      L.R = std::move(R);
      Level next;
      next.A = std::move(Ac);
      H->lev.push_back(std::move(next));
      continue;
    }
    if (L.A.n <= prm->max_coarse || (int)H->lev.size() >= prm->max_levels) break;
    std::vector<int> agg;
    int na = aggregate(L.A, prm->strength, agg);
    if (na == 0 || na >= L.A.n * 0.9) break;   // coarsening stalled
    // tentative prolongator
    std::vector<int> cnt(na, 0);
    for (int i = 0; i < L.A.n; ++i) if (agg[i] >= 0) cnt[agg[i]]++;
    Csr T;
    T.n = L.A.n; T.m = na;
    T.rp.assign(T.n + 1, 0);
    for (int i = 0; i < T.n; ++i) T.rp[i + 1] = T.rp[i] + (agg[i] >= 0 ? 1 : 0);
    T.ci.resize(T.rp[T.n]);
    T.v.resize(T.rp[T.n]);
    for (int i = 0; i < T.n; ++i)
      if (agg[i] >= 0) { T.ci[T.rp[i]] = agg[i]; T.v[T.rp[i]] = 1.0 / sqrt((double)cnt[agg[i]]); }
    // smoothed prolongator P = T - omega D^-1 A T
    double omega = prm->omega_scale / L.rho;
    Csr AT = spgemm(L.A, T);
    Csr P;
    P.n = T.n; P.m = na;
    P.rp.assign(P.n + 1, 0);
    std::vector<std::vector<int>> pc(P.n);
    std::vector<std::vector<double>> pv(P.n);
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < P.n; ++i) {
      // merge row i of T (<= 1 entry) with -omega dinv_i * row i of AT
      std::vector<int>& c = pc[i];
      std::vector<double>& vv = pv[i];
      int tc = T.rp[i + 1] > T.rp[i] ? T.ci[T.rp[i]] : -1;
      double tv = tc >= 0 ? T.v[T.rp[i]] : 0.0;
      bool placed = tc < 0;
      for (int64_t k = AT.rp[i]; k < AT.rp[i + 1]; ++k) {
        int j = AT.ci[k];
        double val = -omega * L.dinv[i] * AT.v[k];
        if (!placed && tc < j) { c.push_back(tc); vv.push_back(tv); placed = true; }
        if (!placed && tc == j) { val += tv; placed = true; }
        c.push_back(j);
        vv.push_back(val);
      }
      if (!placed) { c.push_back(tc); vv.push_back(tv); }
    }
    for (int i = 0; i < P.n; ++i) P.rp[i + 1] = P.rp[i] + (int64_t)pc[i].size();
    P.ci.resize(P.rp[P.n]);
    P.v.resize(P.rp[P.n]);
    for (int i = 0; i < P.n; ++i) {
      std::copy(pc[i].begin(), pc[i].end(), P.ci.begin() + P.rp[i]);
      std::copy(pv[i].begin(), pv[i].end(), P.v.begin() + P.rp[i]);
    }
    Csr R = transpose(P);
    Csr AP = spgemm(L.A, P);
    Csr Ac = spgemm(R, AP);
    L.P = std::move(P);
    L.R = std::move(R);
    Level next;
    next.A = std::move(Ac);
    H->lev.push_back(std::move(next));
  }
  Level& C = H->lev.back();
  H->nc = C.A.n;
  auto tc0 = std::chrono::steady_clock::now();
  const bool ok = cholesky_inverse(C.A, H->coarse_inv);
  if (getenv("MPET_DEBUG"))
    fprintf(stderr, "[amg] coarse inverse n=%d: %.3f s\n", C.A.n,
            std::chrono::duration<double>(std::chrono::steady_clock::now() - tc0).count());
  if (!ok) {
    delete H;
    mpet_set_error("mpet_amg_setup: coarsest operator not SPD");
    return 3;
  }
  *handle = H;
  return 0;
}

// OpenMP threads of the setups that the calling host thread runs next
// (several instances set up at once from a thread pool share the cores).
int mpet_amg_set_threads(int n) {
  if (n < 1) return 1;
  omp_set_num_threads(n);
  return 0;
}

int mpet_amg_num_levels(void* h) { return (int)((Hier*)h)->lev.size(); }

// sizes: out[0]=n, out[1]=nnz(A), out[2]=nc (cols of P, 0 on coarsest), out[3]=nnz(P)
int mpet_amg_level_sizes(void* h, int l, int64_t* out) {
  Hier* H = (Hier*)h;
  if (l < 0 || l >= (int)H->lev.size()) return 1;
  Level& L = H->lev[l];
  out[0] = L.A.n;
  out[1] = L.A.nnz();
  out[2] = L.P.m;
  out[3] = L.P.nnz();
  return 0;
}

static void copy_csr(const Csr& A, int64_t* rp, int* ci, double* v) {
  if (rp) memcpy(rp, A.rp.data(), sizeof(int64_t) * (A.n + 1));
  if (ci) memcpy(ci, A.ci.data(), sizeof(int) * A.nnz());
  if (v) memcpy(v, A.v.data(), sizeof(double) * A.nnz());
}

// Copies level l into caller (host) buffers; any pointer may be NULL.
int mpet_amg_level_copy(void* h, int l, int64_t* A_rp, int* A_ci, double* A_v, double* dinv,
                        double* rho, int64_t* P_rp, int* P_ci, double* P_v, int64_t* R_rp,
                        int* R_ci, double* R_v) {
  Hier* H = (Hier*)h;
  if (l < 0 || l >= (int)H->lev.size()) return 1;
  Level& L = H->lev[l];
  copy_csr(L.A, A_rp, A_ci, A_v);
  if (dinv) memcpy(dinv, L.dinv.data(), sizeof(double) * L.A.n);
  if (rho) *rho = L.rho;
  if (L.P.n) {
    copy_csr(L.P, P_rp, P_ci, P_v);
    copy_csr(L.R, R_rp, R_ci, R_v);
  }
  return 0;
}

int mpet_amg_coarse_inverse(void* h, double* out) {
  Hier* H = (Hier*)h;
  memcpy(out, H->coarse_inv.data(), sizeof(double) * H->coarse_inv.size());
  return H->nc;
}

void mpet_amg_destroy(void* h) { delete (Hier*)h; }

}  // extern "C"
