// Sparse operator application for the MPET system.
//
//  * csr_spmv: scalar CSR (the assembled AIJ matrix of the plain FP64 path),
//    LPR lanes per row (4..64, from the average row length of the matrix).
//  * op_apply: the structured total-pressure operator on the node patterns,
//    values stored structure-of-arrays so that every value load of a wave is
//    one contiguous 256-byte (FP32) or 512-byte (FP64) segment:
//        y_u(a)  = sum_b Kuu_ab u_b + sum_c Dv_ac p0_c
//        y_p0(c) = sum_b Dt_cb . u_b + sum_g C[0][g] (M1 p_g)(c)
//        y_pf(c) = sum_g C[f][g] (M1 p_g)(c) + stiff[f] (St_f p_f)(c)
//    The mass is applied once per field; only St differs per network.
//
// Dirichlet handling: the operator is stored un-eliminated.  Rows flagged in
// dmask produce 0 (Krylov mode: iterates have zero Dirichlet components, so
// Pi A Pi x + (I - Pi) x == Pi A x), and with `b` the kernel returns the
// masked residual Pi (b - A x) used for the lifting r0 = Pi (b - A x0) and for
// the FP64 outer residual of iterative refinement.
#include "mpet_common.h"
#include "mpet_ops.h"
#include "op_rows.h"

namespace mpet {

template <typename T, int LPR>
__global__ void __launch_bounds__(256)
csr_spmv_kernel(long long nrows, const long long* __restrict__ rp, const int* __restrict__ ci,
                const T* __restrict__ val, const T* __restrict__ x, const T* __restrict__ b,
                const uint8_t* __restrict__ dmask, T* __restrict__ y) {
  long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long r = gid / LPR;
  int lane = (int)(gid % LPR);
  if (r >= nrows) return;   // whole groups exit together (LPR divides 256)
  T s = csr_row<T, LPR>(rp, ci, val, x, r, lane);
  s = group_sum<LPR>(s);
  if (lane == 0) {
    T out = b ? b[r] - s : s;
    if (dmask && dmask[r]) out = 0;
    y[r] = out;
  }
}

// ---------------------------------------------------------------------------
// structured operator: GU lanes per P2 node row, GP lanes per P1 node row,
// 256-thread blocks; the first nbu blocks take u rows, the rest p rows.
// ---------------------------------------------------------------------------
template <int G, typename T>
__device__ __forceinline__ T op_allsum(T v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <typename T, int F, int GU, int GP>
__global__ void __launch_bounds__(256)
op_apply_kernel(MpetOpT<T> op, const T* __restrict__ x, const T* __restrict__ b,
                T* __restrict__ y, int nbu) {
  const long long nu = 3ll * op.nP2;
  const int bid = xcd_tile(blockIdx.x, gridDim.x, op.xcd_chunk);
  if (bid < nbu) {
    const int lane = threadIdx.x % GU;
    const int a = bid * (256 / GU) + threadIdx.x / GU;
    if (a >= op.nP2) return;
    T y0 = 0, y1 = 0, y2 = 0;
    urow_apply<T, GU>(op.rp22, op.ci22, op.Kuu, op.nnz22, op.rp21, op.ci21, op.Dv, op.nnz21, x,
                      nu, F, a, lane, y0, y1, y2);
    y0 = op_allsum<GU>(y0);
    y1 = op_allsum<GU>(y1);
    y2 = op_allsum<GU>(y2);
    if (lane < 3) {
      const long long o = 3ll * a + lane;
      T v = lane == 0 ? y0 : (lane == 1 ? y1 : y2);
      if (b) v = b[o] - v;
      if (op.dmask && op.dmask[o]) v = 0;
      y[o] = v;
    }
  } else {
    const int lane = threadIdx.x % GP;
    const int c = (bid - nbu) * (256 / GP) + threadIdx.x / GP;
    if (c >= op.nP1) return;
    T yd = 0, m[F], s[F];
#pragma unroll
    for (int g = 0; g < F; ++g) { m[g] = 0; s[g] = 0; }
    prow_apply<T, GP, F>(op.rp12, op.ci12, op.Dt, op.nnz12, op.rp11, op.ci11, op.M1, op.St,
                         op.nnz11, 0, x, nu, F, F, c, lane, yd, m, s);
    yd = op_allsum<GP>(yd);
    T mine = 0;
#pragma unroll
    for (int g = 0; g < F; ++g) {
      m[g] = op_allsum<GP>(m[g]);
      T sg = op_allsum<GP>(s[g]);
      if (lane == g) mine = (g == 0) ? yd : (T)op.stiff[g] * sg;
    }
    if (lane < F) {
      const long long o = nu + (long long)F * c + lane;
#pragma unroll
      for (int g = 0; g < F; ++g) mine += (T)op.C[lane * MPET_MAXF + g] * m[g];
      if (b) mine = b[o] - mine;
      if (op.dmask && op.dmask[o]) mine = 0;
      y[o] = mine;
    }
  }
}

// AoS FP64 primitives -> SoA values of type T (structured operator storage).
template <typename T>
__global__ void soa_convert_kernel(const double* __restrict__ src, int ncomp, long long n,
                                   T* __restrict__ dst) {
  long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  for (int c = 0; c < ncomp; ++c) dst[c * n + k] = (T)src[k * ncomp + c];
}

template <typename T>
__global__ void copy_convert_kernel(const double* __restrict__ src, long long n,
                                    T* __restrict__ dst) {
  long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dst[k] = (T)src[k];
}

// ---------------------------------------------------------------------------
// launchers (host)
// ---------------------------------------------------------------------------
template <typename T>
int launch_csr_spmv(long long nrows, const long long* rp, const int* ci, const T* val,
                    const T* x, const T* b, const uint8_t* dmask, T* y, int lpr,
                    hipStream_t s) {
  long long threads = nrows * lpr;
  int blocks = (int)((threads + 255) / 256);
  switch (lpr) {
    case 4: hipLaunchKernelGGL((csr_spmv_kernel<T, 4>), dim3(blocks), dim3(256), 0, s, nrows, rp, ci, val, x, b, dmask, y); break;
    case 8: hipLaunchKernelGGL((csr_spmv_kernel<T, 8>), dim3(blocks), dim3(256), 0, s, nrows, rp, ci, val, x, b, dmask, y); break;
    case 16: hipLaunchKernelGGL((csr_spmv_kernel<T, 16>), dim3(blocks), dim3(256), 0, s, nrows, rp, ci, val, x, b, dmask, y); break;
    case 32: hipLaunchKernelGGL((csr_spmv_kernel<T, 32>), dim3(blocks), dim3(256), 0, s, nrows, rp, ci, val, x, b, dmask, y); break;
    case 64: hipLaunchKernelGGL((csr_spmv_kernel<T, 64>), dim3(blocks), dim3(256), 0, s, nrows, rp, ci, val, x, b, dmask, y); break;
    default: return ERR_INVALID;
  }
  return hipGetLastError() == hipSuccess ? OK : ERR_HIP;
}
template int launch_csr_spmv<double>(long long, const long long*, const int*, const double*,
                                     const double*, const double*, const uint8_t*, double*, int,
                                     hipStream_t);
template int launch_csr_spmv<float>(long long, const long long*, const int*, const float*,
                                    const float*, const float*, const uint8_t*, float*, int,
                                    hipStream_t);

template <typename T, int F, int GU>
static void launch_op_gp(const MpetOpT<T>& op, const T* x, const T* b, T* y, int gp,
                         hipStream_t s) {
  int nbu = (op.nP2 + 256 / GU - 1) / (256 / GU);
  int gpp = gp < F ? 8 : gp;
  int nbp = (op.nP1 + 256 / gpp - 1) / (256 / gpp);
  dim3 g(nbu + nbp), blk(256);
  switch (gpp) {
    case 8: hipLaunchKernelGGL((op_apply_kernel<T, F, GU, 8>), g, blk, 0, s, op, x, b, y, nbu); break;
    case 16: hipLaunchKernelGGL((op_apply_kernel<T, F, GU, 16>), g, blk, 0, s, op, x, b, y, nbu); break;
    default: hipLaunchKernelGGL((op_apply_kernel<T, F, GU, 32>), g, blk, 0, s, op, x, b, y, nbu); break;
  }
}

template <typename T, int F>
static void launch_op_f(const MpetOpT<T>& op, const T* x, const T* b, T* y, hipStream_t s) {
  switch (op.lanes_u) {
    case 4: launch_op_gp<T, F, 4>(op, x, b, y, op.lanes_p, s); break;
    case 8: launch_op_gp<T, F, 8>(op, x, b, y, op.lanes_p, s); break;
    case 32: launch_op_gp<T, F, 32>(op, x, b, y, op.lanes_p, s); break;
    default: launch_op_gp<T, F, 16>(op, x, b, y, op.lanes_p, s); break;
  }
}

template <typename T>
int launch_op_apply(const MpetOpT<T>& op, const T* x, const T* b, T* y, hipStream_t s) {
  switch (op.A + 1) {
    case 2: launch_op_f<T, 2>(op, x, b, y, s); break;
    case 3: launch_op_f<T, 3>(op, x, b, y, s); break;
    case 4: launch_op_f<T, 4>(op, x, b, y, s); break;
    case 5: launch_op_f<T, 5>(op, x, b, y, s); break;
    case 6: launch_op_f<T, 6>(op, x, b, y, s); break;
    case 7: launch_op_f<T, 7>(op, x, b, y, s); break;
    case 8: launch_op_f<T, 8>(op, x, b, y, s); break;
    default: return ERR_INVALID;
  }
  return hipGetLastError() == hipSuccess ? OK : ERR_HIP;
}
template int launch_op_apply<double>(const MpetOpT<double>&, const double*, const double*,
                                     double*, hipStream_t);
template int launch_op_apply<float>(const MpetOpT<float>&, const float*, const float*, float*,
                                    hipStream_t);

template <typename T>
int launch_soa_convert(const double* src, int ncomp, long long n, T* dst, hipStream_t s) {
  int blocks = (int)((n + 255) / 256);
  if (ncomp == 1)
    hipLaunchKernelGGL((copy_convert_kernel<T>), dim3(blocks), dim3(256), 0, s, src, n, dst);
  else
    hipLaunchKernelGGL((soa_convert_kernel<T>), dim3(blocks), dim3(256), 0, s, src, ncomp, n,
                       dst);
  return hipGetLastError() == hipSuccess ? OK : ERR_HIP;
}
template int launch_soa_convert<double>(const double*, int, long long, double*, hipStream_t);
template int launch_soa_convert<float>(const double*, int, long long, float*, hipStream_t);

}  // namespace mpet
