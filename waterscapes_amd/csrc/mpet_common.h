// Shared device/host definitions for the MPET HIP library (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MPET_WAVE 64

// Reference-element tables (exact rationals converted to float64), packed
// by waterscapes_amd/refel.py:tables_float64():
//   W[a][b][k][l]  (10x10x4x4)  (1/|K|) int dphi_a/dlam_k dphi_b/dlam_l
//   V[b][c][k]     (10x4x4)     (1/|K|) int psi_c dphi_b/dlam_k
//   M2[a][b]       (10x10)      (1/|K|) int phi_a phi_b
#define MPET_TAB_W 0
#define MPET_TAB_V 1600
#define MPET_TAB_M2 1760
#define MPET_TAB_SIZE 1860

namespace mpet {

// Error codes returned by every C entry point.
enum Status : int {
  OK = 0,
  ERR_INVALID = 1,      // bad argument / shape
  ERR_HIP = 2,          // HIP runtime error
  ERR_NOT_CONVERGED = 3,
  ERR_NAN = 4,
  ERR_OOM = 5,
  ERR_BUDGET = 6,       // admission refused (serving)
};

// NC consecutive values at p (NC * sizeof(T) bytes, element-aligned) as one
// memory instruction where the width allows (FP32 x 3: dwordx3), instead of
// NC separate gathers.
template <typename T, int NC>
__device__ __forceinline__ void load_nc(const T* __restrict__ p, T (&v)[NC]) {
  __builtin_memcpy(v, p, sizeof(T) * NC);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Sum within aligned groups of G lanes (G power of two <= 64); result valid
// in the first lane of each group.
template <int G, typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Workgroups go to the 8 XCDs round-robin (workgroup w on XCD w % 8), each
// XCD with its own L2.  Map hardware tile t of n to a logical tile so that
// every XCD works through runs of `chunk` consecutive logical tiles (rows
// that share gathered vector entries), the runs dealt round-robin over the
// XCDs; tiles past the last full round keep their index.  A bijection on
// [0, n); chunk <= 0 is the identity.
__device__ __forceinline__ int xcd_tile(int t, int n, int chunk) {
  if (chunk <= 0) return t;
  const int round = 8 * chunk;
  const int full = (n / round) * round;
  if (t >= full) return t;
  const int x = t % 8, j = t / 8;           // j-th tile of XCD x
  const int q = j / chunk, r = j % chunk;   // run q, position r
  return (q * 8 + x) * chunk + r;
}

// Lower bound of `key` in sorted cols[lo, hi).
__device__ __forceinline__ int lower_bound_i32(const int* __restrict__ cols, int lo, int hi,
                                               int key) {
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (cols[mid] < key) lo = mid + 1; else hi = mid;
  }
  return lo;
}

}  // namespace mpet
