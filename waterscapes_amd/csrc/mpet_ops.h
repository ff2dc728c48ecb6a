// Internal C++ declarations shared by the .hip translation units.
#pragma once
#include "mpet_common.h"
#include "../../include/mpet_hip.h"

namespace mpet {

// Structured total-pressure operator with values of type T (SoA layout).
template <typename T>
struct MpetOpT {
  int nP2, nP1, A;
  const int *rp22, *ci22, *rp21, *ci21, *rp12, *ci12, *rp11, *ci11;
  int nnz22, nnz21, nnz12, nnz11;
  const T* Kuu;   // 9 * nnz22, component c = 3*i + j at [c * nnz22 + k]
  const T* Dv;    // 3 * nnz21
  const T* Dt;    // 3 * nnz12
  const T* M1;    // nnz11
  const T* St;    // A * nnz11
  double C[MPET_MAXF * MPET_MAXF];
  double stiff[MPET_MAXF];
  const uint8_t* dmask;   // per DOF, 1 = Dirichlet (may be null)
  int lanes_u, lanes_p;   // lanes per P2 / P1 node row (0 = default)
  int xcd_chunk;          // tiles per XCD run (0 = hardware order)
};

template <typename T>
int launch_csr_spmv(long long nrows, const long long* rp, const int* ci, const T* val,
                    const T* x, const T* b, const uint8_t* dmask, T* y, int lpr, hipStream_t s);
template <typename T>
int launch_op_apply(const MpetOpT<T>& op, const T* x, const T* b, T* y, hipStream_t s);
template <typename T>
int launch_soa_convert(const double* src, int ncomp, long long n, T* dst, hipStream_t s);

}  // namespace mpet

extern "C" void mpet_set_error(const char* msg);
