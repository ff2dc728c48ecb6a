// extern "C" entry points for the sparse operators (declared in
// include/mpet_hip.h).
#include "mpet_ops.h"

using namespace mpet;

template <typename T>
static MpetOpT<T> make_op(const MpetOpDesc* d) {
  MpetOpT<T> op;
  op.nP2 = d->nP2; op.nP1 = d->nP1; op.A = d->A;
  op.rp22 = d->pat.rp22; op.ci22 = d->pat.ci22;
  op.rp21 = d->pat.rp21; op.ci21 = d->pat.ci21;
  op.rp12 = d->pat.rp12; op.ci12 = d->pat.ci12;
  op.rp11 = d->pat.rp11; op.ci11 = d->pat.ci11;
  op.nnz22 = d->pat.nnz22; op.nnz21 = d->pat.nnz21;
  op.nnz12 = d->pat.nnz12; op.nnz11 = d->pat.nnz11;
  op.Kuu = (const T*)d->Kuu; op.Dv = (const T*)d->Dv; op.Dt = (const T*)d->Dt;
  op.M1 = (const T*)d->M1; op.St = (const T*)d->St;
  for (int i = 0; i < MPET_MAXF * MPET_MAXF; ++i) op.C[i] = d->coef.C[i];
  for (int i = 0; i < MPET_MAXF; ++i) op.stiff[i] = d->coef.stiff[i];
  op.dmask = d->dmask;
  op.lanes_u = d->lanes_u;
  op.lanes_p = d->lanes_p;
  op.xcd_chunk = d->xcd_chunk;
  return op;
}

template <typename T>
MpetOpT<T> mpet_make_op(const MpetOpDesc* d) { return make_op<T>(d); }
template MpetOpT<double> mpet_make_op<double>(const MpetOpDesc*);
template MpetOpT<float> mpet_make_op<float>(const MpetOpDesc*);

extern "C" int mpet_op_apply(const MpetOpDesc* d, const void* x, const void* b, void* y,
                             hipStream_t stream) {
  if (!d || !x || !y || d->A < 1 || d->A > MPET_MAXF - 1 || d->pat.nnz22 < 1 ||
      d->pat.nnz21 < 1 || d->pat.nnz12 < 1 || d->pat.nnz11 < 1) {
    mpet_set_error("mpet_op_apply: invalid argument");
    return ERR_INVALID;
  }
  int rc;
  if (d->dtype == 0)
    rc = launch_op_apply<double>(make_op<double>(d), (const double*)x, (const double*)b,
                                 (double*)y, stream);
  else
    rc = launch_op_apply<float>(make_op<float>(d), (const float*)x, (const float*)b, (float*)y,
                                stream);
  if (rc) mpet_set_error("mpet_op_apply: launch failed");
  return rc;
}

extern "C" int mpet_csr_spmv(int dtype, long long nrows, const long long* rp, const int* ci,
                             const void* val, const void* x, const void* b,
                             const uint8_t* dmask, void* y, int lpr, hipStream_t stream) {
  if (nrows < 0 || !rp || !ci || !val || !x || !y) {
    mpet_set_error("mpet_csr_spmv: invalid argument");
    return ERR_INVALID;
  }
  int rc = dtype == 0
               ? launch_csr_spmv<double>(nrows, rp, ci, (const double*)val, (const double*)x,
                                         (const double*)b, dmask, (double*)y, lpr, stream)
               : launch_csr_spmv<float>(nrows, rp, ci, (const float*)val, (const float*)x,
                                        (const float*)b, dmask, (float*)y, lpr, stream);
  if (rc) mpet_set_error("mpet_csr_spmv: bad lanes-per-row or launch failure");
  return rc;
}

extern "C" int mpet_soa_convert(int dtype, const double* src, int ncomp, long long n, void* dst,
                                hipStream_t stream) {
  if (!src || !dst || ncomp < 1 || n < 0) {
    mpet_set_error("mpet_soa_convert: invalid argument");
    return ERR_INVALID;
  }
  int rc = dtype == 0 ? launch_soa_convert<double>(src, ncomp, n, (double*)dst, stream)
                      : launch_soa_convert<float>(src, ncomp, n, (float*)dst, stream);
  if (rc) mpet_set_error("mpet_soa_convert: launch failed");
  return rc;
}
