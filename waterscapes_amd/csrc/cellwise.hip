// C entry points of the cell-wise (matrix-free) operator, FP32 and FP64:
// geometry set-up and the plain application (the fused MINRES step and the
// refinement's residuals live in krylov.hip).  The kernels are in
// cellwise.h.
#include "cellwise.h"

namespace mpet {

// One thread per cell: gradients of the barycentric coordinates (rows of
// J^-1, J's columns v_k - v_0) and |T|, in FP64, stored as G (FP32:
// rounded).
template <typename G>
__global__ void __launch_bounds__(256)
cw_geometry_kernel(int nC, const double* __restrict__ X, const int* __restrict__ cells,
                   G* __restrict__ geom) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nC) return;
  const int v0 = cells[4 * e], v1 = cells[4 * e + 1], v2 = cells[4 * e + 2],
            v3 = cells[4 * e + 3];
  double J[3][3];
  for (int r = 0; r < 3; ++r) {
    J[r][0] = X[3 * v1 + r] - X[3 * v0 + r];
    J[r][1] = X[3 * v2 + r] - X[3 * v0 + r];
    J[r][2] = X[3 * v3 + r] - X[3 * v0 + r];
  }
  const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  const double id = 1.0 / det;
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
  const long long n = nC;
  for (int r = 0; r < 3; ++r)
    for (int m = 0; m < 3; ++m) geom[(3 * r + m) * n + e] = (G)Ji[r][m];
  geom[9 * n + e] = (G)(fabs(det) / 6.0);
}

}  // namespace mpet

extern "C" int mpet_cell_geometry(const MpetMeshDesc* mesh, float* geom, hipStream_t stream) {
  if (!mesh || mesh->nC <= 0 || !mesh->coords || !mesh->cells || !geom) {
    mpet_set_error("mpet_cell_geometry: invalid arguments");
    return ERR_INVALID;
  }
  hipLaunchKernelGGL((cw_geometry_kernel<float>), dim3((mesh->nC + 255) / 256), dim3(256), 0,
                     stream, mesh->nC, mesh->coords, mesh->cells, geom);
  HIPCHK(hipGetLastError());
  return OK;
}

extern "C" int mpet_cell_geometry64(const MpetMeshDesc* mesh, double* geom, hipStream_t stream) {
  if (!mesh || mesh->nC <= 0 || !mesh->coords || !mesh->cells || !geom) {
    mpet_set_error("mpet_cell_geometry64: invalid arguments");
    return ERR_INVALID;
  }
  hipLaunchKernelGGL((cw_geometry_kernel<double>), dim3((mesh->nC + 255) / 256), dim3(256), 0,
                     stream, mesh->nC, mesh->coords, mesh->cells, geom);
  HIPCHK(hipGetLastError());
  return OK;
}

template <typename T>
static int cell_op_apply_t(const MpetCellOpDesc* d, const void* x, const void* b, void* y,
                           void* work, hipStream_t stream) {
  CellOpT<T> op;
  int rc = cw_from_desc<T>(d, (T*)work, op);
  if (rc) return rc;
  if (!x || !y) {
    mpet_set_error("mpet_cell_op_apply: null vector");
    return ERR_INVALID;
  }
  CwEpiApplyT<T> epi{(T*)y, (const T*)b, op.dmask};
  cw_launch(op, (const T*)x, epi, (double*)nullptr, (const int*)nullptr, stream);
  HIPCHK(hipGetLastError());
  return OK;
}

extern "C" int mpet_cell_op_apply(const MpetCellOpDesc* d, const void* x, const void* b,
                                  void* y, void* work, hipStream_t stream) {
  if (!d) {
    mpet_set_error("invalid cell-wise operator description");
    return ERR_INVALID;
  }
  return d->fp64 ? cell_op_apply_t<double>(d, x, b, y, work, stream)
                 : cell_op_apply_t<float>(d, x, b, y, work, stream);
}
