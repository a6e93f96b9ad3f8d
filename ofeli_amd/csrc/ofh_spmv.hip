// ofh_spmv.hip -- CSR SpMV for gfx950 (SpMatrix::Mult / MultAdd replacement).
//
// Reference: SpMatrix_impl.h:728-757 -- row loop, y[i] += a_k * x[col_k-1],
// summed left to right.  Design here ("CSR-stream"): each 256-thread
// workgroup owns a run of consecutive rows whose nonzeros fit an LDS tile.
//   phase 1: the workgroup streams val/col of the whole run with fully
//            coalesced loads (4 independent loads in flight per thread),
//            gathers x and stores the products a_k*x_c into LDS;
//   phase 2: one thread per row adds its products from LDS left to right.
// Phase 2 keeps the reference's summation order, so with FP contraction off
// (the library is built with -ffp-contract=off) every y_i is bitwise equal to
// the reference's sequential row sum.  Workgroup ids are remapped so each
// XCD walks a contiguous range of rows (shared x window in its L2).
#include <hip/hip_runtime.h>

#include "ofh_internal.h"
#include "ofh_spmv.h"

namespace ofh {

template <typename T, int TILE, int MODE>
__global__ __launch_bounds__(kBlock) void k_spmv(int nblocks, const int* __restrict__ rb,
                                                 const int* __restrict__ rp,
                                                 ColIndex ci,
                                                 const T* __restrict__ val,
                                                 const T* __restrict__ x, T* __restrict__ y,
                                                 T alpha) {
  __shared__ T lds[TILE];
  const int b = xcd_block(blockIdx.x, nblocks);
  const int r0 = rb[b], r1 = rb[b + 1];
  const int r = r0 + threadIdx.x;
  T init = T(0);
  if (MODE != 0 && r < r1) init = y[r];
  T s = csr_stream_rowsum<T, TILE, MODE == 2>(r0, r1, rp, ci, val, x, lds, init, alpha);
  if (r < r1) y[r] = s;
}

template <typename T, int TILE, int MODE>
static int launch(ofh_csr A, const T* val, const T* x, T* y, T alpha, hipStream_t s) {
  // FP32 runs on its own partition (larger tile) with 32-bit columns: the
  // 16-bit offsets are relative to the FP64 partition's blocks
  const bool f32 = sizeof(T) == 4;
  const int nb = f32 ? A->nblocks32 : A->nblocks;
  hipLaunchKernelGGL((k_spmv<T, TILE, MODE>), dim3(nb), dim3(kBlock), 0, s, nb,
                     f32 ? A->row_blocks32 : A->row_blocks, A->row_ptr,
                     ColIndex{A->col, f32 ? nullptr : A->col16}, val, x, y, alpha);
  OFH_HIP(hipGetLastError());
  return OFH_OK;
}

int launch_spmv64(ofh_csr A, const double* x, double* y, hipStream_t s) {
  return launch<double, kTile64, 0>(A, A->val, x, y, 0.0, s);
}

int launch_spmv32(ofh_csr A, const float* x, float* y, hipStream_t s) {
  int rc = csr_ensure_val32(A);
  if (rc) return rc;
  return launch<float, kTile32, 0>(A, A->val32, x, y, 0.0f, s);
}

}  // namespace ofh

using namespace ofh;

extern "C" {

int ofh_spmv_f64(ofh_csr A, const double* x, double* y) {
  OFH_CHECK(A && x && y, OFH_E_INVALID, "ofh_spmv_f64: null argument");
  return launch_spmv64(A, x, y, A->ctx->stream);
}

int ofh_spmv_add_f64(ofh_csr A, double a, const double* x, double* y) {
  OFH_CHECK(A && x && y, OFH_E_INVALID, "ofh_spmv_add_f64: null argument");
  // MultAdd(x,y) (SpMatrix_impl.h:737-745) accumulates a_k*x onto y;
  // MultAdd(a,x,y) (:748-757) accumulates a*a_k*x.
  if (a == 1.0) return launch<double, kTile64, 1>(A, A->val, x, y, 1.0, A->ctx->stream);
  return launch<double, kTile64, 2>(A, A->val, x, y, a, A->ctx->stream);
}

int ofh_spmv_f32(ofh_csr A, const float* x, float* y) {
  OFH_CHECK(A && x && y, OFH_E_INVALID, "ofh_spmv_f32: null argument");
  return launch_spmv32(A, x, y, A->ctx->stream);
}

}  // extern "C"
