// ofh_solver.h -- device-side solver state and the host chunk loop helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "ofh_internal.h"

namespace ofh {

// Scalars of one Krylov solve, resident on the device.  Kernels read them at
// launch and the last block of each reduction writes them.
struct KrylovState {
  double rho, rho_1, alpha, beta, omega, nrm, res, tol;
  double res0, sconv_res, h_norm, aux;
  int it, done, status, max_it;
  int sconv, cycle_i, halt, pad1;  // halt = done || sconv
};
static_assert(sizeof(KrylovState) <= 512, "state block");

inline KrylovState* state_ptr(ofh_ctx ctx) {
  return reinterpret_cast<KrylovState*>(reinterpret_cast<char*>(ctx->dev_scratch) + 512);
}

template <class S>
inline int read_dev(ofh_ctx ctx, const S* d, S* h) {
  static_assert(sizeof(S) <= 4096, "pinned staging");
  OFH_HIP(hipMemcpyAsync(ctx->host_pinned, d, sizeof(S), hipMemcpyDeviceToHost, ctx->stream));
  OFH_HIP(hipStreamSynchronize(ctx->stream));
  std::memcpy(h, ctx->host_pinned, sizeof(S));
  return OFH_OK;
}

inline int read_state(ofh_ctx ctx, KrylovState* h) { return read_dev(ctx, state_ptr(ctx), h); }

// Iterations to enqueue before the next host check: the observed geometric
// residual reduction predicts the remaining count; clamp to [4, 1024].
inline int next_chunk(int it, double res, double res0, double tol, int remaining) {
  int c = 32;
  if (it > 0 && res0 > 0 && res > 0 && res < res0 && tol > 0 && res > tol) {
    double rate = std::pow(res / res0, 1.0 / it);
    if (rate < 1.0) {
      double pred = std::log(tol / res) / std::log(rate);
      c = (int)std::ceil(pred * 1.05) + 2;
    }
  }
  c = std::max(4, std::min(c, 1024));
  return std::max(1, std::min(c, remaining));
}

int fill_info(ofh_csr A, const double* b, const double* x, const KrylovState& h, double t0,
              ofh_solve_info* info);

}  // namespace ofh
