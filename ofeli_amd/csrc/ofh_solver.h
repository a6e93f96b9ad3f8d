// ofh_solver.h -- device-side solver state and the host chunk loop helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "ofh_internal.h"

namespace ofh {

// Scalars of one Krylov solve, resident on the device.  Block 0 of the kernel
// that folds a reduction writes them; fields read by the blocks of a kernel
// are never written by that same kernel (rho is double-buffered by the
// iteration parity, which the host passes as a launch argument).
struct KrylovState {
  double rho_buf[2], alpha, beta, omega, nrm, res, tol;
  double res0, sconv_res, h_norm, rmax;
  double bnorm, rnorm;  // BiCG: ||b|| as is, ||r|| of the last check
  int it, done, status, max_it;
  int sconv, cycle_i, halt, pause;  // halt = done || sconv || pause
  int replacements;
  int kcur;  // iteration index of the current update kernel, for kernels
             // launched with it = -1 (captured chunks of iterations)
  int pad[2];
  double rscale;  // mixed CG on D^-1/2 A D^-1/2: true relres / ||r32|| at the
                  // last replacement (estimates the relres from the scaled ||r32||)
};
static_assert(sizeof(KrylovState) <= 512, "state block");

// The state scalars a CG decision reads, loaded together at the start of a
// kernel so that they arrive with its other loads instead of one dependent
// load after another behind the partial fold.  rho_prev is rho_{it-1} (an
// unused slot when it == 0).
struct CgSnap {
  double nrm, tol, rho_prev;
  int max_it;
};

__device__ __forceinline__ CgSnap cg_snap(const KrylovState* st, int it) {
  return CgSnap{st->nrm, st->tol, st->rho_buf[(it - 1) & 1], st->max_it};
}

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

// Iterations a solver loop enqueues behind each copy of its state.
constexpr int kLookAhead = 4;

// The state as it is once everything enqueued so far has run, read while
// the launches `more()` enqueues behind its copy keep the GPU busy.
template <class F>
inline int read_state_ahead(ofh_ctx ctx, KrylovState* h, F&& more) {
  if (!ctx->look_ev) OFH_HIP(hipEventCreateWithFlags(&ctx->look_ev, hipEventDisableTiming));
  OFH_HIP(hipMemcpyAsync(ctx->host_pinned, state_ptr(ctx), sizeof(KrylovState),
                         hipMemcpyDeviceToHost, ctx->stream));
  OFH_HIP(hipEventRecord(ctx->look_ev, ctx->stream));
  if (const int rc = more()) return rc;
  OFH_HIP(hipEventSynchronize(ctx->look_ev));
  std::memcpy(h, ctx->host_pinned, sizeof(KrylovState));
  return OFH_OK;
}

// Iterations to enqueue before the next host check.  The residual reduction
// observed since the previous check predicts the remaining count, capped by
// the prediction from the rate since the solve's first check (`res0` after
// `it0` iterations): the norm of a Krylov residual is not monotone, and on a
// plateau the rate over one chunk is ~1, which alone would predict hundreds
// of iterations past the stop.  A chunk is at least 4 and at most as long as
// the iterations already done (>= 16).
inline int next_chunk(int it, double res, int it_prev, double res_prev, double target,
                      int remaining, int it0 = 0, double res0 = 0.0) {
  int c = 8;
  if (it > it_prev && res_prev > 0 && res > 0 && res < res_prev && target > 0 && res > target) {
    const double rate = std::pow(res / res_prev, 1.0 / (it - it_prev));
    double pred = std::log(target / res) / std::log(rate);
    if (it > it0 && res0 > 0 && res < res0)
      pred = std::min(pred, std::log(target / res) * (it - it0) / std::log(res / res0));
    c = (int)std::ceil(pred);
  }
  c = std::max(4, std::min(c, std::max(16, it)));
  return std::max(1, std::min(c, remaining));
}

int fill_info(ofh_csr A, const double* b, const double* x, const KrylovState& h, double t0,
              ofh_solve_info* info);

}  // namespace ofh
