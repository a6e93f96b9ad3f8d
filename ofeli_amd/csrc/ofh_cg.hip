// ofh_cg.hip -- preconditioned CG on the GPU (CG.h:99-157 replacement).
//
// Per iteration three kernels, all stream-ordered and all reading their
// scalars (alpha, beta, the stop flag) from a device-side state block, so the
// host only synchronises once per chunk of iterations:
//   k_cg_dir   p = z + beta p          (z = D^-1 r fused for DIAG_PREC)
//   k_cg_spmv  q = A p, (p,q)  -> alpha = rho/(p,q)     (CSR-stream SpMV)
//   k_cg_upd   x += alpha p, r -= alpha q, (r,r), z = P r, (r,z)
//              -> stop test ||r||/||b|| <= toler, beta for the next step
// SSOR_PREC inserts the level-scheduled sweeps and a (r,z) reduction between
// k_cg_upd and the next k_cg_dir.  Every kernel returns at once when the stop
// flag is set, so a chunk enqueued past convergence costs only launch slots.
// The arithmetic of each vector entry is the reference's (p_i = z_i+beta*p_i,
// x_i += alpha*p_i, r_i -= alpha*q_i, q = A p by sequential row sums); only
// the dot products are summed as fixed-order trees instead of left to right.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "ofh_internal.h"
#include "ofh_solver.h"
#include "ofh_spmv.h"

namespace ofh {

template <int PREC>
__device__ __forceinline__ double prec_z(int i, const double* __restrict__ r,
                                         const double* __restrict__ dinv,
                                         const double* __restrict__ z) {
  if (PREC == OFH_DIAG_PREC) return r[i] * dinv[i];  // Prec_impl.h:245-247
  if (PREC == OFH_IDENT_PREC) return r[i];
  return z[i];
}

// r = b - A x, z = P r, (b,b) (r,r) (r,z); the last block sets up the state.
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_cg_init(int nblocks, const int* __restrict__ rb,
                                                   const int* __restrict__ rp,
                                                   const int* __restrict__ col,
                                                   const double* __restrict__ val,
                                                   const double* __restrict__ b,
                                                   const double* __restrict__ x,
                                                   double* __restrict__ r,
                                                   const double* __restrict__ dinv,
                                                   KrylovState* st, double* partials,
                                                   unsigned* counter) {
  __shared__ double lds[kTile64];
  __shared__ double red[3 * kBlock / kWave];
  __shared__ bool last;
  const int blk = xcd_block(blockIdx.x, nblocks);
  const int r0 = rb[blk], r1 = rb[blk + 1];
  const double s = csr_stream_rowsum<double, kTile64>(r0, r1, rp, col, val, x, lds, 0.0);
  double v[3] = {0.0, 0.0, 0.0};
  const int i = r0 + threadIdx.x;
  const bool own = i < r1;  // a long-row block holds one row, summed in thread 0
  if (own) {
    const double bi = b[i], ri = bi - s;
    r[i] = ri;
    v[0] = bi * bi;
    v[1] = ri * ri;
    if (PREC != OFH_SSOR_PREC) {
      const double zi = PREC == OFH_DIAG_PREC ? ri * dinv[i] : ri;
      v[2] = ri * zi;
    }
  }
  block_sum_k<kBlock, 3>(v, red);
  if (publish_partials<3>(v, partials, counter, blockIdx.x, gridDim.x, &last)) {
    double t[3];
    fold_partials<kBlock, 3>(t, partials, gridDim.x, red);
    if (threadIdx.x == 0) {
      double nrm = sqrt(t[0]);
      if (nrm == 0) nrm = 1;  // CG.h:113-114
      st->nrm = nrm;
      st->res = sqrt(t[1]) / nrm;
      st->res0 = st->res;
      st->rho = t[2];
      st->it = 0;
      st->done = 0;
      if (st->res <= st->tol) {  // CG.h:116
        st->done = 1;
        st->status = 0;
      }
      *counter = 0u;
    }
  }
}

// p = z + beta p (p = z on the first step), CG.h:127-133
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_cg_dir(int n, const double* __restrict__ r,
                                                  const double* __restrict__ dinv,
                                                  const double* __restrict__ z,
                                                  double* __restrict__ p,
                                                  const KrylovState* __restrict__ st) {
  if (st->done) return;
  const bool first = st->it == 0;
  const double beta = st->beta;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double zi = prec_z<PREC>(i, r, dinv, z);
    p[i] = first ? zi : zi + beta * p[i];
  }
}

// q = A p and (p,q); alpha = rho/(p,q), CG.h:134-135
__global__ __launch_bounds__(kBlock) void k_cg_spmv(int nblocks, const int* __restrict__ rb,
                                                   const int* __restrict__ rp,
                                                   const int* __restrict__ col,
                                                   const double* __restrict__ val,
                                                   const double* __restrict__ p,
                                                   double* __restrict__ q, KrylovState* st,
                                                   double* partials, unsigned* counter) {
  if (st->done) return;
  __shared__ double lds[kTile64];
  __shared__ double red[kBlock / kWave];
  __shared__ bool last;
  const int blk = xcd_block(blockIdx.x, nblocks);
  const int r0 = rb[blk], r1 = rb[blk + 1];
  const double s = csr_stream_rowsum<double, kTile64>(r0, r1, rp, col, val, p, lds, 0.0);
  const int i = r0 + threadIdx.x;
  const bool own = i < r1;  // a long-row block holds one row, summed in thread 0
  double v[1] = {0.0};
  if (own) {
    q[i] = s;
    v[0] = p[i] * s;
  }
  block_sum_k<kBlock, 1>(v, red);
  if (publish_partials<1>(v, partials, counter, blockIdx.x, gridDim.x, &last)) {
    double t[1];
    fold_partials<kBlock, 1>(t, partials, gridDim.x, red);
    if (threadIdx.x == 0) {
      st->alpha = st->rho / t[0];
      *counter = 0u;
    }
  }
}

// x += alpha p; r -= alpha q; res = ||r||/||b||; z = P r and (r,z) for the
// next step (CG.h:136-152, 126).
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_cg_upd(int n, double* __restrict__ x,
                                                  double* __restrict__ r,
                                                  const double* __restrict__ p,
                                                  const double* __restrict__ q,
                                                  const double* __restrict__ dinv,
                                                  KrylovState* st, double* partials,
                                                  unsigned* counter) {
  if (st->done) return;
  __shared__ double red[2 * kBlock / kWave];
  __shared__ bool last;
  const double alpha = st->alpha;
  double v[2] = {0.0, 0.0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    v[0] += ri * ri;
    if (PREC != OFH_SSOR_PREC) {
      const double zi = PREC == OFH_DIAG_PREC ? ri * dinv[i] : ri;
      v[1] += ri * zi;
    }
  }
  block_sum_k<kBlock, 2>(v, red);
  if (publish_partials<2>(v, partials, counter, blockIdx.x, gridDim.x, &last)) {
    double t[2];
    fold_partials<kBlock, 2>(t, partials, gridDim.x, red);
    if (threadIdx.x == 0) {
      const int it = st->it + 1;
      st->it = it;
      const double res = sqrt(t[0]) / st->nrm;
      st->res = res;
      if (res <= st->tol) {
        st->done = 1;
        st->status = it;
      } else if (it >= st->max_it || !(res == res)) {
        // CG.h:154-156: the loop counter ends at max_it+1.  A NaN residual
        // can never pass the test, so the reference's outcome is already known.
        st->done = 1;
        st->status = -(st->max_it + 1);
      } else if (PREC != OFH_SSOR_PREC) {
        const double rho = t[1];
        st->beta = rho / st->rho;  // CG.h:131
        st->rho = rho;
      }
      *counter = 0u;
    }
  }
}

// SSOR: rho = (r,z) after the sweeps; beta for the next direction.
__global__ __launch_bounds__(kBlock) void k_cg_rz(int n, const double* __restrict__ r,
                                                 const double* __restrict__ z,
                                                 KrylovState* st, double* partials,
                                                 unsigned* counter) {
  if (st->done) return;
  __shared__ double red[kBlock / kWave];
  __shared__ bool last;
  double v[1] = {0.0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    v[0] += r[i] * z[i];
  block_sum_k<kBlock, 1>(v, red);
  if (publish_partials<1>(v, partials, counter, blockIdx.x, gridDim.x, &last)) {
    double t[1];
    fold_partials<kBlock, 1>(t, partials, gridDim.x, red);
    if (threadIdx.x == 0) {
      if (st->it > 0) st->beta = t[0] / st->rho;
      st->rho = t[0];
      *counter = 0u;
    }
  }
}

template <int PREC>
static int cg_impl(ofh_csr A, const double* b, double* x, int max_it, double toler,
                   ofh_solve_info* info) {
  ofh_ctx ctx = A->ctx;
  hipStream_t s = ctx->stream;
  const int n = A->n;
  const double t0 = now_seconds();
  int rc;
  if (PREC == OFH_DIAG_PREC && (rc = csr_ensure_dinv(A))) return rc;
  if (PREC == OFH_SSOR_PREC && (rc = csr_ensure_levels(A))) return rc;
  void* wsp;
  const int nv = PREC == OFH_SSOR_PREC ? 4 : 3;
  if ((rc = ensure_vws(ctx, sizeof(double) * (size_t)n * nv, &wsp))) return rc;
  double* r = (double*)wsp;
  double* p = r + n;
  double* q = p + n;
  double* z = PREC == OFH_SSOR_PREC ? q + n : nullptr;
  KrylovState* st = state_ptr(ctx);
  KrylovState h0{};
  h0.tol = toler;
  h0.max_it = max_it;
  h0.done = 0;
  OFH_HIP(hipMemcpyAsync(st, &h0, sizeof(h0), hipMemcpyHostToDevice, s));
  unsigned* c0 = ctx->ws.counters;
  double* part = ctx->ws.partials;
  const int g = reduce_grid(n);
  hipLaunchKernelGGL(k_cg_init<PREC>, dim3(A->nblocks), dim3(kBlock), 0, s, A->nblocks,
                     A->row_blocks, A->row_ptr, A->col, A->val, b, x, r, A->dinv, st, part,
                     c0);
  OFH_HIP(hipGetLastError());
  const int* d_done = &st->done;
  if (PREC == OFH_SSOR_PREC) {
    if ((rc = launch_ssor(A, r, z, d_done, s))) return rc;
    hipLaunchKernelGGL(k_cg_rz, dim3(g), dim3(kBlock), 0, s, n, r, z, st, part, c0 + 1);
  }
  KrylovState h{};
  int enq = 0, chunk = 4;
  for (;;) {
    for (int c = 0; c < chunk && enq < max_it; ++c, ++enq) {
      hipLaunchKernelGGL(k_cg_dir<PREC>, dim3(g), dim3(kBlock), 0, s, n, r, A->dinv, z, p, st);
      hipLaunchKernelGGL(k_cg_spmv, dim3(A->nblocks), dim3(kBlock), 0, s, A->nblocks,
                         A->row_blocks, A->row_ptr, A->col, A->val, p, q, st, part, c0 + 2);
      hipLaunchKernelGGL(k_cg_upd<PREC>, dim3(g), dim3(kBlock), 0, s, n, x, r, p, q, A->dinv,
                         st, part, c0 + 3);
      if (PREC == OFH_SSOR_PREC) {
        if ((rc = launch_ssor(A, r, z, d_done, s))) return rc;
        hipLaunchKernelGGL(k_cg_rz, dim3(g), dim3(kBlock), 0, s, n, r, z, st, part, c0 + 1);
      }
    }
    OFH_HIP(hipGetLastError());
    if ((rc = read_state(ctx, &h))) return rc;
    if (h.done || enq >= max_it) break;
    chunk = next_chunk(h.it, h.res, h.res0, toler, max_it - enq);
  }
  return fill_info(A, b, x, h, t0, info);
}

int cg_dispatch(ofh_csr A, int prec, const double* b, double* x, int max_it, double toler,
                ofh_solve_info* info) {
  OFH_CHECK(A && b && x, OFH_E_INVALID, "ofh_cg: null argument");
  OFH_CHECK(max_it >= 0, OFH_E_INVALID, "ofh_cg: negative max_it");
  switch (prec) {
    case OFH_IDENT_PREC: return cg_impl<OFH_IDENT_PREC>(A, b, x, max_it, toler, info);
    case OFH_DIAG_PREC: return cg_impl<OFH_DIAG_PREC>(A, b, x, max_it, toler, info);
    case OFH_SSOR_PREC: return cg_impl<OFH_SSOR_PREC>(A, b, x, max_it, toler, info);
    case OFH_DILU_PREC:
    case OFH_ILU_PREC:
      return fail(OFH_E_UNSUPPORTED, "DILU_PREC/ILU_PREC are not provided by this path");
    default: return fail(OFH_E_INVALID, "unknown preconditioner");
  }
}

}  // namespace ofh

extern "C" int ofh_cg(ofh_csr A, int prec, const double* b, double* x, int max_it,
                      double toler, ofh_solve_info* info) {
  return ofh::cg_dispatch(A, prec, b, x, max_it, toler, info);
}
