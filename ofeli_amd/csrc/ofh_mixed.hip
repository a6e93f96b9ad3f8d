// ofh_mixed.hip -- mixed-precision CG (no reference counterpart; judged against
// the FP64 CG of CG.h:99-157).
//
// Inner iteration in FP32: SpMV on the FP32 copy of the values (8 B/nonzero
// instead of 12), FP32 vectors r, p, q and the correction x32.  FP64 is kept
// for (a) every Krylov dot product (FP32 products accumulated in FP64),
// (b) the solution, x64 += x32 at each residual replacement, and (c) the
// residual replacement itself, r = b - A x64 with the FP64 matrix, after
// which the FP32 residual restarts from the rounded FP64 one while the search
// direction p is kept (reliable updates, Sleijpen & van der Vorst 1996).
// A replacement happens whenever the recursive residual has dropped by
// kDelta since the last one, or claims convergence; the solve stops only when
// the true FP64 residual ||b - A x64|| / ||b|| <= toler.  If the true residual
// stops improving, the FP64 CG finishes from the current x64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "ofh_internal.h"
#include "ofh_solver.h"
#include "ofh_spmv.h"

namespace ofh {

constexpr double kDelta = 0.1;  // recursive-residual drop that triggers a replacement

struct MixedState {
  double rho, rho_1, alpha, beta, nrm, res_rec, res_true, rmax;
  double tol, res0, pad[6];
  int it, done, status, max_it;
  int pause, halt, replacements, pad1;
};
static_assert(sizeof(MixedState) <= 512, "state block");

inline MixedState* mstate_ptr(ofh_ctx ctx) {
  return reinterpret_cast<MixedState*>(reinterpret_cast<char*>(ctx->dev_scratch) + 1024);
}

template <int PREC>
__device__ __forceinline__ float zf(int i, const float* __restrict__ r, const float* __restrict__ d) {
  return PREC == OFH_DIAG_PREC ? r[i] * d[i] : r[i];
}

// FP64 residual r = b - A x64 -> r32; (b,b) | (r,r) in FP64, (r32,z32), (r32,r32).
// INIT: set up the state; otherwise a residual replacement.
template <int PREC, bool INIT>
__global__ __launch_bounds__(kBlock) void k_mx_resid(
    int nblocks, const int* __restrict__ rb, const int* __restrict__ rp,
    const int* __restrict__ col, const double* __restrict__ val, const double* __restrict__ b,
    const double* __restrict__ x, float* __restrict__ r32, float* __restrict__ x32,
    const float* __restrict__ dinv32, MixedState* st, double* partials, unsigned* counter) {
  if (!INIT && st->done) return;
  __shared__ double lds[kTile64];
  __shared__ double red[4 * kBlock / kWave];
  __shared__ bool last;
  const int blk = xcd_block(blockIdx.x, nblocks);
  const int r0 = rb[blk], r1 = rb[blk + 1];
  const double s = csr_stream_rowsum<double, kTile64>(r0, r1, rp, col, val, x, lds, 0.0);
  const int i = r0 + threadIdx.x;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  if (i < r1) {
    const double bi = b[i], ri = bi - s;
    const float rf = (float)ri;
    r32[i] = rf;
    x32[i] = 0.0f;
    const float z = PREC == OFH_DIAG_PREC ? rf * dinv32[i] : rf;
    v[0] = bi * bi;
    v[1] = ri * ri;
    v[2] = (double)rf * (double)z;
    v[3] = (double)rf * (double)rf;
  }
  block_sum_k<kBlock, 4>(v, red);
  if (publish_partials<4>(v, partials, counter, blockIdx.x, gridDim.x, &last)) {
    double t[4];
    fold_partials<kBlock, 4>(t, partials, gridDim.x, red);
    if (threadIdx.x == 0) {
      if (INIT) {
        double nrm = sqrt(t[0]);
        if (nrm == 0) nrm = 1;
        st->nrm = nrm;
        st->it = 0;
        st->replacements = 0;
      } else {
        st->replacements += 1;
      }
      const double res = sqrt(t[1]) / st->nrm;
      st->res_true = res;
      st->res_rec = sqrt(t[3]) / st->nrm;
      if (INIT) st->res0 = res;
      st->rmax = res;
      st->pause = 0;
      if (res <= st->tol) {
        st->done = 1;
        st->status = st->it;
      } else if (!INIT && (st->it >= st->max_it || !(res == res))) {
        st->done = 1;
        st->status = -(st->max_it + 1);
      } else {
        if (!INIT) st->beta = t[2] / st->rho_1;  // keep the direction p
        st->rho = t[2];
      }
      st->halt = st->done;
      *counter = 0u;
    }
  }
}

template <int PREC>
__global__ __launch_bounds__(kBlock) void k_mx_dir(int n, const float* __restrict__ r32,
                                                  const float* __restrict__ dinv32,
                                                  float* __restrict__ p32,
                                                  const MixedState* __restrict__ st) {
  if (st->halt) return;
  const bool first = st->it == 0;
  const float beta = (float)st->beta;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const float z = zf<PREC>(i, r32, dinv32);
    p32[i] = first ? z : z + beta * p32[i];
  }
}

__global__ __launch_bounds__(kBlock) void k_mx_spmv(int nblocks, const int* __restrict__ rb,
                                                   const int* __restrict__ rp,
                                                   const int* __restrict__ col,
                                                   const float* __restrict__ val32,
                                                   const float* __restrict__ p32,
                                                   float* __restrict__ q32, MixedState* st,
                                                   double* partials, unsigned* counter) {
  if (st->halt) return;
  __shared__ float lds[kTile32];
  __shared__ double red[kBlock / kWave];
  __shared__ bool last;
  const int blk = xcd_block(blockIdx.x, nblocks);
  const int r0 = rb[blk], r1 = rb[blk + 1];
  const float s = csr_stream_rowsum<float, kTile32>(r0, r1, rp, col, val32, p32, lds, 0.0f);
  const int i = r0 + threadIdx.x;
  double v[1] = {0.0};
  if (i < r1) {
    q32[i] = s;
    v[0] = (double)p32[i] * (double)s;
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

template <int PREC>
__global__ __launch_bounds__(kBlock) void k_mx_upd(int n, float* __restrict__ x32,
                                                  float* __restrict__ r32,
                                                  const float* __restrict__ p32,
                                                  const float* __restrict__ q32,
                                                  const float* __restrict__ dinv32,
                                                  MixedState* st, double* partials,
                                                  unsigned* counter) {
  if (st->halt) return;
  __shared__ double red[2 * kBlock / kWave];
  __shared__ bool last;
  const float alpha = (float)st->alpha;
  double v[2] = {0.0, 0.0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    x32[i] += alpha * p32[i];
    const float ri = r32[i] - alpha * q32[i];
    r32[i] = ri;
    const float z = PREC == OFH_DIAG_PREC ? ri * dinv32[i] : ri;
    v[0] += (double)ri * (double)ri;
    v[1] += (double)ri * (double)z;
  }
  block_sum_k<kBlock, 2>(v, red);
  if (publish_partials<2>(v, partials, counter, blockIdx.x, gridDim.x, &last)) {
    double t[2];
    fold_partials<kBlock, 2>(t, partials, gridDim.x, red);
    if (threadIdx.x == 0) {
      const int it = st->it + 1;
      st->it = it;
      const double res = sqrt(t[0]) / st->nrm;
      st->res_rec = res;
      st->rho_1 = st->rho;
      st->beta = t[1] / st->rho;
      st->rho = t[1];
      if (res <= kDelta * st->rmax || res <= st->tol || it >= st->max_it || !(res == res)) {
        st->pause = 1;
        st->halt = 1;
      }
      *counter = 0u;
    }
  }
}

// x64 += x32 (the correction accumulated since the last replacement)
__global__ void k_mx_accum(int n, double* __restrict__ x, const float* __restrict__ x32,
                           const MixedState* __restrict__ st) {
  if (st->done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += (double)x32[i];
}

template <int PREC>
static int mixed_impl(ofh_csr A, const double* b, double* x, int max_it, double toler,
                      ofh_solve_info* info) {
  ofh_ctx ctx = A->ctx;
  hipStream_t s = ctx->stream;
  const int n = A->n;
  const double t0 = now_seconds();
  int rc;
  if (PREC == OFH_DIAG_PREC && (rc = csr_ensure_dinv(A))) return rc;
  if ((rc = csr_ensure_val32(A))) return rc;
  void* wsp;
  if ((rc = ensure_vws(ctx, sizeof(float) * (size_t)n * 4, &wsp))) return rc;
  float* r32 = (float*)wsp;
  float* p32 = r32 + n;
  float* q32 = p32 + n;
  float* x32 = q32 + n;
  MixedState* st = mstate_ptr(ctx);
  MixedState h0{};
  h0.tol = toler;
  h0.max_it = max_it;
  OFH_HIP(hipMemcpyAsync(st, &h0, sizeof(h0), hipMemcpyHostToDevice, s));
  unsigned* c0 = ctx->ws.counters + 16;
  double* part = ctx->ws.partials;
  const int g = reduce_grid(n);
  const int ge = (n + 255) / 256;
  hipLaunchKernelGGL((k_mx_resid<PREC, true>), dim3(A->nblocks), dim3(kBlock), 0, s, A->nblocks,
                     A->row_blocks, A->row_ptr, A->col, A->val, b, x, r32, x32, A->dinv32, st,
                     part, c0);
  OFH_HIP(hipGetLastError());
  MixedState h{};
  if ((rc = read_dev(ctx, st, &h))) return rc;
  double last_true = h.res_true;
  int stagnant = 0, fallback = 0;
  while (!h.done) {
    // iterations expected until the next replacement
    int chunk = 32;
    if (h.it > 0 && h.res_rec < h.res0) {
      double rate = std::pow(h.res_rec / h.res0, 1.0 / h.it);
      double target = std::max(kDelta * h.rmax, toler);
      if (rate < 1.0 && target < h.res_rec)
        chunk = (int)std::ceil(std::log(target / h.res_rec) / std::log(rate) * 1.05) + 2;
    }
    chunk = std::max(4, std::min(chunk, 1024));
    chunk = std::max(1, std::min(chunk, max_it - h.it));
    for (int c = 0; c < chunk; ++c) {
      hipLaunchKernelGGL(k_mx_dir<PREC>, dim3(g), dim3(kBlock), 0, s, n, r32, A->dinv32, p32, st);
      hipLaunchKernelGGL(k_mx_spmv, dim3(A->nblocks), dim3(kBlock), 0, s, A->nblocks,
                         A->row_blocks, A->row_ptr, A->col, A->val32, p32, q32, st, part, c0 + 1);
      hipLaunchKernelGGL(k_mx_upd<PREC>, dim3(g), dim3(kBlock), 0, s, n, x32, r32, p32, q32,
                         A->dinv32, st, part, c0 + 2);
    }
    OFH_HIP(hipGetLastError());
    if ((rc = read_dev(ctx, st, &h))) return rc;
    if (!h.pause && h.it < max_it) continue;
    // residual replacement in FP64
    hipLaunchKernelGGL(k_mx_accum, dim3(ge), dim3(256), 0, s, n, x, x32, st);
    hipLaunchKernelGGL((k_mx_resid<PREC, false>), dim3(A->nblocks), dim3(kBlock), 0, s,
                       A->nblocks, A->row_blocks, A->row_ptr, A->col, A->val, b, x, r32, x32,
                       A->dinv32, st, part, c0 + 3);
    OFH_HIP(hipGetLastError());
    if ((rc = read_dev(ctx, st, &h))) return rc;
    if (h.done) break;
    // stagnation: the true residual must keep falling
    stagnant = (h.res_true > 0.5 * last_true) ? stagnant + 1 : 0;
    last_true = std::min(last_true, h.res_true);
    if (stagnant >= 3) {
      fallback = 1;
      break;
    }
  }
  if (fallback) {
    ofh_solve_info sub{};
    const int left = std::max(0, max_it - h.it);
    if ((rc = ofh_cg(A, PREC, b, x, left, toler, &sub))) return rc;
    if (info) {
      *info = sub;
      info->iterations += h.it;
      info->outer = h.replacements;
      info->fallback = 1;
      info->seconds = now_seconds() - t0;
      if (sub.nb_it > 0) info->nb_it = sub.nb_it + h.it;
    }
    return OFH_OK;
  }
  OFH_HIP(hipStreamSynchronize(s));
  const double t1 = now_seconds();
  if (info) {
    info->nb_it = h.done ? h.status : -(max_it + 1);
    info->iterations = h.it;
    info->relres = h.res_true;
    info->outer = h.replacements;
    info->fallback = 0;
    info->seconds = t1 - t0;
    if ((rc = true_relres(A, b, x, &info->true_relres))) return rc;
  }
  return OFH_OK;
}

}  // namespace ofh

extern "C" int ofh_cg_mixed(ofh_csr A, int prec, const double* b, double* x, int max_it,
                            double toler, ofh_solve_info* info) {
  using namespace ofh;
  OFH_CHECK(A && b && x && max_it >= 0, OFH_E_INVALID, "ofh_cg_mixed: bad argument");
  switch (prec) {
    case OFH_IDENT_PREC: return mixed_impl<OFH_IDENT_PREC>(A, b, x, max_it, toler, info);
    case OFH_DIAG_PREC: return mixed_impl<OFH_DIAG_PREC>(A, b, x, max_it, toler, info);
    default:
      return fail(OFH_E_UNSUPPORTED, "ofh_cg_mixed: IDENT_PREC or DIAG_PREC only");
  }
}
