// ofh_krylov.hip -- BiCGStab, GMRes(m), Prec::solve and the LinearSolver
// dispatch on the GPU, plus the solve report shared by every solver.
//
// References: BiCGStab.h:87-162, GMRes.h:67-241, Prec_impl.h:226-255,
// LinearSolver.cpp:194-244.  As in ofh_cg.hip, scalars live on the device,
// reductions are folded by the kernel that consumes them, and every kernel
// returns at once when the stop flag is set; the host synchronises once per
// chunk of iterations (BiCGStab) or once per restart cycle (GMRes).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "ofh_internal.h"
#include "ofh_solver.h"
#include "ofh_spmv.h"

namespace ofh {

// ---------------------------------------------------------------------------
// true residual ||b - A x|| / ||b|| (FP64, no temporary vector)
// ---------------------------------------------------------------------------
template <bool SELL>
__global__ __launch_bounds__(kBlock) void k_true_res(MatRows64 m, const double* __restrict__ b,
                                                    const double* __restrict__ x,
                                                    double* __restrict__ partials) {
  __shared__ double red[2 * kBlock / kWave];
  int i;
  bool valid;
  const double s = mat_rowsum64<SELL>(m, x, i, valid);
  double v[2] = {0.0, 0.0};
  if (valid) {
    const double d = b[i] - s;
    v[0] = d * d;
    v[1] = b[i] * b[i];
  }
  block_sum_k<kBlock, 2>(v, red);
  const int slot[2] = {0, 1};
  write_partials<2>(v, slot, partials, blockIdx.x);
}

__global__ __launch_bounds__(kBlock) void k_true_fold(int nparts, const double* __restrict__ partials,
                                                     double* out) {
  __shared__ double red[2 * kBlock / kWave];
  double t[2];
  const int slot[2] = {0, 1};
  fold_partials<kBlock, 2>(t, slot, partials, nparts, red);
  if (threadIdx.x == 0) {
    double nb = sqrt(t[1]);
    if (nb == 0) nb = 1;
    *out = sqrt(t[0]) / nb;
  }
}

int true_relres(ofh_csr A, const double* b, const double* x, double* out) {
  ofh_ctx ctx = A->ctx;
  double* d_out = ctx->dev_scratch + 16;
  MatRows64 mr;
  bool sell = false;
  int rc = mat_rows64(A, &mr, &sell);
  if (rc) return rc;
  if (sell)
    hipLaunchKernelGGL(k_true_res<true>, dim3(mr.nblocks), dim3(kBlock), 0, ctx->stream, mr, b, x,
                       ctx->ws.partials);
  else
    hipLaunchKernelGGL(k_true_res<false>, dim3(mr.nblocks), dim3(kBlock), 0, ctx->stream, mr, b,
                       x, ctx->ws.partials);
  hipLaunchKernelGGL(k_true_fold, dim3(1), dim3(kBlock), 0, ctx->stream, mr.nblocks,
                     ctx->ws.partials, d_out);
  OFH_HIP(hipGetLastError());
  OFH_HIP(hipMemcpyAsync(ctx->host_pinned, d_out, sizeof(double), hipMemcpyDeviceToHost,
                         ctx->stream));
  OFH_HIP(hipStreamSynchronize(ctx->stream));
  *out = *(double*)ctx->host_pinned;
  return OFH_OK;
}

// The same from the CSR-stream partition only, with caller-provided partial
// slots and result word (nothing is built or allocated).
int true_relres_csr(ofh_csr A, const double* b, const double* x, double* partials, double* d_out,
                    double* out) {
  ofh_ctx ctx = A->ctx;
  const MatRows64 mr{A->nblocks, A->row_blocks, A->row_ptr, ColIndex{A->col, A->col16}, A->val,
                     Sell64{A->n, nullptr, nullptr, nullptr, nullptr, nullptr}};
  hipLaunchKernelGGL(k_true_res<false>, dim3(mr.nblocks), dim3(kBlock), 0, ctx->stream, mr, b, x,
                     partials);
  hipLaunchKernelGGL(k_true_fold, dim3(1), dim3(kBlock), 0, ctx->stream, mr.nblocks, partials,
                     d_out);
  OFH_HIP(hipGetLastError());
  OFH_HIP(hipMemcpyAsync(ctx->host_pinned, d_out, sizeof(double), hipMemcpyDeviceToHost,
                         ctx->stream));
  OFH_HIP(hipStreamSynchronize(ctx->stream));
  *out = *(double*)ctx->host_pinned;
  return OFH_OK;
}

int fill_info(ofh_csr A, const double* b, const double* x, const KrylovState& h, double t0,
              ofh_solve_info* info) {
  OFH_HIP(hipStreamSynchronize(A->ctx->stream));
  const double t1 = now_seconds();
  if (!info) return OFH_OK;
  info->nb_it = h.done ? h.status : -(h.max_it + 1);
  info->iterations = h.it;
  info->relres = h.res;
  info->outer = 0;
  info->fallback = 0;
  info->seconds = t1 - t0;
  return true_relres(A, b, x, &info->true_relres);
}

// Prec::solve(b,x) (Prec_impl.h:226-255) for IDENT and DIAG.
__global__ void k_prec_diag(int n, const double* __restrict__ b, const double* __restrict__ dinv,
                            double* __restrict__ x) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = dinv ? b[i] * dinv[i] : b[i];
}

// ===========================================================================
// BiCGStab
// ===========================================================================
// r = b - A x, rr = r; partials (b,b) -> 0, (r,r) -> 1  (BiCGStab.h:101-103)
__global__ __launch_bounds__(kBlock) void k_bi_init(int nblocks, const int* __restrict__ rb,
                                                   const int* __restrict__ rp,
                                                   ColIndex ci,
                                                   const double* __restrict__ val,
                                                   const double* __restrict__ b,
                                                   const double* __restrict__ x,
                                                   double* __restrict__ r,
                                                   double* __restrict__ rr,
                                                   double* __restrict__ partials) {
  __shared__ double lds[kTile64];
  __shared__ double red[2 * kBlock / kWave];
  const int blk = xcd_block(blockIdx.x, nblocks);
  const int r0 = rb[blk], r1 = rb[blk + 1];
  const double s = csr_stream_rowsum<double, kTile64>(r0, r1, rp, ci, val, x, lds, 0.0);
  const int i = r0 + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (i < r1) {
    const double ri = b[i] - s;
    r[i] = ri;
    rr[i] = ri;
    v[0] = b[i] * b[i];
    v[1] = ri * ri;
  }
  block_sum_k<kBlock, 2>(v, red);
  const int slot[2] = {0, 1};
  write_partials<2>(v, slot, partials, blockIdx.x);
}

// Stop tests closing iteration it-1 and the direction of iteration it:
//   it == 1: ||r0||/||b|| <= toler -> 0 (BiCGStab.h:111-116); rho_1 = (r0,r0)
//   it >  1: ||r||/||b|| < toler -> it-1; omega == 0 -> 3 (BiCGStab.h:149-157);
//            loop end -> -(max_it+1); rho_1 = (rr,r) == 0 -> 2 (:117-120);
//            beta = (rho_1/rho_2) (alpha/omega) (:125)
// then p = r + beta (p - omega v), pp = P p (:122-129).
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_bi_dir(int n, int it, int nparts,
                                                  const double* __restrict__ r,
                                                  const double* __restrict__ v,
                                                  double* __restrict__ p, double* __restrict__ pp,
                                                  const double* __restrict__ dinv,
                                                  KrylovState* st,
                                                  const double* __restrict__ partials) {
  if (st->done) return;
  __shared__ double red[2 * kBlock / kWave];
  double t[2];
  const int slot[2] = {0, 1};
  fold_partials<kBlock, 2>(t, slot, partials, nparts, red);
  bool stop = false;
  int status = 0;
  double nrm = st->nrm, res, rho, beta = 0.0;
  if (it == 1) {
    nrm = sqrt(t[0]);
    if (nrm == 0.0) nrm = 1;
    res = sqrt(t[1]) / nrm;
    rho = t[1];  // (rr, r) with rr == r
    if (res <= st->tol) { stop = true; status = 0; }
    else if (it > st->max_it) { stop = true; status = -(st->max_it + 1); }
    else if (rho == 0.0) { stop = true; status = 2; }
  } else {
    res = sqrt(t[0]) / nrm;
    rho = t[1];
    if (res < st->tol) { stop = true; status = it - 1; }
    else if (st->omega == 0.0) { stop = true; status = 3; }
    else if (it > st->max_it || !(res == res)) { stop = true; status = -(st->max_it + 1); }
    else if (rho == 0.0) { stop = true; status = 2; }
    else beta = (rho / st->rho_buf[(it - 1) & 1]) * (st->alpha / st->omega);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->it = it - 1;
    st->res = res;
    if (it == 1) {
      st->nrm = nrm;
      st->res0 = res;
    }
    if (stop) {
      st->done = 1;
      st->halt = 1;
      st->status = status;
    } else {
      st->rho_buf[it & 1] = rho;
      st->beta = beta;
    }
  }
  if (stop) return;
  const double omega = st->omega;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double pi = it == 1 ? r[i] : r[i] + beta * (p[i] - omega * v[i]);
    p[i] = pi;
    if (PREC == OFH_DIAG_PREC) pp[i] = pi * dinv[i];
    else if (PREC == OFH_IDENT_PREC) pp[i] = pi;
  }
}

// y = A x with fused partial dots:
// MODE 0: v = A pp, (rr,v) -> slot 2                 (BiCGStab.h:130-131)
// MODE 1: t = A ss, (t,s) -> slot 4, (t,t) -> slot 5  (BiCGStab.h:140-141)
template <int MODE, bool SELL>
__global__ __launch_bounds__(kBlock) void k_bi_spmv(MatRows64 m, const double* __restrict__ x,
                                                   double* __restrict__ y,
                                                   const double* __restrict__ u,
                                                   const KrylovState* __restrict__ st,
                                                   double* __restrict__ partials) {
  if (MODE == 0 ? st->done : st->halt) return;
  __shared__ double red[2 * kBlock / kWave];
  int i;
  bool valid;
  const double s = mat_rowsum64<SELL>(m, x, i, valid);
  double v[2] = {0.0, 0.0};
  if (valid) {
    y[i] = s;
    if (MODE == 0) {
      v[0] = u[i] * s;
    } else {
      v[0] = s * u[i];
      v[1] = s * s;
    }
  }
  block_sum_k<kBlock, 2>(v, red);
  const int slot[2] = {MODE == 0 ? 2 : 4, MODE == 0 ? 6 : 5};
  write_partials<2>(v, slot, partials, blockIdx.x);
}

// alpha = rho_1/(rr,v); s = r - alpha v; (s,s) -> slot 3; ss = P s
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_bi_s(int n, int it, int nparts_rv,
                                                const double* __restrict__ r,
                                                const double* __restrict__ v,
                                                double* __restrict__ s, double* __restrict__ ss,
                                                const double* __restrict__ dinv, KrylovState* st,
                                                double* __restrict__ partials) {
  if (st->done) return;
  __shared__ double red[kBlock / kWave];
  double rv[1];
  const int sl[1] = {2};
  fold_partials<kBlock, 1>(rv, sl, partials, nparts_rv, red);
  const double alpha = st->rho_buf[it & 1] / rv[0];
  double acc[1] = {0.0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double si = r[i] - alpha * v[i];
    s[i] = si;
    acc[0] += si * si;
    if (PREC == OFH_DIAG_PREC) ss[i] = si * dinv[i];
    else if (PREC == OFH_IDENT_PREC) ss[i] = si;
  }
  block_sum_k<kBlock, 1>(acc, red);
  const int slot[1] = {3};
  write_partials<1>(acc, slot, partials, blockIdx.x);
  if (blockIdx.x == 0 && threadIdx.x == 0) st->alpha = alpha;
}

// ||s||/||b|| < toler: early exit of BiCGStab.h:132-138 (one block)
__global__ __launch_bounds__(kBlock) void k_bi_scheck(int nparts, KrylovState* st,
                                                     const double* __restrict__ partials) {
  if (st->done) return;
  __shared__ double red[kBlock / kWave];
  double t[1];
  const int slot[1] = {3};
  fold_partials<kBlock, 1>(t, slot, partials, nparts, red);
  if (threadIdx.x != 0) return;
  const double res = sqrt(t[0]) / st->nrm;
  if (res < st->tol) {
    st->sconv = 1;
    st->halt = 1;
    st->sconv_res = res;
  }
}

// x += alpha pp + omega ss; r = s - omega t; (r,r) -> 0, (rr,r) -> 1
__global__ __launch_bounds__(kBlock) void k_bi_upd(int n, int it, int nparts_t,
                                                  double* __restrict__ x, double* __restrict__ r,
                                                  const double* __restrict__ rr,
                                                  const double* __restrict__ pp,
                                                  const double* __restrict__ s,
                                                  const double* __restrict__ ss,
                                                  const double* __restrict__ t, KrylovState* st,
                                                  double* __restrict__ partials) {
  if (st->done) return;
  __shared__ double red[2 * kBlock / kWave];
  const double alpha = st->alpha;
  if (st->sconv) {  // x += alpha pp; return it (BiCGStab.h:133-137)
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
      x[i] += alpha * pp[i];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->it = it;
      st->res = st->sconv_res;
      st->done = 1;
      st->status = it;
    }
    return;
  }
  double ts[2];
  const int sl[2] = {4, 5};
  fold_partials<kBlock, 2>(ts, sl, partials, nparts_t, red);
  const double omega = ts[0] / ts[1];
  double v[2] = {0.0, 0.0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    x[i] += alpha * pp[i] + omega * ss[i];
    const double ri = s[i] - omega * t[i];
    r[i] = ri;
    v[0] += ri * ri;
    v[1] += rr[i] * ri;
  }
  block_sum_k<kBlock, 2>(v, red);
  const int slot[2] = {0, 1};
  write_partials<2>(v, slot, partials, blockIdx.x);
  if (blockIdx.x == 0 && threadIdx.x == 0) st->omega = omega;
}

template <int PREC>
static int bicgstab_impl(ofh_csr A, const double* b, double* x, int max_it, double toler,
                         ofh_solve_info* info) {
  ofh_ctx ctx = A->ctx;
  hipStream_t s = ctx->stream;
  const int n = A->n;
  const double t0 = now_seconds();
  int rc;
  if (PREC == OFH_DIAG_PREC && (rc = csr_ensure_dinv(A))) return rc;
  if (sweep_prec(PREC) && (rc = csr_ensure_prec(A, PREC))) return rc;
  void* wsp;
  if ((rc = ensure_vws(ctx, sizeof(double) * (size_t)n * 8, &wsp))) return rc;
  double* r = (double*)wsp;
  double *rr = r + n, *p = rr + n, *pp = p + n, *v = pp + n, *sv = v + n, *ss = sv + n,
         *t = ss + n;
  KrylovState* st = state_ptr(ctx);
  KrylovState h0{};
  h0.tol = toler;
  h0.max_it = max_it;
  OFH_HIP(hipMemcpyAsync(st, &h0, sizeof(h0), hipMemcpyHostToDevice, s));
  double* part = ctx->ws.partials;
  const int g = reduce_grid(n);
  const int nb = A->nblocks;
  MatRows64 mr;
  bool sell = false;
  if ((rc = mat_rows64(A, &mr, &sell))) return rc;
  const int nbp = mr.nblocks;  // workgroups (= partials) of the iteration's products
  auto spmv = [&](int mode, const double* xv, double* yv, const double* uv) {
#define OFH_BI_SPMV(MD, SL)                                                                   \
  hipLaunchKernelGGL((k_bi_spmv<MD, SL>), dim3(nbp), dim3(kBlock), 0, s, mr, xv, yv, uv, st, \
                     part)
    if (sell) {
      if (mode == 0) OFH_BI_SPMV(0, true);
      else OFH_BI_SPMV(1, true);
    } else {
      if (mode == 0) OFH_BI_SPMV(0, false);
      else OFH_BI_SPMV(1, false);
    }
#undef OFH_BI_SPMV
  };
  hipLaunchKernelGGL(k_bi_init, dim3(nb), dim3(kBlock), 0, s, nb, A->row_blocks, A->row_ptr,
                     ColIndex{A->col, A->col16}, A->val, b, x, r, rr, part);
  auto dir = [&](int it) -> int {
    hipLaunchKernelGGL(k_bi_dir<PREC>, dim3(g), dim3(kBlock), 0, s, n, it, it == 1 ? nb : g, r, v,
                       p, pp, A->dinv, st, part);
    if (sweep_prec(PREC)) return launch_prec(A, PREC, p, pp, &st->done, s);
    return OFH_OK;
  };
  if ((rc = dir(1))) return rc;
  OFH_HIP(hipGetLastError());
  KrylovState h{};
  if ((rc = read_state(ctx, &h))) return rc;
  int enq = 0, it_prev = 0;
  double res_prev = h.res;
  const int it0 = h.it;
  const double res0 = h.res;
  while (!h.done && enq < max_it) {
    const int chunk =
        next_chunk(h.it, h.res, it_prev, res_prev, toler, max_it - enq, it0, res0);
    it_prev = h.it;
    res_prev = h.res;
    for (int c = 0; c < chunk; ++c) {
      const int it = ++enq;
      spmv(0, pp, v, rr);
      hipLaunchKernelGGL(k_bi_s<PREC>, dim3(g), dim3(kBlock), 0, s, n, it, nbp, r, v, sv, ss,
                         A->dinv, st, part);
      hipLaunchKernelGGL(k_bi_scheck, dim3(1), dim3(kBlock), 0, s, g, st, part);
      if (sweep_prec(PREC) && (rc = launch_prec(A, PREC, sv, ss, &st->halt, s))) return rc;
      spmv(1, ss, t, sv);
      hipLaunchKernelGGL(k_bi_upd, dim3(g), dim3(kBlock), 0, s, n, it, nbp, x, r, rr, pp, sv, ss,
                         t, st, part);
      if ((rc = dir(it + 1))) return rc;
    }
    OFH_HIP(hipGetLastError());
    if ((rc = read_state(ctx, &h))) return rc;
  }
  if (!h.done) {
    h.done = 1;
    h.status = -(max_it + 1);
  }
  return fill_info(A, b, x, h, t0, info);
}

// ===========================================================================
// GMRes(m)
// ===========================================================================
// r = P (b - A x); (r,r) -> 0, (b,b) -> 1     (GMRes.h:164-167, 226-229)
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_gm_resid(int nblocks, const int* __restrict__ rb,
                                                    const int* __restrict__ rp,
                                                    ColIndex ci,
                                                    const double* __restrict__ val,
                                                    const double* __restrict__ b,
                                                    const double* __restrict__ x,
                                                    double* __restrict__ r,
                                                    const double* __restrict__ dinv,
                                                    const KrylovState* __restrict__ st,
                                                    double* __restrict__ partials) {
  if (st->done) return;
  __shared__ double lds[kTile64];
  __shared__ double red[2 * kBlock / kWave];
  const int blk = xcd_block(blockIdx.x, nblocks);
  const int r0 = rb[blk], r1 = rb[blk + 1];
  const double s = csr_stream_rowsum<double, kTile64>(r0, r1, rp, ci, val, x, lds, 0.0);
  const int i = r0 + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (i < r1) {
    double ri = b[i] - s;
    if (PREC == OFH_DIAG_PREC) ri *= dinv[i];  // P.solve(r) in place: r_i *= pivot_i
    r[i] = ri;
    v[0] = ri * ri;
    v[1] = b[i] * b[i];
  }
  block_sum_k<kBlock, 2>(v, red);
  const int slot[2] = {0, 1};
  write_partials<2>(v, slot, partials, blockIdx.x);
}

// plain b - A x into w (the SSOR path preconditions it with the sweeps)
__global__ __launch_bounds__(kBlock) void k_resid_plain(int nblocks, const int* __restrict__ rb,
                                                       const int* __restrict__ rp,
                                                       ColIndex ci,
                                                       const double* __restrict__ val,
                                                       const double* __restrict__ b,
                                                       const double* __restrict__ x,
                                                       double* __restrict__ w,
                                                       const KrylovState* __restrict__ st) {
  if (st->done) return;
  __shared__ double lds[kTile64];
  const int blk = xcd_block(blockIdx.x, nblocks);
  const int r0 = rb[blk], r1 = rb[blk + 1];
  const double s = csr_stream_rowsum<double, kTile64>(r0, r1, rp, ci, val, x, lds, 0.0);
  const int i = r0 + threadIdx.x;
  if (i < r1) w[i] = b[i] - s;
}

// SSOR: (r,r) -> 0, (b,b) -> 1 of the preconditioned residual
__global__ __launch_bounds__(kBlock) void k_gm_rnorm(int n, const double* __restrict__ r,
                                                    const double* __restrict__ b,
                                                    const KrylovState* __restrict__ st,
                                                    double* __restrict__ partials) {
  if (st->done) return;
  __shared__ double red[2 * kBlock / kWave];
  double v[2] = {0.0, 0.0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    v[0] += r[i] * r[i];
    v[1] += b[i] * b[i];
  }
  block_sum_k<kBlock, 2>(v, red);
  const int slot[2] = {0, 1};
  write_partials<2>(v, slot, partials, blockIdx.x);
}

// beta = ||r||; stop tests: INIT (GMRes.h:169-179, returns it = 1) or the
// restart test (GMRes.h:230-238).  One block.
__global__ __launch_bounds__(kBlock) void k_gm_rcheck(int init, int nparts, KrylovState* st,
                                                     const double* __restrict__ partials) {
  if (st->done) return;
  __shared__ double red[2 * kBlock / kWave];
  double t[2];
  const int slot[2] = {0, 1};
  fold_partials<kBlock, 2>(t, slot, partials, nparts, red);
  if (threadIdx.x != 0) return;
  if (init) {
    double nb = sqrt(t[1]);
    if (nb == 0.0) nb = 1;
    st->nrm = nb;
    st->it = 1;
  }
  st->h_norm = sqrt(t[0]);
  const double res = st->h_norm / st->nrm;
  st->res = res;
  if (init) {
    st->res0 = res;
    if (res <= st->tol) {
      st->done = 1;
      st->halt = 1;
      st->status = 1;
    }
  } else if (res < st->tol) {
    st->done = 1;
    st->halt = 1;
    st->status = st->it;
  }
}

// v0 = (1/beta) r; s = 0; s[0] = beta (GMRes.h:187-189)
__global__ void k_gm_v0(int n, int m, const double* __restrict__ r, double* __restrict__ v0,
                        double* __restrict__ sv, const KrylovState* __restrict__ st) {
  if (st->done) return;
  const double beta = st->h_norm, sc = 1.0 / beta;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v0[i] = sc * r[i];
  if (i <= m) sv[i] = i == 0 ? beta : 0.0;
}

// w = P (A v_i)   (GMRes.h:191-192)
template <int PREC, bool SELL>
__global__ __launch_bounds__(kBlock) void k_gm_av(MatRows64 m, const double* __restrict__ vi,
                                                 double* __restrict__ w,
                                                 const double* __restrict__ dinv,
                                                 const KrylovState* __restrict__ st) {
  if (st->halt) return;
  int i;
  bool valid;
  const double s = mat_rowsum64<SELL>(m, vi, i, valid);
  if (valid) w[i] = PREC == OFH_DIAG_PREC ? s * dinv[i] : s;
}

// Modified Gram-Schmidt step k of column i (GMRes.h:193-197):
//   k > 0: H(k-1,i) = fold of the previous step; w -= H(k-1,i) v_{k-1}
//   partial (w, v_k) (k <= i) or (w,w) (k == i+1) -> slot k&1
__global__ __launch_bounds__(kBlock) void k_gm_mgs(int n, int k, int i, int m1, int nparts,
                                                  double* __restrict__ w,
                                                  const double* __restrict__ vprev,
                                                  const double* __restrict__ vk,
                                                  double* __restrict__ H,
                                                  const KrylovState* __restrict__ st,
                                                  double* __restrict__ partials) {
  if (st->halt) return;
  __shared__ double red[kBlock / kWave];
  double h = 0.0;
  if (k > 0) {
    double t[1];
    const int sl[1] = {(k - 1) & 1};
    fold_partials<kBlock, 1>(t, sl, partials, nparts, red);
    h = t[0];
    if (blockIdx.x == 0 && threadIdx.x == 0) H[(k - 1) * m1 + i] = h;
  }
  double v[1] = {0.0};
  for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) {
    double wj = w[j];
    if (k > 0) {
      wj -= h * vprev[j];
      w[j] = wj;
    }
    v[0] += vk ? wj * vk[j] : wj * wj;
  }
  block_sum_k<kBlock, 1>(v, red);
  const int slot[1] = {k & 1};
  write_partials<1>(v, slot, partials, blockIdx.x);
}

// H(i+1,i) = ||w||; v_{i+1} = (1/H(i+1,i)) w  (GMRes.h:198-199)
__global__ __launch_bounds__(kBlock) void k_gm_scale(int n, int i, int m1, int nparts,
                                                    const double* __restrict__ w,
                                                    double* __restrict__ vnext,
                                                    double* __restrict__ H,
                                                    const KrylovState* __restrict__ st,
                                                    const double* __restrict__ partials) {
  if (st->halt) return;
  __shared__ double red[kBlock / kWave];
  double t[1];
  const int sl[1] = {(i + 1) & 1};
  fold_partials<kBlock, 1>(t, sl, partials, nparts, red);
  const double h = sqrt(t[0]);
  if (blockIdx.x == 0 && threadIdx.x == 0) H[(i + 1) * m1 + i] = h;
  const double sc = 1.0 / h;
  for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock)
    vnext[j] = sc * w[j];
}

__device__ void gm_rotate(double& dx, double& dy, double cs, double sn) {
  const double temp = cs * dx + sn * dy;  // GMRes.h:67-76
  dy = cs * dy - sn * dx;
  dx = temp;
}

// Givens rotations and the residual estimate (GMRes.h:200-210); one thread.
__global__ void k_gm_givens(int i, int m1, double* H, double* sv, double* cs, double* sn,
                            KrylovState* st) {
  if (st->halt || threadIdx.x != 0) return;
  for (int k = 0; k < i; k++) gm_rotate(H[k * m1 + i], H[(k + 1) * m1 + i], cs[k], sn[k]);
  const double dx = H[i * m1 + i], dy = H[(i + 1) * m1 + i];
  double c, s;
  if (dy == 0.0) {  // set_rotation, GMRes.h:97-115
    c = 1.0;
    s = 0.0;
  } else if (fabs(dy) > fabs(dx)) {
    const double temp = dx / dy;
    s = 1.0 / sqrt(1.0 + temp * temp);
    c = temp * s;
  } else {
    const double temp = dy / dx;
    c = 1.0 / sqrt(1.0 + temp * temp);
    s = temp * c;
  }
  cs[i] = c;
  sn[i] = s;
  gm_rotate(H[i * m1 + i], H[(i + 1) * m1 + i], c, s);
  gm_rotate(sv[i], sv[i + 1], c, s);
  const double res = fabs(sv[i + 1]) / st->nrm;
  st->res = res;
  if (res < st->tol) {
    st->sconv = 1;  // converged inside the cycle: update(x, i) follows
    st->halt = 1;
    st->cycle_i = i;
    st->status = st->it;
  } else {
    st->it = st->it + 1;
    if (!(res == res)) {
      st->done = 1;
      st->halt = 1;
      st->status = -(st->max_it + 1);
    }
  }
}

// update(): back substitution y = H^-1 s (GMRes.h:78-94); k from the state.
__global__ void k_gm_backsolve(int m, int m1, const double* H, const double* sv, double* y,
                               const KrylovState* st) {
  if (st->done || threadIdx.x != 0) return;
  const int k = st->sconv ? st->cycle_i : m - 1;
  for (int q = 0; q < m1; ++q) y[q] = sv[q];
  for (int i = k; i >= 0; i--) {
    y[i] /= H[i * m1 + i];
    for (int j = i - 1; j >= 0; j--) y[j] -= H[j * m1 + i] * y[i];
  }
}

// x += sum_j v_j y_j (j = 0..k, in order, GMRes.h:91-93)
__global__ void k_gm_xupd(int n, int m, const double* __restrict__ V, const double* __restrict__ y,
                          double* __restrict__ x, const KrylovState* __restrict__ st) {
  if (st->done) return;
  const int k = st->sconv ? st->cycle_i : m - 1;
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n) return;
  double xl = x[l];
  for (int j = 0; j <= k; ++j) xl += V[(size_t)j * n + l] * y[j];
  x[l] = xl;
}

__global__ void k_gm_finish_cycle(KrylovState* st) {
  if (threadIdx.x != 0 || st->done) return;
  if (st->sconv) {
    st->done = 1;
    st->halt = 1;
  }
}

template <int PREC>
static int gmres_impl(ofh_csr A, const double* b, double* x, int m, int max_it, double toler,
                      ofh_solve_info* info) {
  ofh_ctx ctx = A->ctx;
  hipStream_t s = ctx->stream;
  const int n = A->n;
  const int m1 = m + 1;
  const double t0 = now_seconds();
  int rc;
  if (PREC == OFH_DIAG_PREC && (rc = csr_ensure_dinv(A))) return rc;
  if (sweep_prec(PREC) && (rc = csr_ensure_prec(A, PREC))) return rc;
  void* wsp;
  const size_t nd = (size_t)n * (m1 + 3) + (size_t)m1 * m1 + 4 * (size_t)m1;
  if ((rc = ensure_vws(ctx, sizeof(double) * nd, &wsp))) return rc;
  double* V = (double*)wsp;
  double* w = V + (size_t)n * m1;
  double* r = w + n;
  double* tmp = r + n;
  double* H = tmp + n;
  double* sv = H + (size_t)m1 * m1;
  double* cs = sv + m1;
  double* sn = cs + m1;
  double* y = sn + m1;
  OFH_HIP(hipMemsetAsync(H, 0, sizeof(double) * ((size_t)m1 * m1 + 4 * (size_t)m1), s));
  KrylovState* st = state_ptr(ctx);
  KrylovState h0{};
  h0.tol = toler;
  h0.max_it = max_it;
  OFH_HIP(hipMemcpyAsync(st, &h0, sizeof(h0), hipMemcpyHostToDevice, s));
  double* part = ctx->ws.partials;
  const int g = reduce_grid(n);
  const int nb = A->nblocks;
  const int ge = (n + 255) / 256;
  MatRows64 mr;
  bool sell = false;
  if ((rc = mat_rows64(A, &mr, &sell))) return rc;
  // w = P (A v_i) for the plain and diagonal preconditioners, A v_i otherwise
  auto av = [&](const double* vi, double* wv, bool diag) {
#define OFH_GM_AV(P, SL)                                                                   \
  hipLaunchKernelGGL((k_gm_av<P, SL>), dim3(mr.nblocks), dim3(kBlock), 0, s, mr, vi, wv, \
                     A->dinv, st)
    if (sell) {
      if (diag) OFH_GM_AV(OFH_DIAG_PREC, true);
      else OFH_GM_AV(OFH_IDENT_PREC, true);
    } else {
      if (diag) OFH_GM_AV(OFH_DIAG_PREC, false);
      else OFH_GM_AV(OFH_IDENT_PREC, false);
    }
#undef OFH_GM_AV
  };
  auto residual = [&](bool init) -> int {
    if (sweep_prec(PREC)) {
      hipLaunchKernelGGL(k_resid_plain, dim3(nb), dim3(kBlock), 0, s, nb, A->row_blocks,
                         A->row_ptr, ColIndex{A->col, A->col16}, A->val, b, x, tmp, st);
      int e = launch_prec(A, PREC, tmp, r, &st->done, s);
      if (e) return e;
      hipLaunchKernelGGL(k_gm_rnorm, dim3(g), dim3(kBlock), 0, s, n, r, b, st, part);
      hipLaunchKernelGGL(k_gm_rcheck, dim3(1), dim3(kBlock), 0, s, (int)init, g, st, part);
    } else {
      hipLaunchKernelGGL(k_gm_resid<PREC>, dim3(nb), dim3(kBlock), 0, s, nb, A->row_blocks,
                         A->row_ptr, ColIndex{A->col, A->col16}, A->val, b, x, r, A->dinv, st, part);
      hipLaunchKernelGGL(k_gm_rcheck, dim3(1), dim3(kBlock), 0, s, (int)init, nb, st, part);
    }
    return OFH_OK;
  };
  if ((rc = residual(true))) return rc;
  KrylovState h{};
  if ((rc = read_state(ctx, &h))) return rc;
  while (!h.done && h.it <= max_it) {
    int it = h.it;
    hipLaunchKernelGGL(k_gm_v0, dim3((std::max(n, m1) + 255) / 256), dim3(256), 0, s, n, m, r, V,
                       sv, st);
    for (int i = 0; i < m && it <= max_it; i++, it++) {
      double* vi = V + (size_t)i * n;
      if (sweep_prec(PREC)) {
        av(vi, tmp, false);
        if ((rc = launch_prec(A, PREC, tmp, w, &st->halt, s))) return rc;
      } else {
        av(vi, w, PREC == OFH_DIAG_PREC);
      }
      for (int k = 0; k <= i + 1; ++k) {
        const double* vprev = k > 0 ? V + (size_t)(k - 1) * n : nullptr;
        const double* vk = k <= i ? V + (size_t)k * n : nullptr;
        hipLaunchKernelGGL(k_gm_mgs, dim3(g), dim3(kBlock), 0, s, n, k, i, m1, g, w, vprev, vk,
                           H, st, part);
      }
      hipLaunchKernelGGL(k_gm_scale, dim3(g), dim3(kBlock), 0, s, n, i, m1, g, w,
                         V + (size_t)(i + 1) * n, H, st, part);
      hipLaunchKernelGGL(k_gm_givens, dim3(1), dim3(64), 0, s, i, m1, H, sv, cs, sn, st);
    }
    hipLaunchKernelGGL(k_gm_backsolve, dim3(1), dim3(64), 0, s, m, m1, H, sv, y, st);
    hipLaunchKernelGGL(k_gm_xupd, dim3(ge), dim3(256), 0, s, n, m, V, y, x, st);
    hipLaunchKernelGGL(k_gm_finish_cycle, dim3(1), dim3(64), 0, s, st);
    if ((rc = residual(false))) return rc;
    OFH_HIP(hipGetLastError());
    if ((rc = read_state(ctx, &h))) return rc;
  }
  if (!h.done) {
    h.status = -h.it;  // GMRes.h:236-240: the counter's value after the loop
    h.done = 1;
  }
  return fill_info(A, b, x, h, t0, info);
}

}  // namespace ofh

using namespace ofh;

extern "C" {

int ofh_bicgstab(ofh_csr A, int prec, const double* b, double* x, int max_it, double toler,
                 ofh_solve_info* info) {
  OFH_CHECK(A && b && x && max_it >= 0, OFH_E_INVALID, "ofh_bicgstab: bad argument");
  switch (prec) {
    case OFH_IDENT_PREC: return bicgstab_impl<OFH_IDENT_PREC>(A, b, x, max_it, toler, info);
    case OFH_DIAG_PREC: return bicgstab_impl<OFH_DIAG_PREC>(A, b, x, max_it, toler, info);
    case OFH_SSOR_PREC: return bicgstab_impl<OFH_SSOR_PREC>(A, b, x, max_it, toler, info);
    case OFH_DILU_PREC:  // the reference's ILU is its DILU
    case OFH_ILU_PREC: return bicgstab_impl<OFH_DILU_PREC>(A, b, x, max_it, toler, info);
    default: return fail(OFH_E_INVALID, "unknown preconditioner");
  }
}

int ofh_gmres(ofh_csr A, int prec, const double* b, double* x, int m, int max_it, double toler,
              ofh_solve_info* info) {
  OFH_CHECK(A && b && x && max_it >= 0 && m >= 1, OFH_E_INVALID, "ofh_gmres: bad argument");
  switch (prec) {
    case OFH_IDENT_PREC: return gmres_impl<OFH_IDENT_PREC>(A, b, x, m, max_it, toler, info);
    case OFH_DIAG_PREC: return gmres_impl<OFH_DIAG_PREC>(A, b, x, m, max_it, toler, info);
    case OFH_SSOR_PREC: return gmres_impl<OFH_SSOR_PREC>(A, b, x, m, max_it, toler, info);
    case OFH_DILU_PREC:  // the reference's ILU is its DILU
    case OFH_ILU_PREC: return gmres_impl<OFH_DILU_PREC>(A, b, x, m, max_it, toler, info);
    default: return fail(OFH_E_INVALID, "unknown preconditioner");
  }
}

int ofh_prec_apply(ofh_csr A, int prec, const double* b, double* x) {
  OFH_CHECK(A && b && x, OFH_E_INVALID, "ofh_prec_apply: null argument");
  const int n = A->n;
  int rc;
  switch (prec) {
    case OFH_IDENT_PREC:
      hipLaunchKernelGGL(k_prec_diag, dim3((n + 255) / 256), dim3(256), 0, A->ctx->stream, n, b,
                         (const double*)nullptr, x);
      break;
    case OFH_DIAG_PREC:
      if ((rc = csr_ensure_dinv(A))) return rc;
      hipLaunchKernelGGL(k_prec_diag, dim3((n + 255) / 256), dim3(256), 0, A->ctx->stream, n, b,
                         A->dinv, x);
      break;
    case OFH_SSOR_PREC:
      OFH_CHECK(b != x, OFH_E_INVALID, "ofh_prec_apply: SSOR needs distinct b and x");
      return launch_ssor(A, b, x, nullptr, A->ctx->stream);
    case OFH_DILU_PREC:
    case OFH_ILU_PREC:
      OFH_CHECK(b != x, OFH_E_INVALID, "ofh_prec_apply: DILU/ILU needs distinct b and x");
      return launch_prec(A, prec, b, x, nullptr, A->ctx->stream);
    default: return fail(OFH_E_INVALID, "unknown preconditioner");
  }
  OFH_HIP(hipGetLastError());
  return OFH_OK;
}

int ofh_linear_solve(ofh_csr A, int solver, int prec, const double* b, double* x, int max_it,
                     double toler, ofh_solve_info* info) {
  OFH_CHECK(A, OFH_E_INVALID, "In LinearSolver::solve(): No matrix has been defined.");
  switch (solver) {
    case OFH_CG_SOLVER: return ofh_cg(A, prec, b, x, max_it, toler, info);
    case OFH_BICG_STAB_SOLVER: return ofh_bicgstab(A, prec, b, x, max_it, toler, info);
    case OFH_GMRES_SOLVER: {
      int m = std::max(1, std::min(A->n / 5, 50));
      return ofh_gmres(A, prec, b, x, m, max_it, toler, info);
    }
    case OFH_CGS_SOLVER: return ofh_cgs(A, prec, b, x, max_it, toler, info);
    case OFH_BICG_SOLVER: return ofh_bicg(A, prec, b, x, max_it, toler, info);
    case OFH_DIRECT_SOLVER:
      return fail(OFH_E_UNSUPPORTED, "DIRECT_SOLVER is not provided by this path");
    default: return fail(OFH_E_INVALID, "unknown solver");
  }
}

}  // extern "C"
