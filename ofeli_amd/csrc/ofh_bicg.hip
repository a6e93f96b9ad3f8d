// ofh_bicg.hip -- CGS and BiCG on the GPU, the transposed product y = A^T x
// and Prec::TransSolve.
//
// References: CGS.h:91-171, BiCG.h:94-170, SpMatrix::TMult
// (SpMatrix_impl.h:761-767), Prec::TransSolve (Prec_impl.h:259-296),
// LinearSolver.cpp:225-231.  The solvers follow ofh_krylov.hip: scalars live
// on the device, the kernel that consumes a reduction folds its per-block
// partials, every kernel returns at once when the stop flag is set, and the
// host reads the state once per chunk of iterations.
//
// CGS is the reference's as written: its vector q is never assigned
// (CGS.h:127-135), so q = 0 in u = r + beta q and p = u + beta (q + beta p).
// BiCG uses the products the reference means: its TMult indexes the column
// of the previous entry (and _col_ind[-1]) and accumulates into an uncleared
// vector, and its ILU transposed solve reads Prec's never-filled
// _row_ptr/_col_ind; neither can be reproduced.  A^T is a CSR copy of the
// transpose (rows of A^T in ascending order of A's rows, so every product
// sums in the order the reference's loop visits the entries).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "ofh_internal.h"
#include "ofh_solver.h"
#include "ofh_spmv.h"

namespace ofh {

// ---------------------------------------------------------------------------
// transposed CSR
// ---------------------------------------------------------------------------
__global__ void k_gather_perm(int64_t nnz, const int* __restrict__ perm,
                              const double* __restrict__ src, double* __restrict__ dst) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nnz; k += stride)
    dst[k] = src[perm[k]];
}

// A->tr: the pattern once (host counting sort of the entries by column, rows
// ascending within a column), the values whenever A's values version moved.
int csr_ensure_transpose(ofh_csr A) {
  hipStream_t s = A->ctx->stream;
  const int n = A->n;
  const int64_t nnz = A->nnz;
  if (!A->tr) {
    std::vector<int> rp(n + 1), col(nnz);
    OFH_HIP(hipMemcpyAsync(rp.data(), A->row_ptr, sizeof(int) * (n + 1), hipMemcpyDeviceToHost, s));
    OFH_HIP(hipMemcpyAsync(col.data(), A->col, sizeof(int) * nnz, hipMemcpyDeviceToHost, s));
    OFH_HIP(hipStreamSynchronize(s));
    std::vector<int> trp(n + 1, 0), tcol(nnz), perm(nnz);
    for (int64_t k = 0; k < nnz; ++k) trp[col[k] + 1]++;
    for (int j = 0; j < n; ++j) trp[j + 1] += trp[j];
    std::vector<int> fill(trp.begin(), trp.end() - 1);
    for (int i = 0; i < n; ++i)
      for (int k = rp[i]; k < rp[i + 1]; ++k) {
        const int p = fill[col[k]]++;
        tcol[p] = i;
        perm[p] = k;
      }
    ofh_csr T = nullptr;
    int rc = ofh_csr_create(A->ctx, n, nnz, trp.data(), tcol.data(), nullptr, &T);
    if (rc) return rc;
    int* d_perm = nullptr;
    if (hipMalloc(&d_perm, sizeof(int) * std::max<int64_t>(nnz, 1)) != hipSuccess) {
      ofh_csr_destroy(T);
      return fail(OFH_E_HIP, "transpose: out of device memory");
    }
    OFH_HIP(hipMemcpyAsync(d_perm, perm.data(), sizeof(int) * nnz, hipMemcpyHostToDevice, s));
    OFH_HIP(hipStreamSynchronize(s));  // perm leaves scope
    A->tr_perm = d_perm;
    A->tr = T;
    A->tr_version = ~0ull;
  }
  if (A->tr_version != A->version) {
    const int g = (int)std::min<int64_t>((nnz + 255) / 256, 65536);
    hipLaunchKernelGGL(k_gather_perm, dim3(std::max(g, 1)), dim3(256), 0, s, nnz, A->tr_perm,
                       A->val, A->tr->val);
    OFH_HIP(hipGetLastError());
    csr_values_changed(A->tr);
    A->tr_version = A->version;
  }
  return OFH_OK;
}

// ---------------------------------------------------------------------------
// shared kernels
// ---------------------------------------------------------------------------
template <int SLOT, bool SELL>
__global__ __launch_bounds__(kBlock) void k_kr_spmv(MatRows64 m, const double* __restrict__ x,
                                                   double* __restrict__ y,
                                                   const double* __restrict__ u,
                                                   const KrylovState* __restrict__ st,
                                                   double* __restrict__ partials);

// y = A x with SLOT's partials, from the matrix's SELL-64 copy when it has one
template <int SLOT>
static void kr_spmv(const MatRows64& m, bool sell, const double* x, double* y, const double* u,
                    const KrylovState* st, double* partials, hipStream_t s) {
  if (sell)
    hipLaunchKernelGGL((k_kr_spmv<SLOT, true>), dim3(m.nblocks), dim3(kBlock), 0, s, m, x, y, u,
                       st, partials);
  else
    hipLaunchKernelGGL((k_kr_spmv<SLOT, false>), dim3(m.nblocks), dim3(kBlock), 0, s, m, x, y,
                       u, st, partials);
}
// r = b - A x; r2 = r; partials (b,b) -> 0, (r,r) -> 1
__global__ __launch_bounds__(kBlock) void k_kr_init(int nblocks, const int* __restrict__ rb,
                                                   const int* __restrict__ rp, ColIndex ci,
                                                   const double* __restrict__ val,
                                                   const double* __restrict__ b,
                                                   const double* __restrict__ x,
                                                   double* __restrict__ r, double* __restrict__ r2,
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
    r2[i] = ri;
    v[0] = b[i] * b[i];
    v[1] = ri * ri;
  }
  block_sum_k<kBlock, 2>(v, red);
  const int slot[2] = {0, 1};
  write_partials<2>(v, slot, partials, blockIdx.x);
}

// y = A x (skipped once the solver stopped); with SLOT >= 0 also the partial
// sums of (u, y) into that slot.
template <int SLOT, bool SELL>
__global__ __launch_bounds__(kBlock) void k_kr_spmv(MatRows64 m, const double* __restrict__ x,
                                                   double* __restrict__ y,
                                                   const double* __restrict__ u,
                                                   const KrylovState* __restrict__ st,
                                                   double* __restrict__ partials) {
  if (st->done) return;
  __shared__ double red[kBlock / kWave];
  int i;
  bool valid;
  const double s = mat_rowsum64<SELL>(m, x, i, valid);
  double v[1] = {0.0};
  if (valid) {
    y[i] = s;
    if (SLOT >= 0) v[0] = u[i] * s;
  }
  if (SLOT >= 0) {
    block_sum_k<kBlock, 1>(v, red);
    const int slot[1] = {SLOT};
    write_partials<1>(v, slot, partials, blockIdx.x);
  }
}

// partials of (u, v) into SLOT (skipped once the solver stopped)
template <int SLOT>
__global__ __launch_bounds__(kBlock) void k_kr_dot(int n, const double* __restrict__ u,
                                                  const double* __restrict__ v,
                                                  const KrylovState* __restrict__ st,
                                                  double* __restrict__ partials) {
  if (st->done) return;
  __shared__ double red[kBlock / kWave];
  double acc[1] = {0.0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    acc[0] += u[i] * v[i];
  block_sum_k<kBlock, 1>(acc, red);
  const int slot[1] = {SLOT};
  write_partials<1>(acc, slot, partials, blockIdx.x);
}

__device__ __forceinline__ void kr_stop(KrylovState* st, int status) {
  st->done = 1;
  st->halt = 1;
  st->status = status;
}

// ===========================================================================
// CGS
// ===========================================================================
// Stop tests closing iteration it-1 and the directions of iteration it:
//   it == 1: nrm = ||b|| (0 -> 1), ||r0||/nrm <= toler -> 0 (CGS.h:110-117)
//   it >  1: ||r||/nrm < toler -> it-1 (:156-162)
//   loop end -> -max_it (:166-169); rho_1 = (s,r) == 0 -> 2 (:120-126)
// then u = r + beta q, p = u + beta (q + beta p) with q = 0 (:127-135) and,
// for IDENT/DIAG, y = P p (:136-137).
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_cgs_dir(int n, int it, int nparts,
                                                   const double* __restrict__ r,
                                                   double* __restrict__ u, double* __restrict__ p,
                                                   double* __restrict__ y,
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
  if (it == 1) {  // partials of the start: (b,b), (r,r) = (s,r)
    nrm = sqrt(t[0]);
    if (nrm == 0) nrm = 1;
    res = sqrt(t[1]) / nrm;
    rho = t[1];
    if (res <= st->tol) { stop = true; status = 0; }
    else if (it > st->max_it) { stop = true; status = -st->max_it; }
    else if (rho == 0.0) { stop = true; status = 2; }
  } else {  // partials of the update: (r,r), (s,r)
    res = sqrt(t[0]) / nrm;
    rho = t[1];
    if (res < st->tol) { stop = true; status = it - 1; }
    else if (it > st->max_it || !(res == res)) { stop = true; status = -st->max_it; }
    else if (rho == 0.0) { stop = true; status = 2; }
    else beta = rho / st->rho_buf[(it - 1) & 1];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->it = it - 1;
    st->res = res;
    if (it == 1) {
      st->nrm = nrm;
      st->res0 = res;
    }
    if (stop) kr_stop(st, status);
    else st->rho_buf[it & 1] = rho;
  }
  if (stop) return;
  const double q = 0.0;  // CGS.h: q is never assigned
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    double ui, pi;
    if (it == 1) {
      ui = r[i];
      pi = ui;
    } else {
      ui = r[i] + beta * q;
      pi = ui + beta * (q + beta * p[i]);
    }
    u[i] = ui;
    p[i] = pi;
    if (PREC == OFH_DIAG_PREC) y[i] = pi * dinv[i];
  }
}

// alpha = rho_1/(s,v); w = (u - alpha v) + u (CGS.h:139-141); for IDENT/DIAG
// w = P w in place (:142).
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_cgs_w(int n, int it, int nparts,
                                                 const double* __restrict__ u,
                                                 const double* __restrict__ v,
                                                 double* __restrict__ w,
                                                 const double* __restrict__ dinv,
                                                 KrylovState* st,
                                                 const double* __restrict__ partials) {
  if (st->done) return;
  __shared__ double red[kBlock / kWave];
  double sv[1];
  const int sl[1] = {2};
  fold_partials<kBlock, 1>(sv, sl, partials, nparts, red);
  const double alpha = st->rho_buf[it & 1] / sv[0];
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    double wi = u[i] - alpha * v[i];
    wi += u[i];
    if (PREC == OFH_DIAG_PREC) wi *= dinv[i];
    w[i] = wi;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) st->alpha = alpha;
}

// x += alpha w; r += (-alpha) z (CGS.h:143-145); partials (r,r) -> 0, (s,r) -> 1
__global__ __launch_bounds__(kBlock) void k_cgs_upd(int n, double* __restrict__ x,
                                                   double* __restrict__ r,
                                                   const double* __restrict__ s,
                                                   const double* __restrict__ w,
                                                   const double* __restrict__ z,
                                                   const KrylovState* __restrict__ st,
                                                   double* __restrict__ partials) {
  if (st->done) return;
  __shared__ double red[2 * kBlock / kWave];
  const double alpha = st->alpha;
  double v[2] = {0.0, 0.0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    x[i] += alpha * w[i];
    const double ri = r[i] + -alpha * z[i];
    r[i] = ri;
    v[0] += ri * ri;
    v[1] += s[i] * ri;
  }
  block_sum_k<kBlock, 2>(v, red);
  const int slot[2] = {0, 1};
  write_partials<2>(v, slot, partials, blockIdx.x);
}

template <int PREC>
static int cgs_impl(ofh_csr A, const double* b, double* x, int max_it, double toler,
                    ofh_solve_info* info) {
  ofh_ctx ctx = A->ctx;
  hipStream_t s = ctx->stream;
  const int n = A->n;
  const double t0 = now_seconds();
  int rc;
  if (PREC == OFH_DIAG_PREC && (rc = csr_ensure_dinv(A))) return rc;
  if (sweep_prec(PREC) && (rc = csr_ensure_prec(A, PREC))) return rc;
  void* wsp;
  if ((rc = ensure_vws(ctx, sizeof(double) * (size_t)n * 9, &wsp))) return rc;
  double* r = (double*)wsp;
  double *sv = r + n, *u = sv + n, *p = u + n, *y = p + n, *v = y + n, *w = v + n, *wp = w + n,
         *z = wp + n;
  if (PREC == OFH_IDENT_PREC) y = p;             // P.solve leaves y = p
  if (!sweep_prec(PREC)) wp = w;                 // P.solve(w) in place
  KrylovState* st = state_ptr(ctx);
  KrylovState h0{};
  h0.tol = toler;
  h0.max_it = max_it;
  OFH_HIP(hipMemcpyAsync(st, &h0, sizeof(h0), hipMemcpyHostToDevice, s));
  double* part = ctx->ws.partials;
  const int g = reduce_grid(n);
  const int nb = A->nblocks;
  const ColIndex ci{A->col, A->col16};
  MatRows64 mr;
  bool sell = false;
  if ((rc = mat_rows64(A, &mr, &sell))) return rc;
  hipLaunchKernelGGL(k_kr_init, dim3(nb), dim3(kBlock), 0, s, nb, A->row_blocks, A->row_ptr, ci,
                     A->val, b, x, r, sv, part);
  KrylovState h{};
  int enq = 0, it_prev = 0, it0 = 0;
  double res_prev = 0.0, res0 = 0.0;
  bool first = true;
  while (!h.done && enq <= max_it) {
    const int chunk = first ? 1
                            : next_chunk(h.it, h.res, it_prev, res_prev, toler, max_it + 1 - enq,
                                         it0, res0);
    if (!first) {
      if (res0 == 0.0) {  // the first look: the rate since the start is measured from here
        it0 = h.it;
        res0 = h.res;
      }
      it_prev = h.it;
      res_prev = h.res;
    }
    for (int c = 0; c < chunk; ++c) {
      const int it = ++enq;
      hipLaunchKernelGGL(k_cgs_dir<PREC>, dim3(g), dim3(kBlock), 0, s, n, it, it == 1 ? nb : g, r,
                         u, p, y, A->dinv, st, part);
      if (sweep_prec(PREC) && (rc = launch_prec(A, PREC, p, y, &st->done, s))) return rc;
      kr_spmv<2>(mr, sell, y, v, sv, st, part, s);
      hipLaunchKernelGGL(k_cgs_w<PREC>, dim3(g), dim3(kBlock), 0, s, n, it, mr.nblocks, u, v, w,
                         A->dinv, st, part);
      if (sweep_prec(PREC) && (rc = launch_prec(A, PREC, w, wp, &st->done, s))) return rc;
      kr_spmv<-1>(mr, sell, wp, z, nullptr, st, part, s);
      hipLaunchKernelGGL(k_cgs_upd, dim3(g), dim3(kBlock), 0, s, n, x, r, sv, wp, z, st, part);
    }
    OFH_HIP(hipGetLastError());
    if ((rc = read_state(ctx, &h))) return rc;
    first = false;
  }
  if (!h.done) {  // only reachable through max_it: the last dir kernel stops
    h.done = 1;
    h.status = -max_it;
  }
  return fill_info(A, b, x, h, t0, info);
}

// ===========================================================================
// BiCG
// ===========================================================================
// Stop tests closing iteration it-1, then, for IDENT/DIAG, z = P r and
// zz = P^T rr with the partials of (z, rr) (BiCG.h:120-122):
//   it == 1: ||r0||/nrm <= toler -> 0 (BiCG.h:110-117), nrm = ||b|| or 1
//   it >  1: ||r||/||b|| < toler -> it-1 (:158-164)
//   loop end -> -(max_it+1) (:166-168)
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_bcg_check(int n, int it, int nparts,
                                                     const double* __restrict__ r,
                                                     const double* __restrict__ rr,
                                                     double* __restrict__ z,
                                                     double* __restrict__ zz,
                                                     const double* __restrict__ dinv,
                                                     KrylovState* st,
                                                     double* __restrict__ partials) {
  if (st->done) return;
  __shared__ double red[2 * kBlock / kWave];
  double t[2];
  const int slot[2] = {0, 1};
  fold_partials<kBlock, 2>(t, slot, partials, nparts, red);
  bool stop = false;
  int status = 0;
  double res, rn;
  if (it == 1) {  // (b,b), (r,r)
    const double bn = sqrt(t[0]);
    double nrm = bn;
    if (nrm == 0.0) nrm = 1;
    rn = sqrt(t[1]);
    res = rn / nrm;
    if (res <= st->tol) { stop = true; status = 0; }
    else if (it > st->max_it) { stop = true; status = -(st->max_it + 1); }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->nrm = nrm;
      st->bnorm = bn;
      st->res0 = res;
    }
  } else {  // (r,r) of the update
    rn = sqrt(t[0]);
    res = rn / st->bnorm;
    if (res < st->tol) { stop = true; status = it - 1; }
    else if (it > st->max_it || !(res == res)) { stop = true; status = -(st->max_it + 1); }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->it = it - 1;
    st->res = res;
    st->rnorm = rn;
    if (stop) kr_stop(st, status);
  }
  if (stop || PREC != OFH_DIAG_PREC) return;
  double acc[1] = {0.0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double zi = r[i] * dinv[i];
    const double zzi = rr[i] * dinv[i];
    z[i] = zi;
    zz[i] = zzi;
    acc[0] += zi * rr[i];
  }
  block_sum_k<kBlock, 1>(acc, red);
  const int s2[1] = {2};
  write_partials<1>(acc, s2, partials, blockIdx.x);
}

// rho_1 = (z, rr); 0 -> it (BiCG.h:123-130); p = z + beta p, pp = zz + beta pp
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_bcg_dir(int n, int it, int nparts,
                                                   const double* __restrict__ z,
                                                   const double* __restrict__ zz,
                                                   double* __restrict__ p,
                                                   double* __restrict__ pp, KrylovState* st,
                                                   const double* __restrict__ partials) {
  if (st->done) return;
  __shared__ double red[kBlock / kWave];
  double t[1];
  const int slot[1] = {2};
  fold_partials<kBlock, 1>(t, slot, partials, nparts, red);
  const double rho = t[0];
  if (rho == 0.0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->res = st->rnorm / st->nrm;
      kr_stop(st, it);
    }
    return;
  }
  const double beta = it == 1 ? 0.0 : rho / st->rho_buf[(it - 1) & 1];
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    if (it == 1) {
      p[i] = z[i];
      pp[i] = zz[i];
    } else {
      p[i] = z[i] + beta * p[i];
      pp[i] = zz[i] + beta * pp[i];
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) st->rho_buf[it & 1] = rho;
}

// alpha = rho_1/(pp,q); x += alpha p; r += (-alpha) q; rr += (-alpha) qq
// (BiCG.h:149-153); partials (r,r) -> 0
__global__ __launch_bounds__(kBlock) void k_bcg_upd(int n, int it, int nparts,
                                                   double* __restrict__ x, double* __restrict__ r,
                                                   double* __restrict__ rr,
                                                   const double* __restrict__ p,
                                                   const double* __restrict__ q,
                                                   const double* __restrict__ qq, KrylovState* st,
                                                   double* __restrict__ partials) {
  if (st->done) return;
  __shared__ double red[kBlock / kWave];
  double pq[1];
  const int sl[1] = {3};
  fold_partials<kBlock, 1>(pq, sl, partials, nparts, red);
  const double alpha = st->rho_buf[it & 1] / pq[0];
  double v[1] = {0.0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    x[i] += alpha * p[i];
    const double ri = r[i] + -alpha * q[i];
    r[i] = ri;
    rr[i] += -alpha * qq[i];
    v[0] += ri * ri;
  }
  block_sum_k<kBlock, 1>(v, red);
  const int slot[1] = {0};
  write_partials<1>(v, slot, partials, blockIdx.x);
}

template <int PREC>
static int bicg_impl(ofh_csr A, const double* b, double* x, int max_it, double toler,
                     ofh_solve_info* info) {
  ofh_ctx ctx = A->ctx;
  hipStream_t s = ctx->stream;
  const int n = A->n;
  const double t0 = now_seconds();
  int rc;
  if (PREC == OFH_DIAG_PREC && (rc = csr_ensure_dinv(A))) return rc;
  if (sweep_prec(PREC) && (rc = csr_ensure_prec(A, PREC))) return rc;
  if ((rc = csr_ensure_transpose(A))) return rc;
  ofh_csr T = A->tr;
  void* wsp;
  if ((rc = ensure_vws(ctx, sizeof(double) * (size_t)n * 8, &wsp))) return rc;
  double* r = (double*)wsp;
  double *rr = r + n, *z = rr + n, *zz = z + n, *p = zz + n, *pp = p + n, *q = pp + n,
         *qq = q + n;
  // P.solve / P.TransSolve that leave their argument unchanged read it in place
  if (PREC == OFH_IDENT_PREC) z = r;
  if (PREC != OFH_DIAG_PREC && PREC != OFH_ILU_PREC) zz = rr;
  KrylovState* st = state_ptr(ctx);
  KrylovState h0{};
  h0.tol = toler;
  h0.max_it = max_it;
  OFH_HIP(hipMemcpyAsync(st, &h0, sizeof(h0), hipMemcpyHostToDevice, s));
  double* part = ctx->ws.partials;
  const int g = reduce_grid(n);
  const int nb = A->nblocks;
  const ColIndex ci{A->col, A->col16};
  MatRows64 mr, mrt;  // A and its transposed copy
  bool sell = false, sellt = false;
  if ((rc = mat_rows64(A, &mr, &sell)) || (rc = mat_rows64(T, &mrt, &sellt))) return rc;
  hipLaunchKernelGGL(k_kr_init, dim3(nb), dim3(kBlock), 0, s, nb, A->row_blocks, A->row_ptr, ci,
                     A->val, b, x, r, rr, part);
  KrylovState h{};
  int enq = 0, it_prev = 0, it0 = 0;
  double res_prev = 0.0, res0 = 0.0;
  bool first = true;
  while (!h.done && enq <= max_it) {
    const int chunk = first ? 1
                            : next_chunk(h.it, h.res, it_prev, res_prev, toler, max_it + 1 - enq,
                                         it0, res0);
    if (!first) {
      if (res0 == 0.0) {  // the first look: the rate since the start is measured from here
        it0 = h.it;
        res0 = h.res;
      }
      it_prev = h.it;
      res_prev = h.res;
    }
    for (int c = 0; c < chunk; ++c) {
      const int it = ++enq;
      hipLaunchKernelGGL(k_bcg_check<PREC>, dim3(PREC == OFH_DIAG_PREC ? g : 1), dim3(kBlock), 0,
                         s, n, it, it == 1 ? nb : g, r, rr, z, zz, A->dinv, st, part);
      if (PREC != OFH_DIAG_PREC) {
        if (sweep_prec(PREC) && (rc = launch_prec(A, PREC, r, z, &st->done, s))) return rc;
        if (PREC == OFH_ILU_PREC && (rc = launch_dilu_trans(A, rr, zz, &st->done, s))) return rc;
        hipLaunchKernelGGL(k_kr_dot<2>, dim3(g), dim3(kBlock), 0, s, n, z, rr, st, part);
      }
      hipLaunchKernelGGL(k_bcg_dir<PREC>, dim3(g), dim3(kBlock), 0, s, n, it, g, z, zz, p, pp, st,
                         part);
      kr_spmv<3>(mr, sell, p, q, pp, st, part, s);
      kr_spmv<-1>(mrt, sellt, pp, qq, nullptr, st, part, s);
      hipLaunchKernelGGL(k_bcg_upd, dim3(g), dim3(kBlock), 0, s, n, it, mr.nblocks, x, r, rr, p,
                         q, qq, st, part);
    }
    OFH_HIP(hipGetLastError());
    if ((rc = read_state(ctx, &h))) return rc;
    first = false;
  }
  if (!h.done) {
    h.done = 1;
    h.status = -(max_it + 1);
  }
  return fill_info(A, b, x, h, t0, info);
}

__global__ void k_scale_copy(int n, const double* __restrict__ b, const double* __restrict__ dinv,
                             double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = dinv ? b[i] * dinv[i] : b[i];
}

}  // namespace ofh

using namespace ofh;

extern "C" {

int ofh_cgs(ofh_csr A, int prec, const double* b, double* x, int max_it, double toler,
            ofh_solve_info* info) {
  OFH_CHECK(A && b && x && max_it >= 0, OFH_E_INVALID, "ofh_cgs: bad argument");
  switch (prec) {
    case OFH_IDENT_PREC: return cgs_impl<OFH_IDENT_PREC>(A, b, x, max_it, toler, info);
    case OFH_DIAG_PREC: return cgs_impl<OFH_DIAG_PREC>(A, b, x, max_it, toler, info);
    case OFH_SSOR_PREC: return cgs_impl<OFH_SSOR_PREC>(A, b, x, max_it, toler, info);
    case OFH_DILU_PREC:  // P.solve of ILU_PREC is DILUSolve
    case OFH_ILU_PREC: return cgs_impl<OFH_DILU_PREC>(A, b, x, max_it, toler, info);
    default: return fail(OFH_E_INVALID, "unknown preconditioner");
  }
}

int ofh_bicg(ofh_csr A, int prec, const double* b, double* x, int max_it, double toler,
             ofh_solve_info* info) {
  OFH_CHECK(A && b && x && max_it >= 0, OFH_E_INVALID, "ofh_bicg: bad argument");
  switch (prec) {
    case OFH_IDENT_PREC: return bicg_impl<OFH_IDENT_PREC>(A, b, x, max_it, toler, info);
    case OFH_DIAG_PREC: return bicg_impl<OFH_DIAG_PREC>(A, b, x, max_it, toler, info);
    case OFH_SSOR_PREC: return bicg_impl<OFH_SSOR_PREC>(A, b, x, max_it, toler, info);
    case OFH_DILU_PREC: return bicg_impl<OFH_DILU_PREC>(A, b, x, max_it, toler, info);
    case OFH_ILU_PREC: return bicg_impl<OFH_ILU_PREC>(A, b, x, max_it, toler, info);
    default: return fail(OFH_E_INVALID, "unknown preconditioner");
  }
}

int ofh_spmv_t_f64(ofh_csr A, const double* x, double* y) {
  OFH_CHECK(A && x && y, OFH_E_INVALID, "ofh_spmv_t_f64: null argument");
  OFH_CHECK(x != y, OFH_E_INVALID, "ofh_spmv_t_f64: x and y must differ");
  int rc = csr_ensure_transpose(A);
  if (rc) return rc;
  return ofh_spmv_f64(A->tr, x, y);
}

int ofh_prec_trans_apply(ofh_csr A, int prec, const double* b, double* x) {
  OFH_CHECK(A && b && x, OFH_E_INVALID, "ofh_prec_trans_apply: null argument");
  const int n = A->n;
  hipStream_t s = A->ctx->stream;
  int rc;
  switch (prec) {
    case OFH_DIAG_PREC:
      if ((rc = csr_ensure_dinv(A))) return rc;
      hipLaunchKernelGGL(k_scale_copy, dim3((n + 255) / 256), dim3(256), 0, s, n, b, A->dinv, x);
      break;
    case OFH_ILU_PREC:
      OFH_CHECK(b != x, OFH_E_INVALID, "ofh_prec_trans_apply: ILU needs distinct b and x");
      return launch_dilu_trans(A, b, x, nullptr, s);
    case OFH_IDENT_PREC:
    case OFH_DILU_PREC:
    case OFH_SSOR_PREC:  // no case in Prec::TransSolve: x = b
      if (b != x)
        hipLaunchKernelGGL(k_scale_copy, dim3((n + 255) / 256), dim3(256), 0, s, n, b,
                           (const double*)nullptr, x);
      break;
    default: return fail(OFH_E_INVALID, "unknown preconditioner");
  }
  OFH_HIP(hipGetLastError());
  return OFH_OK;
}

}  // extern "C"
