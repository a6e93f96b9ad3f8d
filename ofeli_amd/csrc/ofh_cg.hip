// ofh_cg.hip -- preconditioned CG on the GPU (CG.h:99-157 replacement).
//
// Launch sequence, IDENT/DIAG (host enqueues, the device decides), two
// kernels per iteration:
//   k_cg_init           r = b - A x, z = P r; partial (b,b) (r,r) (r,z)
//   k_cg_dirspmv(0)     fold; stop test of the initial guess (CG.h:116);
//                       p = z and q = A p; partial (p,q)
//   per iteration it = 1, 2, ...:
//     k_cg_upd(it)      fold (p,q): alpha = rho/(p,q); x += alpha p;
//                       r -= alpha q; z = P r; partial (r,r) (r,z)
//     k_cg_dirspmv(it)  fold: ||r||/||b|| <= toler -> stop (CG.h:144); else
//                       beta = rho_it/rho_{it-1}; q = A p with the new
//                       direction p = z + beta p (CG.h:127-135) formed at the
//                       SpMV's gather (ping-pong p buffers); partial (p,q)
// Larger systems and SSOR_PREC use three kernels per iteration: k_cg_spmv
// (q = A p), k_cg_upd and k_cg_dir (fold, stop test, p = z + beta p); with
// SSOR, k_cg_check folds (r,r) for the stop decision before the
// level-scheduled sweeps compute z and k_cg_rz writes (r,z).
// A reduction is folded by every block of the kernel that needs it (no
// intra-kernel cross-block traffic); all blocks fold the same partials in the
// same order and take the same decision; block 0 records it.  Once the stop
// flag is set every later kernel returns at once, so the iterations the host
// enqueues past convergence cost only launch slots.
// Entry-wise arithmetic is the reference's (p_i = z_i + beta p_i,
// x_i += alpha p_i, r_i -= alpha q_i, q = A p as sequential row sums); only
// the dot products are summed as fixed-order trees.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "ofh_internal.h"
#include "ofh_solver.h"
#include "ofh_spmv.h"

namespace ofh {

template <int PREC>
__device__ __forceinline__ double zval(int i, double ri, const double* __restrict__ dinv,
                                       const double* __restrict__ z) {
  if (PREC == OFH_DIAG_PREC) return ri * dinv[i];  // Prec_impl.h:245-247
  if (PREC == OFH_IDENT_PREC) return ri;
  return z[i];
}

// r = b - A x; partials (b,b) -> slot 0, (r,r) -> 1, (r,z) -> 2
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_cg_init(int nblocks, const int* __restrict__ rb,
                                                   const int* __restrict__ rp,
                                                   ColIndex ci,
                                                   const double* __restrict__ val,
                                                   const double* __restrict__ b,
                                                   const double* __restrict__ x,
                                                   double* __restrict__ r,
                                                   const double* __restrict__ dinv,
                                                   double* __restrict__ z,
                                                   double* __restrict__ partials) {
  __shared__ double lds[kTile64];
  __shared__ double red[3 * kBlock / kWave];
  const int blk = xcd_block(blockIdx.x, nblocks);
  const int r0 = rb[blk], r1 = rb[blk + 1];
  const double s = csr_stream_rowsum<double, kTile64>(r0, r1, rp, ci, val, x, lds, 0.0);
  double v[3] = {0.0, 0.0, 0.0};
  const int i = r0 + threadIdx.x;  // a long-row block holds one row, summed in thread 0
  if (i < r1) {
    const double bi = b[i], ri = bi - s;  // r = b - A*x (CG.h:115)
    r[i] = ri;
    v[0] = bi * bi;
    v[1] = ri * ri;
    if (!sweep_prec(PREC)) {
      const double zi = zval<PREC>(i, ri, dinv, nullptr);
      if (PREC == OFH_DIAG_PREC && z) z[i] = zi;
      v[2] = ri * zi;
    }
  }
  block_sum_k<kBlock, 3>(v, red);
  const int slot[3] = {0, 1, 2};
  write_partials<3>(v, slot, partials, blockIdx.x);
}

// Stop decision after `it` iterations and the new direction.
//   it == 0: folds the init partials (nparts0 blocks): ||b||, ||r0||, (r0,z0)
//   it >  0: folds (r,r) (slot 0) and (r,z) (slot 1 or, SSOR, slot 2)
// SSOR: the stop decision was taken by k_cg_check before the sweeps.
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_cg_dir(int n, int it_arg, int nparts0, int nparts,
                                                  const double* __restrict__ r,
                                                  const double* __restrict__ dinv,
                                                  const double* __restrict__ z,
                                                  double* __restrict__ p, KrylovState* st,
                                                  const double* __restrict__ partials) {
  if (st->done) return;
  const int it = it_arg >= 0 ? it_arg : st->kcur;
  const CgSnap sn = cg_snap(st, it);
  __shared__ double red[3 * kBlock / kWave];
  // issue this thread's loads first: they are in flight while the partials fold
  const int base = blockIdx.x * (kBlock * kVec) + threadIdx.x;
  double zv[kVec], pv[kVec];
#pragma unroll
  for (int j = 0; j < kVec; ++j) {
    const int i = base + j * kBlock;
    if (i < n) {
      zv[j] = sweep_prec(PREC) ? z[i] : r[i];
      if (PREC == OFH_DIAG_PREC) zv[j] *= dinv[i];  // z = D^-1 r (Prec_impl.h:245-247)
      pv[j] = p[i];
    }
  }
  double t[3];
  double nrm;
  if (sweep_prec(PREC)) {
    const int slot[1] = {2};
    double u[1];
    fold_partials<kBlock, 1>(u, slot, partials, nparts, red);  // written by k_cg_rz
    t[2] = u[0];
    nrm = sn.nrm;
  } else if (it == 0) {
    const int slot[3] = {0, 1, 2};
    fold_partials<kBlock, 3>(t, slot, partials, nparts0, red);
    nrm = sqrt(t[0]);
    if (nrm == 0) nrm = 1;  // CG.h:113-114
  } else {
    const int slot[2] = {0, 1};
    double u[2];
    fold_partials<kBlock, 2>(u, slot, partials, nparts, red);
    t[1] = u[0];
    t[2] = u[1];
    nrm = sn.nrm;
  }
  bool stop = false;
  int status = 0;
  double res = 0;
  if (!sweep_prec(PREC)) {
    res = sqrt(t[1]) / nrm;
    if (res <= sn.tol) {
      stop = true;
      status = it;  // 0 when the initial guess passes (CG.h:116-123)
    } else if (it >= sn.max_it || !(res == res)) {
      stop = true;  // the loop counter ends at max_it+1 (CG.h:154-156)
      status = -(sn.max_it + 1);
    }
  }
  const double rho = t[2];
  const double beta = it > 0 ? rho / sn.rho_prev : 0.0;  // CG.h:131
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->it = it;
    if (!sweep_prec(PREC)) {
      st->res = res;
      if (it == 0) {
        st->nrm = nrm;
        st->res0 = res;
      }
    }
    if (stop) {
      st->done = 1;
      st->status = status;
    } else {
      st->rho_buf[it & 1] = rho;
      st->beta = beta;
    }
  }
  if (stop) return;
#pragma unroll
  for (int j = 0; j < kVec; ++j) {
    const int i = base + j * kBlock;
    if (i < n) p[i] = it == 0 ? zv[j] : zv[j] + beta * pv[j];
  }
}

// The new direction p_c = z_c + beta p_c (CG.h:131-133; p_c = z_c for the
// first direction), formed where the SpMV gathers it.
struct GatherDir {
  const double* __restrict__ z;
  const double* __restrict__ p;
  double beta;
  int first;
  __device__ __forceinline__ double operator()(int c) const {
    return first ? z[c] : z[c] + beta * p[c];
  }
};

// IDENT/DIAG, after `it` iterations: fold (r,r), (r,z) (it > 0: slots 0, 1
// of k_cg_upd; it == 0: slots 0..2 of k_cg_init), stop test, beta; then
// q = A p_new with p_new = z + beta p_old; p_new of the block's rows is
// stored; partial (p,q) -> slot 3.  z is D^-1 r (DIAG) or r (IDENT); the
// directions alternate between p0 and p1 with the parity of `it`
// (it_arg < 0: the iteration index comes from the state).
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_cg_dirspmv(int nblocks, const int* __restrict__ rb,
                                                      const int* __restrict__ rp, ColIndex ci,
                                                      const double* __restrict__ val,
                                                      int it_arg, int nparts0, int nparts,
                                                      const double* __restrict__ z,
                                                      double* __restrict__ p0,
                                                      double* __restrict__ p1,
                                                      double* __restrict__ q, KrylovState* st,
                                                      double* __restrict__ partials) {
  if (st->done) return;
  const int it = it_arg >= 0 ? it_arg : st->kcur;
  const CgSnap sn = cg_snap(st, it);
  const double* __restrict__ p_old = (it & 1) ? p1 : p0;
  double* __restrict__ p_new = (it & 1) ? p0 : p1;
  __shared__ double lds[kTile64];
  __shared__ double red[3 * kBlock / kWave];
  const int blk = xcd_block(blockIdx.x, nblocks);
  const int r0 = rb[blk], r1 = rb[blk + 1];
  double t[3];
  double nrm;
  if (it == 0) {
    const int slot[3] = {0, 1, 2};
    fold_partials<kBlock, 3>(t, slot, partials, nparts0, red);
    nrm = sqrt(t[0]);
    if (nrm == 0) nrm = 1;  // CG.h:113-114
  } else {
    const int slot[2] = {0, 1};
    double u[2];
    fold_partials<kBlock, 2>(u, slot, partials, nparts, red);
    t[1] = u[0];
    t[2] = u[1];
    nrm = sn.nrm;
  }
  bool stop = false;
  int status = 0;
  const double res = sqrt(t[1]) / nrm;
  if (res <= sn.tol) {
    stop = true;
    status = it;  // 0 when the initial guess passes (CG.h:116-123)
  } else if (it >= sn.max_it || !(res == res)) {
    stop = true;  // the loop counter ends at max_it+1 (CG.h:154-156)
    status = -(sn.max_it + 1);
  }
  const double rho = t[2];
  const double beta = it > 0 ? rho / sn.rho_prev : 0.0;  // CG.h:131
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->it = it;
    st->res = res;
    if (it == 0) {
      st->nrm = nrm;
      st->res0 = res;
    }
    if (stop) {
      st->done = 1;
      st->status = status;
    } else {
      st->rho_buf[it & 1] = rho;
      st->beta = beta;
    }
  }
  if (stop) return;
  const GatherDir g{z, p_old, beta, it == 0};
  const double sq = csr_stream_rowsum_g<double, kTile64>(r0, r1, rp, ci, val, g, lds, 0.0);
  const int i = r0 + threadIdx.x;
  double v[1] = {0.0};
  if (i < r1) {
    const double pi = g(i);
    p_new[i] = pi;
    q[i] = sq;
    v[0] = pi * sq;
  }
  block_sum_k<kBlock, 1>(v, red);
  const int slot[1] = {3};
  write_partials<1>(v, slot, partials, blockIdx.x);
}

// q = A p; partial (p,q) -> slot 3 (CG.h:134-135)
__global__ __launch_bounds__(kBlock) void k_cg_spmv(int nblocks, const int* __restrict__ rb,
                                                   const int* __restrict__ rp,
                                                   ColIndex ci,
                                                   const double* __restrict__ val,
                                                   const double* __restrict__ p,
                                                   double* __restrict__ q,
                                                   const KrylovState* __restrict__ st,
                                                   double* __restrict__ partials) {
  if (st->done) return;
  __shared__ double lds[kTile64];
  __shared__ double red[kBlock / kWave];
  const int blk = xcd_block(blockIdx.x, nblocks);
  const int r0 = rb[blk], r1 = rb[blk + 1];
  const double s = csr_stream_rowsum<double, kTile64>(r0, r1, rp, ci, val, p, lds, 0.0);
  const int i = r0 + threadIdx.x;
  double v[1] = {0.0};
  if (i < r1) {
    q[i] = s;
    v[0] = p[i] * s;
  }
  block_sum_k<kBlock, 1>(v, red);
  const int slot[1] = {3};
  write_partials<1>(v, slot, partials, blockIdx.x);
}

// The new direction at a gather of the SELL-64 copy (buffer loads: the
// padding column reads z = p = +0.0, so it contributes -0.0 * +0.0).
struct SellGatherDir {
  XRsrc z, p;
  double beta;
  int first;
  __device__ __forceinline__ double operator()(uint32_t c) const {
    const double zc = x_load<double>(z, c);
    return first ? zc : zc + beta * x_load<double>(p, c);
  }
};

// k_cg_dirspmv with the product from the SELL-64 copy (W = 0) or the FP64
// diagonal copy of W slots per row (SYM: its symmetric form; one row per
// lane, 256 rows per workgroup): the same stop test, beta, p_new = z + beta
// p_old at the gathers, p_new of the block's rows stored, partial (p,q) ->
// slot 3.
template <int PREC, int W, bool SYM>
__global__ __launch_bounds__(kBlock) void k_cg_dirspmv_sell(
    int n, int nblocks, const int* __restrict__ cptr, const int* __restrict__ clen,
    const uint32_t* __restrict__ col, const int16_t* __restrict__ col16,
    const double* __restrict__ val, DiaF64 da, DiaSymRows<double> ds, int it_arg, int nparts0,
    int nparts, const double* __restrict__ z, double* __restrict__ p0, double* __restrict__ p1,
    double* __restrict__ q, KrylovState* st, double* __restrict__ partials) {
  if (st->done) return;
  const int it = it_arg >= 0 ? it_arg : st->kcur;
  const CgSnap sn = cg_snap(st, it);
  const double* __restrict__ p_old = (it & 1) ? p1 : p0;
  double* __restrict__ p_new = (it & 1) ? p0 : p1;
  __shared__ double red[3 * kBlock / kWave];
  double t[3];
  double nrm;
  if (it == 0) {
    const int slot[3] = {0, 1, 2};
    fold_partials<kBlock, 3>(t, slot, partials, nparts0, red);
    nrm = sqrt(t[0]);
    if (nrm == 0) nrm = 1;  // CG.h:113-114
  } else {
    const int slot[2] = {0, 1};
    double u[2];
    fold_partials<kBlock, 2>(u, slot, partials, nparts, red);
    t[1] = u[0];
    t[2] = u[1];
    nrm = sn.nrm;
  }
  bool stop = false;
  int status = 0;
  const double res = sqrt(t[1]) / nrm;
  if (res <= sn.tol) {
    stop = true;
    status = it;
  } else if (it >= sn.max_it || !(res == res)) {
    stop = true;
    status = -(sn.max_it + 1);
  }
  const double rho = t[2];
  const double beta = it > 0 ? rho / sn.rho_prev : 0.0;  // CG.h:131
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->it = it;
    st->res = res;
    if (it == 0) {
      st->nrm = nrm;
      st->res0 = res;
    }
    if (stop) {
      st->done = 1;
      st->status = status;
    } else {
      st->rho_buf[it & 1] = rho;
      st->beta = beta;
    }
  }
  if (stop) return;
  const SellGatherDir g{x_rsrc(z, n), x_rsrc(p_old, n), beta, it == 0};
  const int row = xcd_block(blockIdx.x, nblocks) * kBlock + threadIdx.x;
  double v[1] = {0.0};
  if (W > 0 ? row < n : row / kSellChunk * kSellChunk < n) {  // SELL: whole waves past the end only reduce
    double sq;
    if constexpr (W > 0)
      sq = dia_row<W, SYM>(row, da, ds, g);
    else
      sq = sell_rowsum_g<double, 2>(row, cptr, clen, col, col16, val, g);
    if (row < n) {
      const double pi = it == 0 ? z[row] : z[row] + beta * p_old[row];
      p_new[row] = pi;
      q[row] = sq;
      v[0] = pi * sq;
    }
  }
  block_sum_k<kBlock, 1>(v, red);
  const int slot[1] = {3};
  write_partials<1>(v, slot, partials, blockIdx.x);
}

// q = A p from the SELL-64 copy, partials of (p,q) -> slot 3
__global__ __launch_bounds__(kBlock) void k_cg_spmv_sell(int n, int nblocks,
                                                        const int* __restrict__ cptr,
                                                        const int* __restrict__ clen,
                                                        const uint32_t* __restrict__ col,
                                                        const int16_t* __restrict__ col16,
                                                        const double* __restrict__ val,
                                                        const double* __restrict__ p,
                                                        double* __restrict__ q,
                                                        const KrylovState* __restrict__ st,
                                                        double* __restrict__ partials) {
  if (st->done) return;
  __shared__ double red[kBlock / kWave];
  const int row = xcd_block(blockIdx.x, nblocks) * kBlock + threadIdx.x;
  double v[1] = {0.0};
  if (row / kSellChunk * kSellChunk < n) {  // whole waves past the end only reduce
    const double s = sell_rowsum<double, 2>(row, cptr, clen, col, col16, val, x_rsrc(p, n));
    if (row < n) {
      q[row] = s;
      v[0] = p[row] * s;
    }
  }
  block_sum_k<kBlock, 1>(v, red);
  const int slot[1] = {3};
  write_partials<1>(v, slot, partials, blockIdx.x);
}

// q = A p from the FP64 diagonal copy (W slots per row; SYM: its symmetric
// form), partials of (p,q) -> slot 3
template <int W, bool SYM>
__global__ __launch_bounds__(kBlock) void k_cg_spmv_dia(int nblocks, DiaF64 da,
                                                       DiaSymRows<double> ds,
                                                       const double* __restrict__ p,
                                                       double* __restrict__ q,
                                                       const KrylovState* __restrict__ st,
                                                       double* __restrict__ partials) {
  if (st->done) return;
  __shared__ double red[kBlock / kWave];
  const int row = xcd_block(blockIdx.x, nblocks) * kBlock + threadIdx.x;
  double v[1] = {0.0};
  if (row < da.n) {
    const double s = dia_row<W, SYM>(row, da, ds, SellGather<double>{x_rsrc(p, da.n)});
    q[row] = s;
    v[0] = p[row] * s;
  }
  block_sum_k<kBlock, 1>(v, red);
  const int slot[1] = {3};
  write_partials<1>(v, slot, partials, blockIdx.x);
}

// alpha = rho/(p,q); x += alpha p; r -= alpha q; partials (r,r) -> 0 and
// (r,z) -> 1 (CG.h:135-138)
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_cg_upd(int n, int it_arg, int nparts_pq,
                                                  double* __restrict__ x, double* __restrict__ r,
                                                  const double* __restrict__ p0,
                                                  const double* __restrict__ p1,
                                                  const double* __restrict__ q,
                                                  const double* __restrict__ dinv,
                                                  double* __restrict__ z, KrylovState* st,
                                                  double* __restrict__ partials) {
  if (st->done) return;
  // it_arg < 0: the iteration after the one the last decision recorded
  const int it = it_arg >= 0 ? it_arg : st->it + 1;
  const double rho_prev = st->rho_buf[(it - 1) & 1];  // loaded with the vectors
  const double* __restrict__ p = (it & 1) ? p1 : p0;
  __shared__ double red[2 * kBlock / kWave];
  // the (p,q) partials load first: their fold (and its barriers) runs while
  // this thread's vector loads are in flight
  const int sl[1] = {3};
  PartialsPrefetch<kBlock, 1, 14> pf;
  const bool pre = nparts_pq <= pf.kMax;
  if (pre) pf.load(sl, partials, nparts_pq);
  const int base = blockIdx.x * (kBlock * kVec) + threadIdx.x;
  double xv[kVec], pv[kVec], rv[kVec], qv[kVec], dv[kVec];
#pragma unroll
  for (int j = 0; j < kVec; ++j) {
    const int i = base + j * kBlock;
    if (i < n) {
      xv[j] = x[i];
      pv[j] = p[i];
      rv[j] = r[i];
      qv[j] = q[i];
      if (PREC == OFH_DIAG_PREC) dv[j] = dinv[i];
    }
  }
  double pq[1];
  if (pre)
    pf.fold(pq, nparts_pq, red);
  else
    fold_partials<kBlock, 1>(pq, sl, partials, nparts_pq, red);
  const double alpha = rho_prev / pq[0];
  double v[2] = {0.0, 0.0};
#pragma unroll
  for (int j = 0; j < kVec; ++j) {
    const int i = base + j * kBlock;
    if (i < n) {
      x[i] = xv[j] + alpha * pv[j];
      const double ri = rv[j] - alpha * qv[j];
      r[i] = ri;
      v[0] += ri * ri;
      if (PREC == OFH_DIAG_PREC) {
        const double zi = ri * dv[j];  // Prec_impl.h:245-247
        if (z) z[i] = zi;
        v[1] += ri * zi;
      } else if (PREC == OFH_IDENT_PREC) {
        v[1] += ri * ri;
      }
    }
  }
  block_sum_k<kBlock, 2>(v, red);
  const int slot[2] = {0, 1};
  write_partials<2>(v, slot, partials, blockIdx.x);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->alpha = alpha;
    st->kcur = it;
  }
}

// SSOR: stop decision on (r,r) before the sweeps (one block).
__global__ __launch_bounds__(kBlock) void k_cg_check(int it, int nparts, int init,
                                                    KrylovState* st,
                                                    const double* __restrict__ partials) {
  if (st->done) return;
  __shared__ double red[2 * kBlock / kWave];
  double t[2] = {0.0, 0.0};
  if (init) {
    const int s2[2] = {0, 1};  // (b,b), (r,r) of k_cg_init
    fold_partials<kBlock, 2>(t, s2, partials, nparts, red);
  } else {
    const int s1[1] = {0};  // (r,r) of k_cg_upd
    double u[1];
    fold_partials<kBlock, 1>(u, s1, partials, nparts, red);
    t[1] = u[0];
  }
  if (threadIdx.x != 0) return;
  double nrm = st->nrm;
  if (init) {
    nrm = sqrt(t[0]);
    if (nrm == 0) nrm = 1;
    st->nrm = nrm;
  }
  const double res = sqrt(t[1]) / nrm;
  st->res = res;
  if (init) st->res0 = res;
  st->it = it;
  if (res <= st->tol) {
    st->done = 1;
    st->status = it;
  } else if (it >= st->max_it || !(res == res)) {
    st->done = 1;
    st->status = -(st->max_it + 1);
  }
}

// SSOR: (r,z) after the sweeps -> slot 2
__global__ __launch_bounds__(kBlock) void k_cg_rz(int n, const double* __restrict__ r,
                                                 const double* __restrict__ z,
                                                 const KrylovState* __restrict__ st,
                                                 double* __restrict__ partials) {
  if (st->done) return;
  __shared__ double red[kBlock / kWave];
  double v[1] = {0.0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    v[0] += r[i] * z[i];
  block_sum_k<kBlock, 1>(v, red);
  const int slot[1] = {2};
  write_partials<1>(v, slot, partials, blockIdx.x);
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
  if (sweep_prec(PREC) && (rc = csr_ensure_prec(A, PREC))) return rc;
  void* wsp;
  const int nv = PREC == OFH_IDENT_PREC ? 4 : 5;
  if ((rc = ensure_vws(ctx, sizeof(double) * (size_t)n * nv, &wsp))) return rc;
  double* r = (double*)wsp;
  double* p = r + n;
  double* q = p + n;
  double* p1 = q + n;  // second direction buffer (IDENT/DIAG)
  double* z = PREC == OFH_IDENT_PREC ? r : p1 + n;  // SSOR: p1 + n is the sweep output
  KrylovState* st = state_ptr(ctx);
  KrylovState h0{};
  h0.tol = toler;
  h0.max_it = max_it;
  OFH_HIP(hipMemcpyAsync(st, &h0, sizeof(h0), hipMemcpyHostToDevice, s));
  double* part = ctx->ws.partials;
  const int g = reduce_grid(n);
  const int nb = A->nblocks;
  hipLaunchKernelGGL(k_cg_init<PREC>, dim3(nb), dim3(kBlock), 0, s, nb, A->row_blocks,
                     A->row_ptr, ColIndex{A->col, A->col16}, A->val, b, x, r, A->dinv,
                     PREC == OFH_DIAG_PREC ? z : nullptr, part);
  // The two-kernel form gathers z and the old direction at every nonzero;
  // it is faster only while those vectors stay L2-resident (measured on
  // MI355X: -5% per iteration at 261k rows, +26% at 857k rows).
  constexpr int kFuseRows = 3 << 17;
  if (!sweep_prec(PREC) && n <= kFuseRows) {
    const ColIndex ci{A->col, A->col16};
    double* zw = PREC == OFH_DIAG_PREC ? z : nullptr;
    // the product from the FP64 diagonal copy or the SELL-64 copy when one is
    // in use (one partial per 256 rows)
    bool dia = false, sell = false;
    const int nbs = (n + kBlock - 1) / kBlock;
    if (nbs <= kMaxPartials && (rc = csr_ensure_dia64(A, &dia))) return rc;
    if (!dia && nbs <= kMaxPartials && (rc = csr_ensure_sell(A, false, &sell))) return rc;
    bool sym = false;  // the symmetric form of the diagonal copy
    if (dia && (rc = csr_ensure_dias(A, 0, &sym))) return rc;
    const int npq = dia || sell ? nbs : nb;
    const DiaF64 da = dia ? dia_f64(A) : DiaF64{};
    const DiaSymRows<double> ds = sym ? dias_f64(A) : DiaSymRows<double>{};
    auto dirspmv = [&](int it, hipStream_t st_, int pk = -1) {  // pk: profiling kind
#define OFH_CG_DIRSPMV_S(W, SY)                                                                 \
  prof_kernel(ctx, pk, k_cg_dirspmv_sell<PREC, W, SY>, dim3(nbs), dim3(kBlock), st_, n, nbs,     \
              A->sell_cptr, A->sell_clen, A->sell_col, A->sell_col16, A->sell_val, da, ds, it,  \
              nb, g, z, p, p1, q, st, part)
      if (dia && A->dia_w == 4)
        sym ? OFH_CG_DIRSPMV_S(4, true) : OFH_CG_DIRSPMV_S(4, false);
      else if (dia && A->dia_w == 8)
        sym ? OFH_CG_DIRSPMV_S(8, true) : OFH_CG_DIRSPMV_S(8, false);
      else if (dia)
        sym ? OFH_CG_DIRSPMV_S(16, true) : OFH_CG_DIRSPMV_S(16, false);
      else if (sell)
        OFH_CG_DIRSPMV_S(0, false);
#undef OFH_CG_DIRSPMV_S
      else
        prof_kernel(ctx, pk, k_cg_dirspmv<PREC>, dim3(nb), dim3(kBlock), st_, nb, A->row_blocks,
                    A->row_ptr, ci, A->val, it, nb, g, z, p, p1, q, st, part);
    };
    // one iteration: k_cg_upd(it), k_cg_dirspmv(it); it = -1 takes the
    // index from the state (captured chunks)
    auto iteration = [&](int it, hipStream_t st_) {
      prof_kernel(ctx, OFH_PROF_CG_UPD, k_cg_upd<PREC>, dim3(g), dim3(kBlock), st_, n, it, npq, x,
                  r, p, p1, q, A->dinv, zw, st, part);
      dirspmv(it, st_, OFH_PROF_CG_DIRSPMV);
    };
    dirspmv(0, s);
    OFH_HIP(hipGetLastError());
    hipGraphExec_t chunk_graph = nullptr;
    if (ctx->use_graphs) {
      GraphKey key;
      csr_graph_key(A, 16 * 1 + PREC, (uintptr_t)sell | (uintptr_t)dia << 1, x, wsp, A->val,
                    A->dinv, &key);
      const bool prof = ctx->prof;
      ctx->prof = false;  // no event records inside a captured graph
      rc = ctx_graph(ctx, key, [&](hipStream_t cs) {
        for (int c = 0; c < kGraphIters; ++c) iteration(-1, cs);
        OFH_HIP(hipGetLastError());
        return OFH_OK;
      }, &chunk_graph);
      ctx->prof = prof;
      if (rc) return rc;
    }
    KrylovState h{};
    if ((rc = read_state(ctx, &h))) return rc;
    int enq = 0, it_prev = 0;
    double res_prev = h.res;
    const int it0 = h.it;
    const double res0 = h.res;
    // each look at the state is taken while up to kLookAhead more
    // iterations run (launches past the stop return at once)
    int ahead = 0;
    while (!h.done && enq < max_it) {
      const int chunk = std::max(
          1, next_chunk(h.it, h.res, it_prev, res_prev, toler, max_it - enq, it0, res0) - ahead);
      it_prev = h.it;
      res_prev = h.res;
      // whole captured chunks, then the remainder launched directly
      const int full = chunk_graph ? chunk / kGraphIters : 0;
      for (int c = 0; c < full; ++c) OFH_HIP(hipGraphLaunch(chunk_graph, s));
      enq += full * kGraphIters;
      for (int c = full * kGraphIters; c < chunk && enq < max_it; ++c) iteration(++enq, s);
      OFH_HIP(hipGetLastError());
      ahead = std::min(kLookAhead, max_it - enq);
      if ((rc = read_state_ahead(ctx, &h, [&]() -> int {
             for (int c = 0; c < ahead; ++c) iteration(++enq, s);
             OFH_HIP(hipGetLastError());
             return OFH_OK;
           })))
        return rc;
    }
    if (ahead > 0 && (rc = read_state(ctx, &h))) return rc;  // the final state
    if (!h.done) {  // max_it == 0 with a non-converged guess
      h.done = 1;
      h.status = -(max_it + 1);
    }
    return fill_info(A, b, x, h, t0, info);
  }
  auto ssor_z = [&](int it, bool init) -> int {
    hipLaunchKernelGGL(k_cg_check, dim3(1), dim3(kBlock), 0, s, it, init ? nb : g, (int)init, st,
                       part);
    int e = launch_prec(A, PREC, r, z, &st->done, s);
    if (e) return e;
    hipLaunchKernelGGL(k_cg_rz, dim3(g), dim3(kBlock), 0, s, n, r, z, st, part);
    return OFH_OK;
  };
  if (sweep_prec(PREC) && (rc = ssor_z(0, true))) return rc;
  hipLaunchKernelGGL(k_cg_dir<PREC>, dim3(g), dim3(kBlock), 0, s, n, 0, nb, g, r, A->dinv, z, p,
                     st, part);
  OFH_HIP(hipGetLastError());
  KrylovState h{};
  if ((rc = read_state(ctx, &h))) return rc;
  int enq = 0, it_prev = 0;
  double res_prev = h.res;
  // the SpMV from the FP64 diagonal copy or the SELL-64 copy when one is in
  // use (one partial per 256 rows)
  bool dia = false, sell = false;
  const int nbs = (n + kBlock - 1) / kBlock;
  if (nbs <= kMaxPartials && (rc = csr_ensure_dia64(A, &dia))) return rc;
  if (!dia && nbs <= kMaxPartials && (rc = csr_ensure_sell(A, false, &sell))) return rc;
  bool sym = false;  // the symmetric form of the diagonal copy
  if (dia && (rc = csr_ensure_dias(A, 0, &sym))) return rc;
  const int npq = dia || sell ? nbs : nb;
  const DiaF64 da = dia ? dia_f64(A) : DiaF64{};
  const DiaSymRows<double> ds = sym ? dias_f64(A) : DiaSymRows<double>{};
  // one iteration: k_cg_spmv, k_cg_upd(it), [SSOR sweeps], k_cg_dir(it)
  auto iteration = [&](int it, hipStream_t st_) -> int {
#define OFH_CG_SPMV_DIA(W, SY)                                                                  \
  prof_kernel(ctx, OFH_PROF_CG_SPMV, k_cg_spmv_dia<W, SY>, dim3(nbs), dim3(kBlock), st_, nbs, da, \
              ds, p, q, st, part)
    if (dia && A->dia_w == 4)
      sym ? OFH_CG_SPMV_DIA(4, true) : OFH_CG_SPMV_DIA(4, false);
    else if (dia && A->dia_w == 8)
      sym ? OFH_CG_SPMV_DIA(8, true) : OFH_CG_SPMV_DIA(8, false);
    else if (dia)
      sym ? OFH_CG_SPMV_DIA(16, true) : OFH_CG_SPMV_DIA(16, false);
#undef OFH_CG_SPMV_DIA
    else if (sell)
      prof_kernel(ctx, OFH_PROF_CG_SPMV, k_cg_spmv_sell, dim3(nbs), dim3(kBlock), st_, n, nbs,
                  A->sell_cptr, A->sell_clen, A->sell_col, A->sell_col16, A->sell_val, p, q, st,
                  part);
    else
      prof_kernel(ctx, OFH_PROF_CG_SPMV, k_cg_spmv, dim3(nb), dim3(kBlock), st_, nb,
                  A->row_blocks, A->row_ptr, ColIndex{A->col, A->col16}, A->val, p, q, st, part);
    prof_kernel(ctx, OFH_PROF_CG_UPD, k_cg_upd<PREC>, dim3(g), dim3(kBlock), st_, n, it, npq, x,
                r, p, p, q, A->dinv, (double*)nullptr, st, part);
    if (sweep_prec(PREC)) {
      const int e = ssor_z(it, false);
      if (e) return e;
    }
    prof_kernel(ctx, OFH_PROF_CG_DIR, k_cg_dir<PREC>, dim3(g), dim3(kBlock), st_, n, it, nb, g, r,
                A->dinv, z, p, st, part);
    return OFH_OK;
  };
  hipGraphExec_t chunk_graph = nullptr;
  if (ctx->use_graphs && !sweep_prec(PREC)) {
    GraphKey key;
    csr_graph_key(A, 16 * 2 + PREC, (uintptr_t)sell | (uintptr_t)dia << 1, x, wsp, A->val,
                  A->dinv, &key);
    const bool prof = ctx->prof;
    ctx->prof = false;  // no event records inside a captured graph
    rc = ctx_graph(ctx, key, [&](hipStream_t cs) {
      for (int c = 0; c < kGraphIters; ++c) {
        const int e = iteration(-1, cs);
        if (e) return e;
      }
      OFH_HIP(hipGetLastError());
      return OFH_OK;
    }, &chunk_graph);
    ctx->prof = prof;
    if (rc) return rc;
  }
  const int it0 = h.it;
  const double res0 = h.res;
  int ahead = 0;  // as in the two-kernel loop
  while (!h.done && enq < max_it) {
    const int chunk = std::max(
        1, next_chunk(h.it, h.res, it_prev, res_prev, toler, max_it - enq, it0, res0) - ahead);
    it_prev = h.it;
    res_prev = h.res;
    // whole captured chunks, then the remainder launched directly
    const int full = chunk_graph ? chunk / kGraphIters : 0;
    for (int c = 0; c < full; ++c) OFH_HIP(hipGraphLaunch(chunk_graph, s));
    enq += full * kGraphIters;
    for (int c = full * kGraphIters; c < chunk && enq < max_it; ++c)
      if ((rc = iteration(++enq, s))) return rc;
    OFH_HIP(hipGetLastError());
    ahead = std::min(kLookAhead, max_it - enq);
    if ((rc = read_state_ahead(ctx, &h, [&]() -> int {
           for (int c = 0; c < ahead; ++c)
             if (const int e = iteration(++enq, s)) return e;
           OFH_HIP(hipGetLastError());
           return OFH_OK;
         })))
      return rc;
  }
  if (ahead > 0 && (rc = read_state(ctx, &h))) return rc;  // the final state
  if (!h.done) {  // max_it == 0 with a non-converged guess
    h.done = 1;
    h.status = -(max_it + 1);
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
    case OFH_DILU_PREC:  // the reference's ILU is its DILU
    case OFH_ILU_PREC: return cg_impl<OFH_DILU_PREC>(A, b, x, max_it, toler, info);
    default: return fail(OFH_E_INVALID, "unknown preconditioner");
  }
}

// ---------------------------------------------------------------------------
// CG without allocation (ofh_cg_noalloc): the IDENT/DIAG two-kernel iteration
// above over the CSR-stream partition, with every buffer beyond the caller's
// matrix, b and x in the file-scope device array `scratch`:
//   [partials: kSlots x kMaxPartials | state block | 6 vectors of kNaMaxRows]
// The array exists once per GPU for the life of the process; host code uses
// the address hipGetSymbolAddress returned for the current device at context
// creation (noalloc_setup), never the array's name.  Every call on a GPU
// shares that GPU's array: calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
constexpr int kNaMaxRows = 1 << 18;  // 262,144 rows
constexpr size_t kNaPartials = (size_t)kSlots * kMaxPartials;  // doubles
constexpr size_t kNaStateDoubles = 128;  // KrylovState (<= 512 B), result, pivot flag
constexpr size_t kNaDoubles = kNaPartials + kNaStateDoubles + 6 * (size_t)kNaMaxRows;
static_assert(sizeof(KrylovState) <= 512, "state block");

static __device__ __align__(256) double scratch[kNaDoubles];

constexpr int kNaMaxDevices = 64;
static double* g_na_scratch[kNaMaxDevices] = {};
static std::mutex g_na_mu;

int noalloc_setup(int device) {
  OFH_CHECK(device >= 0 && device < kNaMaxDevices, OFH_E_INVALID,
            "ofh_cg_noalloc: device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (g_na_scratch[device]) return OFH_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(OFH_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_na_scratch[device] = static_cast<double*>(ptr);
  return OFH_OK;
}

// KrylovState of a new solve (a kernel, so that no host copy is needed)
__global__ void k_na_state_init(KrylovState* st, double tol, int max_it, int* bad) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    *st = KrylovState{};
    st->tol = tol;
    st->max_it = max_it;
    *bad = 0x7fffffff;
  }
}

// pivot_i = 1/a_ii (Prec_impl.h:306-316) with the diagonal found by a binary
// search of the sorted row; a missing or zero diagonal is reported in *bad.
__global__ void k_na_inv_diag(int n, const int* __restrict__ rp, const int* __restrict__ col,
                              const double* __restrict__ val, double* __restrict__ dinv,
                              int* __restrict__ bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int lo = rp[i], hi = rp[i + 1] - 1, pos = -1;
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1;
    const int c = col[mid];
    if (c == i) {
      pos = mid;
      break;
    }
    if (c < i) lo = mid + 1;
    else hi = mid - 1;
  }
  const double a = pos < 0 ? 0.0 : val[pos];
  if (a == 0.0) {
    atomicMin(bad, i);
    dinv[i] = 0.0;
  } else {
    dinv[i] = 1.0 / a;
  }
}

template <int PREC>
static int cg_noalloc_impl(ofh_csr A, double* sc, const double* b, double* x, int max_it,
                           double toler, ofh_solve_info* info) {
  ofh_ctx ctx = A->ctx;
  hipStream_t s = ctx->stream;
  const int n = A->n;
  const double t0 = now_seconds();
  double* part = sc;
  KrylovState* st = reinterpret_cast<KrylovState*>(sc + kNaPartials);
  double* d_out = sc + kNaPartials + 64;
  int* bad = reinterpret_cast<int*>(sc + kNaPartials + 72);
  double* r = sc + kNaPartials + kNaStateDoubles;
  double* p0 = r + kNaMaxRows;
  double* p1 = p0 + kNaMaxRows;
  double* q = p1 + kNaMaxRows;
  double* z = q + kNaMaxRows;
  double* dinv = z + kNaMaxRows;
  hipLaunchKernelGGL(k_na_state_init, dim3(1), dim3(64), 0, s, st, toler, max_it, bad);
  if (PREC == OFH_DIAG_PREC) {
    hipLaunchKernelGGL(k_na_inv_diag, dim3((n + 255) / 256), dim3(256), 0, s, n, A->row_ptr,
                       A->col, A->val, dinv, bad);
    int hbad = 0;
    OFH_HIP(hipMemcpyAsync(ctx->host_pinned, bad, sizeof(int), hipMemcpyDeviceToHost, s));
    OFH_HIP(hipStreamSynchronize(s));
    std::memcpy(&hbad, ctx->host_pinned, sizeof(int));
    if (hbad != 0x7fffffff)
      return fail(OFH_E_ZERO_PIVOT,
                  "In Prec::inv_diag(): null diagonal term: " + std::to_string(hbad + 1));
  }
  const int g = reduce_grid(n);
  const int nb = A->nblocks;
  const ColIndex ci{A->col, A->col16};
  double* zw = PREC == OFH_DIAG_PREC ? z : nullptr;
  const double* dv = PREC == OFH_DIAG_PREC ? dinv : nullptr;
  hipLaunchKernelGGL(k_cg_init<PREC>, dim3(nb), dim3(kBlock), 0, s, nb, A->row_blocks, A->row_ptr,
                     ci, A->val, b, x, r, dv, zw, part);
  const double* zsrc = PREC == OFH_DIAG_PREC ? z : r;
  auto dirspmv = [&](int it) {
    hipLaunchKernelGGL(k_cg_dirspmv<PREC>, dim3(nb), dim3(kBlock), 0, s, nb, A->row_blocks,
                       A->row_ptr, ci, A->val, it, nb, g, zsrc, p0, p1, q, st, part);
  };
  dirspmv(0);
  OFH_HIP(hipGetLastError());
  auto read = [&](KrylovState* h) -> int {
    OFH_HIP(hipMemcpyAsync(ctx->host_pinned, st, sizeof(KrylovState), hipMemcpyDeviceToHost, s));
    OFH_HIP(hipStreamSynchronize(s));
    std::memcpy(h, ctx->host_pinned, sizeof(KrylovState));
    return OFH_OK;
  };
  KrylovState h{};
  int rc;
  if ((rc = read(&h))) return rc;
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
      ++enq;
      hipLaunchKernelGGL(k_cg_upd<PREC>, dim3(g), dim3(kBlock), 0, s, n, enq, nb, x, r, p0, p1, q,
                         dv, zw, st, part);
      dirspmv(enq);
    }
    OFH_HIP(hipGetLastError());
    if ((rc = read(&h))) return rc;
  }
  if (!h.done) {  // max_it == 0 with a non-converged guess
    h.done = 1;
    h.status = -(max_it + 1);
  }
  const double t1 = now_seconds();
  if (!info) return OFH_OK;
  info->nb_it = h.status;
  info->iterations = h.it;
  info->relres = h.res;
  info->outer = 0;
  info->fallback = 0;
  if ((rc = true_relres_csr(A, b, x, part, d_out, &info->true_relres))) return rc;
  info->seconds = t1 - t0;
  return OFH_OK;
}

int cg_noalloc(ofh_csr A, int prec, const double* b, double* x, int max_it, double toler,
               ofh_solve_info* info) {
  OFH_CHECK(A && b && x, OFH_E_INVALID, "ofh_cg_noalloc: null argument");
  OFH_CHECK(max_it >= 0, OFH_E_INVALID, "ofh_cg_noalloc: negative max_it");
  OFH_CHECK(A->n <= kNaMaxRows, OFH_E_INVALID,
            "ofh_cg_noalloc: " + std::to_string(A->n) + " rows exceed the scratch maximum of " +
                std::to_string(kNaMaxRows));
  const int dev = A->ctx->device;
  double* sc = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_na_mu);
    sc = dev >= 0 && dev < kNaMaxDevices ? g_na_scratch[dev] : nullptr;
  }
  OFH_CHECK(sc, OFH_E_INVALID, "ofh_cg_noalloc: scratch not set up for this device");
  switch (prec) {
    case OFH_IDENT_PREC: return cg_noalloc_impl<OFH_IDENT_PREC>(A, sc, b, x, max_it, toler, info);
    case OFH_DIAG_PREC: return cg_noalloc_impl<OFH_DIAG_PREC>(A, sc, b, x, max_it, toler, info);
    default: return fail(OFH_E_UNSUPPORTED, "ofh_cg_noalloc: IDENT_PREC or DIAG_PREC only");
  }
}

}  // namespace ofh

extern "C" int ofh_cg(ofh_csr A, int prec, const double* b, double* x, int max_it,
                      double toler, ofh_solve_info* info) {
  return ofh::cg_dispatch(A, prec, b, x, max_it, toler, info);
}

extern "C" int ofh_cg_noalloc(ofh_csr A, int prec, const double* b, double* x, int max_it,
                              double toler, ofh_solve_info* info) {
  return ofh::cg_noalloc(A, prec, b, x, max_it, toler, info);
}
