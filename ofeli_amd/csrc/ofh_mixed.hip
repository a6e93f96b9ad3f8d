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
// stops halving over three replacements, or twice in a row a cycle is slow
// (its true-residual reduction per inner iteration is below a fifth of the
// best cycle's, or the recursive residual needed over 3x the first cycle's
// length: FP32 no longer resolves the slow modes and the host forced the
// replacement), or the very first cycle had to be forced (2,000 inner
// iterations without a kDelta drop), the FP64 CG finishes from the current
// x64.
// Per inner iteration two kernels, k_mx_upd and k_mx_dirspmv[_dia], when the
// product reads the diagonal copy (structured-mesh patterns) or the SELL-64
// copy: there q = A p is carried by recurrence, q_new = A z + beta q_old, so
// the product gathers z alone and runs before the direction's decision.
// Without either copy, L2-resident systems
// (n <= 393,216) form the direction at the product's gathers and larger ones
// use three kernels: k_mx_spmv, k_mx_upd, k_mx_dir.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "ofh_internal.h"
#include "ofh_solver.h"
#include "ofh_spmv.h"

namespace ofh {

constexpr double kDelta = 0.1;  // recursive-residual drop that triggers a replacement
// DIAG_PREC applied as a symmetric scaling: the FP32 inner iteration is CG
// without preconditioner on D^-1/2 A D^-1/2 (the same iterates as Jacobi-
// preconditioned CG on A in exact arithmetic) with r32 = D^-1/2 r and
// x64 += D^-1/2 x32, so z = r and the update reads no pivots and writes no z.
constexpr int kScaledDiag = 100;

// Inner iterations to enqueue before the host looks at the state again.  The
// device pauses by itself at the replacement point and the launches queued
// past it return at once, but each still costs its dispatch; a look costs one
// round trip.  So approach the pause from below: 0.9 of the predicted
// distance from the residual `res` after `it` iterations to `target`, the
// rate taken over the whole cycle so far (from `res_start` at `cycle_start`:
// the norm of a CG residual is not monotone, and a rate over one short chunk
// can be ~1 on a plateau); at the start of a cycle the previous cycle's
// length, scaled by how far this one has to go against a drop of kDelta (the
// last cycle only has to reach the tolerance).  At least 2, at most the
// previous cycle's length, at most max(32, it) while the solve is young.
// (512^2 heat steps with projected starts: 2.73 ms per step against 2.90 ms
// with a rate over the last chunk and 3/4 of the distance, which enqueued up
// to 190 launches past the pause per solve.)
inline int mixed_chunk(int it, double res, int cycle_start, double res_start, double target,
                       int remaining, int last_cycle) {
  int c = 8;  // no usable rate yet
  const double need = std::log(target / res);  // < 0 while res is above target
  if (it == cycle_start) {
    if (last_cycle > 0) c = (int)(0.9 * last_cycle * need / std::log(kDelta));
  } else if (res > 0 && res < res_start && need < 0) {
    const double lr = std::log(res / res_start) / (it - cycle_start);  // < 0
    c = (int)(0.9 * need / lr);
  }
  if (last_cycle > 0) c = std::min(c, last_cycle);
  c = std::max(2, std::min(c, std::max(32, it)));
  return std::max(1, std::min(c, remaining));
}

// z_i of the inner iteration: D^-1 r (DIAG, d = the FP32 pivots), r (IDENT),
// or, for the sweep preconditioners (SSOR, DILU), the z the sweeps computed
// (d = that z vector, see k_mx_rz).
template <int PREC>
__device__ __forceinline__ float zf(int i, float ri, const float* __restrict__ d) {
  if (sweep_prec(PREC)) return d[i];
  return PREC == OFH_DIAG_PREC ? ri * d[i] : ri;
}

// The state scalars a decision reads, loaded together at the start of the
// kernel (one round trip, overlapped with the kernel's other loads) rather
// than one dependent load after another where the decision needs them.
// rho_prev is rho_{k-1} (an unused slot when k == 0).
struct MxSnap {
  double nrm, rmax, tol, rho_prev, rscale;
  int max_it;
};

__device__ __forceinline__ MxSnap mx_snap(const KrylovState* st, int k) {
  return MxSnap{st->nrm, st->rmax, st->tol, st->rho_buf[(k - 1) & 1], st->rscale, st->max_it};
}

// FP64 residual r = b - A x64 -> r32 (rounded), x32 = 0; partials
// (b,b) -> sb, (r,r) -> sb + 1 in FP64, (r32,z32) -> sb + 2, (r32,r32) -> sb + 3.
// (kScaledDiag: r32 = D^-1/2 r, sc = D^-1/2; the FP64 partials stay unscaled.)
template <int PREC, bool SELL>
__global__ __launch_bounds__(kBlock) void k_mx_resid(
    MatRows64 m, const double* __restrict__ b, const double* __restrict__ x,
    float* __restrict__ r32, float* __restrict__ x32, const float* __restrict__ dinv32,
    float* __restrict__ z32, const double* __restrict__ sc, const KrylovState* __restrict__ st,
    double* __restrict__ partials, int sb) {
  if (st->done) return;
  __shared__ double red[4 * kBlock / kWave];
  int i;
  bool valid;
  const double s = mat_rowsum64<SELL>(m, x, i, valid);
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  if (valid) {
    const double bi = b[i], ri = bi - s;
    const float rf = PREC == kScaledDiag ? (float)(sc[i] * ri) : (float)ri;
    r32[i] = rf;
    x32[i] = 0.0f;
    v[0] = bi * bi;
    v[1] = ri * ri;
    if (!sweep_prec(PREC)) {  // sweeps: (r,z) -> slot 2 comes from k_mx_rz
      const float z = zf<PREC>(i, rf, dinv32);
      if (PREC == OFH_DIAG_PREC && z32) z32[i] = z;
      v[2] = (double)rf * (double)z;
    }
    v[3] = (double)rf * (double)rf;
  }
  block_sum_k<kBlock, 4>(v, red);
  const int slot[4] = {sb, sb + 1, sb + 2, sb + 3};
  write_partials<4>(v, slot, partials, blockIdx.x);
}

// Decision and direction after `k` inner iterations.
//   PHASE 0 (after k_mx_upd): recursive residual; pause for a replacement
//           when it dropped by kDelta, claims convergence or k hits max_it;
//           else beta = rho_k/rho_{k-1}, p = z + beta p.
//   PHASE 1 (after the initial k_mx_resid): ||b||, stop test of the guess, p = z.
//   PHASE 2 (after a replacement): true residual stop test; else keep p:
//           beta = (r,z)_new / rho_{k-1}, p = z + beta p.
template <int PREC, int PHASE>
__global__ __launch_bounds__(kBlock) void k_mx_dir(int n, int k_arg, int nparts,
                                                  const float* __restrict__ r32,
                                                  const float* __restrict__ dinv32,
                                                  float* __restrict__ p32, KrylovState* st,
                                                  const double* __restrict__ partials) {
  if (PHASE == 0 ? st->halt : st->done) return;
  const int k = k_arg >= 0 ? k_arg : st->kcur;  // k_arg < 0: index from the state
  const MxSnap sn = mx_snap(st, k);
  __shared__ double red[4 * kBlock / kWave];
  const int base = blockIdx.x * (kBlock * kVec) + threadIdx.x;
  float zv[kVec], pv[kVec];
#pragma unroll
  for (int j = 0; j < kVec; ++j) {  // loads in flight during the fold
    const int i = base + j * kBlock;
    if (i < n) {
      zv[j] = zf<PREC>(i, r32[i], dinv32);
      pv[j] = p32[i];
    }
  }
  double nrm = sn.nrm, res, rz;
  bool stop = false, pause = false;
  int status = 0;
  if (PHASE == 0) {
    double t[2];
    const int slot[2] = {4, 5};
    fold_partials<kBlock, 2>(t, slot, partials, nparts, red);
    res = sqrt(t[0]) / nrm;
    rz = t[1];
    pause = (res <= kDelta * sn.rmax) | (res <= sn.tol) | (k >= sn.max_it) | !(res == res);
  } else {
    double t[4];
    const int slot[4] = {0, 1, 2, 3};
    fold_partials<kBlock, 4>(t, slot, partials, nparts, red);
    if (PHASE == 1) {
      nrm = sqrt(t[0]);
      if (nrm == 0) nrm = 1;
    }
    res = sqrt(t[1]) / nrm;
    rz = t[2];
    if (res <= sn.tol) {
      stop = true;
      status = k;
    } else if (k >= sn.max_it || !(res == res)) {
      stop = true;
      status = -(sn.max_it + 1);
    }
  }
  const double beta = k > 0 ? rz / sn.rho_prev : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->it = k;
    if (PHASE == 0) {
      st->res = res;  // recursive estimate
      if (pause) {
        st->pause = 1;
        st->halt = 1;
      }
    } else {
      st->sconv_res = res;  // true FP64 residual
      st->res = res;
      st->rmax = res;
      if (PHASE == 1) {
        st->nrm = nrm;
        st->res0 = res;
      } else {
        st->replacements += 1;
      }
      st->pause = 0;
      st->halt = stop ? 1 : 0;
      if (stop) {
        st->done = 1;
        st->status = status;
      }
    }
    if (!stop && !pause) {
      st->rho_buf[k & 1] = rz;
      st->beta = beta;
    }
  }
  if (stop || pause) return;
  const float bf = (float)beta;
#pragma unroll
  for (int j = 0; j < kVec; ++j) {
    const int i = base + j * kBlock;
    if (i < n) p32[i] = k == 0 ? zv[j] : zv[j] + bf * pv[j];
  }
}

// q32 = A32 p32 from the SELL-64 copy; (p,q) -> slot 6 (FP64 accumulation)
__global__ __launch_bounds__(kBlock) void k_mx_spmv_sell(int n, int nblocks,
                                                        const int* __restrict__ cptr,
                                                        const int* __restrict__ clen,
                                                        const uint32_t* __restrict__ col,
                                                        const int16_t* __restrict__ col16,
                                                        const float* __restrict__ val32,
                                                        const float* __restrict__ p32,
                                                        float* __restrict__ q32,
                                                        const KrylovState* __restrict__ st,
                                                        double* __restrict__ partials) {
  if (st->halt) return;
  __shared__ double red[kBlock / kWave];
  const int row = xcd_block(blockIdx.x, nblocks) * kBlock + threadIdx.x;
  double v[1] = {0.0};
  if (row / kSellChunk * kSellChunk < n) {  // whole waves past the end only reduce
    const float s = sell_rowsum<float, 4>(row, cptr, clen, col, col16, val32, x_rsrc(p32, n));
    if (row < n) {
      q32[row] = s;
      v[0] = (double)p32[row] * (double)s;
    }
  }
  block_sum_k<kBlock, 1>(v, red);
  const int slot[1] = {6};
  write_partials<1>(v, slot, partials, blockIdx.x);
}

// q32 = A32 p32; (p,q) -> slot 6 (FP64 accumulation)
__global__ __launch_bounds__(kBlock) void k_mx_spmv(int nblocks, const int* __restrict__ rb,
                                                   const int* __restrict__ rp,
                                                   ColIndex ci,
                                                   const float* __restrict__ val32,
                                                   const float* __restrict__ p32,
                                                   float* __restrict__ q32,
                                                   const KrylovState* __restrict__ st,
                                                   double* __restrict__ partials) {
  if (st->halt) return;
  __shared__ float lds[kTile32];
  __shared__ double red[kBlock / kWave];
  const int blk = xcd_block(blockIdx.x, nblocks);
  const int r0 = rb[blk], r1 = rb[blk + 1];
  const float s = csr_stream_rowsum<float, kTile32>(r0, r1, rp, ci, val32, p32, lds, 0.0f);
  const int i = r0 + threadIdx.x;
  double v[1] = {0.0};
  if (i < r1) {
    q32[i] = s;
    v[0] = (double)p32[i] * (double)s;
  }
  block_sum_k<kBlock, 1>(v, red);
  const int slot[1] = {6};
  write_partials<1>(v, slot, partials, blockIdx.x);
}

// The new FP32 direction p_c = z_c + beta p_c (p_c = z_c when k == 0),
// formed where the SpMV gathers it.
struct GatherDirF {
  const float* __restrict__ z;
  const float* __restrict__ p;
  float beta;
  int first;
  __device__ __forceinline__ float operator()(int c) const {
    return first ? z[c] : z[c] + beta * p[c];
  }
};

// The same at a gather of the SELL-64 copy (buffer loads: the padding column
// reads z = p = +0.0, so it contributes -0.0 * +0.0).
struct SellGatherDirF {
  XRsrc z, p;
  float beta;
  int first;
  __device__ __forceinline__ float operator()(uint32_t c) const {
    const float zc = x_load<float>(z, c);
    return first ? zc : zc + beta * x_load<float>(p, c);
  }
};

// The SELL-64 copy's FP32 arrays (SELL kernels only).
struct SellF32 {
  int n;
  const int* __restrict__ cptr;
  const int* __restrict__ clen;
  const uint32_t* __restrict__ col;
  const int16_t* __restrict__ col16;
  const float* __restrict__ val;
};

// The decision of k_mx_dirspmv after `k` inner iterations (PHASE as in
// k_mx_dir): folds the partials (slots 4, 5 from the prefetch `pf` when
// `pre`, else slots sb..sb+3 (PHASE 1, 2) or 4, 5 from memory), records the
// state (block 0), and returns false when the launch stops or pauses; else beta.
template <int PHASE, int PREC, class PF>
__device__ __forceinline__ bool mx_decide(int k, int nparts, bool pre, const PF& pf,
                                          const double* __restrict__ partials, KrylovState* st,
                                          const MxSnap& sn, double* red, double& beta_out,
                                          int sb = 0) {
  double nrm = sn.nrm, res, rz, rs = 0.0;
  bool stop = false, pause = false;
  int status = 0;
  if (PHASE == 0) {
    double t[2];
    const int slot[2] = {4, 5};
    if (pre)
      pf.fold(t, nparts, red);
    else
      fold_partials<kBlock, 2>(t, slot, partials, nparts, red);
    // the recursive residual as a true-relres estimate (scaled: its ratio to
    // the scaled residual at the last replacement)
    res = PREC == kScaledDiag ? sqrt(t[0]) * sn.rscale : sqrt(t[0]) / nrm;
    rz = t[1];
    pause = (res <= kDelta * sn.rmax) | (res <= sn.tol) | (k >= sn.max_it) | !(res == res);
  } else {
    double t[4];
    const int slot[4] = {sb, sb + 1, sb + 2, sb + 3};
    fold_partials<kBlock, 4>(t, slot, partials, nparts, red);
    if (PHASE == 1) {
      nrm = sqrt(t[0]);
      if (nrm == 0) nrm = 1;
    }
    res = sqrt(t[1]) / nrm;
    rz = t[2];
    rs = t[3] > 0 ? res / sqrt(t[3]) : 0.0;
    if (res <= sn.tol) {
      stop = true;
      status = k;
    } else if (k >= sn.max_it || !(res == res)) {
      stop = true;
      status = -(sn.max_it + 1);
    }
  }
  const double beta = k > 0 ? rz / sn.rho_prev : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->it = k;
    if (PHASE == 0) {
      st->res = res;
      if (pause) {
        st->pause = 1;
        st->halt = 1;
      }
    } else {
      st->sconv_res = res;
      st->res = res;
      st->rmax = res;
      if (PREC == kScaledDiag) st->rscale = rs;
      if (PHASE == 1) {
        st->nrm = nrm;
        st->res0 = res;
      } else {
        st->replacements += 1;
      }
      st->pause = 0;
      st->halt = stop ? 1 : 0;
      if (stop) {
        st->done = 1;
        st->status = status;
      }
    }
    if (!stop && !pause) {
      st->rho_buf[k & 1] = rz;
      st->beta = beta;
    }
  }
  beta_out = beta;
  return !(stop || pause);
}

// k_mx_dir<PREC, PHASE> fused with the next SpMV, for L2-resident systems:
// the same decision (and, on pause or stop, the same early return), then
// q32 = A32 p_new with p_new = z + beta p_old formed at the gather; p_new of
// the block's rows is stored; (p,q) -> slot 6.  z32 holds D^-1 r (DIAG) or
// r (IDENT); p_old and p_new are ping-pong buffers.
// SELL: the product from the SELL-64 copy (256 rows per workgroup, one per
// lane), else from the CSR-stream partition.  QREC (SELL only): the product
// by recurrence, q_new = A z + beta q_old (= A p_new in exact arithmetic, as
// q_old = A p_old), so the gather reads z alone and p is never gathered; the
// rounding drift of q against A p lasts one cycle: the launch after each FP64
// residual replacement (PHASE 2) uses the gathering form, q = A p_new.
template <int PREC, int PHASE, bool SELL, bool QREC>
__global__ __launch_bounds__(kBlock) void k_mx_dirspmv(int nblocks, const int* __restrict__ rb,
                                                      const int* __restrict__ rp, ColIndex ci,
                                                      const float* __restrict__ val32,
                                                      int k_arg, int nparts,
                                                      const float* __restrict__ z32,
                                                      float* __restrict__ p0,
                                                      float* __restrict__ p1,
                                                      float* __restrict__ q32, KrylovState* st,
                                                      double* __restrict__ partials, SellF32 sa) {
  if (PHASE == 0 ? st->halt : st->done) return;
  const int k = k_arg >= 0 ? k_arg : st->kcur;  // k_arg < 0: index from the state
  const MxSnap sn = mx_snap(st, k);
  const float* __restrict__ p_old = (k & 1) ? p1 : p0;
  float* __restrict__ p_new = (k & 1) ? p0 : p1;
  __shared__ double red[4 * kBlock / kWave];
  // QREC: A z does not depend on the decision below, so the product runs
  // first, with the update's partials already loading (prefetch) when few
  float az = 0.0f, zi = 0.0f, po = 0.0f, qo = 0.0f;
  int qrow = 0;
  const bool pre = QREC && PHASE == 0 && nparts <= PartialsPrefetch<kBlock, 2>::kMax;
  PartialsPrefetch<kBlock, 2> pf;
  if constexpr (SELL && QREC) {
    const int slot[2] = {4, 5};
    if (pre) pf.load(slot, partials, nparts);
    qrow = xcd_block(blockIdx.x, nblocks) * kBlock + threadIdx.x;
    if (qrow < sa.n) {  // the row's own entries, loaded before the decision
      zi = z32[qrow];
      po = p_old[qrow];  // unused when k == 0
      qo = q32[qrow];
    }
    if (qrow / kSellChunk * kSellChunk < sa.n)  // whole waves past the end skip it
      az = sell_rowsum_g<float, 4>(qrow, sa.cptr, sa.clen, sa.col, sa.col16, sa.val,
                                   SellGather<float>{x_rsrc(z32, sa.n)});
  }
  double beta;
  if (!mx_decide<PHASE, PREC>(k, nparts, pre, pf, partials, st, sn, red, beta)) return;
  double v[1] = {0.0};
  if constexpr (SELL && QREC) {
    const float bf = (float)beta;
    const int row = qrow;
    if (row < sa.n) {
      const float pi = k == 0 ? zi : zi + bf * po;
      const float qi = k == 0 ? az : az + bf * qo;
      p_new[row] = pi;
      q32[row] = qi;
      v[0] = (double)pi * (double)qi;
    }
  } else if constexpr (SELL) {
    const int n = sa.n;
    const float bf = (float)beta;
    const SellGatherDirF g{x_rsrc(z32, n), x_rsrc(p_old, n), bf, k == 0};
    const int row = xcd_block(blockIdx.x, nblocks) * kBlock + threadIdx.x;
    if (row / kSellChunk * kSellChunk < n) {  // whole waves past the end only reduce
      const float sq =
          sell_rowsum_g<float, 4>(row, sa.cptr, sa.clen, sa.col, sa.col16, sa.val, g);
      if (row < n) {
        const float pi = k == 0 ? z32[row] : z32[row] + bf * p_old[row];
        p_new[row] = pi;
        q32[row] = sq;
        v[0] = (double)pi * (double)sq;
      }
    }
  } else {
    __shared__ float lds[kTile32];
    const int blk = xcd_block(blockIdx.x, nblocks);
    const int r0 = rb[blk], r1 = rb[blk + 1];
    const GatherDirF g{z32, p_old, (float)beta, k == 0};
    const float sq = csr_stream_rowsum_g<float, kTile32>(r0, r1, rp, ci, val32, g, lds, 0.0f);
    const int i = r0 + threadIdx.x;
    if (i < r1) {
      const float pi = g(i);
      p_new[i] = pi;
      q32[i] = sq;
      v[0] = (double)pi * (double)sq;
    }
  }
  block_sum_k<kBlock, 1>(v, red);
  const int slot[1] = {6};
  write_partials<1>(v, slot, partials, blockIdx.x);
}

// k_mx_dirspmv from the diagonal copy (W floats per row): the same decision,
// direction and (p,q) partials.  QREC: q_new = A z + beta q_old, the product
// formed before the decision with the update's partials loading beside it;
// otherwise (the launch after a replacement) q = A p_new with p_new formed at
// the gathers.  The row sums are those of the CSR rows bit for bit (SYM: from
// the symmetric form of the copy).  A lane owns R rows, kBlock apart
// (R * kBlock rows per workgroup, one partial each).
template <int PREC, int PHASE, bool QREC, int W, int R, bool SYM>
__global__ __launch_bounds__(kBlock) void k_mx_dirspmv_dia(int nblocks, DiaF32 da,
                                                          DiaSymRows<float> ds, int k_arg,
                                                          int nparts,
                                                          const float* __restrict__ z32,
                                                          float* __restrict__ p0,
                                                          float* __restrict__ p1,
                                                          float* __restrict__ q32, KrylovState* st,
                                                          double* __restrict__ partials) {
  if (PHASE == 0 ? st->halt : st->done) return;
  const int k = k_arg >= 0 ? k_arg : st->kcur;  // k_arg < 0: index from the state
  const MxSnap sn = mx_snap(st, k);
  const float* __restrict__ p_old = (k & 1) ? p1 : p0;
  float* __restrict__ p_new = (k & 1) ? p0 : p1;
  __shared__ double red[4 * kBlock / kWave];
  const int n = da.n;
  const int row0 = xcd_block(blockIdx.x, nblocks) * (kBlock * R) + threadIdx.x;
  const bool pre = QREC && PHASE == 0 && nparts <= PartialsPrefetch<kBlock, 2>::kMax;
  PartialsPrefetch<kBlock, 2> pf;
  // every load of the rows is issued before the decision: the partials, the
  // rows' own z, p_old and q_old, the matrix rows and their gathers
  float az[R], zi[R], po[R], qo[R];
  if constexpr (QREC) {
    const int slot[2] = {4, 5};
    if (pre) pf.load(slot, partials, nparts);
  }
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int row = row0 + j * kBlock;
    az[j] = zi[j] = po[j] = qo[j] = 0.0f;
    if (row < n) {
      zi[j] = z32[row];
      po[j] = p_old[row];  // unused when k == 0
      if constexpr (QREC) {
        qo[j] = q32[row];
        az[j] = dia_row<W, SYM>(row, da, ds, SellGather<float>{x_rsrc(z32, n)});
      }
    }
  }
  double beta;
  if (!mx_decide<PHASE, PREC>(k, nparts, pre, pf, partials, st, sn, red, beta)) return;
  const float bf = (float)beta;
  double v[1] = {0.0};
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int row = row0 + j * kBlock;
    if (row < n) {
      const float pi = k == 0 ? zi[j] : zi[j] + bf * po[j];
      float qi;
      if constexpr (QREC)
        qi = k == 0 ? az[j] : az[j] + bf * qo[j];
      else
        qi = dia_row<W, SYM>(row, da, ds,
                             SellGatherDirF{x_rsrc(z32, n), x_rsrc(p_old, n), bf, k == 0});
      p_new[row] = pi;
      q32[row] = qi;
      v[0] += (double)pi * (double)qi;
    }
  }
  block_sum_k<kBlock, 1>(v, red);
  const int slot[1] = {6};
  write_partials<1>(v, slot, partials, blockIdx.x);
}

// alpha = rho/(p,q); x32 += alpha p; r32 -= alpha q; (r,r) -> 4, (r,z) -> 5
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_mx_upd(int n, int k_arg, int nparts_pq,
                                                  float* __restrict__ x32, float* __restrict__ r32,
                                                  const float* __restrict__ p0,
                                                  const float* __restrict__ p1,
                                                  const float* __restrict__ q32,
                                                  const float* __restrict__ dinv32,
                                                  float* __restrict__ z32, KrylovState* st,
                                                  double* __restrict__ partials) {
  if (st->halt) return;
  // k_arg < 0: the iteration after the one the last decision recorded
  const int k = k_arg >= 0 ? k_arg : st->it + 1;
  const double rho_prev = st->rho_buf[(k - 1) & 1];  // loaded with the vectors
  const float* __restrict__ p32 = (k & 1) ? p1 : p0;
  __shared__ double red[2 * kBlock / kWave];
  // the (p,q) partials load first: their fold (and its barriers) runs while
  // this thread's vector loads are in flight
  const int sl[1] = {6};
  PartialsPrefetch<kBlock, 1, 14> pf;
  const bool pre = nparts_pq <= pf.kMax;
  if (pre) pf.load(sl, partials, nparts_pq);
  const int base = blockIdx.x * (kBlock * kVec) + threadIdx.x;
  float xv[kVec], pv[kVec], rv[kVec], qv[kVec], dv[kVec];
#pragma unroll
  for (int j = 0; j < kVec; ++j) {
    const int i = base + j * kBlock;
    if (i < n) {
      xv[j] = x32[i];
      pv[j] = p32[i];
      rv[j] = r32[i];
      qv[j] = q32[i];
      dv[j] = PREC == OFH_DIAG_PREC ? dinv32[i] : 1.0f;  // (unused by the sweeps)
    }
  }
  double pq[1];
  if (pre)
    pf.fold(pq, nparts_pq, red);
  else
    fold_partials<kBlock, 1>(pq, sl, partials, nparts_pq, red);
  const double alpha = rho_prev / pq[0];
  const float af = (float)alpha;
  double v[2] = {0.0, 0.0};
#pragma unroll
  for (int j = 0; j < kVec; ++j) {
    const int i = base + j * kBlock;
    if (i < n) {
      x32[i] = xv[j] + af * pv[j];
      const float ri = rv[j] - af * qv[j];
      r32[i] = ri;
      v[0] += (double)ri * (double)ri;
      if (!sweep_prec(PREC)) {  // sweeps: (r,z) -> slot 5 comes from k_mx_rz
        const float z = PREC == OFH_DIAG_PREC ? ri * dv[j] : ri;
        if (PREC == OFH_DIAG_PREC && z32) z32[i] = z;
        v[1] += (double)ri * (double)z;
      }
    }
  }
  block_sum_k<kBlock, 2>(v, red);
  const int slot[2] = {4, 5};
  write_partials<2>(v, slot, partials, blockIdx.x);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->alpha = alpha;
    st->kcur = k;
  }
}

// Sweep preconditioners, PHASE 0: the pause decision of k_mx_dir<.., 0> on
// the recursive residual (slot 4), taken before the sweeps so that a pause
// skips them; the same fold, the same numbers, the same state as k_mx_dir
// records when it pauses (k_mx_dir then returns at once on halt).
__global__ __launch_bounds__(kBlock) void k_mx_check(int k_arg, int nparts, KrylovState* st,
                                                    const double* __restrict__ partials) {
  if (st->halt) return;
  const int k = k_arg >= 0 ? k_arg : st->kcur;
  __shared__ double red[kBlock / kWave];
  double t[1];
  const int slot[1] = {4};
  fold_partials<kBlock, 1>(t, slot, partials, nparts, red);
  const double res = sqrt(t[0]) / st->nrm;
  const bool pause = res <= kDelta * st->rmax || res <= st->tol || k >= st->max_it || !(res == res);
  if (pause && threadIdx.x == 0) {
    st->it = k;
    st->res = res;
    st->pause = 1;
    st->halt = 1;
  }
}

// r64 = r32 (the FP64 input of the sweeps), unless *skip
__global__ void k_mx_to64(int n, const int* __restrict__ skip, const float* __restrict__ r32,
                          double* __restrict__ r64) {
  if (*skip) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r64[i] = (double)r32[i];
}

// z32 = z64 (the sweeps' result, rounded) and the partial (r32, z32) -> slot,
// unless *skip
__global__ __launch_bounds__(kBlock) void k_mx_rz(int n, const int* __restrict__ skip,
                                                 const double* __restrict__ z64,
                                                 const float* __restrict__ r32,
                                                 float* __restrict__ z32, int slot_id,
                                                 double* __restrict__ partials) {
  if (*skip) return;
  __shared__ double red[kBlock / kWave];
  double v[1] = {0.0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const float z = (float)z64[i];
    z32[i] = z;
    v[0] += (double)r32[i] * (double)z;
  }
  block_sum_k<kBlock, 1>(v, red);
  const int slot[1] = {slot_id};
  write_partials<1>(v, slot, partials, blockIdx.x);
}

// ---------------------------------------------------------------------------
// One kernel per inner iteration, for z = r (IDENT, or DIAG as the
// iteration on D^-1/2 A D^-1/2) on a diagonal copy whose rows and vectors
// stay in the L2 caches: the single-reduction CG of Chronopoulos & Gear
// (1989).  Launch j holds r_j, w_j = A r_j, s_{j-1} = A p_{j-1}, p_{j-1},
// x_j and the dots (r_j,r_j), (w_j,r_j), (r_j,s_{j-1}), (p_{j-1},s_{j-1})
// of the launch before it, and
//   takes k_mx_dirspmv's PHASE 0 decision on r_j (pause, or go on);
//   beta = (r_j,r_j) / (r_{j-1},r_{j-1}),  p_j = r_j + beta p_{j-1},
//   s_j = w_j + beta s_{j-1} (= A p_j);
//   (p_j,s_j) = (w_j,r_j) + 2 beta (r_j,s_{j-1}) + beta^2 (p_{j-1},s_{j-1})
//   (A symmetric; formed from three dots, FP32 directions stay conjugate,
//   where the two-term recurrence (w,r) - beta (r,r) / alpha loses them);
//   alpha = (r_j,r_j) / (p_j,s_j),  x_{j+1} = x_j + alpha p_j,
//   r_{j+1} = r_j - alpha s_j,  w_{j+1} = A r_{j+1}: each neighbour's r_{j+1}
//   is formed at the gathers from its r_j, w_j and s_{j-1};
//   the four dots of launch j + 1.
// r, w and s alternate between two buffers (launch j reads set j & 1 and
// writes the other), the dots between partial slots 0-3 and 4-7 (launch j
// reads slots 4 (j & 1)).  x and p are updated in place, row by row.
// A lane owns R rows, kBlock apart.  The matrix comes from the full diagonal
// copy even where a symmetric form exists: this kernel is latency-bound at
// the sizes it runs at, and the symmetric form's partner loads made it
// slower (512^2 heat matrix, mixed CG from zero: 7.0 vs 6.7 ms).

// The new direction's product and residual of one row from its r_j, w_j and
// s_{j-1}: the same arithmetic for a row's own values and for the copies its
// neighbours form at their gathers.
struct Cg1Step {
  float alpha, beta;
  bool first;  // j == 0: p_0 = r_0, s_0 = w_0
  __device__ __forceinline__ float s(float w, float s_old) const {
    return first ? w : w + beta * s_old;
  }
  __device__ __forceinline__ float r(float r_old, float s_new) const {
    return r_old - alpha * s_new;
  }
};

template <int PREC, int W, int R>
__global__ __launch_bounds__(kBlock) void k_mx_cg1(int nblocks, DiaF32 da, int j, int nparts,
                                                  const float* __restrict__ r_in,
                                                  const float* __restrict__ w_in,
                                                  const float* __restrict__ s_in,
                                                  float* __restrict__ r_out,
                                                  float* __restrict__ w_out,
                                                  float* __restrict__ s_out,
                                                  float* __restrict__ x32,
                                                  float* __restrict__ p32, KrylovState* st,
                                                  double* __restrict__ partials) {
  if (st->halt) return;
  const MxSnap sn = mx_snap(st, j);
  __shared__ double red[4 * kBlock / kWave];
  const int n = da.n;
  const int row0 = xcd_block(blockIdx.x, nblocks) * (kBlock * R) + threadIdx.x;
  const int sb = 4 * (j & 1);
  const int slot[4] = {sb, sb + 1, sb + 2, sb + 3};
  PartialsPrefetch<kBlock, 4> pf;
  const bool pre = nparts <= pf.kMax;
  if (pre) pf.load(slot, partials, nparts);
  // every load is issued before the decision: the dots, the rows' own
  // values, the matrix rows and the neighbours' r, w, s
  const XRsrc rr = x_rsrc(r_in, n), wr = x_rsrc(w_in, n), sr = x_rsrc(s_in, n);
  float a[R][W], rg[R][W - 1], wg[R][W - 1], sg[R][W - 1];
  float ri[R], wi[R], si[R], xi[R], pi[R];
#pragma unroll
  for (int u = 0; u < R; ++u) {
    const int row = row0 + u * kBlock;
    ri[u] = wi[u] = si[u] = xi[u] = pi[u] = 0.0f;
#pragma unroll
    for (int d = 0; d < W; ++d) a[u][d] = 0.0f;
#pragma unroll
    for (int d = 0; d < W - 1; ++d) rg[u][d] = wg[u][d] = sg[u][d] = 0.0f;
    if (row < n) {
      ri[u] = r_in[row];
      wi[u] = w_in[row];
      si[u] = s_in[row];  // unused when j == 0
      xi[u] = x32[row];
      pi[u] = p32[row];   // unused when j == 0
      const float* __restrict__ base =
          da.v + (size_t)(row / kDiaChunk) * (W * kDiaChunk) + (row % kDiaChunk);
#pragma unroll
      for (int d = 0; d < W; ++d) a[u][d] = base[d * kDiaChunk];
#pragma unroll
      for (int d = 0; d < W - 1; ++d) {
        const uint32_t c = (uint32_t)(row + da.off[d]);
        rg[u][d] = x_load<float>(rr, c);
        wg[u][d] = x_load<float>(wr, c);
        sg[u][d] = x_load<float>(sr, c);
      }
    }
  }
  double t[4];
  if (pre)
    pf.fold(t, nparts, red);
  else
    fold_partials<kBlock, 4>(t, slot, partials, nparts, red);
  const double res = PREC == kScaledDiag ? sqrt(t[0]) * sn.rscale : sqrt(t[0]) / sn.nrm;
  const bool pause =
      (res <= kDelta * sn.rmax) | (res <= sn.tol) | (j >= sn.max_it) | !(res == res);
  const double beta = j > 0 ? t[0] / sn.rho_prev : 0.0;
  const double pap = j > 0 ? t[1] + beta * (2.0 * t[2] + beta * t[3]) : t[1];
  const double alpha = t[0] / pap;
  // st->it counts the steps taken: j on a pause (x_j is the iterate the
  // replacement accumulates), j + 1 once this launch has formed x_{j+1}
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->res = res;
    if (pause) {
      st->it = j;
      st->pause = 1;
      st->halt = 1;
    } else {
      st->it = j + 1;
      st->rho_buf[j & 1] = t[0];
      st->beta = beta;
      st->alpha = alpha;
    }
  }
  if (pause) return;
  const Cg1Step f{(float)alpha, (float)beta, j == 0};
  double v[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int u = 0; u < R; ++u) {
    const int row = row0 + u * kBlock;
    if (row < n) {
      const uint32_t m = __builtin_bit_cast(uint32_t, a[u][W - 1]);
      float w = 0.0f;  // A r_{j+1}, the present slots left to right
#pragma unroll
      for (int d = 0; d < W - 1; ++d) {
        const float rn = f.r(rg[u][d], f.s(wg[u][d], sg[u][d]));
        const float t1 = w + a[u][d] * rn;
        w = (m >> d & 1u) ? t1 : w;
      }
      const float sj = f.s(wi[u], si[u]);
      const float rn = f.r(ri[u], sj);
      const float pj = f.first ? ri[u] : ri[u] + f.beta * pi[u];
      x32[row] = xi[u] + f.alpha * pj;
      p32[row] = pj;
      s_out[row] = sj;
      r_out[row] = rn;
      w_out[row] = w;
      v[0] += (double)rn * (double)rn;
      v[1] += (double)w * (double)rn;
      v[2] += (double)rn * (double)sj;
      v[3] += (double)pj * (double)sj;
    }
  }
  block_sum_k<kBlock, 4>(v, red);
  const int wslot[4] = {sb ^ 4, (sb ^ 4) + 1, (sb ^ 4) + 2, (sb ^ 4) + 3};
  write_partials<4>(v, wslot, partials, blockIdx.x);
}

// The launch after k_mx_resid, ahead of launch j of k_mx_cg1 (PHASE 1: the
// initial residual, j = 0; PHASE 2: a replacement after j inner
// iterations): k_mx_dirspmv's PHASE 1/2 decision on the FP64 residual (the
// partials k_mx_resid wrote to slots 4 (j & 1) ^ 4), then w = A r and, in
// PHASE 2, s = A p_{j-1} from the diagonal copy (the product re-formed from
// p once per cycle, so the rounding drift of the recurrence-carried s does
// not outlive a cycle), and the four dots launch j folds (slots 4 (j & 1)).
// r, w and s are the buffers of set j & 1.
template <int PREC, int PHASE, int W, int R>
__global__ __launch_bounds__(kBlock) void k_mx_cg1_start(int nblocks, DiaF32 da, int j, int nparts,
                                                        const float* __restrict__ r,
                                                        const float* __restrict__ p32,
                                                        float* __restrict__ w,
                                                        float* __restrict__ s, KrylovState* st,
                                                        double* __restrict__ partials) {
  if (st->done) return;
  const MxSnap sn = mx_snap(st, j);
  __shared__ double red[4 * kBlock / kWave];
  const int n = da.n;
  const int row0 = xcd_block(blockIdx.x, nblocks) * (kBlock * R) + threadIdx.x;
  const bool gp = PHASE == 2 && j > 0;  // p_{j-1} exists
  float ri[R], pi[R], wr[R], sp[R];
#pragma unroll
  for (int u = 0; u < R; ++u) {
    const int row = row0 + u * kBlock;
    ri[u] = pi[u] = wr[u] = sp[u] = 0.0f;
    if (row < n) {
      ri[u] = r[row];
      wr[u] = dia_rowsum_g<W>(row, da, SellGather<float>{x_rsrc(r, n)});
      if (gp) {
        pi[u] = p32[row];
        sp[u] = dia_rowsum_g<W>(row, da, SellGather<float>{x_rsrc(p32, n)});
      }
    }
  }
  const int sb = 4 * (j & 1);
  double beta;
  PartialsPrefetch<kBlock, 2> pf;  // not used: PHASE 1, 2 fold from memory
  if (!mx_decide<PHASE, PREC>(j, nparts, false, pf, partials, st, sn, red, beta, sb ^ 4)) return;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int u = 0; u < R; ++u) {
    const int row = row0 + u * kBlock;
    if (row < n) {
      w[row] = wr[u];
      if (gp) s[row] = sp[u];
      v[0] += (double)ri[u] * (double)ri[u];
      v[1] += (double)wr[u] * (double)ri[u];
      if (gp) {
        v[2] += (double)ri[u] * (double)sp[u];
        v[3] += (double)pi[u] * (double)sp[u];
      }
    }
  }
  block_sum_k<kBlock, 4>(v, red);
  const int slot[4] = {sb, sb + 1, sb + 2, sb + 3};
  write_partials<4>(v, slot, partials, blockIdx.x);
}

// x64 += x32 (the correction accumulated since the last replacement), or
// x64 += D^-1/2 x32 on the scaled system (sc = D^-1/2)
__global__ void k_mx_accum(int n, double* __restrict__ x, const float* __restrict__ x32,
                           const double* __restrict__ sc, const KrylovState* __restrict__ st) {
  if (st->done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += sc ? sc[i] * (double)x32[i] : (double)x32[i];
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
  if (sweep_prec(PREC) && (rc = csr_ensure_prec(A, PREC))) return rc;
  if ((rc = csr_ensure_val32(A))) return rc;
  // Two kernels per inner iteration (k_mx_upd, k_mx_dirspmv) when the
  // product reads the SELL-64 copy (q by recurrence: one gather, the
  // update's partials loading during the product) or the system is
  // L2-resident (direction formed at the gathers); otherwise three
  // (k_mx_spmv, k_mx_upd, k_mx_dir).
  constexpr int kFuseRows = 3 << 17;
  // The sweep preconditioners (SSOR, DILU) take the three-kernel form with
  // z = M^-1 r between the update and the direction: k_mx_check (pause
  // decision), the FP64 level-scheduled sweeps of the FP32 residual, k_mx_rz.
  // The product reads the diagonal copy when the pattern has one, else the
  // SELL-64 copy when its padding is small.
  bool fdia = false, fsell = false;
  const int nbsf = (n + kBlock - 1) / kBlock;
  if (!sweep_prec(PREC) && nbsf <= kMaxPartials && (rc = csr_ensure_dia(A, &fdia))) return rc;
  if (!sweep_prec(PREC) && !fdia && nbsf <= kMaxPartials &&
      (rc = csr_ensure_sell(A, true, &fsell)))
    return rc;
  const bool qrec = fsell || fdia;
  const bool fused = !sweep_prec(PREC) && (n <= kFuseRows || qrec);
  // One kernel per inner iteration (k_mx_cg1) where z = r and A is
  // symmetric bit for bit (its (p, Ap) comes from three dot products through
  // (p_old, A r) = (A p_old, r)), up to the size where its gathers of three
  // vectors per neighbour stop hitting the L2 caches.  Measured per inner
  // iteration, two kernels vs one (mixed CG+DIAG from zero): 16 slots per row
  // (3D P1), 250,047 rows 15.4 vs 12.4 us, 357,911 rows 16.6 vs 16.9 us,
  // 493,039 rows 19.1 vs 20.8 us; 8 slots (2D P1), 261,121 rows 12.9 vs
  // 9.3 us, 522,729 rows 14.6 vs 10.4 us, 808,201 rows 18.7 vs 18.5 us.
  // ofh_ctx_mixed_form overrides the choice.
  const int one_rows = A->dia_w == 16 ? 300000 : 650000;
  bool symA = false;
  if (fdia && (PREC == kScaledDiag || PREC == OFH_IDENT_PREC) && ctx->mixed_form != 1 &&
      (rc = csr_dia_symmetric(A, &symA)))
    return rc;
  const bool one = fdia && symA && (PREC == kScaledDiag || PREC == OFH_IDENT_PREC) &&
                   ctx->mixed_form != 1 && (ctx->mixed_form == 2 || n <= one_rows);
  void* wsp;
  const size_t vbytes = sizeof(float) * (size_t)n * (one ? 8 : 6) +
                        (sweep_prec(PREC) ? sizeof(double) * 2 * (size_t)n : 0);
  if ((rc = ensure_vws(ctx, vbytes, &wsp))) return rc;
  float* r32 = (float*)wsp;
  float* p32 = r32 + n;
  float* q32 = p32 + n;
  float* x32 = q32 + n;
  float* p32b = x32 + n;                                     // second direction buffer
  // z: D^-1 r (fused DIAG form), the sweeps' result (SSOR, DILU), else r
  float* z32 = (PREC == OFH_DIAG_PREC || sweep_prec(PREC)) ? p32b + n : r32;
  double* r64 = sweep_prec(PREC) ? (double*)((char*)wsp + sizeof(float) * (size_t)n * 6) : nullptr;
  double* z64 = r64 ? r64 + n : nullptr;
  // what k_mx_dir / k_mx_resid read as `dinv32`: the pivots, or z for the sweeps
  const float* dz32 = sweep_prec(PREC) ? z32 : A->dinv32;
  float* pb[2] = {p32, p32b};
  // the one-kernel form's alternating r, w and s (z = r; x32 and p32 as above)
  float* const ext = r32 + 5 * (size_t)n;
  float* const r1[2] = {r32, ext};
  float* const w1[2] = {q32, ext + n};
  float* const s1[2] = {p32b, ext + 2 * (size_t)n};
  float* zw = fused && PREC == OFH_DIAG_PREC ? z32 : nullptr;  // z written by resid/upd
  const ColIndex ci{A->col, nullptr};  // FP32 partition: 32-bit columns
  const SellF32 sa{n, A->sell_cptr, A->sell_clen, A->sell_col, A->sell_col16, A->sell_val32};
  // rows per lane of the diagonal product (96^3 mixed solve: 1 row 15.6 ms,
  // 2 rows 14.7 ms, 4 rows 14.9 ms), and its workgroups (= partials)
  constexpr int kDiaRows = 2;
  const int nbd = (n + kBlock * kDiaRows - 1) / (kBlock * kDiaRows);
  const int npq_fused = fdia ? nbd : qrec ? nbsf : A->nblocks32;  // partials of the fused product
  DiaF32 da = fdia ? dia_f32(A) : DiaF32{};
  // the scaled system's copy and D^-1/2 (kScaledDiag: the dispatch checked both)
  const double* sc = PREC == kScaledDiag ? A->dscale : nullptr;
  if (PREC == kScaledDiag) {
    if (!fdia || !sc) return fail(OFH_E_INVALID, "ofh_cg_mixed: no scaled diagonal copy");
    da.v = A->dia_val32s;
  }
  // the symmetric form of that copy, for the two-kernel product
  bool fsym = false;
  if (fdia && !one && (rc = csr_ensure_dias(A, PREC == kScaledDiag ? 2 : 1, &fsym))) return rc;
  const DiaSymRows<float> dsy = fsym ? dias_f32(A, PREC == kScaledDiag) : DiaSymRows<float>{};
  // pk: the profiling kind of this launch (-1: not timed)
  auto dirspmv = [&](int phase, int k, int nparts, hipStream_t st_, int pk = -1) {
    if constexpr (!sweep_prec(PREC)) {
      if (fdia) {
#define OFH_MX_DIA_S(PH, QR, W, SY)                                                             \
  prof_kernel(ctx, pk, k_mx_dirspmv_dia<PREC, PH, QR, W, kDiaRows, SY>, dim3(nbd), dim3(kBlock),  \
              st_, nbd, da, dsy, k, nparts, z32, pb[0], pb[1], q32, state_ptr(ctx),             \
              ctx->ws.partials)
#define OFH_MX_DIA(PH, QR, W) \
  fsym ? OFH_MX_DIA_S(PH, QR, W, true) : OFH_MX_DIA_S(PH, QR, W, false)
#define OFH_MX_DIA_W(W)                       \
  if (phase == 0) OFH_MX_DIA(0, true, W);     \
  else if (phase == 1) OFH_MX_DIA(1, true, W); \
  else OFH_MX_DIA(2, false, W);
        if (A->dia_w == 4) {
          OFH_MX_DIA_W(4)
        } else if (A->dia_w == 8) {
          OFH_MX_DIA_W(8)
        } else {
          OFH_MX_DIA_W(16)
        }
#undef OFH_MX_DIA_W
#undef OFH_MX_DIA
#undef OFH_MX_DIA_S
        return;
      }
    }
    const int nbf = fsell ? nbsf : A->nblocks32;
#define OFH_MX_DIRSPMV(PH, SL, QR)                                                              \
  prof_kernel(ctx, pk, k_mx_dirspmv<PREC, PH, SL, QR>, dim3(nbf), dim3(kBlock), st_, nbf,        \
              A->row_blocks32, A->row_ptr, ci, A->val32, k, nparts, z32, pb[0], pb[1], q32,      \
              state_ptr(ctx), ctx->ws.partials, sa)
    if (qrec) {
      // After a residual replacement (phase 2) the product is formed from
      // the gathered direction, q = A (z + beta p_old): this resets the drift
      // of the recurrence-carried q against A p once per cycle.
      if (phase == 0) OFH_MX_DIRSPMV(0, true, true);
      else if (phase == 1) OFH_MX_DIRSPMV(1, true, true);
      else OFH_MX_DIRSPMV(2, true, false);
    } else if (fsell) {
      if (phase == 0) OFH_MX_DIRSPMV(0, true, false);
      else if (phase == 1) OFH_MX_DIRSPMV(1, true, false);
      else OFH_MX_DIRSPMV(2, true, false);
    } else {
      if (phase == 0) OFH_MX_DIRSPMV(0, false, false);
      else if (phase == 1) OFH_MX_DIRSPMV(1, false, false);
      else OFH_MX_DIRSPMV(2, false, false);
    }
#undef OFH_MX_DIRSPMV
  };
  KrylovState* st = state_ptr(ctx);
  KrylovState h0{};
  h0.tol = toler;
  h0.max_it = max_it;
  OFH_HIP(hipMemcpyAsync(st, &h0, sizeof(h0), hipMemcpyHostToDevice, s));
  double* part = ctx->ws.partials;
  const int g = reduce_grid(n);
  const int nb32 = A->nblocks32;  // FP32 partition (inner SpMV)
  const ColIndex ci32{A->col, nullptr};  // 16-bit offsets follow the FP64 partition
  const int ge = (n + 255) / 256;
  // the unfused inner SpMV from the SELL-64 copy when it is in use
  bool sell = false;
  const int nbs = (n + kBlock - 1) / kBlock;
  if (!fused && nbs <= kMaxPartials && (rc = csr_ensure_sell(A, true, &sell))) return rc;
  // the FP64 residual b - A x, from the SELL-64 copy when it is in use
  MatRows64 mr;
  bool rsell = false;
  if ((rc = mat_rows64(A, &mr, &rsell))) return rc;
  const int nbr = mr.nblocks;  // its workgroups (= partials)
  // r = b - A x into rd, its partials to slots sb..sb+3
  auto resid = [&](float* rd, int sb) {
    if (rsell)
      hipLaunchKernelGGL((k_mx_resid<PREC, true>), dim3(nbr), dim3(kBlock), 0, s, mr, b, x, rd,
                         x32, A->dinv32, zw, sc, st, part, sb);
    else
      hipLaunchKernelGGL((k_mx_resid<PREC, false>), dim3(nbr), dim3(kBlock), 0, s, mr, b, x, rd,
                         x32, A->dinv32, zw, sc, st, part, sb);
  };
  // the one-kernel form: launch j of k_mx_cg1, and k_mx_cg1_start ahead of
  // launch j (rows per lane: 4, 2 with 16 slots per row)
  const int nb1 = (n + kBlock * (A->dia_w == 16 ? 2 : 4) - 1) / (kBlock * (A->dia_w == 16 ? 2 : 4));
  auto cg1 = [&](int j, hipStream_t st_) {
    if constexpr (PREC == kScaledDiag || PREC == OFH_IDENT_PREC) {
      const int u = j & 1;
#define OFH_CG1(W, R)                                                                            \
  prof_kernel(ctx, OFH_PROF_MX_ITER, k_mx_cg1<PREC, W, R>, dim3(nb1), dim3(kBlock), st_, nb1, da, \
              j, nb1, r1[u], w1[u], s1[u], r1[u ^ 1], w1[u ^ 1], s1[u ^ 1], x32, p32, st, part)
      if (A->dia_w == 4)
        OFH_CG1(4, 4);
      else if (A->dia_w == 8)
        OFH_CG1(8, 4);
      else
        OFH_CG1(16, 2);
#undef OFH_CG1
    }
  };
  auto cg1_start = [&](int phase, int j) {
    if constexpr (PREC == kScaledDiag || PREC == OFH_IDENT_PREC) {
      const int u = j & 1;
#define OFH_CG1S(PH, W, R)                                                                   \
  hipLaunchKernelGGL((k_mx_cg1_start<PREC, PH, W, R>), dim3(nb1), dim3(kBlock), 0, s, nb1, da, j, \
                     nbr, r1[u], p32, w1[u], s1[u], st, part)
#define OFH_CG1S_W(PH)        \
  if (A->dia_w == 4)          \
    OFH_CG1S(PH, 4, 4);       \
  else if (A->dia_w == 8)     \
    OFH_CG1S(PH, 8, 4);       \
  else                        \
    OFH_CG1S(PH, 16, 2);
      if (phase == 1) {
        OFH_CG1S_W(1)
      } else {
        OFH_CG1S_W(2)
      }
#undef OFH_CG1S_W
#undef OFH_CG1S
    }
  };
  // z32 = M^-1 r32 by the FP64 sweeps and (r32, z32) -> slot (sweep precs)
  auto sweep_z = [&](const int* skip, int nparts, int slot_id, hipStream_t st_) -> int {
    hipLaunchKernelGGL(k_mx_to64, dim3(ge), dim3(256), 0, st_, n, skip, r32, r64);
    const int e = launch_prec(A, PREC, r64, z64, skip, st_);
    if (e) return e;
    hipLaunchKernelGGL(k_mx_rz, dim3(nparts), dim3(kBlock), 0, st_, n, skip, z64, r32, z32,
                       slot_id, part);
    return OFH_OK;
  };
  if (one) {  // the residual into set 0, its partials to the slots launch 0 does not read
    resid(r1[0], 4);
    cg1_start(1, 0);
  } else {
    resid(r32, 0);
    if (sweep_prec(PREC) && (rc = sweep_z(&st->done, nbr, 2, s))) return rc;
    if (fused)
      dirspmv(1, 0, nbr, s);
    else
      hipLaunchKernelGGL((k_mx_dir<PREC, 1>), dim3(g), dim3(kBlock), 0, s, n, 0, nbr, r32,
                         dz32, p32, st, part);
  }
  OFH_HIP(hipGetLastError());
  // one inner iteration; k = -1 takes the index from the state (captured chunks)
  auto iteration = [&](int k, hipStream_t st_) -> int {
    if (one) {  // launch k - 1 decides on r_{k-1} and forms r_k, x_k (state it = k)
      cg1(k - 1, st_);
      return OFH_OK;
    }
    if (fused) {
      prof_kernel(ctx, OFH_PROF_MX_UPD, k_mx_upd<PREC>, dim3(g), dim3(kBlock), st_, n, k,
                  npq_fused, x32, r32, pb[0], pb[1], q32, A->dinv32, zw, st, part);
      dirspmv(0, k, g, st_, OFH_PROF_MX_DIRSPMV);
      return OFH_OK;
    }
    if (sell)
      hipLaunchKernelGGL(k_mx_spmv_sell, dim3(nbs), dim3(kBlock), 0, st_, n, nbs, A->sell_cptr,
                         A->sell_clen, A->sell_col, A->sell_col16, A->sell_val32, p32, q32, st,
                         part);
    else
      hipLaunchKernelGGL(k_mx_spmv, dim3(nb32), dim3(kBlock), 0, st_, nb32, A->row_blocks32,
                         A->row_ptr, ci32, A->val32, p32, q32, st, part);
    hipLaunchKernelGGL(k_mx_upd<PREC>, dim3(g), dim3(kBlock), 0, st_, n, k, sell ? nbs : nb32, x32,
                       r32, p32, p32, q32, A->dinv32, nullptr, st, part);
    if (sweep_prec(PREC)) {
      hipLaunchKernelGGL(k_mx_check, dim3(1), dim3(kBlock), 0, st_, k, g, st, part);
      const int e = sweep_z(&st->halt, g, 5, st_);
      if (e) return e;
    }
    hipLaunchKernelGGL((k_mx_dir<PREC, 0>), dim3(g), dim3(kBlock), 0, st_, n, k, g, r32,
                       dz32, p32, st, part);
    return OFH_OK;
  };
  hipGraphExec_t chunk_graph = nullptr;
  if (ctx->use_graphs && !sweep_prec(PREC) && !one) {
    GraphKey key;
    csr_graph_key(A, 16 * (fused ? 3 : 4) + (PREC == kScaledDiag ? 15 : PREC),
                  (uintptr_t)fsell | (uintptr_t)qrec << 1 | (uintptr_t)sell << 2 |
                      (uintptr_t)fdia << 3,
                  x, wsp,
                  A->val32, PREC == kScaledDiag ? (const void*)A->dscale : A->dinv32, &key);
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
  KrylovState h{};
  if ((rc = read_state(ctx, &h))) return rc;
  double last_true = h.res;
  int stagnant = 0, slow = 0, fallback = 0, it_repl = 0, first_cycle = 0;
  int last_cycle = 0;
  double best_rate = 0.0;
  // Each look at the state is taken while kLookAhead more iterations run: the
  // state is copied after the chunk, those launches are enqueued behind
  // the copy, and the host waits for the copy only, so the GPU does not idle
  // while the host decides.  Launches past a pause return at once.
  int ahead = 0;  // iterations enqueued after the copy h was taken
  while (!h.done) {
    int k = h.it + ahead;
    const double target = std::max(kDelta * h.rmax, toler);
    const int chunk = std::max(
        1, mixed_chunk(h.it, h.res, it_repl, h.rmax, target, max_it - h.it, last_cycle) - ahead);
    // whole captured chunks, then the remainder launched directly
    const int full = chunk_graph ? chunk / kGraphIters : 0;
    for (int c = 0; c < full; ++c) OFH_HIP(hipGraphLaunch(chunk_graph, s));
    k += full * kGraphIters;
    for (int c = full * kGraphIters; c < chunk; ++c)
      if ((rc = iteration(++k, s))) return rc;
    OFH_HIP(hipGetLastError());
    ahead = kLookAhead;  // past max_it a launch pauses (the first) or returns
    if ((rc = read_state_ahead(ctx, &h, [&]() -> int {
           for (int c = 0; c < ahead; ++c)
             if (const int e = iteration(++k, s)) return e;
           OFH_HIP(hipGetLastError());
           return OFH_OK;
         })))
      return rc;
    if (h.done) break;
    // A cycle far longer than the first one means FP32 can no longer lower
    // the recursive residual: replace it now (the direction p_{k+1} and
    // rho_k are already in place, so the replacement acts as iteration k+1).
    const int cap = first_cycle > 0 ? std::max(200, 3 * first_cycle) : 2000;
    if (!h.pause && ahead > 0 && h.it + ahead - it_repl >= cap) {
      // the iterations run ahead may reach the cap: the exact state decides
      if ((rc = read_state(ctx, &h))) return rc;
      ahead = 0;
      if (h.done) break;
    }
    const int cyc_now = h.it - it_repl;
    const bool forced = !h.pause && cyc_now >= cap;
    if (!h.pause && !forced) continue;
    ahead = 0;  // paused: the iterations run ahead returned at once
    if (h.pause && first_cycle == 0) first_cycle = std::max(1, cyc_now);
    // (the one-kernel form has already taken the step along p_{it-1}: x32 is
    // x_it, so its replacement acts as iteration it)
    const int kr = forced && !one ? h.it + 1 : h.it;
    last_cycle = kr - it_repl;
    // residual replacement in FP64 after kr inner iterations
    hipLaunchKernelGGL(k_mx_accum, dim3(ge), dim3(256), 0, s, n, x, x32, sc, st);
    if (one) {
      resid(r1[kr & 1], 4 * (kr & 1) ^ 4);
      cg1_start(2, kr);
    } else {
      resid(r32, 0);
      if (sweep_prec(PREC) && (rc = sweep_z(&st->done, nbr, 2, s))) return rc;
      if (fused)
        dirspmv(2, kr, nbr, s);
      else
        hipLaunchKernelGGL((k_mx_dir<PREC, 2>), dim3(g), dim3(kBlock), 0, s, n, kr, nbr, r32,
                           dz32, p32, st, part);
    }
    OFH_HIP(hipGetLastError());
    if ((rc = read_state(ctx, &h))) return rc;
    if (h.done) break;
    // stagnation: the true residual must keep falling ...
    stagnant = (h.res > 0.5 * last_true) ? stagnant + 1 : 0;
    // ... at a rate per inner iteration that does not collapse: when FP32
    // can no longer resolve the slow modes (very ill-conditioned systems)
    // the cycles grow long while the true residual barely moves
    const int cyc = std::max(1, h.it - it_repl);
    const double rate = (h.res > 0 && last_true > 0) ? std::log(last_true / h.res) / cyc : 0.0;
    best_rate = std::max(best_rate, rate);
    slow = (forced || rate < 0.2 * best_rate) ? slow + 1 : 0;
    it_repl = h.it;
    last_true = std::min(last_true, h.res);
    // a first cycle that had to be forced (FP32 never lowered the recursive
    // residual by kDelta within the cap) ends the FP32 phase at once
    if (stagnant >= 3 || slow >= 2 || (forced && first_cycle == 0)) {
      fallback = 1;
      break;
    }
  }
  if (fallback) {
    ofh_solve_info sub{};
    const int left = std::max(0, max_it - h.it);
    const int prec64 = PREC == kScaledDiag ? OFH_DIAG_PREC : PREC;
    if ((rc = ofh_cg(A, prec64, b, x, left, toler, &sub))) return rc;
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
    info->nb_it = h.status;
    info->iterations = h.it;
    info->relres = h.res;
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
    case OFH_DIAG_PREC: {
      // with a diagonal copy and positive pivots: the inner iteration on
      // D^-1/2 A D^-1/2 (no pivot read, no z written per iteration)
      bool scaled = false;
      if (A->n <= kMaxPartials * kBlock) {
        const int rc = csr_ensure_dia_scaled(A, &scaled);
        if (rc) return rc;
      }
      if (scaled) return mixed_impl<kScaledDiag>(A, b, x, max_it, toler, info);
      return mixed_impl<OFH_DIAG_PREC>(A, b, x, max_it, toler, info);
    }
    case OFH_SSOR_PREC: return mixed_impl<OFH_SSOR_PREC>(A, b, x, max_it, toler, info);
    case OFH_DILU_PREC:  // the reference's ILU is its DILU
    case OFH_ILU_PREC: return mixed_impl<OFH_DILU_PREC>(A, b, x, max_it, toler, info);
    default: return fail(OFH_E_INVALID, "ofh_cg_mixed: unknown preconditioner");
  }
}
