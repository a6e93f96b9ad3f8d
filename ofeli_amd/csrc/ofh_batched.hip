// ofh_batched.hip -- many independent systems solved in one persistent launch.
//
// No reference counterpart as a routine: the reference solves such a set
// with a loop of LinearSolver::solve calls (LinearSolver.cpp:194-244, one
// CG.h:99-157 run per system).  Here the systems are the diagonal blocks of
// one block-diagonal CSR (e.g. the assembly of several disjoint meshes in one
// plan).  A grid of persistent workgroups pulls systems from an on-device
// work queue, largest first, and each workgroup runs the complete
// preconditioned CG of its system with workgroup-level reductions only: no
// host round trip and no grid-wide synchronisation per iteration.
//   FP64 variant: the reference's CG (CG.h:99-157) per system.
//   Mixed variant: FP32 SpMV / vectors, FP64 dot accumulation, FP64 solution
//   and FP64 residual replacement (as ofh_cg_mixed), stopping on the true
//   FP64 residual.
// Launch shape (workgroup size, threads per matrix row) is a template choice;
// ofh_batch_autotune times the candidates on the actual batch and keeps the
// fastest.  TPR = 1 sums each row sequentially (bitwise the reference SpMV).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <string>
#include <vector>

#include "ofh_internal.h"
#include "ofh_solver.h"

// What a captured batched launch depends on; a change of any field recaptures.
struct BatchGraphKey {
  const void* fn;
  const double* b;
  double* x;
  int max_it;
  double tol;
  const void* dinv;
  const void* val32;
  const void* work;
  int grid;
  bool operator==(const BatchGraphKey& o) const {
    return fn == o.fn && b == o.b && x == o.x && max_it == o.max_it && tol == o.tol &&
           dinv == o.dinv && val32 == o.val32 && work == o.work && grid == o.grid;
  }
};

struct ofh_batch_s {
  ofh_csr A = nullptr;
  int nsys = 0;
  int* sys_rows = nullptr;  // device, nsys+1
  int* order = nullptr;     // device, systems by decreasing size
  int* counter = nullptr;   // device work-queue counter
  int* iters = nullptr;     // device, nsys
  int* status = nullptr;    // device, nsys
  double* relres = nullptr; // device, nsys
  void* work = nullptr;     // device work vectors
  size_t work_bytes = 0;
  int block = 256, tpr = 1;  // launch shape (tpr = 0: resident form)
  int max_rows = 0;          // largest system
  bool resident_ok = false;  // every system fits the resident form
  bool tuned = false;
  bool use_graph = true;         // replay the launch from a captured graph
  hipGraphExec_t gexec = nullptr;
  hipStream_t cap = nullptr;     // capture-only stream
  BatchGraphKey gkey{};
  std::vector<int> h_rows;
};

namespace ofh {

constexpr double kBatchDelta = 0.1;

template <int BLOCK>
__device__ __forceinline__ double bsum(double v, double* red) {
  return block_sum<BLOCK>(v, red);
}

// sum over the TPR lanes of a row group (lanes are consecutive in a wave)
template <int TPR, typename T>
__device__ __forceinline__ T row_reduce(T v) {
#pragma unroll
  for (int o = TPR / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, TPR);
  return v;
}

// y_i = sum_k a_k x_c for the rows [a,e) of one system, TPR lanes per row.
template <int BLOCK, int TPR, typename T, typename TX>
__device__ __forceinline__ void sys_spmv(int a, int e, const int* __restrict__ rp,
                                         const int* __restrict__ col, const T* __restrict__ val,
                                         const TX* __restrict__ x, TX* __restrict__ y) {
  const int lane = threadIdx.x % TPR, grp = threadIdx.x / TPR;
  constexpr int G = BLOCK / TPR;
  for (int i = a + grp; i < e; i += G) {
    T s = T(0);
    for (int k = rp[i] + lane; k < rp[i + 1]; k += TPR) s += val[k] * (T)x[col[k]];
    if (TPR > 1) s = row_reduce<TPR>(s);
    if (lane == 0) y[i] = (TX)s;
  }
}

struct BatchArgs {
  int nsys;
  const int* sys_rows;
  const int* order;
  int* counter;
  const int* rp;
  const int* col;
  const double* val;
  const float* val32;
  const double* dinv;
  const float* dinv32;
  const double* b;
  double* x;
  void* work;   // r, p, q (T each) and, mixed, x32 (float) per row
  int nrows;
  int max_it;
  double tol;
  int* iters;
  int* status;
  double* relres;
};

// FP64: CG.h:99-157 with DIAG/IDENT for each system.
template <int BLOCK, int TPR, int PREC>
__global__ __launch_bounds__(BLOCK) void k_batch_cg64(BatchArgs A) {
  __shared__ double red[BLOCK / kWave];
  __shared__ int s_w;
  double* r = (double*)A.work;
  double* p = r + A.nrows;
  double* q = p + A.nrows;
  for (;;) {
    if (threadIdx.x == 0) s_w = atomicAdd(A.counter, 1);
    __syncthreads();
    const int w = s_w;
    __syncthreads();
    if (w >= A.nsys) break;
    const int sys = A.order[w];
    const int a = A.sys_rows[sys], e = A.sys_rows[sys + 1];
    // r = b - A x
    sys_spmv<BLOCK, TPR, double, double>(a, e, A.rp, A.col, A.val, A.x, q);
    __syncthreads();
    double bb = 0, rr = 0, rz = 0;
    for (int i = a + threadIdx.x; i < e; i += BLOCK) {
      const double ri = A.b[i] - q[i];
      r[i] = ri;
      bb += A.b[i] * A.b[i];
      rr += ri * ri;
      rz += ri * (PREC == OFH_DIAG_PREC ? ri * A.dinv[i] : ri);
    }
    bb = bsum<BLOCK>(bb, red);
    rr = bsum<BLOCK>(rr, red);
    rz = bsum<BLOCK>(rz, red);
    double nrm = sqrt(bb);
    if (nrm == 0) nrm = 1;
    double res = sqrt(rr) / nrm, rho = rz, rho_1 = 1.0;
    int it = 0, st;
    if (res <= A.tol) {
      st = 0;
    } else {
      st = -(A.max_it + 1);
      for (it = 1; it <= A.max_it; ++it) {
        const double beta = it == 1 ? 0.0 : rho / rho_1;
        for (int i = a + threadIdx.x; i < e; i += BLOCK) {
          const double zi = PREC == OFH_DIAG_PREC ? r[i] * A.dinv[i] : r[i];
          p[i] = it == 1 ? zi : zi + beta * p[i];
        }
        __syncthreads();
        sys_spmv<BLOCK, TPR, double, double>(a, e, A.rp, A.col, A.val, p, q);
        __syncthreads();
        double pq = 0;
        for (int i = a + threadIdx.x; i < e; i += BLOCK) pq += p[i] * q[i];
        pq = bsum<BLOCK>(pq, red);
        const double alpha = rho / pq;
        double r2 = 0, z2 = 0;
        for (int i = a + threadIdx.x; i < e; i += BLOCK) {
          A.x[i] += alpha * p[i];
          const double ri = r[i] - alpha * q[i];
          r[i] = ri;
          r2 += ri * ri;
          z2 += ri * (PREC == OFH_DIAG_PREC ? ri * A.dinv[i] : ri);
        }
        r2 = bsum<BLOCK>(r2, red);
        z2 = bsum<BLOCK>(z2, red);
        res = sqrt(r2) / nrm;
        if (res <= A.tol) {
          st = it;
          break;
        }
        if (!(res == res)) break;
        rho_1 = rho;
        rho = z2;
      }
      if (st < 0) it = A.max_it;
    }
    if (threadIdx.x == 0) {
      A.iters[sys] = it;
      A.status[sys] = st;
      A.relres[sys] = res;
    }
    __syncthreads();
  }
}

// Mixed: FP32 inner iteration with FP64 residual replacement, per system.
template <int BLOCK, int TPR, int PREC>
__global__ __launch_bounds__(BLOCK) void k_batch_cgmx(BatchArgs A) {
  __shared__ double red[BLOCK / kWave];
  __shared__ int s_w;
  float* r = (float*)A.work;
  float* p = r + A.nrows;
  float* q = p + A.nrows;
  float* x32 = q + A.nrows;
  for (;;) {
    if (threadIdx.x == 0) s_w = atomicAdd(A.counter, 1);
    __syncthreads();
    const int w = s_w;
    __syncthreads();
    if (w >= A.nsys) break;
    const int sys = A.order[w];
    const int a = A.sys_rows[sys], e = A.sys_rows[sys + 1];
    double nrm = 0, res = 0, rho = 0, rho_1 = 1.0, rmax = 0;
    int it = 0, st = 0;
    bool first = true, done = false;
    for (;;) {  // residual replacement (the first pass sets up the system)
      __syncthreads();
      // r = b - A x64 in FP64 (q as scratch, reinterpreted: one FP64 row sum per row)
      double rr = 0, bb = 0, rz = 0;
      {
        const int lane = threadIdx.x % TPR, grp = threadIdx.x / TPR;
        for (int i = a + grp; i < e; i += BLOCK / TPR) {
          double s = 0;
          for (int k = A.rp[i] + lane; k < A.rp[i + 1]; k += TPR) s += A.val[k] * A.x[A.col[k]];
          if (TPR > 1) s = row_reduce<TPR>(s);
          if (lane == 0) {
            const double ri = A.b[i] - s;
            const float rf = (float)ri;
            r[i] = rf;
            x32[i] = 0.0f;
            rr += ri * ri;
            bb += A.b[i] * A.b[i];
            const float zf = PREC == OFH_DIAG_PREC ? rf * A.dinv32[i] : rf;
            rz += (double)rf * (double)zf;
          }
        }
      }
      rr = bsum<BLOCK>(rr, red);
      bb = bsum<BLOCK>(bb, red);
      rz = bsum<BLOCK>(rz, red);
      if (first) {
        nrm = sqrt(bb);
        if (nrm == 0) nrm = 1;
      }
      res = sqrt(rr) / nrm;
      rmax = res;
      if (res <= A.tol) {
        st = it;
        done = true;
      } else if (it >= A.max_it || !(res == res)) {
        st = -(A.max_it + 1);
        done = true;
      }
      if (done) break;
      const double beta = first ? 0.0 : rz / rho_1;  // keep the direction p
      rho = rz;
      for (int i = a + threadIdx.x; i < e; i += BLOCK) {
        const float zi = PREC == OFH_DIAG_PREC ? r[i] * A.dinv32[i] : r[i];
        p[i] = first ? zi : zi + (float)beta * p[i];
      }
      first = false;
      // FP32 inner iterations until the recursive residual drops by kBatchDelta
      for (;;) {
        __syncthreads();
        ++it;
        sys_spmv<BLOCK, TPR, float, float>(a, e, A.rp, A.col, A.val32, p, q);
        __syncthreads();
        double pq = 0;
        for (int i = a + threadIdx.x; i < e; i += BLOCK) pq += (double)p[i] * (double)q[i];
        pq = bsum<BLOCK>(pq, red);
        const float alpha = (float)(rho / pq);
        double r2 = 0, z2 = 0;
        for (int i = a + threadIdx.x; i < e; i += BLOCK) {
          x32[i] += alpha * p[i];
          const float ri = r[i] - alpha * q[i];
          r[i] = ri;
          const float zi = PREC == OFH_DIAG_PREC ? ri * A.dinv32[i] : ri;
          r2 += (double)ri * (double)ri;
          z2 += (double)ri * (double)zi;
        }
        r2 = bsum<BLOCK>(r2, red);
        z2 = bsum<BLOCK>(z2, red);
        const double rec = sqrt(r2) / nrm;
        rho_1 = rho;
        if (rec <= kBatchDelta * rmax || rec <= A.tol || it >= A.max_it || !(rec == rec)) break;
        const double beta2 = z2 / rho;
        rho = z2;
        for (int i = a + threadIdx.x; i < e; i += BLOCK) {
          const float zi = PREC == OFH_DIAG_PREC ? r[i] * A.dinv32[i] : r[i];
          p[i] = zi + (float)beta2 * p[i];
        }
      }
      __syncthreads();
      for (int i = a + threadIdx.x; i < e; i += BLOCK) A.x[i] += (double)x32[i];
    }
    if (threadIdx.x == 0) {
      A.iters[sys] = it;
      A.status[sys] = st;
      A.relres[sys] = res;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Resident form: one 1024-thread workgroup per system of at most
// 1024 x kResRpt rows.  Slot j of thread t is the system's local row
// t + 1024 j; the thread keeps the slots' r and q in registers, the direction
// p lives in LDS, where the SpMV gathers it, and only the matrix (and, FP64,
// x) is streamed from L2 / MALL in each iteration.  Global data is read and
// written through per-system buffer descriptors (scalar registers) with
// 32-bit offsets: slot j's offset is t*size + j*1024*size, the second term a
// scalar, so no per-slot address registers are held, and a read outside the
// system's range returns 0.  Each slot's row start and length are packed
// into LDS once per system.  The SpMV walks groups of slots in "row
// positions": one round loads positions s..s+SU-1 of every row of the group
// at once (interior rows of P1 meshes have 7 entries: one round).
// Same arithmetic per system as the streamed form (the reference's
// sequential row sums, CG.h:99-157), dot products as block trees.
// ---------------------------------------------------------------------------
constexpr int kResBlock = 1024;
constexpr int kResRpt = 12;  // rows per thread: systems up to 12,288 rows (96 KB of FP64 p)
// slots per group and row positions per round of loads, FP64 / FP32 SpMV
constexpr int kResGrp64 = 2, kResSu64 = 8;
constexpr int kResGrp32 = 2, kResSu32 = 7;
// each slot's row is packed in LDS as (start - system start) << 8 | length
constexpr int kResMaxLen = 255, kResMaxNnz = 1 << 24;

using Rsrc = __amdgpu_buffer_rsrc_t;

// descriptor of count elements at p (raw buffer, gfx9 data format word)
template <typename T>
__device__ __forceinline__ Rsrc make_rsrc(const T* p, int count) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(p), (short)0, count * (int)sizeof(T),
                                           0x00020000);
}
template <typename T>
__device__ __forceinline__ T bld(Rsrc r, int voff, int soff);
template <>
__device__ __forceinline__ double bld<double>(Rsrc r, int voff, int soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
template <>
__device__ __forceinline__ float bld<float>(Rsrc r, int voff, int soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
template <>
__device__ __forceinline__ int bld<int>(Rsrc r, int voff, int soff) {
  return __builtin_bit_cast(int, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ void bst(double v, Rsrc r, int voff, int soff) {
  typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, 0);
}

// one system's descriptors (uniform)
struct ResSys {
  int a, n_s, k_base;
  Rsrc rp, col, val, b, x, dinv;
};

template <typename TV>
__device__ __forceinline__ ResSys res_sys(const BatchArgs& A, int a, int n_s, const TV* val,
                                          const void* dinv, int dinv_size) {
  ResSys S;
  S.a = a;
  S.n_s = n_s;
  S.k_base = __builtin_amdgcn_readfirstlane(A.rp[a]);
  const int nnz_s = __builtin_amdgcn_readfirstlane(A.rp[a + n_s]) - S.k_base;
  S.rp = make_rsrc(A.rp + a, n_s + 1);
  S.col = make_rsrc(A.col + S.k_base, nnz_s);
  S.val = make_rsrc(val + S.k_base, nnz_s);
  S.b = make_rsrc(A.b + a, n_s);
  S.x = make_rsrc(A.x + a, n_s);
  S.dinv = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<char*>((const char*)dinv) + (size_t)a * dinv_size, (short)0, n_s * dinv_size,
      0x00020000);
  return S;
}

// packed row starts and lengths of this thread's slots (read back only by
// the same thread, so no barrier is needed)
__device__ __forceinline__ void res_rows(const ResSys& S, unsigned* rinfo) {
  const int t = threadIdx.x;
#pragma unroll
  for (int j = 0; j < kResRpt; ++j) {
    unsigned ri = 0;
    if (t < S.n_s - kResBlock * j) {
      const int so = kResBlock * j * 4;
      const int e0 = bld<int>(S.rp, t * 4, so), e1 = bld<int>(S.rp, t * 4 + 4, so);
      ri = ((unsigned)(e0 - S.k_base) << 8) | (unsigned)(e1 - e0);
    }
    rinfo[kResBlock * j + t] = ri;
  }
}

template <int K>
__device__ __forceinline__ void res_sum(double (&v)[K], double* red) {
  block_sum_k<kResBlock, K>(v, red);
}

// q_j = sum_k a_k x[c_k] for the slots of this thread; x is the system's
// vector, local index c - a, read from LDS (xl) or through a descriptor (xg).
// GRP slots at a time, SU row positions per round of loads.
template <int GRP, int SU, typename T, typename TV, bool LDSX>
__device__ __forceinline__ void res_spmv(const ResSys& S, const unsigned* __restrict__ rinfo,
                                         const T* __restrict__ xl, T (&q)[kResRpt]) {
  const int t = threadIdx.x;
#pragma unroll
  for (int j = 0; j < kResRpt; ++j) q[j] = T(0);
#pragma unroll
  for (int g = 0; g < kResRpt; g += GRP) {
    if (S.n_s <= kResBlock * g) break;  // uniform: no rows left in this group
    int k0[GRP], len[GRP], mx = 0;
#pragma unroll
    for (int j = 0; j < GRP; ++j) {
      const unsigned ri = rinfo[kResBlock * (g + j) + t];  // 0 for a slot beyond the system
      k0[j] = (int)(ri >> 8);
      len[j] = (int)(ri & 255u);
      mx = max(mx, len[j]);
    }
    // every load is issued (past the row's end it reads other entries of the
    // system, or 0 past the system), the sum takes only the row's own
    // entries, in row order
    for (int s = 0; s < mx; s += SU) {
      TV v[SU][GRP];
      int c[SU][GRP];
#pragma unroll
      for (int u = 0; u < SU; ++u)
#pragma unroll
        for (int j = 0; j < GRP; ++j) {
          v[u][j] = bld<TV>(S.val, (k0[j] + s + u) * (int)sizeof(TV), 0);
          c[u][j] = bld<int>(S.col, (k0[j] + s + u) * 4, 0);
        }
#pragma unroll
      for (int u = 0; u < SU; ++u)
#pragma unroll
        for (int j = 0; j < GRP; ++j) {
          const bool on = s + u < len[j];
          const int ci = on ? c[u][j] - S.a : 0;
          const T xv = LDSX ? xl[ci] : (T)bld<double>(S.x, ci * 8, 0);
          const T prod = (T)v[u][j] * xv;
          q[g + j] = on ? q[g + j] + prod : q[g + j];
        }
    }
  }
}

template <int PREC>
__global__ __launch_bounds__(kResBlock) void k_batch_cg64_res(BatchArgs A) {
  __shared__ double p[kResBlock * kResRpt];
  __shared__ unsigned rinfo[kResBlock * kResRpt];
  __shared__ double red[3 * kResBlock / kWave];
  __shared__ int s_w;
  const int t = threadIdx.x;
  for (;;) {
    if (t == 0) s_w = atomicAdd(A.counter, 1);
    __syncthreads();
    // wave-uniform: the system's descriptors stay in scalar registers
    const int w = __builtin_amdgcn_readfirstlane(s_w);
    if (w >= A.nsys) break;
    const int sys = __builtin_amdgcn_readfirstlane(A.order[w]);
    const int a = __builtin_amdgcn_readfirstlane(A.sys_rows[sys]);
    const int n_s = __builtin_amdgcn_readfirstlane(A.sys_rows[sys + 1]) - a;
    const ResSys S = res_sys<double>(A, a, n_s, A.val, A.dinv, 8);
    res_rows(S, rinfo);
    const auto own = [&](int j) { return t < n_s - kResBlock * j; };
    const auto dinv = [&](int j) { return bld<double>(S.dinv, t * 8, kResBlock * j * 8); };
    double r[kResRpt], q[kResRpt];
    // r = b - A x (CG.h:115), x gathered through its descriptor
    res_spmv<kResGrp64, kResSu64, double, double, false>(S, rinfo, nullptr, q);
    double v[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < kResRpt; ++j) {
      r[j] = 0.0;
      if (own(j)) {
        const double bi = bld<double>(S.b, t * 8, kResBlock * j * 8), ri = bi - q[j];
        r[j] = ri;
        v[0] += bi * bi;
        v[1] += ri * ri;
        v[2] += ri * (PREC == OFH_DIAG_PREC ? ri * dinv(j) : ri);
      }
    }
    res_sum<3>(v, red);
    double nrm = sqrt(v[0]);
    if (nrm == 0) nrm = 1;
    double res = sqrt(v[1]) / nrm, rho = v[2], rho_1 = 1.0;
    int it = 0, st;
    if (res <= A.tol) {
      st = 0;
    } else {
      st = -(A.max_it + 1);
      for (it = 1; it <= A.max_it; ++it) {
        const double beta = it == 1 ? 0.0 : rho / rho_1;
        // the reductions end in barriers: nobody reads the old p any more
#pragma unroll
        for (int j = 0; j < kResRpt; ++j) {
          if (own(j)) {
            const double zi = PREC == OFH_DIAG_PREC ? r[j] * dinv(j) : r[j];
            double& pj = p[kResBlock * j + t];
            pj = it == 1 ? zi : zi + beta * pj;
          }
        }
        __syncthreads();
        res_spmv<kResGrp64, kResSu64, double, double, true>(S, rinfo, p, q);
        double pq[1] = {0.0};
#pragma unroll
        for (int j = 0; j < kResRpt; ++j)
          if (own(j)) pq[0] += p[kResBlock * j + t] * q[j];
        res_sum<1>(pq, red);
        const double alpha = rho / pq[0];
        double u[2] = {0.0, 0.0};
#pragma unroll
        for (int j = 0; j < kResRpt; ++j) {
          if (own(j)) {
            const int so = kResBlock * j * 8;
            bst(bld<double>(S.x, t * 8, so) + alpha * p[kResBlock * j + t], S.x, t * 8, so);
            const double ri = r[j] - alpha * q[j];
            r[j] = ri;
            u[0] += ri * ri;
            u[1] += ri * (PREC == OFH_DIAG_PREC ? ri * dinv(j) : ri);
          }
        }
        res_sum<2>(u, red);
        res = sqrt(u[0]) / nrm;
        if (res <= A.tol) {
          st = it;
          break;
        }
        if (!(res == res)) break;
        rho_1 = rho;
        rho = u[1];
      }
      if (st < 0) it = A.max_it;
    }
    if (t == 0) {
      A.iters[sys] = it;
      A.status[sys] = st;
      A.relres[sys] = res;
    }
    __syncthreads();
  }
}

// Mixed resident form: FP32 p in LDS, FP32 r, correction and q in registers,
// FP64 residual replacement with x64 read from and written to global memory.
template <int PREC>
__global__ __launch_bounds__(kResBlock) void k_batch_cgmx_res(BatchArgs A) {
  __shared__ float p[kResBlock * kResRpt];
  __shared__ unsigned rinfo[kResBlock * kResRpt];
  __shared__ double red[3 * kResBlock / kWave];
  __shared__ int s_w;
  const int t = threadIdx.x;
  for (;;) {
    if (t == 0) s_w = atomicAdd(A.counter, 1);
    __syncthreads();
    const int w = __builtin_amdgcn_readfirstlane(s_w);
    if (w >= A.nsys) break;
    const int sys = __builtin_amdgcn_readfirstlane(A.order[w]);
    const int a = __builtin_amdgcn_readfirstlane(A.sys_rows[sys]);
    const int n_s = __builtin_amdgcn_readfirstlane(A.sys_rows[sys + 1]) - a;
    const ResSys S = res_sys<float>(A, a, n_s, A.val32, A.dinv32, 4);
    const ResSys S64 = res_sys<double>(A, a, n_s, A.val, A.dinv32, 4);
    res_rows(S, rinfo);
    const auto own = [&](int j) { return t < n_s - kResBlock * j; };
    const auto dinv = [&](int j) { return bld<float>(S.dinv, t * 4, kResBlock * j * 4); };
    float r[kResRpt], x32[kResRpt], q[kResRpt];
    double nrm = 0, res = 0, rho = 0, rho_1 = 1.0, rmax = 0;
    int it = 0, st = 0;
    bool first = true;
    for (;;) {  // residual replacement (the first pass sets up the system)
      double v[3] = {0.0, 0.0, 0.0};
      {
        double rq[kResRpt];  // FP64 row sums of A x64
        res_spmv<kResGrp64, kResSu64, double, double, false>(S64, rinfo, nullptr, rq);
#pragma unroll
        for (int j = 0; j < kResRpt; ++j) {
          r[j] = 0.0f;
          x32[j] = 0.0f;
          if (own(j)) {
            const double bi = bld<double>(S.b, t * 8, kResBlock * j * 8), ri = bi - rq[j];
            const float rf = (float)ri;
            r[j] = rf;
            v[0] += bi * bi;
            v[1] += ri * ri;
            const float zf = PREC == OFH_DIAG_PREC ? rf * dinv(j) : rf;
            v[2] += (double)rf * (double)zf;
          }
        }
      }
      res_sum<3>(v, red);
      if (first) {
        nrm = sqrt(v[0]);
        if (nrm == 0) nrm = 1;
      }
      res = sqrt(v[1]) / nrm;
      rmax = res;
      bool done = false;
      if (res <= A.tol) {
        st = it;
        done = true;
      } else if (it >= A.max_it || !(res == res)) {
        st = -(A.max_it + 1);
        done = true;
      }
      if (done) break;
      const double beta = first ? 0.0 : v[2] / rho_1;  // keep the direction p
      rho = v[2];
#pragma unroll
      for (int j = 0; j < kResRpt; ++j) {
        if (own(j)) {
          const float zi = PREC == OFH_DIAG_PREC ? r[j] * dinv(j) : r[j];
          float& pj = p[kResBlock * j + t];
          pj = first ? zi : zi + (float)beta * pj;
        }
      }
      first = false;
      for (;;) {  // FP32 inner iterations until the recursive residual drops by kBatchDelta
        __syncthreads();
        ++it;
        res_spmv<kResGrp32, kResSu32, float, float, true>(S, rinfo, p, q);
        double pq[1] = {0.0};
#pragma unroll
        for (int j = 0; j < kResRpt; ++j)
          if (own(j)) pq[0] += (double)p[kResBlock * j + t] * (double)q[j];
        res_sum<1>(pq, red);
        const float alpha = (float)(rho / pq[0]);
        double u[2] = {0.0, 0.0};
#pragma unroll
        for (int j = 0; j < kResRpt; ++j) {
          if (own(j)) {
            x32[j] += alpha * p[kResBlock * j + t];
            const float ri = r[j] - alpha * q[j];
            r[j] = ri;
            const float zi = PREC == OFH_DIAG_PREC ? ri * dinv(j) : ri;
            u[0] += (double)ri * (double)ri;
            u[1] += (double)ri * (double)zi;
          }
        }
        res_sum<2>(u, red);
        const double rec = sqrt(u[0]) / nrm;
        rho_1 = rho;
        if (rec <= kBatchDelta * rmax || rec <= A.tol || it >= A.max_it || !(rec == rec)) break;
        const double beta2 = u[1] / rho;
        rho = u[1];
#pragma unroll
        for (int j = 0; j < kResRpt; ++j) {
          if (own(j)) {
            const float zi = PREC == OFH_DIAG_PREC ? r[j] * dinv(j) : r[j];
            float& pj = p[kResBlock * j + t];
            pj = zi + (float)beta2 * pj;
          }
        }
      }
      // x64 += correction (own rows); the barrier makes the new x64 visible to
      // the whole workgroup before the next pass gathers it
#pragma unroll
      for (int j = 0; j < kResRpt; ++j) {
        if (own(j)) {
          const int so = kResBlock * j * 8;
          bst(bld<double>(S.x, t * 8, so) + (double)x32[j], S.x, t * 8, so);
        }
      }
      __syncthreads();
    }
    if (t == 0) {
      A.iters[sys] = it;
      A.status[sys] = st;
      A.relres[sys] = res;
    }
    __syncthreads();
  }
}

__global__ void k_fill_int(int* p, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

using LaunchFn = void (*)(const BatchArgs&, int grid, hipStream_t);

template <int BLOCK, int TPR, int PREC, bool MIXED>
static void launch_batch(const BatchArgs& a, int grid, hipStream_t s) {
  if (MIXED)
    hipLaunchKernelGGL((k_batch_cgmx<BLOCK, TPR, PREC>), dim3(grid), dim3(BLOCK), 0, s, a);
  else
    hipLaunchKernelGGL((k_batch_cg64<BLOCK, TPR, PREC>), dim3(grid), dim3(BLOCK), 0, s, a);
}

template <int PREC, bool MIXED>
static void launch_res(const BatchArgs& a, int grid, hipStream_t s) {
  if (MIXED)
    hipLaunchKernelGGL((k_batch_cgmx_res<PREC>), dim3(grid), dim3(kResBlock), 0, s, a);
  else
    hipLaunchKernelGGL((k_batch_cg64_res<PREC>), dim3(grid), dim3(kResBlock), 0, s, a);
}

// tpr = 0 selects the resident form (1024 threads)
template <int PREC, bool MIXED>
static LaunchFn pick(int block, int tpr) {
  if (block == kResBlock && tpr == 0) return &launch_res<PREC, MIXED>;
#define OFH_CASE(B, T) \
  if (block == B && tpr == T) return &launch_batch<B, T, PREC, MIXED>;
  OFH_CASE(256, 1) OFH_CASE(256, 2) OFH_CASE(256, 4) OFH_CASE(256, 8)
  OFH_CASE(512, 1) OFH_CASE(512, 2) OFH_CASE(512, 4) OFH_CASE(512, 8)
  OFH_CASE(1024, 1) OFH_CASE(1024, 2) OFH_CASE(1024, 4) OFH_CASE(1024, 8)
#undef OFH_CASE
  return nullptr;
}

static LaunchFn pick_fn(int prec, bool mixed, int block, int tpr) {
  if (prec == OFH_DIAG_PREC) return mixed ? pick<OFH_DIAG_PREC, true>(block, tpr) : pick<OFH_DIAG_PREC, false>(block, tpr);
  return mixed ? pick<OFH_IDENT_PREC, true>(block, tpr) : pick<OFH_IDENT_PREC, false>(block, tpr);
}

static int batch_run(ofh_batch B, int prec, int mixed, const double* b, double* x, int max_it,
                     double toler, int block, int tpr, hipStream_t s, bool graph) {
  ofh_csr M = B->A;
  ofh_ctx ctx = M->ctx;
  int rc;
  if (prec == OFH_DIAG_PREC && (rc = csr_ensure_dinv(M))) return rc;
  if (mixed && (rc = csr_ensure_val32(M))) return rc;
  const size_t need = (size_t)M->n * 4 * sizeof(double);
  if (B->work_bytes < need) {
    if (B->work) OFH_HIP(hipFree(B->work));
    OFH_HIP(hipMalloc(&B->work, need));
    B->work_bytes = need;
  }
  LaunchFn fn = pick_fn(prec, mixed != 0, block, tpr);
  OFH_CHECK(fn, OFH_E_INVALID, "batched CG: unsupported launch shape");
  OFH_CHECK(tpr != 0 || B->resident_ok, OFH_E_INVALID,
            "batched CG: the resident form needs systems of at most 12288 rows, 2^24 "
            "nonzeros and rows of at most 255 entries");
  BatchArgs a{};
  a.nsys = B->nsys;
  a.sys_rows = B->sys_rows;
  a.order = B->order;
  a.counter = B->counter;
  a.rp = M->row_ptr;
  a.col = M->col;
  a.val = M->val;
  a.val32 = M->val32;
  a.dinv = M->dinv;
  a.dinv32 = M->dinv32;
  a.b = b;
  a.x = x;
  a.work = B->work;
  a.nrows = M->n;
  a.max_it = max_it;
  a.tol = toler;
  a.iters = B->iters;
  a.status = B->status;
  a.relres = B->relres;
  // persistent grid: every workgroup keeps pulling systems until the queue is
  // empty, so all of them reach the exit
  // a resident workgroup holds 128 VGPRs x 16 waves: one per CU
  const int per_cu = tpr == 0 ? 1 : block >= 1024 ? 2 : (block >= 512 ? 4 : 8);
  const int grid = std::max(1, std::min(B->nsys, ctx->num_cu * per_cu));
  if (!graph) {
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, B->counter, 0);
    fn(a, grid, s);
    OFH_HIP(hipGetLastError());
    return OFH_OK;
  }
  // The queue reset and the persistent launch are captured once per set of
  // arguments into a graph and replayed; the reset is a kernel, not a memset
  // node.
  const BatchGraphKey key{(const void*)fn, b, x, max_it, toler, M->dinv, M->val32, B->work, grid};
  if (!B->gexec || !(key == B->gkey)) {
    if (B->gexec) {
      OFH_HIP(hipGraphExecDestroy(B->gexec));
      B->gexec = nullptr;
    }
    if (!B->cap) OFH_HIP(hipStreamCreateWithFlags(&B->cap, hipStreamNonBlocking));
    OFH_HIP(hipStreamBeginCapture(B->cap, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, B->cap, B->counter, 0);
    fn(a, grid, B->cap);
    hipError_t le = hipGetLastError();
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(B->cap, &g);
    OFH_HIP(le);
    OFH_HIP(ce);
    hipError_t ie = hipGraphInstantiate(&B->gexec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    OFH_HIP(ie);
    B->gkey = key;
  }
  OFH_HIP(hipGraphLaunch(B->gexec, s));
  return OFH_OK;
}

}  // namespace ofh

using namespace ofh;

extern "C" {

int ofh_batch_create(ofh_csr A, int nsys, const int32_t* h_sys_rows, ofh_batch* out) {
  OFH_CHECK(A && h_sys_rows && out && nsys > 0, OFH_E_INVALID, "ofh_batch_create: bad argument");
  OFH_CHECK(h_sys_rows[0] == 0 && h_sys_rows[nsys] == A->n, OFH_E_INVALID,
            "ofh_batch_create: system rows must cover the matrix");
  for (int s = 0; s < nsys; ++s)
    OFH_CHECK(h_sys_rows[s + 1] > h_sys_rows[s], OFH_E_INVALID, "ofh_batch_create: empty system");
  // every row's columns must stay inside its own system (block-diagonal)
  std::vector<int> rp(A->n + 1), col(A->nnz);
  int rc = ofh_csr_get_pattern(A, rp.data(), col.data());
  if (rc) return rc;
  for (int s = 0; s < nsys; ++s)
    for (int i = h_sys_rows[s]; i < h_sys_rows[s + 1]; ++i)
      for (int k = rp[i]; k < rp[i + 1]; ++k)
        OFH_CHECK(col[k] >= h_sys_rows[s] && col[k] < h_sys_rows[s + 1], OFH_E_PATTERN,
                  "ofh_batch_create: the matrix is not block-diagonal over the systems");
  std::vector<int> order(nsys);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int u, int v) {
    return h_sys_rows[u + 1] - h_sys_rows[u] > h_sys_rows[v + 1] - h_sys_rows[v];
  });
  ofh_batch B = new ofh_batch_s();
  B->A = A;
  B->nsys = nsys;
  B->h_rows.assign(h_sys_rows, h_sys_rows + nsys + 1);
  int max_len = 0, max_nnz = 0;
  for (int k = 0; k < nsys; ++k) {
    B->max_rows = std::max(B->max_rows, h_sys_rows[k + 1] - h_sys_rows[k]);
    max_nnz = std::max(max_nnz, rp[h_sys_rows[k + 1]] - rp[h_sys_rows[k]]);
  }
  for (int i = 0; i < A->n; ++i) max_len = std::max(max_len, rp[i + 1] - rp[i]);
  B->resident_ok = B->max_rows <= kResBlock * kResRpt && max_len <= kResMaxLen && max_nnz < kResMaxNnz;

  hipStream_t s = A->ctx->stream;
  OFH_HIP(hipMalloc(&B->sys_rows, sizeof(int) * (nsys + 1)));
  OFH_HIP(hipMalloc(&B->order, sizeof(int) * nsys));
  OFH_HIP(hipMalloc(&B->counter, sizeof(int)));
  OFH_HIP(hipMalloc(&B->iters, sizeof(int) * nsys));
  OFH_HIP(hipMalloc(&B->status, sizeof(int) * nsys));
  OFH_HIP(hipMalloc(&B->relres, sizeof(double) * nsys));
  OFH_HIP(hipMemcpyAsync(B->sys_rows, h_sys_rows, sizeof(int) * (nsys + 1), hipMemcpyHostToDevice, s));
  OFH_HIP(hipMemcpyAsync(B->order, order.data(), sizeof(int) * nsys, hipMemcpyHostToDevice, s));

  OFH_HIP(hipStreamSynchronize(s));
  *out = B;
  return OFH_OK;
}

int ofh_batch_destroy(ofh_batch B) {
  if (!B) return OFH_OK;
  hipDeviceSynchronize();
  void* ptrs[] = {B->sys_rows, B->order, B->counter, B->iters, B->status, B->relres, B->work};
  for (void* p : ptrs) hipFree(p);
  if (B->gexec) hipGraphExecDestroy(B->gexec);
  if (B->cap) hipStreamDestroy(B->cap);
  delete B;
  return OFH_OK;
}

int ofh_batch_set_config(ofh_batch B, int block_threads, int threads_per_row) {
  OFH_CHECK(B, OFH_E_INVALID, "null batch");
  OFH_CHECK(pick_fn(OFH_DIAG_PREC, false, block_threads, threads_per_row), OFH_E_INVALID,
            "ofh_batch_set_config: block in {256,512,1024}, threads per row in {1,2,4,8}, "
            "or (1024, 0) for the resident form");
  OFH_CHECK(threads_per_row != 0 || B->resident_ok, OFH_E_INVALID,
            "ofh_batch_set_config: the resident form needs systems of at most 12288 rows, "
            "2^24 nonzeros and rows of at most 255 entries");
  B->block = block_threads;
  B->tpr = threads_per_row;
  B->tuned = true;
  return OFH_OK;
}

int ofh_batch_use_graph(ofh_batch B, int enable) {
  OFH_CHECK(B, OFH_E_INVALID, "null batch");
  B->use_graph = enable != 0;
  return OFH_OK;
}

int ofh_batch_get_config(ofh_batch B, int* block_threads, int* threads_per_row) {
  OFH_CHECK(B, OFH_E_INVALID, "null batch");
  if (block_threads) *block_threads = B->block;
  if (threads_per_row) *threads_per_row = B->tpr;
  return OFH_OK;
}

int ofh_batch_cg(ofh_batch B, int prec, int mixed, const double* d_b, double* d_x, int max_it,
                 double toler, int32_t* h_iters, int32_t* h_status, double* h_relres,
                 ofh_solve_info* total) {
  OFH_CHECK(B && d_b && d_x && max_it >= 0, OFH_E_INVALID, "ofh_batch_cg: bad argument");
  OFH_CHECK(prec == OFH_DIAG_PREC || prec == OFH_IDENT_PREC, OFH_E_UNSUPPORTED,
            "ofh_batch_cg: IDENT_PREC or DIAG_PREC");
  hipStream_t s = B->A->ctx->stream;
  const double t0 = now_seconds();
  int rc = batch_run(B, prec, mixed, d_b, d_x, max_it, toler, B->block, B->tpr, s, B->use_graph);
  if (rc) return rc;
  std::vector<int> it(B->nsys), st(B->nsys);
  std::vector<double> rr(B->nsys);
  OFH_HIP(hipMemcpyAsync(it.data(), B->iters, sizeof(int) * B->nsys, hipMemcpyDeviceToHost, s));
  OFH_HIP(hipMemcpyAsync(st.data(), B->status, sizeof(int) * B->nsys, hipMemcpyDeviceToHost, s));
  OFH_HIP(hipMemcpyAsync(rr.data(), B->relres, sizeof(double) * B->nsys, hipMemcpyDeviceToHost, s));
  OFH_HIP(hipStreamSynchronize(s));
  const double t1 = now_seconds();
  if (h_iters) std::copy(it.begin(), it.end(), h_iters);
  if (h_status) std::copy(st.begin(), st.end(), h_status);
  if (h_relres) std::copy(rr.begin(), rr.end(), h_relres);
  if (total) {
    *total = ofh_solve_info{};
    int worst = 0, conv = 1;
    double res = 0;
    for (int k = 0; k < B->nsys; ++k) {
      worst = std::max(worst, it[k]);
      conv = conv && st[k] >= 0;
      res = std::max(res, rr[k]);
    }
    total->iterations = worst;
    total->nb_it = conv ? worst : -(max_it + 1);
    total->relres = res;
    total->seconds = t1 - t0;
  }
  return OFH_OK;
}

int ofh_batch_autotune(ofh_batch B, int prec, int mixed, const double* d_b, double* d_x,
                       int max_it, double toler, int* best_block, int* best_tpr,
                       double* best_ms) {
  OFH_CHECK(B && d_b && d_x, OFH_E_INVALID, "ofh_batch_autotune: bad argument");
  ofh_csr M = B->A;
  hipStream_t s = M->ctx->stream;
  const int n = M->n;
  double* x0;
  OFH_HIP(hipMalloc(&x0, sizeof(double) * n));
  OFH_HIP(hipMemcpyAsync(x0, d_x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  hipEvent_t e0, e1;
  OFH_HIP(hipEventCreate(&e0));
  OFH_HIP(hipEventCreate(&e1));
  // candidate shapes: 3 block sizes x 4 threads per row, then the resident form
  // when every system fits on chip
  std::vector<std::pair<int, int>> shapes;
  for (int blk : {256, 512, 1024})
    for (int t : {1, 2, 4, 8}) shapes.emplace_back(blk, t);
  if (B->resident_ok) shapes.emplace_back(kResBlock, 0);
  float best = 1e30f;
  int bb = B->block, bt = B->tpr;
  for (const auto& sh : shapes) {
    {
      float ms_min = 1e30f;
      for (int rep = 0; rep < 2; ++rep) {  // first run warms caches and code
        OFH_HIP(hipMemcpyAsync(d_x, x0, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        OFH_HIP(hipEventRecord(e0, s));
        int rc = batch_run(B, prec, mixed, d_b, d_x, max_it, toler, sh.first, sh.second, s, false);
        if (rc) { hipFree(x0); return rc; }
        OFH_HIP(hipEventRecord(e1, s));
        OFH_HIP(hipEventSynchronize(e1));
        float ms = 0;
        OFH_HIP(hipEventElapsedTime(&ms, e0, e1));
        ms_min = std::min(ms_min, ms);
      }
      if (ms_min < best) {
        best = ms_min;
        bb = sh.first;
        bt = sh.second;
      }
    }
  }
  OFH_HIP(hipMemcpyAsync(d_x, x0, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  OFH_HIP(hipStreamSynchronize(s));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(x0);
  B->block = bb;
  B->tpr = bt;
  B->tuned = true;
  if (best_block) *best_block = bb;
  if (best_tpr) *best_tpr = bt;
  if (best_ms) *best_ms = best;
  return OFH_OK;
}

}  // extern "C"
