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
#include <map>
#include <mutex>
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
  // resident form: ELL copy of the matrix, one width per system, 64-row
  // chunks stored position-major
  int* ell_e0 = nullptr;      // device, nsys + 1: first entry of each system
  int* ell_w = nullptr;       // device, nsys: width (longest row) of each system
  int* ell_col = nullptr;     // device: local column x 8 (bytes); padding -> kResPad x 8
  int* ell_src = nullptr;     // device: CSR entry of each slot, -1 for padding
  double* ell_val = nullptr;  // padding -0.0
  float* ell_val32 = nullptr;
  int64_t ell_n = 0;
  uint64_t ell_version = ~0ull;  // CSR values version gathered into ell_val*
  std::vector<int> h_ell_e0, h_ell_w;  // host copies of the ELL offsets and widths
  bool tuned = false;
  bool use_graph = true;         // replay the launch from a captured graph
  hipGraphExec_t gexec = nullptr;
  hipStream_t cap = nullptr;     // capture-only stream
  BatchGraphKey gkey{};
  std::vector<int> h_rows;
  std::vector<int> h_nnz;  // nonzeros of each system (autotuner cache key)
};

namespace ofh {

constexpr double kBatchDelta = 0.1;

// Autotuner choices of this process, keyed by what the timing depends on:
// the solver variant, the matrix size and every system's rows and nonzeros.
struct TuneEntry {
  int block, tpr;
  double ms;
};
static std::mutex g_tune_mu;
static std::map<std::string, TuneEntry> g_tune_cache;

static std::string tune_key(const ofh_batch_s* B, int prec, int mixed) {
  std::string k;
  auto put = [&k](int64_t v) { k.append(reinterpret_cast<const char*>(&v), sizeof(v)); };
  put(prec);
  put(mixed);
  put(B->A->n);
  put(B->A->nnz);
  put(B->nsys);
  for (int v : B->h_rows) put(v);
  for (int v : B->h_nnz) put(v);
  return k;
}

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
  const int* ell_e0;        // resident form: first ELL entry of each system
  const int* ell_w;         // ELL width of each system
  const int* ell_col;       // byte offset of the local column
  const double* ell_val;
  const float* ell_val32;
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
// x) is streamed from L2 / MALL in each iteration.
//
// The matrix is read from an ELL copy built at batch creation with one width
// W per system (its longest row): local rows 64c..64c+63 form chunk c (one
// wave's rows of one slot), stored position-major, entry s of row `lane` at
// e0 + (c W + s) 64 + lane, so every load instruction reads 64 consecutive
// entries, and the offset of a load is a per-slot scalar plus an immediate.
// Columns are stored as byte offsets (local column x 8).  A padding entry has
// the value -0.0 and points at a slot holding +0.0 (LDS word kResPad; past
// the end of the system for the global vector, where the buffer read returns
// 0): its product is -0.0, and q + (-0.0) == q exactly, so the row sums equal
// the sums over the row's own entries in column order -- the streamed form's
// arithmetic (the reference's sequential row sums, CG.h:99-157) -- without a
// compare or a select.  Global data goes through per-system buffer
// descriptors (scalar registers).  Dot products are block trees whose LDS
// buffers alternate, so each needs one barrier.
// ---------------------------------------------------------------------------
constexpr int kResBlock = 1024;
constexpr int kResRpt = 12;  // rows per thread: systems up to 12,288 rows (96 KB of FP64 p)
constexpr int kResChunk = 64;
constexpr int kResPad = kResBlock * kResRpt;  // LDS word holding +0.0
constexpr int kResSu = 8;  // row positions per round of loads
constexpr int kResGrp = 2;  // slots per round of loads
constexpr int kResRed = 3 * kResBlock / kWave;  // doubles per reduction buffer

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
  int a, n_s, w;
  Rsrc col, val, b, x, dinv;
};

template <typename TV>
__device__ __forceinline__ ResSys res_sys(const BatchArgs& A, int sys, int a, int n_s,
                                          const TV* ell_val, const void* dinv, int dinv_size) {
  ResSys S;
  S.a = a;
  S.n_s = n_s;
  S.w = __builtin_amdgcn_readfirstlane(A.ell_w[sys]);
  const int e0 = __builtin_amdgcn_readfirstlane(A.ell_e0[sys]);
  const int ne = __builtin_amdgcn_readfirstlane(A.ell_e0[sys + 1]) - e0;
  S.col = make_rsrc(A.ell_col + e0, ne);
  S.val = make_rsrc(ell_val + e0, ne);
  S.b = make_rsrc(A.b + a, n_s);
  S.x = make_rsrc(A.x + a, n_s);
  S.dinv = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<char*>((const char*)dinv) + (size_t)a * dinv_size, (short)0, n_s * dinv_size,
      0x00020000);
  return S;
}

// block tree of K sums into one of two alternating LDS buffers: the buffer
// written here was last read before the previous reduction's barrier
template <int K>
__device__ __forceinline__ void res_sum(double (&v)[K], double* red, int& parity) {
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
  double* buf = red + parity * kResRed;
  parity ^= 1;
  const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
  if (l == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) buf[k * (kResBlock / kWave) + w] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < kResBlock / kWave; ++i) s += buf[k * (kResBlock / kWave) + i];
    v[k] = s;
  }
}

// q_j = sum_k a_k x[c_k] for the slots of this thread.  x is the system's
// vector: LDS words of stride 8 bytes at xl (LDSX) or the global vector S.x.
template <typename T, typename TV, bool LDSX>
__device__ __forceinline__ void res_spmv(const ResSys& S, const char* __restrict__ xl,
                                         T (&q)[kResRpt]) {
  const int lane = threadIdx.x % kResChunk;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kResChunk);
  constexpr int kWaves = kResBlock / kResChunk;
#pragma unroll
  for (int j = 0; j < kResRpt; ++j) q[j] = T(0);
#pragma unroll
  for (int g = 0; g < kResRpt; g += kResGrp) {
    if (S.n_s <= kResBlock * g) break;  // uniform: no rows left
    for (int s0 = 0; s0 < S.w; s0 += kResSu) {
      TV v[kResGrp][kResSu];
      int c[kResGrp][kResSu];
#pragma unroll
      for (int j = 0; j < kResGrp; ++j) {
        const int cl = wv + kWaves * (g + j);  // this wave's chunk for slot g + j
        const bool live = kResChunk * cl < S.n_s;
        const int base = (cl * S.w + s0) * kResChunk;  // entry offset, uniform
#pragma unroll
        for (int u = 0; u < kResSu; ++u) {
          v[j][u] = TV(-0.0);  // a position past the width adds -0.0 as padding does
          c[j][u] = kResPad * 8;
          if (live && s0 + u < S.w) {
            v[j][u] = bld<TV>(S.val, lane * (int)sizeof(TV), (base + u * kResChunk) * (int)sizeof(TV));
            c[j][u] = bld<int>(S.col, lane * 4, (base + u * kResChunk) * 4);
          }
        }
      }
#pragma unroll
      for (int j = 0; j < kResGrp; ++j)
#pragma unroll
        for (int u = 0; u < kResSu; ++u) {
          const T xv = LDSX ? *(const T*)(xl + c[j][u]) : (T)bld<double>(S.x, c[j][u], 0);
          q[g + j] += (T)v[j][u] * xv;  // padding: -0.0 * +0.0 = -0.0
        }
    }
  }
}

template <int PREC>
__global__ __launch_bounds__(kResBlock) void k_batch_cg64_res(BatchArgs A) {
  __shared__ double p[kResPad + 1];
  __shared__ double red[2 * kResRed];
  __shared__ int s_w;
  OFH_LDS_BUDGET(sizeof(double) * (kResPad + 1) + sizeof(double) * 2 * kResRed + sizeof(int));
  const int t = threadIdx.x;
  int parity = 0;
  if (t == 0) p[kResPad] = 0.0;
  for (;;) {
    if (t == 0) s_w = atomicAdd(A.counter, 1);
    __syncthreads();
    // wave-uniform: the system's descriptors stay in scalar registers
    const int w = __builtin_amdgcn_readfirstlane(s_w);
    if (w >= A.nsys) break;
    const int sys = __builtin_amdgcn_readfirstlane(A.order[w]);
    const int a = __builtin_amdgcn_readfirstlane(A.sys_rows[sys]);
    const int n_s = __builtin_amdgcn_readfirstlane(A.sys_rows[sys + 1]) - a;
    const ResSys S = res_sys<double>(A, sys, a, n_s, A.ell_val, A.dinv, 8);
    const auto own = [&](int j) { return t < n_s - kResBlock * j; };
    const auto dinv = [&](int j) { return bld<double>(S.dinv, t * 8, kResBlock * j * 8); };
    double r[kResRpt], q[kResRpt];
    // r = b - A x (CG.h:115), x gathered through its descriptor
    res_spmv<double, double, false>(S, nullptr, q);
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
    res_sum<3>(v, red, parity);
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
        // every SpMV read of the old p came before the last reduction's barrier
#pragma unroll
        for (int j = 0; j < kResRpt; ++j) {
          if (own(j)) {
            const double zi = PREC == OFH_DIAG_PREC ? r[j] * dinv(j) : r[j];
            double& pj = p[kResBlock * j + t];
            pj = it == 1 ? zi : zi + beta * pj;
          }
        }
        __syncthreads();
        res_spmv<double, double, true>(S, (const char*)p, q);
        double pq[1] = {0.0};
#pragma unroll
        for (int j = 0; j < kResRpt; ++j)
          if (own(j)) pq[0] += p[kResBlock * j + t] * q[j];
        res_sum<1>(pq, red, parity);
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
        res_sum<2>(u, red, parity);
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

// Mixed resident form: FP32 p in LDS (one word per 8 bytes, matching the
// byte-offset columns), FP32 r, correction and q in registers, FP64 residual
// replacement with x64 read from and written to global memory.
template <int PREC>
__global__ __launch_bounds__(kResBlock) void k_batch_cgmx_res(BatchArgs A) {
  __shared__ float p2[2 * (kResPad + 1)];
  __shared__ double red[2 * kResRed];
  __shared__ int s_w;
  OFH_LDS_BUDGET(sizeof(float) * 2 * (kResPad + 1) + sizeof(double) * 2 * kResRed + sizeof(int));
  const int t = threadIdx.x;
  int parity = 0;
  if (t == 0) p2[2 * kResPad] = 0.0f;
  for (;;) {
    if (t == 0) s_w = atomicAdd(A.counter, 1);
    __syncthreads();
    const int w = __builtin_amdgcn_readfirstlane(s_w);
    if (w >= A.nsys) break;
    const int sys = __builtin_amdgcn_readfirstlane(A.order[w]);
    const int a = __builtin_amdgcn_readfirstlane(A.sys_rows[sys]);
    const int n_s = __builtin_amdgcn_readfirstlane(A.sys_rows[sys + 1]) - a;
    const ResSys S = res_sys<float>(A, sys, a, n_s, A.ell_val32, A.dinv32, 4);
    const ResSys S64 = res_sys<double>(A, sys, a, n_s, A.ell_val, A.dinv32, 4);
    const auto own = [&](int j) { return t < n_s - kResBlock * j; };
    const auto dinv = [&](int j) { return bld<float>(S.dinv, t * 4, kResBlock * j * 4); };
    const auto pw = [&](int j) -> float& { return p2[2 * (kResBlock * j + t)]; };
    float r[kResRpt], x32[kResRpt], q[kResRpt];
    double nrm = 0, res = 0, rho = 0, rho_1 = 1.0, rmax = 0;
    int it = 0, st = 0;
    bool first = true;
    for (;;) {  // residual replacement (the first pass sets up the system)
      double v[3] = {0.0, 0.0, 0.0};
      {
        double rq[kResRpt];  // FP64 row sums of A x64
        res_spmv<double, double, false>(S64, nullptr, rq);
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
      res_sum<3>(v, red, parity);
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
          pw(j) = first ? zi : zi + (float)beta * pw(j);
        }
      }
      first = false;
      for (;;) {  // FP32 inner iterations until the recursive residual drops by kBatchDelta
        __syncthreads();
        ++it;
        res_spmv<float, float, true>(S, (const char*)p2, q);
        double pq[1] = {0.0};
#pragma unroll
        for (int j = 0; j < kResRpt; ++j)
          if (own(j)) pq[0] += (double)pw(j) * (double)q[j];
        res_sum<1>(pq, red, parity);
        const float alpha = (float)(rho / pq[0]);
        double u[2] = {0.0, 0.0};
#pragma unroll
        for (int j = 0; j < kResRpt; ++j) {
          if (own(j)) {
            x32[j] += alpha * pw(j);
            const float ri = r[j] - alpha * q[j];
            r[j] = ri;
            const float zi = PREC == OFH_DIAG_PREC ? ri * dinv(j) : ri;
            u[0] += (double)ri * (double)ri;
            u[1] += (double)ri * (double)zi;
          }
        }
        res_sum<2>(u, red, parity);
        const double rec = sqrt(u[0]) / nrm;
        rho_1 = rho;
        if (rec <= kBatchDelta * rmax || rec <= A.tol || it >= A.max_it || !(rec == rec)) break;
        const double beta2 = u[1] / rho;
        rho = u[1];
#pragma unroll
        for (int j = 0; j < kResRpt; ++j) {
          if (own(j)) {
            const float zi = PREC == OFH_DIAG_PREC ? r[j] * dinv(j) : r[j];
            pw(j) = zi + (float)beta2 * pw(j);
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

// ELL values from the CSR values (padding -0.0); the FP32 copy rounds
// exactly as the CSR's FP32 mirror does
__global__ void k_ell_gather(int64_t n, const int* __restrict__ src, const double* __restrict__ val,
                             double* __restrict__ sv, float* __restrict__ sv32) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int k = src[i];
    const double v = k >= 0 ? val[k] : -0.0;
    sv[i] = v;
    sv32[i] = (float)v;
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

// The ELL copy of the resident form: columns (byte offsets of the local
// column, padding -> kResPad) and the CSR entry of every slot; the values are
// gathered from the CSR values by k_ell_gather whenever they may have changed.
static int ell_build(ofh_batch B) {
  if (B->ell_col) return OFH_OK;
  // drop what an earlier, failed build left
  for (void** p : {(void**)&B->ell_e0, (void**)&B->ell_w, (void**)&B->ell_src,
                   (void**)&B->ell_val, (void**)&B->ell_val32}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  ofh_csr A = B->A;
  const int nsys = B->nsys;
  const std::vector<int>& sr = B->h_rows;
  std::vector<int> rp(A->n + 1), col(A->nnz);
  int rc = ofh_csr_get_pattern(A, rp.data(), col.data());
  if (rc) return rc;
  std::vector<int> h_col(B->ell_n, kResPad * 8), h_src(B->ell_n, -1);
  for (int k = 0; k < nsys; ++k) {
    const int a = sr[k], n_s = sr[k + 1] - a, w = B->h_ell_w[k];
    for (int l = 0; l < n_s; ++l) {
      const int64_t base = B->h_ell_e0[k] + (int64_t)(l / kResChunk) * kResChunk * w + l % kResChunk;
      for (int q = rp[a + l]; q < rp[a + l + 1]; ++q) {
        const int64_t at = base + (int64_t)kResChunk * (q - rp[a + l]);
        h_col[at] = (col[q] - a) * 8;
        h_src[at] = q;
      }
    }
  }
  hipStream_t s = A->ctx->stream;
  const size_t ne = std::max<size_t>(B->ell_n, 1);
  OFH_HIP(hipMalloc(&B->ell_e0, sizeof(int) * (nsys + 1)));
  OFH_HIP(hipMalloc(&B->ell_w, sizeof(int) * nsys));
  OFH_HIP(hipMalloc(&B->ell_src, sizeof(int) * ne));
  OFH_HIP(hipMalloc(&B->ell_val, sizeof(double) * ne));
  OFH_HIP(hipMalloc(&B->ell_val32, sizeof(float) * ne));
  OFH_HIP(hipMemcpyAsync(B->ell_e0, B->h_ell_e0.data(), sizeof(int) * (nsys + 1),
                         hipMemcpyHostToDevice, s));
  OFH_HIP(hipMemcpyAsync(B->ell_w, B->h_ell_w.data(), sizeof(int) * nsys, hipMemcpyHostToDevice, s));
  OFH_HIP(hipMemcpyAsync(B->ell_src, h_src.data(), sizeof(int) * B->ell_n, hipMemcpyHostToDevice, s));
  int* dcol = nullptr;
  OFH_HIP(hipMalloc(&dcol, sizeof(int) * ne));
  OFH_HIP(hipMemcpyAsync(dcol, h_col.data(), sizeof(int) * B->ell_n, hipMemcpyHostToDevice, s));
  OFH_HIP(hipStreamSynchronize(s));  // the host vectors are released on return
  B->ell_col = dcol;  // set last: marks the copy complete
  B->ell_version = ~0ull;
  return OFH_OK;
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
            "batched CG: the resident form needs systems of at most 12288 rows");
  if (tpr == 0 && (rc = ell_build(B))) return rc;
  if (tpr == 0 && B->ell_version != M->version) {  // refresh the ELL values
    const int g = (int)std::min<int64_t>((B->ell_n + 255) / 256, 65535);
    hipLaunchKernelGGL(k_ell_gather, dim3(std::max(g, 1)), dim3(256), 0, s, B->ell_n,
                       B->ell_src, M->val, B->ell_val, B->ell_val32);
    OFH_HIP(hipGetLastError());
    B->ell_version = M->version;
  }
  BatchArgs a{};
  a.nsys = B->nsys;
  a.sys_rows = B->sys_rows;
  a.order = B->order;
  a.counter = B->counter;
  a.rp = M->row_ptr;
  a.col = M->col;
  a.val = M->val;
  a.val32 = M->val32;
  a.ell_e0 = B->ell_e0;
  a.ell_w = B->ell_w;
  a.ell_col = B->ell_col;
  a.ell_val = B->ell_val;
  a.ell_val32 = B->ell_val32;
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
  B->h_nnz.resize(nsys);
  for (int k = 0; k < nsys; ++k) B->h_nnz[k] = rp[h_sys_rows[k + 1]] - rp[h_sys_rows[k]];
  for (int k = 0; k < nsys; ++k)
    B->max_rows = std::max(B->max_rows, h_sys_rows[k + 1] - h_sys_rows[k]);
  // ELL layout of the resident form (built on its first use, ell_build):
  // system k has width w_k (its longest row); chunk c holds local rows
  // 64c..64c+63, entry s of row `lane` at e0_k + (c w_k + s) 64 + lane
  if (B->max_rows <= kResBlock * kResRpt) {
    B->h_ell_e0.resize(nsys + 1);
    B->h_ell_w.resize(nsys);
    int64_t e = 0;
    for (int k = 0; k < nsys; ++k) {
      const int a = h_sys_rows[k], n_s = h_sys_rows[k + 1] - a;
      int w = 0;
      for (int l = 0; l < n_s; ++l) w = std::max(w, rp[a + l + 1] - rp[a + l]);
      B->h_ell_e0[k] = (int)std::min<int64_t>(e, INT32_MAX);
      B->h_ell_w[k] = w;
      e += (int64_t)((n_s + kResChunk - 1) / kResChunk) * kResChunk * w;
    }
    B->h_ell_e0[nsys] = (int)std::min<int64_t>(e, INT32_MAX);
    B->resident_ok = e < INT32_MAX;
    B->ell_n = e;
  }

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
  void* ptrs[] = {B->sys_rows, B->order,  B->counter, B->iters,   B->status,  B->relres,
                  B->work,     B->ell_e0, B->ell_w,   B->ell_col, B->ell_src, B->ell_val,
                  B->ell_val32};
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
            "ofh_batch_set_config: the resident form needs systems of at most 12288 rows");
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
  // a batch of the same shape (system sizes, nonzeros, solver variant) was
  // tuned before in this process: reuse its choice without timing again
  const std::string key = tune_key(B, prec, mixed);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    const auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      B->block = it->second.block;
      B->tpr = it->second.tpr;
      B->tuned = true;
      if (best_block) *best_block = B->block;
      if (best_tpr) *best_tpr = B->tpr;
      if (best_ms) *best_ms = it->second.ms;
      return OFH_OK;
    }
  }
  double* x0;
  OFH_HIP(hipMalloc(&x0, sizeof(double) * n));
  OFH_HIP(hipMemcpyAsync(x0, d_x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  // every return path below releases x0 and both events
  auto release = [&]() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(x0);
  };
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    release();
    return fail(OFH_E_HIP, "ofh_batch_autotune: hipEventCreate failed");
  }
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
        float ms = 0;
        hipError_t he = hipMemcpyAsync(d_x, x0, sizeof(double) * n, hipMemcpyDeviceToDevice, s);
        if (he == hipSuccess) he = hipEventRecord(e0, s);
        if (he != hipSuccess) {
          release();
          return fail(OFH_E_HIP, std::string("ofh_batch_autotune: ") + hipGetErrorString(he));
        }
        const int rc =
            batch_run(B, prec, mixed, d_b, d_x, max_it, toler, sh.first, sh.second, s, false);
        if (rc) {
          release();
          return rc;
        }
        he = hipEventRecord(e1, s);
        if (he == hipSuccess) he = hipEventSynchronize(e1);
        if (he == hipSuccess) he = hipEventElapsedTime(&ms, e0, e1);
        if (he != hipSuccess) {
          release();
          return fail(OFH_E_HIP, std::string("ofh_batch_autotune: ") + hipGetErrorString(he));
        }
        ms_min = std::min(ms_min, ms);
      }
      if (ms_min < best) {
        best = ms_min;
        bb = sh.first;
        bt = sh.second;
      }
    }
  }
  hipError_t he = hipMemcpyAsync(d_x, x0, sizeof(double) * n, hipMemcpyDeviceToDevice, s);
  if (he == hipSuccess) he = hipStreamSynchronize(s);
  release();
  if (he != hipSuccess)
    return fail(OFH_E_HIP, std::string("ofh_batch_autotune: ") + hipGetErrorString(he));
  B->block = bb;
  B->tpr = bt;
  B->tuned = true;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = TuneEntry{bb, bt, (double)best};
  }
  if (best_block) *best_block = bb;
  if (best_tpr) *best_tpr = bt;
  if (best_ms) *best_ms = best;
  return OFH_OK;
}

}  // extern "C"
