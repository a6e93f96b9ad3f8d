// ofh_proj.hip -- start vectors from the span of previous solutions.
//
// No reference routine: the reference starts each solve from zero or from
// the previous step's solution (TimeStepping.cpp:207).  For a sequence of
// solves A x_k = b_k with one SPD matrix (time steps, load cases) the
// solutions seen so far span a space that holds most of the next solution.
// A basis V of the last M new directions is kept A-orthonormal
// (V^T A V = I) together with AV; the start vector
//     x0 = V V^T b
// is then the vector of span(V) closest to A^-1 b in the A-norm (Fischer,
// "Projection techniques for iterative solution of Ax = b with successive
// right-hand sides", Comput. Methods Appl. Mech. Engrg. 163 (1998) 193-204).
// After a solve, the solution x is A-orthogonalised against V (classical
// Gram-Schmidt, two passes).  A direction whose norm is below 1e-6 of x's
// is dropped: x is already in the span up to the previous solves' residual
// noise, and Gram-Schmidt could not keep such a direction orthogonal.
// Otherwise the direction is A-normalised and appended, and x joins a ring of
// the last M accepted solutions; when V is full, it is rebuilt from the last
// M/2 - 1 of them and x.  The span therefore always contains the newest
// solution (up to that noise): the guess is never worse than the warm start
// in the A-norm.
// Minimum-residual mode: the same with the Euclidean inner product of AV
// (W = AV orthonormal): x0 = V W^T b minimises ||b - A x0||_2 over span(V),
// the quantity of the solvers' stop test (Fischer's second variant).
//
// Device: the basis is M x n FP64 (M in {8, 16, 32}; unused slots are zero
// vectors).  Every coefficient set comes from per-workgroup partial dot
// products that the consuming kernel folds in fixed order (no atomics).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "ofh_internal.h"

struct ofh_proj_s {
  ofh_csr A = nullptr;
  int M = 0;       // basis slots
  int mode = 0;    // 0: A-norm (Galerkin), 1: minimum residual
  int count = 0;   // directions stored so far (<= M)
  int head = 0;    // slot of the next direction
  int n = 0;
  int nparts = 0;
  double* V = nullptr;     // M x n
  double* AV = nullptr;    // M x n
  double* X = nullptr;     // M x n ring of the last solutions
  int xhead = 0;           // ring slot of the next solution
  int xcount = 0;
  double* d = nullptr;     // n: direction under construction
  double* Ad = nullptr;    // n
  double* part = nullptr;  // (M + 1) x nparts partial sums
  double* part2 = nullptr; // 2 x nparts
  int* flag = nullptr;     // device: direction accepted
  int* h_flag = nullptr;   // pinned host copy
};

namespace ofh {
namespace pj {

constexpr int kVecP = 4;  // entries per thread

// part[k][blk] = sum over the block's entries of y_i W[k][i], k < M.
template <int M>
__global__ __launch_bounds__(kBlock) void k_pj_mdot(int n, const double* __restrict__ y,
                                                    const double* __restrict__ W,
                                                    double* __restrict__ part, int nparts) {
  __shared__ double red[M * kBlock / kWave];
  double acc[M];
#pragma unroll
  for (int k = 0; k < M; ++k) acc[k] = 0.0;
  const int base = blockIdx.x * (kBlock * kVecP) + threadIdx.x;
#pragma unroll
  for (int j = 0; j < kVecP; ++j) {
    const int i = base + j * kBlock;
    if (i < n) {
      const double yi = y[i];
#pragma unroll
      for (int k = 0; k < M; ++k) acc[k] += yi * W[(size_t)k * n + i];
    }
  }
  block_sum_k<kBlock, M>(acc, red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < M; ++k) part[(size_t)k * nparts + blockIdx.x] = acc[k];
  }
}

// c[k] = sum_p part[k][p] (every workgroup folds the same values).
template <int M>
__device__ __forceinline__ void fold_coef(const double* __restrict__ part, int nparts,
                                          double (&c)[M], double* red) {
#pragma unroll
  for (int k = 0; k < M; ++k) {
    double s = 0.0;
    for (int p = threadIdx.x; p < nparts; p += kBlock) s += part[(size_t)k * nparts + p];
    c[k] = s;
  }
  block_sum_k<kBlock, M>(c, red);
}

// x0 = V c with c = V^T b folded from the partials.
template <int M>
__global__ __launch_bounds__(kBlock) void k_pj_guess(int n, const double* __restrict__ part,
                                                     int nparts, const double* __restrict__ V,
                                                     double* __restrict__ x0) {
  __shared__ double red[M * kBlock / kWave];
  double c[M];
  fold_coef<M>(part, nparts, c, red);
  const int base = blockIdx.x * (kBlock * kVecP) + threadIdx.x;
#pragma unroll
  for (int j = 0; j < kVecP; ++j) {
    const int i = base + j * kBlock;
    if (i < n) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < M; ++k) s += c[k] * V[(size_t)k * n + i];
      x0[i] = s;
    }
  }
}

// One Gram-Schmidt pass: c folded from the partials; d = src - V c,
// Ad -= AV c.  Partials (squared norms in the mode's inner product): pass 1
// stores that of x (the values before the update) in part2[0], pass 2 that
// of d in part2[1]: x.Ax and d.Ad (MR = 0) or Ax.Ax and Ad.Ad (MR = 1).
template <int M, int PASS, int MR>
__global__ __launch_bounds__(kBlock) void k_pj_orth(int n, const double* __restrict__ part,
                                                    int nparts, const double* __restrict__ V,
                                                    const double* __restrict__ AV,
                                                    const double* __restrict__ src,
                                                    double* __restrict__ d, double* __restrict__ Ad,
                                                    double* __restrict__ part2) {
  __shared__ double red[M * kBlock / kWave];
  double c[M];
  fold_coef<M>(part, nparts, c, red);
  double v[1] = {0.0};
  const int base = blockIdx.x * (kBlock * kVecP) + threadIdx.x;
#pragma unroll
  for (int j = 0; j < kVecP; ++j) {
    const int i = base + j * kBlock;
    if (i < n) {
      const double si = src[i], ai = Ad[i];
      double sv = 0.0, sa = 0.0;
#pragma unroll
      for (int k = 0; k < M; ++k) {
        sv += c[k] * V[(size_t)k * n + i];
        sa += c[k] * AV[(size_t)k * n + i];
      }
      const double di = si - sv, adi = ai - sa;
      d[i] = di;
      Ad[i] = adi;
      if (MR)
        v[0] += PASS == 1 ? ai * ai : adi * adi;
      else
        v[0] += PASS == 1 ? si * ai : di * adi;
    }
  }
  block_sum_k<kBlock, 1>(v, red);
  if (threadIdx.x == 0) part2[(PASS - 1) * nparts + blockIdx.x] = v[0];
}

// Folds |x|^2 and |d|^2 (mode norm).  A direction with |d|^2 > 1e-12 |x|^2
// is accepted: d/|d| and Ad/|d| go to the slot (when Vs is not null) and
// block 0 sets *accepted; otherwise nothing is written but *accepted = 0.
__global__ __launch_bounds__(kBlock) void k_pj_insert(int n, const double* __restrict__ part2,
                                                      int nparts, const double* __restrict__ d,
                                                      const double* __restrict__ Ad,
                                                      double* __restrict__ Vs,
                                                      double* __restrict__ AVs, int* accepted) {
  __shared__ double red[2 * kBlock / kWave];
  double t[2] = {0.0, 0.0};
  for (int p = threadIdx.x; p < nparts; p += kBlock) {
    t[0] += part2[p];
    t[1] += part2[nparts + p];
  }
  block_sum_k<kBlock, 2>(t, red);
  const double xax = t[0], dad = t[1];
  const bool ok = dad > 1e-12 * xax && dad > 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) *accepted = ok ? 1 : 0;
  if (!ok || !Vs) return;
  const double scale = 1.0 / sqrt(dad);
  const int base = blockIdx.x * (kBlock * kVecP) + threadIdx.x;
#pragma unroll
  for (int j = 0; j < kVecP; ++j) {
    const int i = base + j * kBlock;
    if (i < n) {
      Vs[i] = d[i] * scale;
      AVs[i] = Ad[i] * scale;
    }
  }
}

__global__ void k_pj_zero(size_t count, double* p) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0.0;
}

template <int M>
static int guess_m(ofh_proj P, const double* b, double* x0, hipStream_t s) {
  const int g = P->nparts;
  hipLaunchKernelGGL(k_pj_mdot<M>, dim3(g), dim3(kBlock), 0, s, P->n, b,
                     P->mode ? P->AV : P->V, P->part, g);
  hipLaunchKernelGGL(k_pj_guess<M>, dim3(g), dim3(kBlock), 0, s, P->n, P->part, g, P->V, x0);
  OFH_HIP(hipGetLastError());
  return OFH_OK;
}

// Orthogonalises x against the basis and appends it to slot `head` if its
// new direction is not negligible (dry: only decides).  *accepted on return.
template <int M>
static int insert_m(ofh_proj P, const double* x, hipStream_t s, bool dry, bool* accepted) {
  const int g = P->nparts, n = P->n;
  int rc = launch_spmv64(P->A, x, P->Ad, s);
  if (rc) return rc;
  // coefficients: A-norm (AV)^T x, or Euclidean (AV)^T (A x)
  hipLaunchKernelGGL(k_pj_mdot<M>, dim3(g), dim3(kBlock), 0, s, n, P->mode ? P->Ad : x, P->AV,
                     P->part, g);
  if (P->mode)
    hipLaunchKernelGGL((k_pj_orth<M, 1, 1>), dim3(g), dim3(kBlock), 0, s, n, P->part, g, P->V,
                       P->AV, x, P->d, P->Ad, P->part2);
  else
    hipLaunchKernelGGL((k_pj_orth<M, 1, 0>), dim3(g), dim3(kBlock), 0, s, n, P->part, g, P->V,
                       P->AV, x, P->d, P->Ad, P->part2);
  hipLaunchKernelGGL(k_pj_mdot<M>, dim3(g), dim3(kBlock), 0, s, n, P->mode ? P->Ad : P->d, P->AV,
                     P->part, g);
  if (P->mode)
    hipLaunchKernelGGL((k_pj_orth<M, 2, 1>), dim3(g), dim3(kBlock), 0, s, n, P->part, g, P->V,
                       P->AV, P->d, P->d, P->Ad, P->part2);
  else
    hipLaunchKernelGGL((k_pj_orth<M, 2, 0>), dim3(g), dim3(kBlock), 0, s, n, P->part, g, P->V,
                       P->AV, P->d, P->d, P->Ad, P->part2);
  double* Vs = dry ? nullptr : P->V + (size_t)P->head * n;
  double* AVs = dry ? nullptr : P->AV + (size_t)P->head * n;
  hipLaunchKernelGGL(k_pj_insert, dim3(g), dim3(kBlock), 0, s, n, P->part2, g, P->d, P->Ad, Vs,
                     AVs, P->flag);
  OFH_HIP(hipGetLastError());
  OFH_HIP(hipMemcpyAsync(P->h_flag, P->flag, sizeof(int), hipMemcpyDeviceToHost, s));
  OFH_HIP(hipStreamSynchronize(s));
  *accepted = *P->h_flag != 0;
  if (*accepted && !dry) {
    P->head = P->head + 1;
    P->count = P->count + 1;
  }
  return OFH_OK;
}

__global__ void k_pj_copy(int n, const double* __restrict__ x, double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = x[i];
}

static int zero_basis(ofh_proj P, hipStream_t s) {
  const size_t cnt = (size_t)P->M * P->n;
  hipLaunchKernelGGL(k_pj_zero, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, cnt, P->V);
  hipLaunchKernelGGL(k_pj_zero, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, cnt, P->AV);
  OFH_HIP(hipGetLastError());
  P->count = 0;
  P->head = 0;
  return OFH_OK;
}

template <int M>
static int update_m(ofh_proj P, const double* x, hipStream_t s) {
  const int n = P->n;
  bool acc = false;
  int rc;
  if (P->count < M) {
    if ((rc = insert_m<M>(P, x, s, false, &acc))) return rc;
  } else {
    // full: if x brings a new direction, rebuild from the last M/2 - 1
    // accepted solutions (oldest first) and x
    if ((rc = insert_m<M>(P, x, s, true, &acc))) return rc;
    if (acc) {
      if ((rc = zero_basis(P, s))) return rc;
      const int R = std::min(M / 2 - 1, P->xcount);
      for (int j = R; j >= 1; --j) {
        const int k = ((P->xhead - j) % M + M) % M;
        bool a2 = false;
        if ((rc = insert_m<M>(P, P->X + (size_t)k * n, s, false, &a2))) return rc;
      }
      if ((rc = insert_m<M>(P, x, s, false, &acc))) return rc;
    }
  }
  if (acc) {  // remember the solution for a later rebuild
    hipLaunchKernelGGL(k_pj_copy, dim3((n + 255) / 256), dim3(256), 0, s, n, x,
                       P->X + (size_t)P->xhead * n);
    OFH_HIP(hipGetLastError());
    P->xhead = (P->xhead + 1) % M;
    P->xcount = std::min(P->xcount + 1, M);
  }
  return OFH_OK;
}

}  // namespace pj
}  // namespace ofh

using namespace ofh;
using namespace ofh::pj;

extern "C" {

int ofh_proj_create(ofh_csr A, int max_vectors, int mode, ofh_proj* out) {
  OFH_CHECK(A && out, OFH_E_INVALID, "ofh_proj_create: bad argument");
  OFH_CHECK(mode == 0 || mode == 1, OFH_E_INVALID, "ofh_proj_create: mode is 0 or 1");
  OFH_CHECK(max_vectors == 8 || max_vectors == 16 || max_vectors == 32, OFH_E_INVALID,
            "ofh_proj_create: max_vectors is 8, 16 or 32");
  OFH_CHECK(A->n > 0 && A->n <= kMaxRows, OFH_E_INVALID, "ofh_proj_create: matrix size");
  ofh_proj P = new ofh_proj_s();
  P->A = A;
  P->M = max_vectors;
  P->mode = mode;
  P->n = A->n;
  P->nparts = (A->n + kBlock * kVecP - 1) / (kBlock * kVecP);
  const size_t n = A->n;
  hipError_t e = hipMalloc(&P->V, sizeof(double) * n * P->M);
  if (e == hipSuccess) e = hipMalloc(&P->AV, sizeof(double) * n * P->M);
  if (e == hipSuccess) e = hipMalloc(&P->X, sizeof(double) * n * P->M);
  if (e == hipSuccess) e = hipMalloc(&P->d, sizeof(double) * n);
  if (e == hipSuccess) e = hipMalloc(&P->Ad, sizeof(double) * n);
  if (e == hipSuccess) e = hipMalloc(&P->part, sizeof(double) * (P->M + 1) * P->nparts);
  if (e == hipSuccess) e = hipMalloc(&P->part2, sizeof(double) * 2 * P->nparts);
  if (e == hipSuccess) e = hipMalloc(&P->flag, sizeof(int));
  if (e == hipSuccess) e = hipHostMalloc((void**)&P->h_flag, sizeof(int), hipHostMallocDefault);
  if (e != hipSuccess) {
    void* ptrs[] = {P->V, P->AV, P->X, P->d, P->Ad, P->part, P->part2, P->flag};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    if (P->h_flag) (void)hipHostFree(P->h_flag);
    delete P;
    return fail(OFH_E_NOMEM, std::string("ofh_proj_create: ") + hipGetErrorString(e));
  }
  *out = P;
  return ofh_proj_reset(P);
}

int ofh_proj_destroy(ofh_proj P) {
  if (!P) return OFH_OK;
  (void)hipDeviceSynchronize();
  void* ptrs[] = {P->V, P->AV, P->X, P->d, P->Ad, P->part, P->part2, P->flag};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (P->h_flag) (void)hipHostFree(P->h_flag);
  delete P;
  return OFH_OK;
}

int ofh_proj_reset(ofh_proj P) {
  OFH_CHECK(P, OFH_E_INVALID, "null projection");
  P->xhead = 0;
  P->xcount = 0;
  return zero_basis(P, P->A->ctx->stream);
}

int ofh_proj_size(ofh_proj P, int* count) {
  OFH_CHECK(P && count, OFH_E_INVALID, "ofh_proj_size: bad argument");
  *count = P->count;
  return OFH_OK;
}

int ofh_proj_guess(ofh_proj P, const double* d_b, double* d_x0) {
  OFH_CHECK(P && d_b && d_x0, OFH_E_INVALID, "ofh_proj_guess: bad argument");
  hipStream_t s = P->A->ctx->stream;
  switch (P->M) {
    case 8: return guess_m<8>(P, d_b, d_x0, s);
    case 16: return guess_m<16>(P, d_b, d_x0, s);
    default: return guess_m<32>(P, d_b, d_x0, s);
  }
}

int ofh_proj_update(ofh_proj P, const double* d_x) {
  OFH_CHECK(P && d_x, OFH_E_INVALID, "ofh_proj_update: bad argument");
  hipStream_t s = P->A->ctx->stream;
  int rc;
  switch (P->M) {
    case 8: rc = update_m<8>(P, d_x, s); break;
    case 16: rc = update_m<16>(P, d_x, s); break;
    default: rc = update_m<32>(P, d_x, s); break;
  }
  return rc;
}

}  // extern "C"
