// ofh_ssor.hip -- SSOR_PREC on the GPU: SpMatrix::SSORSolve (SpMatrix_impl.h:913-942).
//
// The reference applies symmetric Gauss-Seidel (omega = 1):
//   forward   z_i = (b_i - sum_{j<i} a_ij z_j) / a_ii        i = 1..n
//   backward  x_i = z_i - (sum_{j>i} a_ij x_j) / a_ii        i = n..1
// i.e. M = (D+L) D^-1 (D+U).  Both sweeps are executed exactly, row by row,
// in dependency order: rows are grouped into level sets (a row's level is one
// more than the deepest row it reads), and every row of a level is computed
// by one thread with the reference's left-to-right sums, so the result is
// bitwise the reference's.  Runs of narrow levels are executed by a single
// persistent 1024-thread workgroup with one barrier per level (no launch per
// level); wide levels get a grid of their own.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "ofh_internal.h"

namespace ofh {

constexpr int kSgsWg = 1024;       // threads of the single-workgroup sweep
constexpr int kWideLevel = 8192;   // rows per level above which a grid is launched

// one row of the forward sweep
__device__ __forceinline__ void sgs_fwd_row(int i, const int* __restrict__ rp,
                                            const int* __restrict__ col,
                                            const double* __restrict__ val,
                                            const int* __restrict__ dpos,
                                            const double* __restrict__ b, double* z) {
  double s = 0;
  const int d = dpos[i];
  for (int k = rp[i]; k < d; ++k) s += val[k] * z[col[k]];
  z[i] = (b[i] - s) / val[d];
}

__device__ __forceinline__ void sgs_bwd_row(int i, const int* __restrict__ rp,
                                            const int* __restrict__ col,
                                            const double* __restrict__ val,
                                            const int* __restrict__ dpos,
                                            const double* __restrict__ z, double* x) {
  double s = 0;
  const int d = dpos[i];
  for (int k = d + 1; k < rp[i + 1]; ++k) s += val[k] * x[col[k]];
  x[i] = z[i] - s / val[d];
}

template <bool FWD>
__global__ __launch_bounds__(kSgsWg) void k_sgs_levels_wg(
    int l0, int l1, const int* __restrict__ lptr, const int* __restrict__ lrows,
    const int* __restrict__ rp, const int* __restrict__ col, const double* __restrict__ val,
    const int* __restrict__ dpos, const double* __restrict__ in, double* out,
    const int* __restrict__ done) {
  if (done && *done) return;
  for (int l = l0; l < l1; ++l) {
    const int a = lptr[l], e = lptr[l + 1];
    for (int t = a + threadIdx.x; t < e; t += kSgsWg) {
      const int i = lrows[t];
      if (FWD) sgs_fwd_row(i, rp, col, val, dpos, in, out);
      else sgs_bwd_row(i, rp, col, val, dpos, in, out);
    }
    __syncthreads();
  }
}

template <bool FWD>
__global__ __launch_bounds__(256) void k_sgs_level_grid(
    int a, int e, const int* __restrict__ lrows, const int* __restrict__ rp,
    const int* __restrict__ col, const double* __restrict__ val, const int* __restrict__ dpos,
    const double* __restrict__ in, double* out, const int* __restrict__ done) {
  if (done && *done) return;
  const int t = a + blockIdx.x * 256 + threadIdx.x;
  if (t >= e) return;
  const int i = lrows[t];
  if (FWD) sgs_fwd_row(i, rp, col, val, dpos, in, out);
  else sgs_bwd_row(i, rp, col, val, dpos, in, out);
}

// Level sets of the lower (forward) and upper (backward) triangles, computed
// once per pattern on the host.
int csr_ensure_levels(ofh_csr A) {
  if (A->lvl_rows_fwd) return OFH_OK;
  int rc = csr_ensure_dinv(A);  // also locates the diagonal; zero pivot -> error
  if (rc) return rc;
  const int n = A->n;
  hipStream_t s = A->ctx->stream;
  std::vector<int> rp(n + 1), col(A->nnz), dpos(n);
  OFH_HIP(hipMemcpyAsync(rp.data(), A->row_ptr, sizeof(int) * (n + 1), hipMemcpyDeviceToHost, s));
  OFH_HIP(hipMemcpyAsync(col.data(), A->col, sizeof(int) * A->nnz, hipMemcpyDeviceToHost, s));
  OFH_HIP(hipMemcpyAsync(dpos.data(), A->diag_pos, sizeof(int) * n, hipMemcpyDeviceToHost, s));
  OFH_HIP(hipStreamSynchronize(s));
  for (int i = 0; i < n; ++i)
    OFH_CHECK(dpos[i] >= 0, OFH_E_ZERO_PIVOT,
              "In SpMatrix::SSORSolve(b,x): Zero pivot detected in row " + std::to_string(i + 1));
  auto build = [&](bool fwd, std::vector<int>& ptr, std::vector<int>& rows) {
    std::vector<int> lev(n, 0);
    int nl = 0;
    if (fwd) {
      for (int i = 0; i < n; ++i) {
        int l = 0;
        for (int k = rp[i]; k < dpos[i]; ++k) l = std::max(l, lev[col[k]] + 1);
        lev[i] = l;
        nl = std::max(nl, l + 1);
      }
    } else {
      for (int i = n - 1; i >= 0; --i) {
        int l = 0;
        for (int k = dpos[i] + 1; k < rp[i + 1]; ++k) l = std::max(l, lev[col[k]] + 1);
        lev[i] = l;
        nl = std::max(nl, l + 1);
      }
    }
    ptr.assign(nl + 1, 0);
    for (int i = 0; i < n; ++i) ptr[lev[i] + 1]++;
    for (int l = 0; l < nl; ++l) ptr[l + 1] += ptr[l];
    rows.resize(n);
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int i = 0; i < n; ++i) rows[fill[lev[i]]++] = i;
    return nl;
  };
  std::vector<int> pf, rf, pb, rb;
  A->nlev_fwd = build(true, pf, rf);
  A->nlev_bwd = build(false, pb, rb);
  int w = 0;
  for (int l = 0; l < A->nlev_fwd; ++l) w = std::max(w, pf[l + 1] - pf[l]);
  for (int l = 0; l < A->nlev_bwd; ++l) w = std::max(w, pb[l + 1] - pb[l]);
  A->max_lev_width = w;
  OFH_HIP(hipMalloc(&A->lvl_ptr_fwd, sizeof(int) * pf.size()));
  OFH_HIP(hipMalloc(&A->lvl_rows_fwd, sizeof(int) * n));
  OFH_HIP(hipMalloc(&A->lvl_ptr_bwd, sizeof(int) * pb.size()));
  OFH_HIP(hipMalloc(&A->lvl_rows_bwd, sizeof(int) * n));
  OFH_HIP(hipMemcpyAsync(A->lvl_ptr_fwd, pf.data(), sizeof(int) * pf.size(), hipMemcpyHostToDevice, s));
  OFH_HIP(hipMemcpyAsync(A->lvl_rows_fwd, rf.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
  OFH_HIP(hipMemcpyAsync(A->lvl_ptr_bwd, pb.data(), sizeof(int) * pb.size(), hipMemcpyHostToDevice, s));
  OFH_HIP(hipMemcpyAsync(A->lvl_rows_bwd, rb.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
  A->h_lvl_ptr_fwd = new int[pf.size()];
  A->h_lvl_ptr_bwd = new int[pb.size()];
  std::copy(pf.begin(), pf.end(), A->h_lvl_ptr_fwd);
  std::copy(pb.begin(), pb.end(), A->h_lvl_ptr_bwd);
  OFH_HIP(hipStreamSynchronize(s));
  return OFH_OK;
}

template <bool FWD>
static int sweep(ofh_csr A, const double* in, double* out, const int* done, hipStream_t s) {
  const int nl = FWD ? A->nlev_fwd : A->nlev_bwd;
  const int* hp = FWD ? A->h_lvl_ptr_fwd : A->h_lvl_ptr_bwd;
  const int* dp = FWD ? A->lvl_ptr_fwd : A->lvl_ptr_bwd;
  const int* dr = FWD ? A->lvl_rows_fwd : A->lvl_rows_bwd;
  int l = 0;
  while (l < nl) {
    if (hp[l + 1] - hp[l] > kWideLevel) {
      const int a = hp[l], e = hp[l + 1];
      hipLaunchKernelGGL(k_sgs_level_grid<FWD>, dim3((e - a + 255) / 256), dim3(256), 0, s, a, e,
                         dr, A->row_ptr, A->col, A->val, A->diag_pos, in, out, done);
      ++l;
    } else {
      int l1 = l;
      while (l1 < nl && hp[l1 + 1] - hp[l1] <= kWideLevel) ++l1;
      hipLaunchKernelGGL(k_sgs_levels_wg<FWD>, dim3(1), dim3(kSgsWg), 0, s, l, l1, dp, dr,
                         A->row_ptr, A->col, A->val, A->diag_pos, in, out, done);
      l = l1;
    }
  }
  OFH_HIP(hipGetLastError());
  return OFH_OK;
}

// z = M^-1 r with M = (D+L) D^-1 (D+U); the forward result goes to a
// scratch vector placed after the four solver vectors.
int launch_ssor(ofh_csr A, const double* r, double* z, const int* d_done, hipStream_t s) {
  int rc = csr_ensure_levels(A);
  if (rc) return rc;
  double* tmp = A->ssor_tmp;
  if (!tmp) {
    OFH_HIP(hipMalloc(&A->ssor_tmp, sizeof(double) * A->n));
    tmp = A->ssor_tmp;
  }
  if ((rc = sweep<true>(A, r, tmp, d_done, s))) return rc;
  return sweep<false>(A, tmp, z, d_done, s);
}

}  // namespace ofh

extern "C" int ofh_ssor_solve(ofh_csr A, const double* b, double* x) {
  OFH_CHECK(A && b && x, OFH_E_INVALID, "ofh_ssor_solve: null argument");
  OFH_CHECK(b != x, OFH_E_INVALID, "ofh_ssor_solve: b and x must differ");
  return ofh::launch_ssor(A, b, x, nullptr, A->ctx->stream);
}
