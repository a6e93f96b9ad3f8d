// ofh_ssor.hip -- preconditioners applied by triangular sweeps on the GPU:
// SSOR_PREC (SpMatrix::SSORSolve, SpMatrix_impl.h:913-942) and
// DILU_PREC / ILU_PREC (SpMatrix::DILUFactorize / DILUSolve,
// SpMatrix_impl.h:834-900; the reference's ILU routines are its DILU ones).
//
// SSOR, symmetric Gauss-Seidel (omega = 1):
//   forward   z_i = (b_i - sum_{j<i} a_ij z_j) / a_ii        i = 1..n
//   backward  x_i = z_i - (sum_{j>i} a_ij x_j) / a_ii        i = n..1
// D-ILU: pivot_i = 1/D_i with D_l = a_ll - sum_{i<l} (a_li pivot_i) a_il over
// the i whose entries a_li and a_il are both present, subtracted in
// increasing i; then
//   forward   z_i = pivot_i (b_i - sum_{j<i} a_ij z_j)
//   backward  x_i = z_i - pivot_i sum_{j>i} a_ij x_j
// i.e. M = (D+L) D^-1 (D+U) in both cases.  Every sweep (the factorization
// too: row l reads the pivots of the rows of its lower part) is executed
// exactly, row by row, in dependency order: rows are grouped into level sets
// (a row's level is one more than the deepest row it reads), and every row
// of a level is computed by one thread with the reference's left-to-right
// sums, so the result is bitwise the reference's.  Runs of narrow levels are
// executed by a single persistent 1024-thread workgroup with one barrier per
// level (no launch per level); wide levels get a grid of their own.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "ofh_internal.h"

namespace ofh {

constexpr int kSgsWg = 1024;       // threads of the single-workgroup sweep
constexpr int kWideLevel = 8192;   // rows per level above which a grid is launched

enum SweepOp { kSgsFwd = 0, kSgsBwd = 1, kDiluFact = 2, kDiluFwd = 3, kDiluBwd = 4 };
constexpr bool op_forward(int op) { return op != kSgsBwd && op != kDiluBwd; }

struct SweepArgs {
  const int* __restrict__ rp;
  const int* __restrict__ col;
  const double* __restrict__ val;
  const int* __restrict__ dpos;    // position of a_ii
  const int* __restrict__ tpos;    // DILU factorization: position of a_ji, -1 if absent
  const double* __restrict__ inv;  // DILU solves: 1/D
  const double* in;
  double* out;                     // the factorization writes 1/D here
  const int* done;                 // stop flag of the calling solver (may be null)
};

// one row of a sweep (the reference's operation order)
template <int OP>
__device__ __forceinline__ void sweep_row(int i, const SweepArgs& a) {
  const int d = a.dpos[i];
  double s = 0;
  if (OP == kDiluFact) {
    double p = a.val[d];
    for (int k = a.rp[i]; k < d; ++k) {
      const int t = a.tpos[k];
      if (t >= 0) p -= a.val[k] * a.out[a.col[k]] * a.val[t];
    }
    a.out[i] = 1.0 / p;
  } else if (op_forward(OP)) {
    for (int k = a.rp[i]; k < d; ++k) s += a.val[k] * a.out[a.col[k]];
    a.out[i] = OP == kSgsFwd ? (a.in[i] - s) / a.val[d] : a.inv[i] * (a.in[i] - s);
  } else {
    for (int k = d + 1; k < a.rp[i + 1]; ++k) s += a.val[k] * a.out[a.col[k]];
    a.out[i] = OP == kSgsBwd ? a.in[i] - s / a.val[d] : a.in[i] - a.inv[i] * s;
  }
}

template <int OP>
__global__ __launch_bounds__(kSgsWg) void k_sweep_levels_wg(int l0, int l1,
                                                           const int* __restrict__ lptr,
                                                           const int* __restrict__ lrows,
                                                           SweepArgs a) {
  if (a.done && *a.done) return;
  for (int l = l0; l < l1; ++l) {
    const int b = lptr[l], e = lptr[l + 1];
    for (int t = b + threadIdx.x; t < e; t += kSgsWg) sweep_row<OP>(lrows[t], a);
    __syncthreads();
  }
}

template <int OP>
__global__ __launch_bounds__(256) void k_sweep_level_grid(int b, int e,
                                                         const int* __restrict__ lrows,
                                                         SweepArgs a) {
  if (a.done && *a.done) return;
  const int t = b + blockIdx.x * 256 + threadIdx.x;
  if (t >= e) return;
  sweep_row<OP>(lrows[t], a);
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

template <int OP>
static int sweep(ofh_csr A, SweepArgs a, hipStream_t s) {
  constexpr bool FWD = op_forward(OP);
  const int nl = FWD ? A->nlev_fwd : A->nlev_bwd;
  const int* hp = FWD ? A->h_lvl_ptr_fwd : A->h_lvl_ptr_bwd;
  const int* dp = FWD ? A->lvl_ptr_fwd : A->lvl_ptr_bwd;
  const int* dr = FWD ? A->lvl_rows_fwd : A->lvl_rows_bwd;
  a.rp = A->row_ptr;
  a.col = A->col;
  a.val = A->val;
  a.dpos = A->diag_pos;
  int l = 0;
  while (l < nl) {
    if (hp[l + 1] - hp[l] > kWideLevel) {
      const int b = hp[l], e = hp[l + 1];
      hipLaunchKernelGGL(k_sweep_level_grid<OP>, dim3((e - b + 255) / 256), dim3(256), 0, s, b, e,
                         dr, a);
      ++l;
    } else {
      int l1 = l;
      while (l1 < nl && hp[l1 + 1] - hp[l1] <= kWideLevel) ++l1;
      hipLaunchKernelGGL(k_sweep_levels_wg<OP>, dim3(1), dim3(kSgsWg), 0, s, l, l1, dp, dr, a);
      l = l1;
    }
  }
  OFH_HIP(hipGetLastError());
  return OFH_OK;
}

static int ensure_tmp(ofh_csr A) {
  if (!A->ssor_tmp) OFH_HIP(hipMalloc(&A->ssor_tmp, sizeof(double) * std::max(A->n, 1)));
  return OFH_OK;
}

// z = M^-1 r with M = (D+L) D^-1 (D+U); the forward result goes to a
// per-matrix scratch vector.
int launch_ssor(ofh_csr A, const double* r, double* z, const int* d_done, hipStream_t s) {
  int rc = csr_ensure_levels(A);
  if (rc || (rc = ensure_tmp(A))) return rc;
  SweepArgs a{};
  a.done = d_done;
  a.in = r;
  a.out = A->ssor_tmp;
  if ((rc = sweep<kSgsFwd>(A, a, s))) return rc;
  a.in = A->ssor_tmp;
  a.out = z;
  return sweep<kSgsBwd>(A, a, s);
}

// The D-ILU pivots, recomputed whenever the values may have changed.  The
// map from each lower entry a_ij to its transposed entry a_ji (the entries
// whose product enters D) depends on the pattern only and is built once.
static int csr_ensure_dilu(ofh_csr A) {
  int rc = csr_ensure_levels(A);  // diagonal positions, zero pivots, level sets
  if (rc) return rc;
  hipStream_t s = A->ctx->stream;
  if (!A->dilu_tpos) {
    const int n = A->n;
    std::vector<int> rp(n + 1), col(A->nnz), dpos(n), tpos(A->nnz, -1);
    OFH_HIP(hipMemcpyAsync(rp.data(), A->row_ptr, sizeof(int) * (n + 1), hipMemcpyDeviceToHost, s));
    OFH_HIP(hipMemcpyAsync(col.data(), A->col, sizeof(int) * A->nnz, hipMemcpyDeviceToHost, s));
    OFH_HIP(hipMemcpyAsync(dpos.data(), A->diag_pos, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    OFH_HIP(hipStreamSynchronize(s));
    for (int l = 0; l < n; ++l)
      for (int k = rp[l]; k < dpos[l]; ++k) {  // a_li, i < l: find a_il in row i's upper part
        const int i = col[k];
        const int* b = col.data() + dpos[i] + 1;
        const int* e = col.data() + rp[i + 1];
        const int* f = std::lower_bound(b, e, l);
        if (f != e && *f == l) tpos[k] = (int)(f - col.data());
      }
    int* d_tpos = nullptr;
    OFH_HIP(hipMalloc(&d_tpos, sizeof(int) * std::max<int64_t>(A->nnz, 1)));
    OFH_HIP(hipMemcpyAsync(d_tpos, tpos.data(), sizeof(int) * A->nnz, hipMemcpyHostToDevice, s));
    if (!A->dilu_inv) OFH_HIP(hipMalloc(&A->dilu_inv, sizeof(double) * std::max(n, 1)));
    OFH_HIP(hipStreamSynchronize(s));  // tpos leaves scope
    A->dilu_tpos = d_tpos;
    A->dilu_version = ~0ull;
  }
  if (A->dilu_version != A->version) {
    SweepArgs a{};
    a.tpos = A->dilu_tpos;
    a.out = A->dilu_inv;
    if ((rc = sweep<kDiluFact>(A, a, s))) return rc;
    A->dilu_version = A->version;
  }
  return OFH_OK;
}

int csr_ensure_prec(ofh_csr A, int prec) {
  if (prec == OFH_SSOR_PREC) return csr_ensure_levels(A);
  if (prec == OFH_DILU_PREC || prec == OFH_ILU_PREC) return csr_ensure_dilu(A);
  return OFH_OK;
}

// z = M^-1 r for a sweep preconditioner
int launch_prec(ofh_csr A, int prec, const double* r, double* z, const int* d_done,
                hipStream_t s) {
  if (prec == OFH_SSOR_PREC) return launch_ssor(A, r, z, d_done, s);
  int rc = csr_ensure_dilu(A);
  if (rc || (rc = ensure_tmp(A))) return rc;
  SweepArgs a{};
  a.done = d_done;
  a.inv = A->dilu_inv;
  a.in = r;
  a.out = A->ssor_tmp;
  if ((rc = sweep<kDiluFwd>(A, a, s))) return rc;
  a.in = A->ssor_tmp;
  a.out = z;
  return sweep<kDiluBwd>(A, a, s);
}

}  // namespace ofh

extern "C" int ofh_ssor_solve(ofh_csr A, const double* b, double* x) {
  OFH_CHECK(A && b && x, OFH_E_INVALID, "ofh_ssor_solve: null argument");
  OFH_CHECK(b != x, OFH_E_INVALID, "ofh_ssor_solve: b and x must differ");
  return ofh::launch_ssor(A, b, x, nullptr, A->ctx->stream);
}
