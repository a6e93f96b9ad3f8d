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
// exactly: every row is computed by one thread with the reference's
// left-to-right sums, so the result is bitwise the reference's.
//
// Schedule (synchronization-free, one launch per sweep): the rows are listed
// by level (a row's level is one more than the deepest row it reads), each
// level starting a new wave, so every row reads only rows of earlier waves.
// A persistent grid takes one-wave chunks of that list through an atomic
// ticket, so a wave only ever waits for waves that are already running.  The
// output vector is first filled with a signalling-NaN pattern that no
// arithmetic produces; a thread waits until the entries it reads no longer
// read as that pattern (64-bit agent-scope atomic loads: value and readiness
// in one word, no fences) and publishes its own result with one 64-bit atomic
// store.  The wait is bounded; a row whose inputs never arrive publishes a
// NaN, which ends the calling solver loudly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "ofh_internal.h"

namespace ofh {

constexpr int kSweepWg = 64;       // threads (list entries) per chunk of a sweep: one wave
constexpr int kSweepGridMin = 16;  // persistent workgroups of a sweep: 3 x the mean level
constexpr int kSweepGridMax = 192; // width in waves, within these bounds
constexpr int kDepBatch = 8;       // entries of a row polled together
constexpr unsigned long long kPending = 0x7FF4000000005EEDull;  // signalling NaN
constexpr int kSpinLimit = 1 << 20;  // polls before a row gives up (about a second)

// kTDiluFwd / kTDiluBwd run on the transposed matrix (Prec::TransSolve).
enum SweepOp {
  kSgsFwd = 0, kSgsBwd = 1, kDiluFact = 2, kDiluFwd = 3, kDiluBwd = 4, kTDiluFwd = 5, kTDiluBwd = 6
};
constexpr bool op_forward(int op) { return op != kSgsBwd && op != kDiluBwd && op != kTDiluBwd; }

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

__device__ __forceinline__ unsigned long long load_pub(const double* p) {
  return __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void store_pub(double* p, double v) {
  __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), __double_as_longlong(v),
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(kSweepWg) void k_fill_pending(double* __restrict__ out, int n,
                                                          const int* done) {
  if (done && *done) return;  // the sweep that follows is skipped too
  const int i = blockIdx.x * kSweepWg + threadIdx.x;
  if (i < n) reinterpret_cast<unsigned long long*>(out)[i] = kPending;
}

// One row of a sweep, in the reference's operation order.  The entries it
// reads are taken kDepBatch at a time: their published values are loaded
// together and re-polled until none reads as pending, then folded in order.
// Rows of one wave never read each other (see csr_ensure_levels), so a lane
// may simply wait.  Per op, with v the published value of the entry's column
// c and w the entry's value:
//   kSgsFwd/kDiluFwd/kSgsBwd/kDiluBwd  s = 0, s += w v (ascending c)
//   kDiluFact  s = a_ii, s -= (w v) a_ci      (a_ci: the transposed entry)
//   kTDiluFwd  s = in_i, s -= (inv_c v) w     (Prec_impl.h:273-278)
//   kTDiluBwd  s = in_i, s -= w v, descending c (Prec_impl.h:279-283)
template <int OP>
__device__ __forceinline__ void sweep_row(int i, const SweepArgs& a) {
  constexpr bool FWD = op_forward(OP);
  constexpr bool DESC = OP == kTDiluBwd;
  const int d = a.dpos[i];
  const int kb = FWD ? a.rp[i] : d + 1;
  const int ke = FWD ? d : a.rp[i + 1];
  double s = OP == kDiluFact ? a.val[d] : (OP == kTDiluFwd || OP == kTDiluBwd) ? a.in[i] : 0.0;
  int spins = 0;
  bool ok = true;
  for (int b0 = 0; b0 < ke - kb && ok; b0 += kDepBatch) {
    int c[kDepBatch];
    double w[kDepBatch], wt[kDepBatch];
#pragma unroll
    for (int j = 0; j < kDepBatch; ++j) {
      const int k = DESC ? ke - 1 - (b0 + j) : kb + b0 + j;
      c[j] = -1;
      if (b0 + j < ke - kb) {
        if (OP == kDiluFact) {
          const int t = a.tpos[k];
          if (t >= 0) {
            c[j] = a.col[k];
            wt[j] = a.val[t];
          }
        } else {
          c[j] = a.col[k];
          if (OP == kTDiluFwd) wt[j] = a.inv[c[j]];
        }
        w[j] = a.val[k];
      }
    }
    unsigned long long u[kDepBatch];
    for (;;) {
#pragma unroll
      for (int j = 0; j < kDepBatch; ++j) u[j] = c[j] >= 0 ? load_pub(a.out + c[j]) : 0ull;
      bool ready = true;
#pragma unroll
      for (int j = 0; j < kDepBatch; ++j) ready &= u[j] != kPending;
      if (ready) break;
      if (++spins > kSpinLimit) {
        ok = false;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
#pragma unroll
    for (int j = 0; j < kDepBatch; ++j) {
      if (c[j] < 0) continue;
      const double v = __longlong_as_double(u[j]);
      if (OP == kDiluFact)
        s -= w[j] * v * wt[j];
      else if (OP == kTDiluFwd)
        s -= wt[j] * v * w[j];
      else if (OP == kTDiluBwd)
        s -= w[j] * v;
      else
        s += w[j] * v;
    }
  }
  double r;
  if (OP == kDiluFact)
    r = 1.0 / s;
  else if (OP == kSgsFwd)
    r = (a.in[i] - s) / a.val[d];
  else if (OP == kDiluFwd)
    r = a.inv[i] * (a.in[i] - s);
  else if (OP == kSgsBwd)
    r = a.in[i] - s / a.val[d];
  else if (OP == kDiluBwd)
    r = a.in[i] - a.inv[i] * s;
  else if (OP == kTDiluFwd)
    r = s;
  else
    r = a.inv[i] * s;
  store_pub(a.out + i, ok ? r : __builtin_nan(""));
}

// A persistent grid: each workgroup takes chunks of kSweepWg list entries in
// ticket order until the list is exhausted.  Few workgroups in flight keep the
// waiting rows from crowding the memory pipes of the rows being computed.
// ctr[0]: next chunk ticket, ctr[1]: finished workgroups; the last workgroup
// to finish resets both, so every sweep starts from zero.
template <int OP>
__global__ __launch_bounds__(kSweepWg) void k_sweep(int nlist, const int* __restrict__ order,
                                                    int* ctr, SweepArgs a) {
  if (a.done && *a.done) return;  // read alike by every workgroup: none takes a ticket
  __shared__ int chunk;
  const int nchunk = (nlist + kSweepWg - 1) / kSweepWg;
  for (;;) {
    if (threadIdx.x == 0) chunk = atomicAdd(&ctr[0], 1);
    __syncthreads();
    const int c = chunk;
    __syncthreads();  // chunk is rewritten by the next ticket
    if (c >= nchunk) break;
    const int t = c * kSweepWg + threadIdx.x;
    const int i = t < nlist ? order[t] : -1;  // -1: padding after the last row of a level
    if (i >= 0) sweep_row<OP>(i, a);
  }
  if (threadIdx.x == 0 && atomicAdd(&ctr[1], 1) == (int)gridDim.x - 1) {
    __hip_atomic_store(&ctr[0], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&ctr[1], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Rows of the lower (forward) and upper (backward) triangles listed by level,
// computed once per pattern on the host.  Every level starts a new wave (its
// tail is padded with -1 to a multiple of 64 entries), so the rows of one
// wave never read each other: a wave waits on earlier waves only, whatever
// the compiler does with the lanes that are ready first.
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
  auto build = [&](bool fwd, std::vector<int>& list) {
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
    std::vector<int64_t> start(nl + 1, 0);  // padded start of each level in the list
    std::vector<int> width(nl, 0);
    for (int i = 0; i < n; ++i) width[lev[i]]++;
    for (int l = 0; l < nl; ++l) start[l + 1] = start[l] + (width[l] + kWave - 1) / kWave * kWave;
    OFH_CHECK(start[nl] <= INT32_MAX, OFH_E_INVALID, "SSOR/DILU sweep: too many rows");
    list.assign(start[nl], -1);
    for (int i = 0; i < n; ++i) list[start[lev[i]]++] = i;
    return nl;
  };
  std::vector<int> rf, rb;
  A->nlev_fwd = build(true, rf);
  A->nlev_bwd = build(false, rb);
  A->nlist_fwd = (int)rf.size();
  A->nlist_bwd = (int)rb.size();
  int *d_ctr = nullptr, *d_rf = nullptr, *d_rb = nullptr;
  OFH_HIP(hipMalloc(&d_ctr, sizeof(int) * 2));
  OFH_HIP(hipMalloc(&d_rf, sizeof(int) * std::max<size_t>(rf.size(), 1)));
  OFH_HIP(hipMalloc(&d_rb, sizeof(int) * std::max<size_t>(rb.size(), 1)));
  OFH_HIP(hipMemsetAsync(d_ctr, 0, sizeof(int) * 2, s));
  OFH_HIP(hipMemcpyAsync(d_rf, rf.data(), sizeof(int) * rf.size(), hipMemcpyHostToDevice, s));
  OFH_HIP(hipMemcpyAsync(d_rb, rb.data(), sizeof(int) * rb.size(), hipMemcpyHostToDevice, s));
  OFH_HIP(hipStreamSynchronize(s));  // rf, rb leave scope
  A->sweep_ctr = d_ctr;
  A->lvl_rows_bwd = d_rb;
  A->lvl_rows_fwd = d_rf;  // set last: marks the level lists complete
  return OFH_OK;
}

// One sweep of the whole matrix on stream s.  Sweeps of one matrix share its
// ticket counters, so they must not run concurrently (every caller uses the
// matrix's context stream).
template <int OP>
static int sweep(ofh_csr A, SweepArgs a, hipStream_t s) {
  const int n = A->n;
  if (n == 0) return OFH_OK;
  a.rp = A->row_ptr;
  a.col = A->col;
  a.val = A->val;
  a.dpos = A->diag_pos;
  constexpr bool FWD = op_forward(OP);
  const int nlist = FWD ? A->nlist_fwd : A->nlist_bwd;
  hipLaunchKernelGGL(k_fill_pending, dim3((n + kSweepWg - 1) / kSweepWg), dim3(kSweepWg), 0, s,
                     a.out, n, a.done);
  // Waves in flight: enough for about three levels.  Every extra wave only
  // polls, and the polling slows the rows being computed (measured on the
  // 96^3 P1 Poisson matrix, 282 levels of ~47 waves: 128-192 workgroups
  // 1.16-1.21 ms per DILU application, 512: 1.95 ms, 1024: 2.99 ms; on 2D
  // matrices with levels of ~4 waves, 16 workgroups are best).
  const int nlev = FWD ? A->nlev_fwd : A->nlev_bwd;
  const int nchunk = (nlist + kSweepWg - 1) / kSweepWg;
  const int want = std::max(kSweepGridMin, std::min(kSweepGridMax, 3 * ((nchunk + nlev - 1) / nlev)));
  const int grid = std::min(nchunk, want);
  hipLaunchKernelGGL(k_sweep<OP>, dim3(grid), dim3(kSweepWg), 0, s,
                     nlist, FWD ? A->lvl_rows_fwd : A->lvl_rows_bwd, A->sweep_ctr, a);
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

// Prec::TransSolve for ILU_PREC (Prec_impl.h:269-284): z = M^-T r, the
// forward sweep over the lower triangle of A^T (t_j -= (pivot_i z_i) a_ij,
// i ascending) and the backward sweep over its upper triangle (z_j -= a_ij
// x_i, i descending, then x_j = pivot_j z_j), with A's pivots.
int launch_dilu_trans(ofh_csr A, const double* r, double* z, const int* d_done, hipStream_t s) {
  int rc = csr_ensure_dilu(A);
  if (rc || (rc = csr_ensure_transpose(A))) return rc;
  ofh_csr T = A->tr;
  if ((rc = csr_ensure_levels(T)) || (rc = ensure_tmp(T))) return rc;
  SweepArgs a{};
  a.done = d_done;
  a.inv = A->dilu_inv;
  a.in = r;
  a.out = T->ssor_tmp;
  if ((rc = sweep<kTDiluFwd>(T, a, s))) return rc;
  a.in = T->ssor_tmp;
  a.out = z;
  return sweep<kTDiluBwd>(T, a, s);
}

}  // namespace ofh

extern "C" int ofh_ssor_solve(ofh_csr A, const double* b, double* x) {
  OFH_CHECK(A && b && x, OFH_E_INVALID, "ofh_ssor_solve: null argument");
  OFH_CHECK(b != x, OFH_E_INVALID, "ofh_ssor_solve: b and x must differ");
  return ofh::launch_ssor(A, b, x, nullptr, A->ctx->stream);
}
