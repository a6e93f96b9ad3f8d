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
// Schedule (one launch per sweep): rows are listed by level (a row's level is
// one more than the deepest row it reads), and consecutive levels are packed
// into chunks of at most 1024 rows (a level wider than that is split over
// several chunks of its own).  A persistent grid takes chunks through an
// atomic ticket, so a workgroup only ever waits for chunks already taken.
// A workgroup loads every row of its chunk at once, then runs the chunk's
// levels in order, one barrier apart: an entry whose column lies in the same
// chunk is read from LDS, where an earlier level of the chunk left it; an
// entry from an earlier chunk is read from the output vector, which is first
// filled with a signalling-NaN pattern that no arithmetic produces: the
// thread waits until the entry no longer reads as that pattern (64-bit
// agent-scope atomic loads: value and readiness in one word, no fences).
// Every result is published with one 64-bit atomic store.  A chain of narrow
// levels thus costs one barrier per level inside a chunk and one hand-off
// through L2 per chunk.  The wait is bounded; a row whose inputs never
// arrive publishes a NaN, which ends the calling solver loudly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "ofh_internal.h"

namespace ofh {

// Rows of a chunk (one per thread of a workgroup): 512 where levels are
// narrow, so that a chunk holds several levels and the hand-offs through L2
// are few; 64 where they are wide (a level spans many chunks, and small
// chunks keep many workgroups waiting on few rows each).
constexpr int kChunkNarrow = 512;
constexpr int kChunkWide = 64;
constexpr int kWideLevel = 256;  // mean level width from which chunks are kChunkWide rows
constexpr int kLagNarrow = 1;    // chunks a workgroup may run ahead of the finished ones
constexpr int kLagWide = 0;      // (0: no throttle) before it polls
// Entries of a row loaded when its chunk is taken (the lower or upper part
// of a 2D Q4 elasticity row has up to 9, of a 3D P1 row up to 7), and
// further entries, loaded and polled together, by chunk size.
template <int K> constexpr int kPre = K == 64 ? 8 : 12;
template <int K> constexpr int kDepBatch = K == 64 ? 4 : 8;
constexpr int kSweepGridMin = 16;  // persistent workgroups of a sweep: at least this many,
constexpr int kSweepGridMax = 192; // at most this many
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

__global__ __launch_bounds__(256) void k_fill_pending(double* __restrict__ out, int n,
                                                     const int* done) {
  if (done && *done) return;  // the sweep that follows is skipped too
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) reinterpret_cast<unsigned long long*>(out)[i] = kPending;
}

// Entry k of a row (in the row's processing order) for op OP: its column c
// (-1: an entry the op skips, or use = false), the LDS slot of that
// column's value when it lies in the row's chunk (lc >= 0), the entry's
// value w and, for the factorization and the transposed forward sweep, the
// second factor wt.  k must be a valid entry even when use is false; every
// load is issued unconditionally, so the loads of a row's entries overlap.
template <int OP>
__device__ __forceinline__ void sweep_entry(const SweepArgs& a, const int* __restrict__ loc,
                                            int k, bool use, int& c, int& lc, double& w,
                                            double& wt) {
  const int col = a.col[k];
  const int l = loc[k];
  w = a.val[k];
  wt = 0.0;
  if (OP == kDiluFact) {
    const int t = a.tpos[k];
    wt = a.val[t >= 0 ? t : k];
    use = use && t >= 0;
  }
  if (OP == kTDiluFwd) wt = a.inv[col];
  c = use ? col : -1;
  lc = use ? l : -1;
}

// The term entry (w, wt) contributes with column value v: the product the
// row's sum adds (sweeps) or subtracts (factorization, transposed sweeps).
template <int OP>
__device__ __forceinline__ double sweep_term(double w, double wt, double v) {
  if (OP == kDiluFact) return w * v * wt;
  if (OP == kTDiluFwd) return wt * v * w;
  return w * v;
}
constexpr bool op_subtracts(int op) { return op == kDiluFact || op == kTDiluFwd || op == kTDiluBwd; }

// Waits until the first P entries of a row that lie in earlier chunks are
// published and turns each into its term (term[j]); false when the wait
// gives up.  Run when the chunk is taken, so the chunk's levels read LDS only.
template <int OP, int P>
__device__ __forceinline__ bool sweep_remote(const SweepArgs& a, const int (&c)[P],
                                             const int (&lc)[P], const double (&w)[P],
                                             const double (&wt)[P], double (&term)[P],
                                             int& spins) {
  unsigned long long u[P];
  for (;;) {
    bool ready = true;
#pragma unroll
    for (int j = 0; j < P; ++j) {
      const bool remote = c[j] >= 0 && lc[j] < 0;
      u[j] = load_pub(a.out + (remote ? c[j] : 0));
      ready &= !remote || u[j] != kPending;
    }
    if (ready) break;
    if (++spins > kSpinLimit) return false;
    __builtin_amdgcn_s_sleep(1);
  }
#pragma unroll
  for (int j = 0; j < P; ++j) term[j] = sweep_term<OP>(w[j], wt[j], __longlong_as_double(u[j]));
  return true;
}

// Waits until the batch's entries from earlier chunks are published (the
// LDS ones are: they were written before a barrier), then folds the batch,
// in order, into s.  Per op, with v the value of the entry's column:
//   kSgsFwd/kDiluFwd/kSgsBwd/kDiluBwd  s += w v (ascending c)
//   kDiluFact  s -= (w v) wt                 (wt = a_ci, the transposed entry)
//   kTDiluFwd  s -= (wt v) w                 (wt = inv_c, Prec_impl.h:273-278)
//   kTDiluBwd  s -= w v, descending c        (Prec_impl.h:279-283)
template <int OP, int B>
__device__ __forceinline__ bool sweep_fold(const SweepArgs& a, const double* sval, const int (&c)[B],
                                           const int (&lc)[B], const double (&w)[B],
                                           const double (&wt)[B], double& s, int& spins) {
  unsigned long long u[B];
  for (;;) {
    bool ready = true;
#pragma unroll
    for (int j = 0; j < B; ++j) {
      u[j] = c[j] < 0 ? 0ull
             : lc[j] >= 0 ? (unsigned long long)__double_as_longlong(sval[lc[j]])
                          : load_pub(a.out + c[j]);
      ready &= u[j] != kPending;
    }
    if (ready) break;
    if (++spins > kSpinLimit) return false;
    __builtin_amdgcn_s_sleep(1);
  }
#pragma unroll
  for (int j = 0; j < B; ++j) {
    if (c[j] < 0) continue;
    const double x = sweep_term<OP>(w[j], wt[j], __longlong_as_double(u[j]));
    if (op_subtracts(OP))
      s -= x;
    else
      s += x;
  }
  return true;
}

// A persistent grid: each workgroup takes chunks in ticket order until none
// is left; see the schedule above.  chunks[c] = {first list entry, rows,
// levels}.  ctr[0]: next chunk ticket, ctr[1]: finished workgroups, ctr[2]:
// finished chunks; the last
// workgroup to finish resets both, so every sweep starts from zero.
template <int OP, int K>
__global__ __launch_bounds__(K) void k_sweep(int nchunk, const int4* __restrict__ chunks,
                                                      const int* __restrict__ list,
                                                      const int* __restrict__ ilev,
                                                      const int* __restrict__ loc, int* ctr,
                                                      int lag, SweepArgs a) {
  constexpr bool FWD = op_forward(OP);
  constexpr bool DESC = OP == kTDiluBwd;
  constexpr int P = kPre<K>, B = kDepBatch<K>;
  if (a.done && *a.done) return;  // read alike by every workgroup: none takes a ticket
  __shared__ double sval[K];  // results of the chunk's rows, by position
  __shared__ int ticket;
  OFH_LDS_BUDGET(sizeof(double) * K + sizeof(int));
  const int t = threadIdx.x;
  for (;;) {
    if (t == 0) ticket = atomicAdd(&ctr[0], 1);
    __syncthreads();
    const int ch_id = ticket;
    __syncthreads();  // ticket is rewritten by the next chunk
    if (ch_id >= nchunk) break;
    const int4 ch = chunks[ch_id];
    const int nlev = __builtin_amdgcn_readfirstlane(ch.z);
    // the row of this thread and its first P entries, loaded at once
    int lv = -1, i = 0, kb = 0, ne = 0, d = 0;
    int c[P], lc[P];
    double w[P], wt[P];
    double s = 0.0, in_i = 0.0, piv = 0.0;
    if (t < ch.y) {
      lv = ilev[ch.x + t];
      i = list[ch.x + t];
      d = a.dpos[i];
      kb = FWD ? a.rp[i] : d + 1;
      ne = (FWD ? d : a.rp[i + 1]) - kb;
#pragma unroll
      for (int j = 0; j < P; ++j) {  // past the row's entries: the diagonal, unused
        const bool use = j < ne;
        sweep_entry<OP>(a, loc, !use ? d : DESC ? kb + ne - 1 - j : kb + j, use, c[j], lc[j], w[j],
                        wt[j]);
      }
      if (OP == kDiluFact) s = a.val[d];
      if (OP == kTDiluFwd || OP == kTDiluBwd) s = a.in[i];
      if (OP == kSgsFwd || OP == kDiluFwd || OP == kSgsBwd || OP == kDiluBwd) in_i = a.in[i];
      if (OP == kSgsFwd || OP == kSgsBwd) piv = a.val[d];
      if (OP == kDiluFwd || OP == kDiluBwd || OP == kTDiluBwd) piv = a.inv[i];
    }
    // Throttle: start polling only once all but `lag` of the earlier chunks
    // are done (ctr[2] counts finished chunks), so that the workgroups far
    // down the line do not load the lines the running chunk is publishing to.
    if (lag > 0) {
      if (t == 0)
        for (int sp = 0; __hip_atomic_load(&ctr[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
                             ch_id - lag && sp < kSpinLimit;
             ++sp)
          __builtin_amdgcn_s_sleep(2);
      __syncthreads();
    }
    int spins = 0;
    double term[P];
    bool ok = true;
    if (t < ch.y) ok = sweep_remote<OP, P>(a, c, lc, w, wt, term, spins);
    // From here on w[j] holds what entry j adds when it is not in LDS: its
    // term when it comes from an earlier chunk, +0 when it is skipped (a sum
    // that starts at +0 and only adds is never -0 in round-to-nearest, and
    // x - (+0) = x, so +0 leaves every sum unchanged).
#pragma unroll
    for (int j = 0; j < P; ++j)
      if (lc[j] < 0) w[j] = c[j] < 0 ? 0.0 : term[j];
    // Every load above has landed (vmcnt 0; expcnt and lgkmcnt untouched):
    // without this the compiler may wait for one of them inside the levels,
    // and a vmcnt wait there also waits for the previous level's publishing
    // store to reach L2.
    __builtin_amdgcn_s_waitcnt(0x0F70);
    for (int l = 0; l < nlev; ++l) {
      if (lv == l) {
        // the reference's order, local and remote entries mixed
        double v[P];
#pragma unroll
        for (int j = 0; j < P; ++j) v[j] = sval[lc[j] >= 0 ? lc[j] : 0];
#pragma unroll
        for (int j = 0; j < P; ++j) {
          const double x = lc[j] >= 0 ? sweep_term<OP>(w[j], wt[j], v[j]) : w[j];
          if (op_subtracts(OP))
            s -= x;
          else
            s += x;
        }
        for (int b0 = P; b0 < ne && ok; b0 += B) {  // rows with more than P entries
          int cb[B], lb[B];
          double wb[B], wtb[B];
#pragma unroll
          for (int j = 0; j < B; ++j) {
            const bool use = b0 + j < ne;
            sweep_entry<OP>(a, loc, !use ? d : DESC ? kb + ne - 1 - (b0 + j) : kb + b0 + j, use,
                            cb[j], lb[j], wb[j], wtb[j]);
          }
          ok = sweep_fold<OP, B>(a, sval, cb, lb, wb, wtb, s, spins);
        }
        double r;
        if (OP == kDiluFact)
          r = 1.0 / s;
        else if (OP == kSgsFwd)
          r = (in_i - s) / piv;
        else if (OP == kDiluFwd)
          r = piv * (in_i - s);
        else if (OP == kSgsBwd)
          r = in_i - s / piv;
        else if (OP == kDiluBwd)
          r = in_i - piv * s;
        else if (OP == kTDiluFwd)
          r = s;
        else
          r = piv * s;
        if (!ok) r = __builtin_nan("");
        sval[t] = r;
        store_pub(a.out + i, r);
      }
      __syncthreads();
    }
    if (t == 0) atomicAdd(&ctr[2], 1);
  }
  if (t == 0 && atomicAdd(&ctr[1], 1) == (int)gridDim.x - 1) {
    __hip_atomic_store(&ctr[0], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&ctr[1], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&ctr[2], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// The chunk schedules of the forward (lower triangle) and backward (upper
// triangle) sweeps, computed once per pattern on the host: rows listed by
// level, ascending within a level; chunks of consecutive whole levels up to
// K rows (kChunkNarrow or kChunkWide, by the mean level width), a wider
// level split into chunks of its own; per listed
// row its level within the chunk; per entry a sweep reads, the position of
// the entry's column in the reading row's chunk (-1: an earlier chunk).
int csr_ensure_levels(ofh_csr A) {
  if (A->sw_list[0]) return OFH_OK;
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
  std::vector<int> loc(A->nnz, -1);
  std::vector<int> list[2], ilev[2];
  std::vector<int4> chunks[2];
  for (int dir = 0; dir < 2; ++dir) {
    const bool fwd = dir == 0;
    auto lo = [&](int i) { return fwd ? rp[i] : dpos[i] + 1; };
    auto hi = [&](int i) { return fwd ? dpos[i] : rp[i + 1]; };
    std::vector<int> lev(n, 0);
    int nl = 0;
    for (int q = 0; q < n; ++q) {
      const int i = fwd ? q : n - 1 - q;
      int l = 0;
      for (int k = lo(i); k < hi(i); ++k) l = std::max(l, lev[col[k]] + 1);
      lev[i] = l;
      nl = std::max(nl, l + 1);
    }
    const int K = (int64_t)n >= (int64_t)kWideLevel * nl ? kChunkWide : kChunkNarrow;
    A->sw_rows[dir] = K;
    std::vector<int> start(nl + 1, 0);
    for (int i = 0; i < n; ++i) start[lev[i] + 1]++;
    for (int l = 0; l < nl; ++l) start[l + 1] += start[l];
    std::vector<int>& L = list[dir];
    L.assign(n, 0);
    {
      std::vector<int> pos(start.begin(), start.end() - 1);
      for (int i = 0; i < n; ++i) L[pos[lev[i]]++] = i;
    }
    std::vector<int> chunk_of(n), pos_of(n);
    std::vector<int>& IL = ilev[dir];
    IL.assign(n, 0);
    std::vector<int4>& C = chunks[dir];
    int first = 0, rows = 0, levs = 0, groups = 0;  // groups: chunks that run one after another
    auto close = [&]() {
      if (rows) {
        C.push_back(make_int4(first, rows, levs, 0));
        groups++;
      }
      first += rows;
      rows = levs = 0;
    };
    for (int l = 0; l < nl; ++l) {
      const int q0 = start[l], w = start[l + 1] - q0;
      if (w > K) {  // a wide level: chunks of its own, one level each
        close();
        for (int o = 0; o < w; o += K) {
          const int m = std::min(K, w - o);
          for (int q = q0 + o; q < q0 + o + m; ++q) {
            chunk_of[L[q]] = (int)C.size();
            pos_of[L[q]] = q - q0 - o;
          }
          C.push_back(make_int4(q0 + o, m, 1, 0));
        }
        first = q0 + w;
        groups++;  // the chunks of one level run side by side
        continue;
      }
      if (rows + w > K) close();
      for (int q = q0; q < q0 + w; ++q) {
        IL[q] = levs;
        chunk_of[L[q]] = (int)C.size();
        pos_of[L[q]] = rows + q - q0;
      }
      rows += w;
      levs++;
    }
    close();
    for (int i = 0; i < n; ++i)
      for (int k = lo(i); k < hi(i); ++k)
        if (chunk_of[col[k]] == chunk_of[i]) loc[k] = pos_of[col[k]];
    A->sw_nlev[dir] = nl;
    A->sw_par[dir] = (int)((C.size() + std::max(groups, 1) - 1) / std::max(groups, 1));
  }
  int* d_ctr = nullptr;
  int* d_loc = nullptr;
  OFH_HIP(hipMalloc(&d_ctr, sizeof(int) * 3));
  OFH_HIP(hipMemsetAsync(d_ctr, 0, sizeof(int) * 3, s));
  OFH_HIP(hipMalloc(&d_loc, sizeof(int) * std::max<int64_t>(A->nnz, 1)));
  OFH_HIP(hipMemcpyAsync(d_loc, loc.data(), sizeof(int) * A->nnz, hipMemcpyHostToDevice, s));
  A->sweep_ctr = d_ctr;
  A->sw_loc = d_loc;
  int* d_list[2] = {nullptr, nullptr};
  for (int dir = 0; dir < 2; ++dir) {
    int* d_ilev = nullptr;
    int4* d_chunks = nullptr;
    OFH_HIP(hipMalloc(&d_list[dir], sizeof(int) * std::max(n, 1)));
    OFH_HIP(hipMalloc(&d_ilev, sizeof(int) * std::max(n, 1)));
    OFH_HIP(hipMalloc(&d_chunks, sizeof(int4) * std::max<size_t>(chunks[dir].size(), 1)));
    OFH_HIP(hipMemcpyAsync(d_list[dir], list[dir].data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    OFH_HIP(hipMemcpyAsync(d_ilev, ilev[dir].data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    OFH_HIP(hipMemcpyAsync(d_chunks, chunks[dir].data(), sizeof(int4) * chunks[dir].size(),
                           hipMemcpyHostToDevice, s));
    A->sw_ilev[dir] = d_ilev;
    A->sw_chunks[dir] = d_chunks;
    A->sw_nchunk[dir] = (int)chunks[dir].size();
  }
  OFH_HIP(hipStreamSynchronize(s));  // the host vectors leave scope
  A->sw_list[1] = d_list[1];
  A->sw_list[0] = d_list[0];  // set last: marks the schedules complete
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
  const int dir = op_forward(OP) ? 0 : 1;
  hipLaunchKernelGGL(k_fill_pending, dim3((n + 255) / 256), dim3(256), 0, s, a.out, n, a.done);
  // Workgroups in flight.  Wide levels: about three levels' worth of chunks
  // (every further workgroup only polls, and the polling slows the rows being
  // computed).  Narrow levels (one chunk at a time on the chain): 8 x, so the
  // chunks after the running one have loaded their rows when their turn comes.
  const int nchunk = A->sw_nchunk[dir];
  const int par = A->sw_par[dir];
  const int want = A->sw_rows[dir] == kChunkWide ? 3 * par : std::min(64, 8 * par);
  const int grid = std::min(nchunk, std::max(kSweepGridMin, std::min(kSweepGridMax, want)));
  if (A->sw_rows[dir] == kChunkWide)
    hipLaunchKernelGGL((k_sweep<OP, kChunkWide>), dim3(grid), dim3(kChunkWide), 0, s, nchunk,
                       A->sw_chunks[dir], A->sw_list[dir], A->sw_ilev[dir], A->sw_loc, A->sweep_ctr,
                       kLagWide, a);
  else
    hipLaunchKernelGGL((k_sweep<OP, kChunkNarrow>), dim3(grid), dim3(kChunkNarrow), 0, s, nchunk,
                       A->sw_chunks[dir], A->sw_list[dir], A->sw_ilev[dir], A->sw_loc, A->sweep_ctr,
                       kLagNarrow, a);
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
