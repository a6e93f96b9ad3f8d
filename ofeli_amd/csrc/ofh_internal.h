// ofh_internal.h -- shared internals of libofeli_hip (HIP, gfx950 only).
//
// Device conventions: CSR with int32 row_ptr (0-based) and int32 col (0-based),
// FP64 values plus an optional FP32 mirror.  Every reduction is a fixed-order
// two-stage tree (wave shuffle -> LDS -> per-block partial -> the last block
// to finish folds the partials in block order), so a given launch shape
// always produces bitwise identical results.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include <functional>
#include <string>
#include <utility>
#include <vector>

#include "../../include/ofeli_hip.h"

namespace ofh {

// ---------------------------------------------------------------------------
// error handling
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define OFH_HIP(call)                                                          \
  do {                                                                         \
    hipError_t _e = (call);                                                    \
    if (_e != hipSuccess)                                                      \
      return ::ofh::fail(OFH_E_HIP, std::string(#call) + ": " +              \
                                        hipGetErrorString(_e));               \
  } while (0)

#define OFH_CHECK(cond, code, msg)                                             \
  do {                                                                         \
    if (!(cond)) return ::ofh::fail((code), (msg));                            \
  } while (0)

// ---------------------------------------------------------------------------
// launch geometry
// ---------------------------------------------------------------------------
constexpr int kBlock = 256;           // threads per workgroup for stream kernels
constexpr int kWave = 64;             // CDNA wavefront
constexpr int kMaxPartials = 1 << 16; // per-reduction partial slots (max workgroups of a reduction)
// LDS product tiles of the CSR-stream SpMV: 16 KB in both precisions
// (measured on MI355X at 96^3: FP64 31.8 us with 2048 vs 35.0 us with 4096;
// FP32 26.6 us with 4096 vs 29.4 us with 2048).
constexpr int kTile64 = 2048;         // FP64
constexpr int kTile32 = 4096;         // FP32
constexpr int kNumXcd = 8;            // MI355X: 8 XCDs, workgroups dealt round-robin
constexpr int kVec = 4;               // vector entries per thread in element-wise solver kernels
constexpr int kMaxRows = kMaxPartials * kBlock * kVec;  // largest n the reductions cover
// CSR value/column arrays carry kPad zeroed trailing elements, so that
// kernels may load nonzeros in aligned pairs or vectors past an odd end.
constexpr int kPad = 2;
// LDS a gfx950 workgroup may allocate (160 KiB).  A kernel whose static LDS
// exceeds it does not launch (the queue reports an invalid allocation), so
// every kernel whose __shared__ arrays are sized by constants or template
// parameters checks its total at compile time with OFH_LDS_BUDGET.
constexpr size_t kLdsBytes = 160 * 1024;
#define OFH_LDS_BUDGET(bytes) \
  static_assert((size_t)(bytes) <= ::ofh::kLdsBytes, "static LDS exceeds the 160 KiB per workgroup")
// the CSR-stream kernels: one product tile plus at most four reduction rows
OFH_LDS_BUDGET(sizeof(double) * kTile64 + 4 * sizeof(double) * kBlock / kWave);
OFH_LDS_BUDGET(sizeof(float) * kTile32 + 4 * sizeof(double) * kBlock / kWave);

// Workgroups are dispatched round-robin over the XCDs; remap the hardware
// block id so that each XCD walks one contiguous range of row blocks, which
// keeps the gathered x window of neighbouring rows inside one L2.
__device__ __forceinline__ int xcd_block(int bid, int nblocks) {
  int per = (nblocks + kNumXcd - 1) / kNumXcd;
  int xcd = bid % kNumXcd, slot = bid / kNumXcd;
  int full = nblocks - (per - 1) * kNumXcd;  // XCDs that own `per` blocks
  int base = xcd < full ? xcd * per : full * per + (xcd - full) * (per - 1);
  return base + slot;
}

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block-wide sum of BLOCK threads; result valid in every thread.
template <int BLOCK>
__device__ __forceinline__ double block_sum(double v, double* lds /*BLOCK/64*/) {
  v = wave_sum(v);
  const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
  __syncthreads();
  if (l == 0) lds[w] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < BLOCK / kWave; ++i) s += lds[i];
  return s;
}

// Three sums at once (CG update kernel).
template <int BLOCK, int K>
__device__ __forceinline__ void block_sum_k(double (&v)[K], double* lds /*K*BLOCK/64*/) {
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
  const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
  __syncthreads();
  if (l == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) lds[k * (BLOCK / kWave) + w] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < BLOCK / kWave; ++i) s += lds[k * (BLOCK / kWave) + i];
    v[k] = s;
  }
}

// Cross-workgroup reductions never communicate inside a kernel: each block
// writes its partial sums (one slot per quantity, one entry per block) and
// the NEXT kernel on the stream, whose launch makes them visible, folds them
// in block order.  Every consuming block folds the same numbers in the same
// order, so all blocks agree bitwise on the result (and on any decision taken
// from it) without fences or atomics.
template <int K>
__device__ __forceinline__ void write_partials(const double (&v)[K], const int (&slot)[K],
                                               double* partials, int bid) {
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) partials[slot[k] * kMaxPartials + bid] = v[k];
  }
}

// Each thread loads eight partials per step (independent loads in flight,
// summed as a fixed tree), so the fold is not a chain of dependent
// load-add latencies; every workgroup runs the same order.
template <int BLOCK, int K>
__device__ __forceinline__ void fold_partials(double (&out)[K], const int (&slot)[K],
                                              const double* __restrict__ partials, int nparts,
                                              double* lds) {
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double* __restrict__ p = partials + slot[k] * kMaxPartials;
    double s = 0.0;
    int i = threadIdx.x;
    for (; i + 7 * BLOCK < nparts; i += 8 * BLOCK) {
      double v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] = p[i + j * BLOCK];
      s += ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
    }
    for (; i < nparts; i += BLOCK) s += p[i];
    out[k] = s;
  }
  block_sum_k<BLOCK, K>(out, lds);
}

// fold_partials in two steps for nparts <= PER * BLOCK: the loads first (in
// flight while the caller starts other work), then the sums, in the same
// order in every workgroup: per thread over j, then the block tree.
template <int BLOCK, int K, int PER = 7>
struct PartialsPrefetch {
  static constexpr int kMax = PER * BLOCK;
  double v[K][PER];
  __device__ __forceinline__ void load(const int (&slot)[K], const double* __restrict__ partials,
                                       int nparts) {
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int j = 0; j < PER; ++j) {
        const int i = threadIdx.x + j * BLOCK;
        v[k][j] = i < nparts ? partials[slot[k] * kMaxPartials + i] : 0.0;
      }
  }
  __device__ __forceinline__ void fold(double (&out)[K], int nparts, double* lds) const {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < PER; ++j)
        if (threadIdx.x + j * BLOCK < nparts) s += v[k][j];
      out[k] = s;
    }
    block_sum_k<BLOCK, K>(out, lds);
  }
};

// ---------------------------------------------------------------------------
// objects behind the opaque handles
// ---------------------------------------------------------------------------
constexpr int kSlots = 8;  // partial-sum slots of the workspace

struct Workspace {
  double* partials = nullptr;   // kSlots * kMaxPartials doubles
};

// Identity of a captured launch sequence: which sequence and every pointer
// and size it was recorded with.
struct GraphKey {
  static constexpr int kN = 16;
  uintptr_t v[kN] = {};
  bool operator==(const GraphKey& o) const {
    for (int i = 0; i < kN; ++i)
      if (v[i] != o.v[i]) return false;
    return true;
  }
};

// Iterations per captured chunk of a Krylov solve.
constexpr int kGraphIters = 8;

// One timed launch of a solver kernel (ofh_ctx_profile): events recorded on
// the launching stream right before and right after it.
struct ProfRec {
  int kind;
  hipEvent_t a, b;
};

}  // namespace ofh

struct ofh_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int num_cu = 0;
  ofh::Workspace ws;
  void* host_pinned = nullptr;  // 4 KiB staging for scalars
  hipEvent_t look_ev = nullptr;  // marks the state copy of a look taken ahead
  double* dev_scratch = nullptr;  // small per-solve state (4 KiB)
  void* vws = nullptr;            // solver work vectors, grown on demand
  size_t vws_bytes = 0;
  // Chunks of solver iterations can be replayed from captured graphs
  // (ofh_ctx_use_graphs).  Off by default: measured on MI355X it is 2-3 %
  // slower than direct launches at 261k and 857k rows.
  bool use_graphs = false;
  // form of ofh_cg_mixed's inner iteration (ofh_ctx_mixed_form): 0 by size,
  // 1 two kernels, 2 one kernel wherever z = r on a diagonal copy
  int mixed_form = 0;
  hipStream_t cap = nullptr;  // capture-only stream
  std::vector<std::pair<ofh::GraphKey, hipGraphExec_t>> graphs;  // small cache, oldest first
  // per-kernel timing of the solver loops (ofh_ctx_profile), off by default
  bool prof = false;
  std::vector<ofh::ProfRec> prof_recs;
  std::vector<hipEvent_t> prof_pool;  // events of finished records, reused
};

struct ofh_csr_s {
  ofh_ctx ctx = nullptr;
  uint64_t serial = 0;      // unique per matrix of the process (captured-graph keys)
  int n = 0;
  int64_t nnz = 0;
  int* row_ptr = nullptr;   // device, n+1
  int* col = nullptr;       // device, nnz (0-based)
  int16_t* col16 = nullptr; // device, nnz: col - first row of its SpMV block (null if out of range)
  double* val = nullptr;    // device, nnz
  float* val32 = nullptr;   // device, nnz (lazy FP32 mirror)
  int* row_blocks = nullptr;  // device, nblocks+1 (CSR-stream row partition, FP64 tile)
  int nblocks = 0;
  int* row_blocks32 = nullptr;  // device, nblocks32+1 (partition for the FP32 tile)
  int nblocks32 = 0;
  int max_row_nnz = 0;
  double* dinv = nullptr;     // device, n (DIAG_PREC pivots; lazy)
  float* dinv32 = nullptr;
  int* diag_pos = nullptr;    // device, n (position of a_ii; lazy)
  bool dinv_valid = false, val32_valid = false;
  uint64_t version = 0;  // bumped whenever the values may have changed
  // chunk schedules of the triangular sweeps (lazy), [0] forward (lower
  // triangle), [1] backward (upper triangle)
  int* sw_list[2] = {nullptr, nullptr};      // rows listed by level
  int* sw_ilev[2] = {nullptr, nullptr};      // level of each listed row within its chunk
  int4* sw_chunks[2] = {nullptr, nullptr};   // per chunk: first list entry, rows, levels
  int sw_nchunk[2] = {0, 0}, sw_nlev[2] = {0, 0};
  int sw_par[2] = {1, 1};        // chunks that can run side by side, on average
  int sw_rows[2] = {0, 0};       // rows per chunk
  int* sw_loc = nullptr;         // nnz: position of an entry's column in its row's chunk, -1
  int* sweep_ctr = nullptr;      // 3 ints: chunk ticket, finished workgroups, finished chunks
  double* ssor_tmp = nullptr;    // forward-sweep result of the SSOR / DILU application
  int* dilu_tpos = nullptr;      // nnz: position of a_ji for each lower entry a_ij (-1: none)
  double* dilu_inv = nullptr;    // n: 1/D of the DILU factorization
  uint64_t dilu_version = ~0ull; // values version dilu_inv was computed from
  // SELL-64 copy for the streaming SpMVs (lazy): rows in chunks of 64 (one
  // wave), each chunk padded to its longest row and stored position-major
  int sell_state = 0;            // 0: not built, 1: in use, -1: too much padding
  int sell_nchunks = 0;
  int64_t sell_total = 0;        // stored entries (padding included)
  int* sell_cptr = nullptr;      // nchunks: first stored entry of each chunk
  int* sell_clen = nullptr;      // nchunks: stored entries per row of the chunk
  uint32_t* sell_col = nullptr;  // column index (padding: kSellPadCol)
  int16_t* sell_col16 = nullptr;  // column - first row of the chunk, when every one fits
                                  // (padding: kSellPad16); null otherwise
  double* sell_val = nullptr;    // padding: -0.0
  float* sell_val32 = nullptr;
  uint64_t sell_version = ~0ull, sell_version32 = ~0ull;
  // Diagonal copies of the values (lazy), for patterns whose entries lie on
  // at most kDiaMaxDiag diagonals (structured-mesh numberings): row i has W
  // slots, slot d the value at column i + dia_off[d] (offsets ascending),
  // slot W-1 the bit mask of the slots present in the row; W = 4, 8 or 16,
  // the smallest that holds the diagonals and the mask; rows stored by
  // chunks of 64, slot-major inside a chunk
  int dia_state = 0;             // 0: not detected, 1: pattern fits, -1: it does not
  int dia_nd = 0;                // diagonals
  int dia_w = 0;                 // slots per row: 4, 8 or 16
  int dia_off[16] = {};          // column - row of each diagonal (unused slots 0)
  bool dia_use32 = false;        // the FP32 copy takes no more bytes than SELL-64 FP32
  bool dia_use64 = false;        // the FP64 copy takes no more bytes than SELL-64 FP64
  float* dia_val32 = nullptr;    // ceil(n / 64) * 64 * dia_w
  double* dia_val64 = nullptr;
  uint64_t dia_version32 = ~0ull, dia_version64 = ~0ull;
  // The symmetrically scaled matrix D^-1/2 A D^-1/2 (the mixed CG's FP32
  // inner operator for DIAG_PREC) in the FP32 copy's layout, and D^-1/2
  float* dia_val32s = nullptr;
  double* dscale = nullptr;      // n: 1 / sqrt(a_ii)
  bool dscale_ok = false;        // every a_ii > 0 and finite
  uint64_t dscale_version = ~0ull;
  // Symmetric diagonal copies (DiaSymRows): the pattern's diagonals pair up
  // (+-o) with as many entries above the diagonal as below (dias_pat, found
  // with the diagonals), and the values are symmetric bit for bit (dias_ok,
  // checked on the GPU once per value version)
  bool dias_pat = false;
  int dias_up[8] = {};           // up[0] = 0, up[k] the k-th upper offset (unused 0)
  bool dias_ok = false;
  uint64_t dias_check_version = ~0ull;
  uint16_t* dias_mask = nullptr;  // n
  double* dias_val64 = nullptr;   // ceil(n / 64) * 64 * dia_w / 2
  float* dias_val32 = nullptr;
  float* dias_val32s = nullptr;   // of D^-1/2 A D^-1/2
  uint64_t dias_version64 = ~0ull, dias_version32 = ~0ull, dias_version32s = ~0ull;
  ofh_csr_s* tr = nullptr;       // transpose (BiCG, TMult, TransSolve; lazy)
  int* tr_perm = nullptr;        // nnz: position in A of each entry of tr
  uint64_t tr_version = ~0ull;   // values version tr->val was gathered from
};

namespace ofh {
int ensure_vws(ofh_ctx ctx, size_t bytes, void** out);
// Launches `kernel` on s.  With profiling on and kind >= 0 (OFH_PROF_*) the
// launch carries a start and a stop event in its own dispatch
// (hipExtLaunchKernelGGL), so the recorded time is the kernel's execution,
// and the pair is kept as one timed launch of `kind`.
template <class F, class... Args>
inline void prof_kernel(ofh_ctx ctx, int kind, F kernel, dim3 grid, dim3 block, hipStream_t s,
                        Args... args) {
  hipEvent_t ev[2] = {nullptr, nullptr};
  if (ctx->prof && kind >= 0) {
    for (auto& e : ev) {
      if (!ctx->prof_pool.empty()) {
        e = ctx->prof_pool.back();
        ctx->prof_pool.pop_back();
      } else if (hipEventCreate(&e) != hipSuccess) {
        e = nullptr;
      }
    }
    if (!ev[0] || !ev[1]) {
      for (auto& e : ev)
        if (e) ctx->prof_pool.push_back(e);
      ev[0] = ev[1] = nullptr;
    }
  }
  hipExtLaunchKernelGGL(kernel, grid, block, 0, s, ev[0], ev[1], 0, args...);
  if (ev[0]) ctx->prof_recs.push_back(ProfRec{kind, ev[0], ev[1]});
}
// The executable graph of `record` (launches on the given stream, kernels
// only) for `key`, captured on first use and cached in the context.
int ctx_graph(ofh_ctx ctx, const GraphKey& key, const std::function<int(hipStream_t)>& record,
              hipGraphExec_t* out);
// Key of a captured solver chunk over A: the sequence id and kernel variant,
// the matrix's process-unique serial (a matrix created later at the same
// address never matches), the solution and work buffers, the value and
// pivot arrays it reads, every CSR array, the SELL-64 copy's arrays and the diagonal copy.
inline void csr_graph_key(ofh_csr A, uintptr_t kind, uintptr_t variant, const void* x,
                          const void* work, const void* val, const void* dinv, GraphKey* key) {
  const uintptr_t v[GraphKey::kN] = {
      kind, variant, (uintptr_t)A->serial, (uintptr_t)x, (uintptr_t)work,
      ((uintptr_t)(uint32_t)A->n << 32) | (uint32_t)A->nblocks, (uintptr_t)val,
      (uintptr_t)dinv, (uintptr_t)A->col16, (uintptr_t)A->col, (uintptr_t)A->row_blocks,
      (uintptr_t)A->row_ptr, (uintptr_t)A->sell_cptr, (uintptr_t)A->sell_col16,
      (uintptr_t)A->sell_col ^ ((uintptr_t)A->sell_val << 1) ^ (uintptr_t)A->sell_val32 ^
          ((uintptr_t)A->dia_val32 << 2) ^ ((uintptr_t)A->dia_val32s << 3) ^
          ((uintptr_t)A->dia_val64 << 4) ^ ((uintptr_t)A->dias_val32 << 5) ^
          ((uintptr_t)A->dias_val32s << 6) ^ ((uintptr_t)A->dias_val64 << 7),
      (uintptr_t)kGraphIters};
  for (int i = 0; i < GraphKey::kN; ++i) key->v[i] = v[i];
}
int reduce_grid(int n);
int launch_dot(ofh_ctx ctx, int n, const double* x, const double* y, double* d_out, bool sqrt_out);
int csr_ensure_val32(ofh_csr A);
// 16-bit block-relative column offsets when every column fits (see ColIndex).
int csr_build_rel16(ofh_csr A);
// Preconditioners applied by level-scheduled triangular sweeps (SSOR_PREC,
// DILU_PREC and ILU_PREC, the latter two the same D-ILU as in the
// reference); csr_ensure_prec builds what the first application needs.
constexpr bool sweep_prec(int p) {
  return p == OFH_SSOR_PREC || p == OFH_DILU_PREC || p == OFH_ILU_PREC;
}
int csr_ensure_prec(ofh_csr A, int prec);
int launch_prec(ofh_csr A, int prec, const double* r, double* z, const int* d_done,
                hipStream_t s);
// z = M^-T r for ILU_PREC (Prec::TransSolve): two sweeps over A's transpose
// with A's D-ILU pivots.
int launch_dilu_trans(ofh_csr A, const double* r, double* z, const int* d_done, hipStream_t s);
// The SELL-64 copy with current FP64 (fp32: FP32) values; *ok is false when
// the pattern pads too much (or 16-bit columns are on) and the CSR-stream
// kernels are to be used.
int csr_ensure_sell(ofh_csr A, bool fp32, bool* ok);
// The diagonal copy with current FP32 values; *ok is false when the pattern
// has more than kDiaMaxDiag diagonals, a row whose columns do not ascend,
// or the copy would take more bytes than the SELL-64 FP32 copy with 16-bit
// columns (6 B per entry).
int csr_ensure_dia(ofh_csr A, bool* ok);
// The same with FP64 values (no more bytes than SELL-64 FP64: 10 B per entry).
int csr_ensure_dia64(ofh_csr A, bool* ok);
// The diagonal copy of D^-1/2 A D^-1/2 (values a_ij (s_i s_j), s = D^-1/2,
// rounded once to FP32) and A->dscale = s in FP64, current with the values;
// *ok is false when A has no diagonal copy or a diagonal entry is not
// positive and finite.
int csr_ensure_dia_scaled(ofh_csr A, bool* ok);
// *sym when the pattern has diagonals (a diagonal copy fits), they pair up
// and the values are symmetric bit for bit (checked once per value version).
int csr_dia_symmetric(ofh_csr A, bool* sym);
// The symmetric form of a diagonal copy that is in use (kind 0: FP64, 1:
// FP32, 2: FP32 of D^-1/2 A D^-1/2, after csr_ensure_dia_scaled), current
// with the values; *ok when the matrix is symmetric bit for bit and the copy
// is made.
int csr_ensure_dias(ofh_csr A, int kind, bool* ok);
// A->tr, the CSR transpose of A, with values current (built on first use).
int csr_ensure_transpose(ofh_csr A);
// Everything derived from the values (FP32 copy, pivots, version) is stale.
void csr_values_changed(ofh_csr A);
int csr_ensure_dinv(ofh_csr A);
int csr_ensure_levels(ofh_csr A);
// Both CSR-stream row partitions (FP64 and FP32 tiles) from the host row_ptr.
int csr_set_row_blocks(ofh_csr A, const int* h_row_ptr);

// SpMV launchers (ofh_spmv.hip)
int launch_spmv64(ofh_csr A, const double* x, double* y, hipStream_t s);
int launch_spmv32(ofh_csr A, const float* x, float* y, hipStream_t s);
// SSOR sweeps z = M^-1 r (ofh_ssor.hip); skipped when *d_done != 0.
int launch_ssor(ofh_csr A, const double* r, double* z, const int* d_done, hipStream_t s);
// x <- b - A x style true residual ||b-Ax||/||b|| in FP64 (host result).
int true_relres(ofh_csr A, const double* b, const double* x, double* out);
// The same over the CSR-stream partition with the given partial slots and
// result word; builds and allocates nothing.
int true_relres_csr(ofh_csr A, const double* b, const double* x, double* partials, double* d_out,
                    double* out);
// Looks up, once per device and with that device current, the address of
// the scratch array of ofh_cg_noalloc (called by ofh_ctx_create).
int noalloc_setup(int device);
double now_seconds();
// A process-unique id for each new matrix (ofh_csr_s::serial).
uint64_t next_csr_serial();
}  // namespace ofh
