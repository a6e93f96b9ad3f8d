// ofh_core.hip -- context, device memory, CSR storage and BLAS-1 of libofeli_hip.
//
// Replaces the storage side of SpMatrix<T_> (SpMatrix.h:78-611) and the Vect
// BLAS-1 routines (Vect_impl.h:660-668, 2143-2225) with device-resident data.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ofh_internal.h"

namespace ofh {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }
int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

// ---------------------------------------------------------------------------
// row partition of the CSR-stream SpMV: consecutive rows until either kBlock
// rows or `tile` nonzeros; a row longer than `tile` gets a block of its own.
// ---------------------------------------------------------------------------
static std::vector<int> build_row_blocks(const int* rp, int n, int tile) {
  std::vector<int> rb;
  rb.push_back(0);
  int r = 0;
  while (r < n) {
    int start = r;
    int64_t base = rp[r];
    while (r < n && r - start < kBlock && (int64_t)rp[r + 1] - base <= tile) ++r;
    if (r == start) ++r;  // one long row
    rb.push_back(r);
  }
  return rb;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__global__ void k_find_diag(int n, const int* __restrict__ rp, const int* __restrict__ col,
                            int* __restrict__ diag_pos) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int lo = rp[i], hi = rp[i + 1] - 1, pos = -1;
  while (lo <= hi) {
    int mid = (lo + hi) >> 1;
    int c = col[mid];
    if (c == i) { pos = mid; break; }
    if (c < i) lo = mid + 1; else hi = mid - 1;
  }
  diag_pos[i] = pos;
}

// Prec::inv_diag (Prec_impl.h:306-316): pivot_i = 1/a_ii; flags zero pivots.
__global__ void k_inv_diag(int n, const int* __restrict__ diag_pos, const double* __restrict__ val,
                           double* __restrict__ dinv, float* __restrict__ dinv32,
                           int* __restrict__ bad) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int p = diag_pos[i];
  double a = p < 0 ? 0.0 : val[p];
  if (a == 0.0) {
    atomicMin(bad, i);
    dinv[i] = 0.0;
    dinv32[i] = 0.0f;
    return;
  }
  double d = 1.0 / a;
  dinv[i] = d;
  dinv32[i] = (float)d;
}

__global__ void k_to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) b[i] = (float)a[i];
}

// Fixed-order dot product: grid-stride per thread and block tree into slot 0
// of the partials; k_fold_one folds them in block order.
__global__ __launch_bounds__(kBlock) void k_dot(int n, const double* __restrict__ x,
                                               const double* __restrict__ y,
                                               double* __restrict__ partials) {
  __shared__ double lds[kBlock / kWave];
  double v[1] = {0.0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    v[0] += x[i] * y[i];
  block_sum_k<kBlock, 1>(v, lds);
  const int slot[1] = {0};
  write_partials<1>(v, slot, partials, blockIdx.x);
}

__global__ __launch_bounds__(kBlock) void k_fold_one(const double* __restrict__ partials,
                                                    int nparts, double* __restrict__ out,
                                                    int sqrt_out) {
  __shared__ double lds[kBlock / kWave];
  double s[1];
  const int slot[1] = {0};
  fold_partials<kBlock, 1>(s, slot, partials, nparts, lds);
  if (threadIdx.x == 0) *out = sqrt_out ? sqrt(s[0]) : s[0];
}

__global__ void k_axpy(int n, double a, const double* __restrict__ x, double* __restrict__ y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += a * x[i];
}

// Grid of the element-wise solver kernels: exactly kVec entries per thread
// (the load-first kernels rely on it), a fixed function of n so the partial
// sums always have the same layout.
int reduce_grid(int n) {
  int g = (n + kBlock * kVec - 1) / (kBlock * kVec);
  return std::max(1, g);
}

int launch_dot(ofh_ctx ctx, int n, const double* x, const double* y, double* d_out,
               bool sqrt_out) {
  // k_dot strides over the grid, so any n fits the kMaxPartials slots of slot 0
  const int g = std::min(reduce_grid(n), kMaxPartials);
  hipLaunchKernelGGL(k_dot, dim3(g), dim3(kBlock), 0, ctx->stream, n, x, y, ctx->ws.partials);
  hipLaunchKernelGGL(k_fold_one, dim3(1), dim3(kBlock), 0, ctx->stream, ctx->ws.partials, g,
                     d_out, (int)sqrt_out);
  OFH_HIP(hipGetLastError());
  return OFH_OK;
}

int ctx_graph(ofh_ctx ctx, const GraphKey& key, const std::function<int(hipStream_t)>& record,
              hipGraphExec_t* out) {
  for (auto& e : ctx->graphs)
    if (e.first == key) {
      *out = e.second;
      return OFH_OK;
    }
  if (!ctx->cap) OFH_HIP(hipStreamCreateWithFlags(&ctx->cap, hipStreamNonBlocking));
  OFH_HIP(hipStreamBeginCapture(ctx->cap, hipStreamCaptureModeThreadLocal));
  const int rc = record(ctx->cap);
  hipGraph_t g = nullptr;
  const hipError_t ce = hipStreamEndCapture(ctx->cap, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  OFH_HIP(ce);
  hipGraphExec_t ge = nullptr;
  const hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  OFH_HIP(ie);
  if (ctx->graphs.size() >= 16) {
    (void)hipGraphExecDestroy(ctx->graphs.front().second);
    ctx->graphs.erase(ctx->graphs.begin());
  }
  ctx->graphs.emplace_back(key, ge);
  *out = ge;
  return OFH_OK;
}

int csr_set_row_blocks(ofh_csr A, const int* rp) {
  const std::vector<int> rb = build_row_blocks(rp, A->n, kTile64);
  const std::vector<int> rb32 = build_row_blocks(rp, A->n, kTile32);
  A->nblocks = (int)rb.size() - 1;
  A->nblocks32 = (int)rb32.size() - 1;
  OFH_CHECK(A->nblocks <= kMaxPartials && A->nblocks32 <= kMaxPartials, OFH_E_INVALID,
            "matrix too large for the reduction slots");
  OFH_HIP(hipMalloc(&A->row_blocks, sizeof(int) * rb.size()));
  OFH_HIP(hipMalloc(&A->row_blocks32, sizeof(int) * rb32.size()));
  OFH_HIP(hipMemcpyAsync(A->row_blocks, rb.data(), sizeof(int) * rb.size(), hipMemcpyHostToDevice,
                         A->ctx->stream));
  OFH_HIP(hipMemcpyAsync(A->row_blocks32, rb32.data(), sizeof(int) * rb32.size(),
                         hipMemcpyHostToDevice, A->ctx->stream));
  OFH_HIP(hipStreamSynchronize(A->ctx->stream));
  return OFH_OK;
}

int ensure_vws(ofh_ctx ctx, size_t bytes, void** out) {
  if (ctx->vws_bytes < bytes) {
    if (ctx->vws) {
      OFH_HIP(hipStreamSynchronize(ctx->stream));
      OFH_HIP(hipFree(ctx->vws));
    }
    ctx->vws = nullptr;
    ctx->vws_bytes = 0;
    OFH_HIP(hipMalloc(&ctx->vws, bytes));
    ctx->vws_bytes = bytes;
  }
  *out = ctx->vws;
  return OFH_OK;
}

uint64_t next_csr_serial() {
  static std::atomic<uint64_t> next{1};
  return next++;
}

double now_seconds() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

__global__ void k_make_rel16(const int* __restrict__ rb, const int* __restrict__ rp,
                             const int* __restrict__ col, int16_t* __restrict__ rel, int* bad) {
  const int r0 = rb[blockIdx.x], r1 = rb[blockIdx.x + 1];
  for (int k = rp[r0] + threadIdx.x; k < rp[r1]; k += blockDim.x) {
    const int d = col[k] - r0;
    if (d < -32768 || d > 32767) {
      *bad = 1;
      rel[k] = 0;
    } else {
      rel[k] = (int16_t)d;
    }
  }
}

int csr_build_rel16(ofh_csr A) {
  ofh_ctx ctx = A->ctx;
  if (A->col16) return OFH_OK;
  OFH_HIP(hipMalloc(&A->col16, sizeof(int16_t) * (std::max<int64_t>(A->nnz, 1) + kPad)));
  OFH_HIP(hipMemsetAsync(A->col16 + A->nnz, 0, sizeof(int16_t) * kPad, A->ctx->stream));
  int* bad = (int*)ctx->dev_scratch + 4;
  OFH_HIP(hipMemsetAsync(bad, 0, sizeof(int), ctx->stream));
  hipLaunchKernelGGL(k_make_rel16, dim3(A->nblocks), dim3(256), 0, ctx->stream, A->row_blocks,
                     A->row_ptr, A->col, A->col16, bad);
  OFH_HIP(hipGetLastError());
  int h = 0;
  OFH_HIP(hipMemcpyAsync(&h, bad, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  OFH_HIP(hipStreamSynchronize(ctx->stream));
  if (h) {  // some column is too far from its block: keep the int32 form only
    OFH_HIP(hipFree(A->col16));
    A->col16 = nullptr;
  }
  return OFH_OK;
}

int csr_ensure_val32(ofh_csr A) {
  if (A->val32_valid) return OFH_OK;
  if (!A->val32) {
    OFH_HIP(hipMalloc(&A->val32, sizeof(float) * (std::max<int64_t>(A->nnz, 1) + kPad)));
    OFH_HIP(hipMemsetAsync(A->val32 + A->nnz, 0, sizeof(float) * kPad, A->ctx->stream));
  }
  int g = (int)std::min<int64_t>((A->nnz + 255) / 256, 65535);
  hipLaunchKernelGGL(k_to_f32, dim3(std::max(g, 1)), dim3(256), 0, A->ctx->stream, A->nnz,
                     A->val, A->val32);
  OFH_HIP(hipGetLastError());
  A->val32_valid = true;
  return OFH_OK;
}

int csr_ensure_dinv(ofh_csr A) {
  if (A->dinv_valid) return OFH_OK;
  ofh_ctx ctx = A->ctx;
  int n = A->n;
  if (!A->diag_pos) {
    OFH_HIP(hipMalloc(&A->diag_pos, sizeof(int) * std::max(n, 1)));
    hipLaunchKernelGGL(k_find_diag, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n,
                       A->row_ptr, A->col, A->diag_pos);
    OFH_HIP(hipGetLastError());
  }
  if (!A->dinv) OFH_HIP(hipMalloc(&A->dinv, sizeof(double) * std::max(n, 1)));
  if (!A->dinv32) OFH_HIP(hipMalloc(&A->dinv32, sizeof(float) * std::max(n, 1)));
  int* bad = (int*)ctx->dev_scratch;
  int big = 0x7fffffff;
  OFH_HIP(hipMemcpyAsync(bad, &big, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_inv_diag, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n,
                     A->diag_pos, A->val, A->dinv, A->dinv32, bad);
  OFH_HIP(hipGetLastError());
  int hbad = 0;
  OFH_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  OFH_HIP(hipStreamSynchronize(ctx->stream));
  if (hbad != big)
    return fail(OFH_E_ZERO_PIVOT,
                "In Prec::inv_diag(): null diagonal term: " + std::to_string(hbad + 1));
  A->dinv_valid = true;
  return OFH_OK;
}

void csr_values_changed(ofh_csr A) {
  A->val32_valid = false;
  A->dinv_valid = false;
  ++A->version;
}

}  // namespace ofh

using namespace ofh;

extern "C" {

const char* ofh_last_error(void) { return g_last_error.c_str(); }

int ofh_ctx_create(int device, ofh_ctx* out) {
  OFH_CHECK(out, OFH_E_INVALID, "ofh_ctx_create: null out");
  int count = 0;
  OFH_HIP(hipGetDeviceCount(&count));
  OFH_CHECK(device >= 0 && device < count, OFH_E_INVALID, "ofh_ctx_create: bad device");
  OFH_HIP(hipSetDevice(device));
  {
    const int rc = noalloc_setup(device);
    if (rc) return rc;
  }
  ofh_ctx c = new ofh_ctx_s();
  c->device = device;
  hipDeviceProp_t prop;
  OFH_HIP(hipGetDeviceProperties(&prop, device));
  c->num_cu = prop.multiProcessorCount;
  OFH_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  c->own_stream = true;
  OFH_HIP(hipMalloc(&c->ws.partials, sizeof(double) * kSlots * kMaxPartials));
  OFH_HIP(hipMalloc(&c->dev_scratch, 4096));
  OFH_HIP(hipMemset(c->dev_scratch, 0, 4096));
  OFH_HIP(hipHostMalloc(&c->host_pinned, 4096, hipHostMallocDefault));
  *out = c;
  return OFH_OK;
}

int ofh_ctx_destroy(ofh_ctx c) {
  if (!c) return OFH_OK;
  hipSetDevice(c->device);
  hipDeviceSynchronize();
  hipFree(c->ws.partials);
  hipFree(c->dev_scratch);
  hipHostFree(c->host_pinned);
  if (c->look_ev) (void)hipEventDestroy(c->look_ev);
  if (c->vws) hipFree(c->vws);
  for (auto& e : c->graphs) (void)hipGraphExecDestroy(e.second);
  for (auto& r : c->prof_recs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (auto e : c->prof_pool) (void)hipEventDestroy(e);
  if (c->cap) (void)hipStreamDestroy(c->cap);
  if (c->own_stream) hipStreamDestroy(c->stream);
  delete c;
  return OFH_OK;
}

int ofh_ctx_use_graphs(ofh_ctx c, int enable) {
  OFH_CHECK(c, OFH_E_INVALID, "ofh_ctx_use_graphs: null ctx");
  c->use_graphs = enable != 0;
  return OFH_OK;
}

int ofh_ctx_mixed_form(ofh_ctx c, int form) {
  OFH_CHECK(c && form >= 0 && form <= 2, OFH_E_INVALID, "ofh_ctx_mixed_form: bad argument");
  c->mixed_form = form;
  return OFH_OK;
}

int ofh_ctx_profile(ofh_ctx c, int enable) {
  OFH_CHECK(c, OFH_E_INVALID, "ofh_ctx_profile: null ctx");
  OFH_HIP(hipStreamSynchronize(c->stream));
  for (auto& r : c->prof_recs) {
    c->prof_pool.push_back(r.a);
    c->prof_pool.push_back(r.b);
  }
  c->prof_recs.clear();
  c->prof = enable != 0;
  return OFH_OK;
}

int ofh_ctx_profile_read(ofh_ctx c, int kind, double* h_ms, int cap, int* count) {
  OFH_CHECK(c && count && cap >= 0 && (cap == 0 || h_ms), OFH_E_INVALID,
            "ofh_ctx_profile_read: bad argument");
  OFH_HIP(hipStreamSynchronize(c->stream));
  int k = 0;
  for (const auto& r : c->prof_recs) {
    if (r.kind != kind) continue;
    if (k < cap) {
      float ms = 0.f;
      OFH_HIP(hipEventElapsedTime(&ms, r.a, r.b));
      h_ms[k] = ms;
    }
    ++k;
  }
  *count = k;
  return OFH_OK;
}

int ofh_ctx_set_stream(ofh_ctx c, void* s) {
  OFH_CHECK(c, OFH_E_INVALID, "ofh_ctx_set_stream: null ctx");
  OFH_HIP(hipStreamSynchronize(c->stream));
  if (c->own_stream) OFH_HIP(hipStreamDestroy(c->stream));
  c->stream = (hipStream_t)s;
  c->own_stream = false;
  return OFH_OK;
}

int ofh_ctx_synchronize(ofh_ctx c) {
  OFH_CHECK(c, OFH_E_INVALID, "null ctx");
  OFH_HIP(hipStreamSynchronize(c->stream));
  return OFH_OK;
}

int ofh_malloc(ofh_ctx c, size_t bytes, void** p) {
  OFH_CHECK(c && p, OFH_E_INVALID, "ofh_malloc: null argument");
  OFH_HIP(hipMalloc(p, std::max<size_t>(bytes, 1)));
  return OFH_OK;
}
int ofh_free(ofh_ctx c, void* p) {
  OFH_CHECK(c, OFH_E_INVALID, "ofh_free: null ctx");
  if (p) OFH_HIP(hipFree(p));
  return OFH_OK;
}
int ofh_memcpy_h2d(ofh_ctx c, void* d, const void* h, size_t bytes) {
  OFH_CHECK(c, OFH_E_INVALID, "null ctx");
  OFH_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c->stream));
  OFH_HIP(hipStreamSynchronize(c->stream));
  return OFH_OK;
}
int ofh_memcpy_d2h(ofh_ctx c, void* h, const void* d, size_t bytes) {
  OFH_CHECK(c, OFH_E_INVALID, "null ctx");
  OFH_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, c->stream));
  OFH_HIP(hipStreamSynchronize(c->stream));
  return OFH_OK;
}
int ofh_memcpy_d2d(ofh_ctx c, void* d, const void* s, size_t bytes) {
  OFH_CHECK(c, OFH_E_INVALID, "null ctx");
  OFH_HIP(hipMemcpyAsync(d, s, bytes, hipMemcpyDeviceToDevice, c->stream));
  return OFH_OK;
}

// ---------------------------------------------------------------- CSR
int ofh_csr_create(ofh_ctx c, int n, int64_t nnz, const int32_t* rp, const int32_t* col,
                   const double* val, ofh_csr* out) {
  OFH_CHECK(c && rp && col && out, OFH_E_INVALID, "ofh_csr_create: null argument");
  OFH_CHECK(n > 0 && n <= kMaxRows && nnz >= 0 && nnz < (int64_t)0x7fffffff, OFH_E_INVALID,
            "ofh_csr_create: bad size");
  OFH_CHECK(rp[0] == 0 && rp[n] == nnz, OFH_E_INVALID,
            "ofh_csr_create: row_ptr must start at 0 and end at nnz");
  int maxr = 0;
  for (int i = 0; i < n; ++i) {
    OFH_CHECK(rp[i + 1] >= rp[i], OFH_E_INVALID, "ofh_csr_create: row_ptr not monotone");
    maxr = std::max(maxr, rp[i + 1] - rp[i]);
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      OFH_CHECK(col[k] >= 0 && col[k] < n, OFH_E_INVALID, "ofh_csr_create: column out of range");
      OFH_CHECK(k == rp[i] || col[k] > col[k - 1], OFH_E_INVALID,
                "ofh_csr_create: columns must be sorted ascending and unique per row");
    }
  }
  ofh_csr A = new ofh_csr_s();
  A->ctx = c;
  A->serial = next_csr_serial();
  A->n = n;
  A->nnz = nnz;
  A->max_row_nnz = maxr;
  OFH_HIP(hipSetDevice(c->device));
  OFH_HIP(hipMalloc(&A->row_ptr, sizeof(int) * (n + 1)));
  OFH_HIP(hipMalloc(&A->col, sizeof(int) * (std::max<int64_t>(nnz, 1) + kPad)));
  OFH_HIP(hipMalloc(&A->val, sizeof(double) * (std::max<int64_t>(nnz, 1) + kPad)));
  OFH_HIP(hipMemsetAsync(A->col + nnz, 0, sizeof(int) * kPad, c->stream));
  OFH_HIP(hipMemsetAsync(A->val + nnz, 0, sizeof(double) * kPad, c->stream));
  OFH_HIP(hipMemcpyAsync(A->row_ptr, rp, sizeof(int) * (n + 1), hipMemcpyHostToDevice, c->stream));
  OFH_HIP(hipMemcpyAsync(A->col, col, sizeof(int) * nnz, hipMemcpyHostToDevice, c->stream));
  if (val)
    OFH_HIP(hipMemcpyAsync(A->val, val, sizeof(double) * nnz, hipMemcpyHostToDevice, c->stream));
  else
    OFH_HIP(hipMemsetAsync(A->val, 0, sizeof(double) * std::max<int64_t>(nnz, 1), c->stream));
  {
    const int rc = csr_set_row_blocks(A, rp);
    if (rc) return rc;
  }
  OFH_HIP(hipStreamSynchronize(c->stream));
  *out = A;
  return OFH_OK;
}

int ofh_csr_compress_columns(ofh_csr A, int enable) {
  OFH_CHECK(A, OFH_E_INVALID, "null csr");
  if (enable) return csr_build_rel16(A);
  OFH_HIP(hipStreamSynchronize(A->ctx->stream));
  if (A->col16) OFH_HIP(hipFree(A->col16));
  A->col16 = nullptr;
  return OFH_OK;
}

int ofh_csr_destroy(ofh_csr A) {
  if (!A) return OFH_OK;
  hipDeviceSynchronize();  // the context may already be gone: do not touch it
  hipFree(A->row_ptr);
  hipFree(A->col);
  hipFree(A->col16);
  hipFree(A->val);
  hipFree(A->val32);
  hipFree(A->row_blocks);
  if (A->row_blocks32) hipFree(A->row_blocks32);
  hipFree(A->dinv);
  hipFree(A->dinv32);
  hipFree(A->diag_pos);
  for (int d = 0; d < 2; ++d) {
    hipFree(A->sw_list[d]);
    hipFree(A->sw_ilev[d]);
    hipFree(A->sw_chunks[d]);
  }
  hipFree(A->sw_loc);
  hipFree(A->sweep_ctr);
  hipFree(A->ssor_tmp);
  hipFree(A->dilu_tpos);
  hipFree(A->dilu_inv);
  hipFree(A->tr_perm);
  ofh_csr_destroy(A->tr);
  hipFree(A->sell_cptr);
  hipFree(A->sell_clen);
  hipFree(A->sell_col);
  hipFree(A->sell_col16);
  hipFree(A->sell_val);
  hipFree(A->sell_val32);
  hipFree(A->dia_val32);
  hipFree(A->dia_val64);
  hipFree(A->dia_val32s);
  hipFree(A->dscale);
  hipFree(A->dias_mask);
  hipFree(A->dias_val64);
  hipFree(A->dias_val32);
  hipFree(A->dias_val32s);
  delete A;
  return OFH_OK;
}

int ofh_csr_info(ofh_csr A, int* n, int64_t* nnz, int* maxr) {
  OFH_CHECK(A, OFH_E_INVALID, "null csr");
  if (n) *n = A->n;
  if (nnz) *nnz = A->nnz;
  if (maxr) *maxr = A->max_row_nnz;
  return OFH_OK;
}

int ofh_csr_get_pattern(ofh_csr A, int32_t* rp, int32_t* col) {
  OFH_CHECK(A, OFH_E_INVALID, "null csr");
  hipStream_t s = A->ctx->stream;
  if (rp) OFH_HIP(hipMemcpyAsync(rp, A->row_ptr, sizeof(int) * (A->n + 1), hipMemcpyDeviceToHost, s));
  if (col) OFH_HIP(hipMemcpyAsync(col, A->col, sizeof(int) * A->nnz, hipMemcpyDeviceToHost, s));
  OFH_HIP(hipStreamSynchronize(s));
  return OFH_OK;
}

int ofh_csr_set_values(ofh_csr A, const double* h) {
  OFH_CHECK(A && h, OFH_E_INVALID, "null argument");
  OFH_HIP(hipMemcpyAsync(A->val, h, sizeof(double) * A->nnz, hipMemcpyHostToDevice, A->ctx->stream));
  OFH_HIP(hipStreamSynchronize(A->ctx->stream));
  csr_values_changed(A);
  return OFH_OK;
}

int ofh_csr_set_values_device(ofh_csr A, const double* d) {
  OFH_CHECK(A && d, OFH_E_INVALID, "null argument");
  OFH_HIP(hipMemcpyAsync(A->val, d, sizeof(double) * A->nnz, hipMemcpyDeviceToDevice, A->ctx->stream));
  csr_values_changed(A);
  return OFH_OK;
}

int ofh_csr_get_values(ofh_csr A, double* h) {
  OFH_CHECK(A && h, OFH_E_INVALID, "null argument");
  OFH_HIP(hipMemcpyAsync(h, A->val, sizeof(double) * A->nnz, hipMemcpyDeviceToHost, A->ctx->stream));
  OFH_HIP(hipStreamSynchronize(A->ctx->stream));
  return OFH_OK;
}

int ofh_csr_values_device(ofh_csr A, double** d) {
  OFH_CHECK(A && d, OFH_E_INVALID, "null argument");
  *d = A->val;
  csr_values_changed(A);  // the caller may write through the pointer
  return OFH_OK;
}

// ---------------------------------------------------------------- BLAS-1
int ofh_dot_f64(ofh_ctx c, int n, const double* x, const double* y, double* h_out) {
  OFH_CHECK(c && h_out && n >= 0 && (n == 0 || (x && y)), OFH_E_INVALID,
            "ofh_dot_f64: bad argument");
  double* d_out = c->dev_scratch + 8;
  int rc = launch_dot(c, n, x, y, d_out, false);
  if (rc) return rc;
  OFH_HIP(hipMemcpyAsync(h_out, d_out, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  OFH_HIP(hipStreamSynchronize(c->stream));
  return OFH_OK;
}

int ofh_norm2_f64(ofh_ctx c, int n, const double* x, double* h_out) {
  OFH_CHECK(c && h_out && n >= 0 && (n == 0 || x), OFH_E_INVALID, "ofh_norm2_f64: bad argument");
  double* d_out = c->dev_scratch + 8;
  int rc = launch_dot(c, n, x, x, d_out, true);
  if (rc) return rc;
  OFH_HIP(hipMemcpyAsync(h_out, d_out, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  OFH_HIP(hipStreamSynchronize(c->stream));
  return OFH_OK;
}

int ofh_axpy_f64(ofh_ctx c, int n, double a, const double* x, double* y) {
  OFH_CHECK(c && n >= 0 && (n == 0 || (x && y)), OFH_E_INVALID, "ofh_axpy_f64: bad argument");
  if (n == 0) return OFH_OK;
  hipLaunchKernelGGL(k_axpy, dim3((n + 255) / 256), dim3(256), 0, c->stream, n, a, x, y);
  OFH_HIP(hipGetLastError());
  return OFH_OK;
}

}  // extern "C"
