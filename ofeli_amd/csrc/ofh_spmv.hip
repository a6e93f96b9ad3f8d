// ofh_spmv.hip -- CSR SpMV for gfx950 (SpMatrix::Mult / MultAdd replacement).
//
// Reference: SpMatrix_impl.h:728-757 -- row loop, y[i] += a_k * x[col_k-1],
// summed left to right.  Design here ("CSR-stream"): each 256-thread
// workgroup owns a run of consecutive rows whose nonzeros fit an LDS tile.
//   phase 1: the workgroup streams val/col of the whole run with fully
//            coalesced loads (4 independent loads in flight per thread),
//            gathers x and stores the products a_k*x_c into LDS;
//   phase 2: one thread per row adds its products from LDS left to right.
// Phase 2 keeps the reference's summation order, so with FP contraction off
// (the library is built with -ffp-contract=off) every y_i is bitwise equal to
// the reference's sequential row sum.  Workgroup ids are remapped so each
// XCD walks a contiguous range of rows (shared x window in its L2).
//
// For y = A x the matrix is streamed instead from a SELL-64 copy (see
// ofh_spmv.h) whenever its padding stays within a quarter of nnz: one row
// per lane, coalesced loads of 64 rows' entries per instruction, no LDS and
// no barrier (measured on MI355X, 96^3 P1 matrix: 24.9-25.3 us against 31.8 us
// for the CSR-stream kernel, bitwise equal).  y += A x keeps the CSR-stream
// kernel: with a scaled product the padding could turn a -0.0 sum into +0.0.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "ofh_internal.h"
#include "ofh_spmv.h"

namespace ofh {

// ---------------------------------------------------------------------------
// SELL-64 copy
// ---------------------------------------------------------------------------
// lane = row of its chunk: columns of the row's stored slots, as 32-bit
// indices and as 16-bit offsets from the chunk's first row (*bad = 1 when one
// does not fit)
__global__ __launch_bounds__(kBlock) void k_sell_cols(int n, const int* __restrict__ rp,
                                                      const int* __restrict__ col,
                                                      const int* __restrict__ cptr,
                                                      const int* __restrict__ clen,
                                                      uint32_t* __restrict__ scol,
                                                      int16_t* __restrict__ scol16,
                                                      int* __restrict__ bad) {
  const int row = blockIdx.x * kBlock + threadIdx.x;
  const int q = row / kSellChunk;
  if (q * kSellChunk >= n) return;
  const int base = cptr[q] + (row & (kSellChunk - 1)), len = clen[q];
  const int k0 = row < n ? rp[row] : 0, nr = row < n ? rp[row + 1] - k0 : 0;
  const int row0 = q * kSellChunk;
  bool fits = true;
  for (int j = 0; j < len; ++j) {
    const int e = base + j * kSellChunk;
    if (j < nr) {
      const int c = col[k0 + j];
      const int o = c - row0;
      fits = fits && o >= -32767 && o <= 32767;
      scol[e] = (uint32_t)c;
      scol16[e] = (int16_t)(fits ? o : 0);
    } else {
      scol[e] = kSellPadCol;
      scol16[e] = (int16_t)kSellPad16;
    }
  }
  if (!fits) *bad = 1;  // every writer stores the same value
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_sell_vals(int n, const int* __restrict__ rp,
                                                      const double* __restrict__ val,
                                                      const int* __restrict__ cptr,
                                                      const int* __restrict__ clen,
                                                      T* __restrict__ sval) {
  const int row = blockIdx.x * kBlock + threadIdx.x;
  const int q = row / kSellChunk;
  if (q * kSellChunk >= n) return;
  const int base = cptr[q] + (row & (kSellChunk - 1)), len = clen[q];
  const int k0 = row < n ? rp[row] : 0, nr = row < n ? rp[row + 1] - k0 : 0;
  for (int j = 0; j < len; ++j) sval[base + j * kSellChunk] = j < nr ? (T)val[k0 + j] : T(-0.0);
}

int csr_ensure_sell(ofh_csr A, bool fp32, bool* ok) {
  *ok = false;
  if (A->sell_state < 0 || A->col16) return OFH_OK;
  hipStream_t s = A->ctx->stream;
  const int n = A->n;
  const int grid = (n + kBlock - 1) / kBlock;
  if (A->sell_state == 0) {
    std::vector<int> rp(n + 1);
    OFH_HIP(hipMemcpyAsync(rp.data(), A->row_ptr, sizeof(int) * (n + 1), hipMemcpyDeviceToHost, s));
    OFH_HIP(hipStreamSynchronize(s));
    const int nch = (n + kSellChunk - 1) / kSellChunk;
    std::vector<int> cptr(nch), clen(nch);
    int64_t total = 0;
    for (int q = 0; q < nch; ++q) {
      int len = 0;
      for (int r = q * kSellChunk; r < std::min(n, (q + 1) * kSellChunk); ++r)
        len = std::max(len, rp[r + 1] - rp[r]);
      cptr[q] = (int)std::min<int64_t>(total, INT32_MAX);
      clen[q] = len;
      total += (int64_t)kSellChunk * len;
    }
    if (total > A->nnz + A->nnz / 4 || total >= INT32_MAX) {
      A->sell_state = -1;
      return OFH_OK;
    }
    int *d_cptr = nullptr, *d_clen = nullptr, *d_bad = nullptr;
    uint32_t* d_col = nullptr;
    int16_t* d_col16 = nullptr;
    double* d_val = nullptr;
    const size_t tot = (size_t)std::max<int64_t>(total, 1);
    OFH_HIP(hipMalloc(&d_cptr, sizeof(int) * nch));
    OFH_HIP(hipMalloc(&d_clen, sizeof(int) * nch));
    OFH_HIP(hipMalloc(&d_bad, sizeof(int)));
    OFH_HIP(hipMalloc(&d_col, sizeof(uint32_t) * tot));
    OFH_HIP(hipMalloc(&d_col16, sizeof(int16_t) * tot));
    OFH_HIP(hipMalloc(&d_val, sizeof(double) * tot));
    OFH_HIP(hipMemcpyAsync(d_cptr, cptr.data(), sizeof(int) * nch, hipMemcpyHostToDevice, s));
    OFH_HIP(hipMemcpyAsync(d_clen, clen.data(), sizeof(int) * nch, hipMemcpyHostToDevice, s));
    OFH_HIP(hipMemsetAsync(d_bad, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_sell_cols, dim3(grid), dim3(kBlock), 0, s, n, A->row_ptr, A->col, d_cptr,
                       d_clen, d_col, d_col16, d_bad);
    OFH_HIP(hipGetLastError());
    int bad = 0;
    OFH_HIP(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, s));
    OFH_HIP(hipStreamSynchronize(s));  // cptr, clen leave scope
    OFH_HIP(hipFree(d_bad));
    if (bad) {  // some column lies too far from its chunk: 32-bit columns only
      OFH_HIP(hipFree(d_col16));
      d_col16 = nullptr;
    }
    A->sell_nchunks = nch;
    A->sell_total = total;
    A->sell_cptr = d_cptr;
    A->sell_clen = d_clen;
    A->sell_col = d_col;
    A->sell_col16 = d_col16;
    A->sell_val = d_val;
    A->sell_version = A->sell_version32 = ~0ull;
    A->sell_state = 1;
  }
  if (!fp32 && A->sell_version != A->version) {
    hipLaunchKernelGGL(k_sell_vals<double>, dim3(grid), dim3(kBlock), 0, s, n, A->row_ptr, A->val,
                       A->sell_cptr, A->sell_clen, A->sell_val);
    OFH_HIP(hipGetLastError());
    A->sell_version = A->version;
  }
  if (fp32 && A->sell_version32 != A->version) {
    if (!A->sell_val32)
      OFH_HIP(hipMalloc(&A->sell_val32, sizeof(float) * std::max<int64_t>(A->sell_total, 1)));
    hipLaunchKernelGGL(k_sell_vals<float>, dim3(grid), dim3(kBlock), 0, s, n, A->row_ptr, A->val,
                       A->sell_cptr, A->sell_clen, A->sell_val32);
    OFH_HIP(hipGetLastError());
    A->sell_version32 = A->version;
  }
  *ok = true;
  return OFH_OK;
}

// ---------------------------------------------------------------------------
// Diagonal copies (FP32, FP64)
// ---------------------------------------------------------------------------
struct DiaOffsets {
  int nd;
  int off[kDiaMaxDiag];
};

template <typename T>
__device__ __forceinline__ T dia_mask_value(uint32_t m) {
  using U = typename DiaMaskBits<T>::U;
  return __builtin_bit_cast(T, (U)m);
}

// thread = row: the row's values into their diagonal slots, the mask last;
// with `scale` (D^-1/2) the values of D^-1/2 A D^-1/2, a_ij (s_i s_j)
template <int W, typename T>
__global__ __launch_bounds__(kBlock) void k_dia_vals(int n, const int* __restrict__ rp,
                                                     const int* __restrict__ col,
                                                     const double* __restrict__ val, DiaOffsets o,
                                                     const double* __restrict__ scale,
                                                     T* __restrict__ out) {
  const int row = blockIdx.x * kBlock + threadIdx.x;
  if (row >= n) return;
  T r[W];
#pragma unroll
  for (int d = 0; d < W - 1; ++d) r[d] = T(-0.0);
  uint32_t m = 0;
  const double si = scale ? scale[row] : 1.0;
  for (int k = rp[row]; k < rp[row + 1]; ++k) {
    const int c = col[k] - row;
    const T v = scale ? (T)(val[k] * (si * scale[col[k]])) : (T)val[k];
#pragma unroll
    for (int d = 0; d < W - 1; ++d)
      if (d < o.nd && c == o.off[d]) {
        r[d] = v;
        m |= 1u << d;
      }
  }
  r[W - 1] = dia_mask_value<T>(m);
  T* base = out + (size_t)(row / kDiaChunk) * (W * kDiaChunk) + (row % kDiaChunk);
#pragma unroll
  for (int d = 0; d < W; ++d) base[d * kDiaChunk] = r[d];
}

static size_t dia_elems(ofh_csr A) {
  return (size_t)((A->n + kDiaChunk - 1) / kDiaChunk) * kDiaChunk * A->dia_w;
}

// values of a diagonal copy (scale non-null: of D^-1/2 A D^-1/2) into out
template <typename T>
static int launch_dia_vals(ofh_csr A, const double* scale, T* out) {
  DiaOffsets o{A->dia_nd, {}};
  for (int d = 0; d < kDiaMaxDiag; ++d) o.off[d] = A->dia_off[d];
  const int n = A->n;
  const int grid = (n + kBlock - 1) / kBlock;
  hipStream_t s = A->ctx->stream;
  if (A->dia_w == 4)
    hipLaunchKernelGGL((k_dia_vals<4, T>), dim3(grid), dim3(kBlock), 0, s, n, A->row_ptr, A->col,
                       A->val, o, scale, out);
  else if (A->dia_w == 8)
    hipLaunchKernelGGL((k_dia_vals<8, T>), dim3(grid), dim3(kBlock), 0, s, n, A->row_ptr, A->col,
                       A->val, o, scale, out);
  else
    hipLaunchKernelGGL((k_dia_vals<16, T>), dim3(grid), dim3(kBlock), 0, s, n, A->row_ptr,
                       A->col, A->val, o, scale, out);
  OFH_HIP(hipGetLastError());
  return OFH_OK;
}

// The pattern's diagonals, once per pattern (host pass over row_ptr / col).
static int dia_detect(ofh_csr A) {
  if (A->dia_state != 0) return OFH_OK;
  hipStream_t s = A->ctx->stream;
  const int n = A->n;
  std::vector<int> rp(n + 1), col((size_t)std::max<int64_t>(A->nnz, 1));
  OFH_HIP(hipMemcpyAsync(rp.data(), A->row_ptr, sizeof(int) * (n + 1), hipMemcpyDeviceToHost, s));
  OFH_HIP(hipMemcpyAsync(col.data(), A->col, sizeof(int) * A->nnz, hipMemcpyDeviceToHost, s));
  OFH_HIP(hipStreamSynchronize(s));
  // distinct column - row offsets, kept sorted; within a row the columns
  // must ascend, so its offsets are met in ascending order and one cursor
  // walks the sorted list
  std::vector<int> offs;
  bool fits = A->nnz > 0;
  int64_t n_lower = 0, n_upper = 0;
  for (int i = 0; i < n && fits; ++i) {
    size_t j = 0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      if (k > rp[i] && col[k] <= col[k - 1]) {
        fits = false;
        break;
      }
      const int o = col[k] - i;
      n_lower += o < 0;
      n_upper += o > 0;
      while (j < offs.size() && offs[j] < o) ++j;
      if (j == offs.size() || offs[j] != o) {
        if ((int)offs.size() == kDiaMaxDiag) {
          fits = false;
          break;
        }
        offs.insert(offs.begin() + j, o);
      }
      ++j;
    }
  }
  if (!fits) {
    A->dia_state = -1;
    return OFH_OK;
  }
  const int nd = (int)offs.size();
  A->dia_nd = nd;
  A->dia_w = nd <= 3 ? 4 : nd <= 7 ? 8 : 16;
  for (int d = 0; d < 16; ++d) A->dia_off[d] = d < nd ? offs[d] : 0;
  // the symmetric form needs the diagonals in pairs +-o around the main one
  // and as many entries above it as below (with every upper entry's mirror
  // present, checked with the values, the patterns then match)
  const int h = (nd - 1) / 2;
  bool pairs = nd % 2 == 1 && offs[h] == 0 && n_lower == n_upper;
  for (int d = 0; d < h && pairs; ++d) pairs = offs[d] == -offs[nd - 1 - d];
  A->dias_pat = pairs;
  for (int k = 0; k < 8; ++k) A->dias_up[k] = pairs && k >= 1 && k <= h ? offs[h + k] : 0;
  // each copy only where it takes no more bytes than the SELL-64 copy of the
  // same precision with 16-bit columns (6 and 10 bytes per entry)
  const int64_t rows = (int64_t)((n + kDiaChunk - 1) / kDiaChunk) * kDiaChunk;
  A->dia_use32 = 4 * A->dia_w * rows <= 6 * A->nnz;
  A->dia_use64 = 8 * A->dia_w * rows <= 10 * A->nnz;
  A->dia_state = A->dia_use32 || A->dia_use64 ? 1 : -1;
  return OFH_OK;
}

int csr_ensure_dia(ofh_csr A, bool* ok) {
  *ok = false;
  int rc = dia_detect(A);
  if (rc || A->dia_state < 0 || !A->dia_use32) return rc;
  if (!A->dia_val32) OFH_HIP(hipMalloc(&A->dia_val32, sizeof(float) * dia_elems(A)));
  if (A->dia_version32 != A->version) {
    if ((rc = launch_dia_vals<float>(A, nullptr, A->dia_val32))) return rc;
    A->dia_version32 = A->version;
  }
  *ok = true;
  return OFH_OK;
}

int csr_ensure_dia64(ofh_csr A, bool* ok) {
  *ok = false;
  int rc = dia_detect(A);
  if (rc || A->dia_state < 0 || !A->dia_use64) return rc;
  if (!A->dia_val64) OFH_HIP(hipMalloc(&A->dia_val64, sizeof(double) * dia_elems(A)));
  if (A->dia_version64 != A->version) {
    if ((rc = launch_dia_vals<double>(A, nullptr, A->dia_val64))) return rc;
    A->dia_version64 = A->version;
  }
  *ok = true;
  return OFH_OK;
}

// s_i = 1 / sqrt(a_ii); *bad = 1 when a pivot is not positive and finite
__global__ void k_dscale(int n, const int* __restrict__ diag_pos, const double* __restrict__ val,
                         double* __restrict__ s, int* __restrict__ bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int p = diag_pos[i];
  const double a = p < 0 ? 0.0 : val[p];
  if (!(a > 0.0) || !(a < INFINITY)) {
    *bad = 1;  // every writer stores the same value
    s[i] = 0.0;
    return;
  }
  s[i] = 1.0 / sqrt(a);
}

int csr_ensure_dia_scaled(ofh_csr A, bool* ok) {
  *ok = false;
  bool dia = false;
  int rc = csr_ensure_dia(A, &dia);
  if (rc || !dia) return rc;
  if (A->dscale_version != A->version) {
    if ((rc = csr_ensure_dinv(A))) return rc;  // diag_pos; a zero pivot fails as in DIAG_PREC
    hipStream_t s = A->ctx->stream;
    const int n = A->n;
    if (!A->dscale) OFH_HIP(hipMalloc(&A->dscale, sizeof(double) * n));
    if (!A->dia_val32s) OFH_HIP(hipMalloc(&A->dia_val32s, sizeof(float) * dia_elems(A)));
    int* bad = (int*)A->ctx->dev_scratch;
    OFH_HIP(hipMemsetAsync(bad, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_dscale, dim3((n + 255) / 256), dim3(256), 0, s, n, A->diag_pos, A->val,
                       A->dscale, bad);
    OFH_HIP(hipGetLastError());
    int hbad = 0;
    OFH_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
    OFH_HIP(hipStreamSynchronize(s));
    A->dscale_ok = hbad == 0;
    if (A->dscale_ok && (rc = launch_dia_vals<float>(A, A->dscale, A->dia_val32s))) return rc;
    A->dscale_version = A->version;
  }
  *ok = A->dscale_ok;
  return OFH_OK;
}

// *bad = 1 unless every entry above the diagonal has its mirror a_ji with
// the same bits (thread = row; rows of a diagonal copy hold <= 15 entries)
__global__ void k_dias_check(int n, const int* __restrict__ rp, const int* __restrict__ col,
                             const double* __restrict__ val, int* __restrict__ bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int k = rp[i]; k < rp[i + 1]; ++k) {
    const int c = col[k];
    if (c <= i) continue;
    bool same = false;
    for (int q = rp[c]; q < rp[c + 1]; ++q)
      if (col[q] == i) {
        same = __builtin_bit_cast(uint64_t, val[q]) == __builtin_bit_cast(uint64_t, val[k]);
        break;
      }
    if (!same) *bad = 1;  // every writer stores the same value
  }
}

struct DiaSymOffsets {
  int up[8];
};

// thread = row: the diagonal and the upper values into the K slots of the
// symmetric copy (scale: of D^-1/2 A D^-1/2), and the row's mask
template <int K, typename T>
__global__ __launch_bounds__(kBlock) void k_dias_vals(int n, const int* __restrict__ rp,
                                                      const int* __restrict__ col,
                                                      const double* __restrict__ val,
                                                      DiaSymOffsets o,
                                                      const double* __restrict__ scale,
                                                      T* __restrict__ out,
                                                      uint16_t* __restrict__ mask) {
  const int row = blockIdx.x * kBlock + threadIdx.x;
  if (row >= n) return;
  T r[K];
#pragma unroll
  for (int k = 0; k < K; ++k) r[k] = T(0);
  uint32_t m = 0;
  const double si = scale ? scale[row] : 1.0;
  for (int q = rp[row]; q < rp[row + 1]; ++q) {
    const int c = col[q] - row;
    const T v = scale ? (T)(val[q] * (si * scale[col[q]])) : (T)val[q];
    if (c == 0) {
      r[0] = v;
      m |= 1u << (K - 1);
    }
#pragma unroll
    for (int k = 1; k < K; ++k) {
      if (c == o.up[k]) {
        r[k] = v;
        m |= 1u << (K - 1 + k);
      }
      if (c == -o.up[k]) m |= 1u << (K - 1 - k);  // read from the partner row
    }
  }
  T* base = out + (size_t)(row / kDiaChunk) * (K * kDiaChunk) + (row % kDiaChunk);
#pragma unroll
  for (int k = 0; k < K; ++k) base[k * kDiaChunk] = r[k];
  mask[row] = (uint16_t)m;
}

template <typename T>
static int launch_dias_vals(ofh_csr A, const double* scale, T* out) {
  DiaSymOffsets o{};
  for (int k = 0; k < 8; ++k) o.up[k] = A->dias_up[k];
  const int n = A->n;
  const int grid = (n + kBlock - 1) / kBlock;
  hipStream_t s = A->ctx->stream;
  const int K = A->dia_w / 2;
  if (K == 2)
    hipLaunchKernelGGL((k_dias_vals<2, T>), dim3(grid), dim3(kBlock), 0, s, n, A->row_ptr, A->col,
                       A->val, o, scale, out, A->dias_mask);
  else if (K == 4)
    hipLaunchKernelGGL((k_dias_vals<4, T>), dim3(grid), dim3(kBlock), 0, s, n, A->row_ptr, A->col,
                       A->val, o, scale, out, A->dias_mask);
  else
    hipLaunchKernelGGL((k_dias_vals<8, T>), dim3(grid), dim3(kBlock), 0, s, n, A->row_ptr, A->col,
                       A->val, o, scale, out, A->dias_mask);
  OFH_HIP(hipGetLastError());
  return OFH_OK;
}

int csr_dia_symmetric(ofh_csr A, bool* sym) {
  *sym = false;
  int rc = dia_detect(A);
  if (rc || A->dia_state < 0 || !A->dias_pat) return rc;
  hipStream_t s = A->ctx->stream;
  const int n = A->n;
  if (A->dias_check_version != A->version) {
    int* bad = (int*)A->ctx->dev_scratch;
    OFH_HIP(hipMemsetAsync(bad, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_dias_check, dim3((n + 255) / 256), dim3(256), 0, s, n, A->row_ptr,
                       A->col, A->val, bad);
    OFH_HIP(hipGetLastError());
    int hbad = 0;
    OFH_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
    OFH_HIP(hipStreamSynchronize(s));
    A->dias_ok = hbad == 0;
    A->dias_check_version = A->version;
  }
  *sym = A->dias_ok;
  return OFH_OK;
}

int csr_ensure_dias(ofh_csr A, int kind, bool* ok) {
  *ok = false;
  bool sym = false;
  int rc = csr_dia_symmetric(A, &sym);
  if (rc || !sym) return rc;
  if (kind == 0 ? !A->dia_use64 : !A->dia_use32) return OFH_OK;
  if (kind == 2 && !(A->dscale_ok && A->dscale_version == A->version)) return OFH_OK;
  const int n = A->n;
  const size_t elems = dia_elems(A) / 2;
  if (!A->dias_mask) OFH_HIP(hipMalloc(&A->dias_mask, sizeof(uint16_t) * (size_t)n));
  if (kind == 0) {
    if (!A->dias_val64) OFH_HIP(hipMalloc(&A->dias_val64, sizeof(double) * elems));
    if (A->dias_version64 != A->version) {
      if ((rc = launch_dias_vals<double>(A, nullptr, A->dias_val64))) return rc;
      A->dias_version64 = A->version;
    }
  } else if (kind == 1) {
    if (!A->dias_val32) OFH_HIP(hipMalloc(&A->dias_val32, sizeof(float) * elems));
    if (A->dias_version32 != A->version) {
      if ((rc = launch_dias_vals<float>(A, nullptr, A->dias_val32))) return rc;
      A->dias_version32 = A->version;
    }
  } else {
    if (!A->dias_val32s) OFH_HIP(hipMalloc(&A->dias_val32s, sizeof(float) * elems));
    if (A->dias_version32s != A->version) {
      if ((rc = launch_dias_vals<float>(A, A->dscale, A->dias_val32s))) return rc;
      A->dias_version32s = A->version;
    }
  }
  *ok = true;
  return OFH_OK;
}

DiaSymRows<double> dias_f64(ofh_csr A) {
  DiaSymRows<double> d{A->n, A->dias_val64, A->dias_mask, {}};
  for (int k = 0; k < 8; ++k) d.up[k] = A->dias_up[k];
  return d;
}

DiaSymRows<float> dias_f32(ofh_csr A, bool scaled) {
  DiaSymRows<float> d{A->n, scaled ? A->dias_val32s : A->dias_val32, A->dias_mask, {}};
  for (int k = 0; k < 8; ++k) d.up[k] = A->dias_up[k];
  return d;
}

DiaF32 dia_f32(ofh_csr A) {
  DiaF32 d{A->n, A->dia_val32, {}};
  for (int k = 0; k < kDiaMaxDiag; ++k) d.off[k] = A->dia_off[k];
  return d;
}

DiaF64 dia_f64(ofh_csr A) {
  DiaF64 d{A->n, A->dia_val64, {}};
  for (int k = 0; k < kDiaMaxDiag; ++k) d.off[k] = A->dia_off[k];
  return d;
}

// y = A x from a diagonal copy (SYM: its symmetric form)
template <int W, typename T, bool SYM>
__global__ __launch_bounds__(kBlock) void k_spmv_dia(int nblocks, DiaRows<T> da,
                                                     DiaSymRows<T> ds, const T* x,
                                                     T* __restrict__ y) {
  const int row = xcd_block(blockIdx.x, nblocks) * kBlock + threadIdx.x;
  if (row >= da.n) return;
  y[row] = dia_row<W, SYM>(row, da, ds, SellGather<T>{x_rsrc(x, da.n)});
}

template <typename T>
static int launch_spmv_dia(ofh_csr A, const DiaRows<T>& da, const DiaSymRows<T>* ds, const T* x,
                           T* y, int kind, hipStream_t s) {
  const int nb = (A->n + kBlock - 1) / kBlock;
  const DiaSymRows<T> dsv = ds ? *ds : DiaSymRows<T>{};
#define OFH_SPMV_DIA(W)                                                                         \
  if (ds)                                                                                       \
    prof_kernel(A->ctx, kind, k_spmv_dia<W, T, true>, dim3(nb), dim3(kBlock), s, nb, da, dsv, x, \
                y);                                                                             \
  else                                                                                          \
    prof_kernel(A->ctx, kind, k_spmv_dia<W, T, false>, dim3(nb), dim3(kBlock), s, nb, da, dsv, \
                x, y);
  if (A->dia_w == 4) {
    OFH_SPMV_DIA(4)
  } else if (A->dia_w == 8) {
    OFH_SPMV_DIA(8)
  } else {
    OFH_SPMV_DIA(16)
  }
#undef OFH_SPMV_DIA
  OFH_HIP(hipGetLastError());
  return OFH_OK;
}

// y = A x from the SELL-64 copy
template <typename T, int U>
__global__ __launch_bounds__(kBlock) void k_spmv_sell(int n, int nblocks,
                                                      const int* __restrict__ cptr,
                                                      const int* __restrict__ clen,
                                                      const uint32_t* __restrict__ col,
                                                      const int16_t* __restrict__ col16,
                                                      const T* __restrict__ val, const T* x,
                                                      T* __restrict__ y) {
  const int row = xcd_block(blockIdx.x, nblocks) * kBlock + threadIdx.x;
  if (row / kSellChunk * kSellChunk >= n) return;  // whole wave past the end
  const T s = sell_rowsum<T, U>(row, cptr, clen, col, col16, val, x_rsrc(x, n));
  if (row < n) y[row] = s;
}

template <typename T, int TILE, int MODE>
__global__ __launch_bounds__(kBlock) void k_spmv(int nblocks, const int* __restrict__ rb,
                                                 const int* __restrict__ rp,
                                                 ColIndex ci,
                                                 const T* __restrict__ val,
                                                 const T* __restrict__ x, T* __restrict__ y,
                                                 T alpha) {
  __shared__ T lds[TILE];
  OFH_LDS_BUDGET(sizeof(T) * TILE);
  const int b = xcd_block(blockIdx.x, nblocks);
  const int r0 = rb[b], r1 = rb[b + 1];
  const int r = r0 + threadIdx.x;
  T init = T(0);
  if (MODE != 0 && r < r1) init = y[r];
  T s = csr_stream_rowsum<T, TILE, MODE == 2>(r0, r1, rp, ci, val, x, lds, init, alpha);
  if (r < r1) y[r] = s;
}

template <typename T, int TILE, int MODE>
static int launch(ofh_csr A, const T* val, const T* x, T* y, T alpha, hipStream_t s) {
  // FP32 runs on its own partition (larger tile) with 32-bit columns: the
  // 16-bit offsets are relative to the FP64 partition's blocks
  const bool f32 = sizeof(T) == 4;
  const int nb = f32 ? A->nblocks32 : A->nblocks;
  hipLaunchKernelGGL((k_spmv<T, TILE, MODE>), dim3(nb), dim3(kBlock), 0, s, nb,
                     f32 ? A->row_blocks32 : A->row_blocks, A->row_ptr,
                     ColIndex{A->col, f32 ? nullptr : A->col16}, val, x, y, alpha);
  OFH_HIP(hipGetLastError());
  return OFH_OK;
}

}  // namespace ofh

extern "C" int ofh_csr_dia_info(ofh_csr A, int* in_use32, int* in_use64, int* diagonals,
                                int* row_slots, int* symmetric) {
  using namespace ofh;
  OFH_CHECK(A && in_use32 && in_use64 && diagonals && row_slots && symmetric, OFH_E_INVALID,
            "ofh_csr_dia_info: null argument");
  bool ok32 = false, ok64 = false, sym = false;
  int rc = csr_ensure_dia(A, &ok32);
  if (!rc) rc = csr_ensure_dia64(A, &ok64);
  if (!rc && (ok32 || ok64)) rc = csr_ensure_dias(A, ok64 ? 0 : 1, &sym);
  if (rc) return rc;
  OFH_HIP(hipStreamSynchronize(A->ctx->stream));
  *in_use32 = ok32 ? 1 : 0;
  *in_use64 = ok64 ? 1 : 0;
  *diagonals = A->dia_state > 0 ? A->dia_nd : 0;
  *row_slots = A->dia_state > 0 ? A->dia_w : 0;
  *symmetric = sym ? 1 : 0;
  return OFH_OK;
}

extern "C" int ofh_csr_sell_info(ofh_csr A, int* in_use, int64_t* stored, int* col_bytes) {
  using namespace ofh;
  OFH_CHECK(A && in_use && stored && col_bytes, OFH_E_INVALID, "ofh_csr_sell_info: null argument");
  bool ok = false;
  const int rc = csr_ensure_sell(A, false, &ok);
  if (rc) return rc;
  OFH_HIP(hipStreamSynchronize(A->ctx->stream));
  *in_use = ok ? 1 : 0;
  *stored = ok ? A->sell_total : 0;
  *col_bytes = ok ? (A->sell_col16 ? 2 : 4) : 0;
  return OFH_OK;
}

namespace ofh {

int launch_spmv64(ofh_csr A, const double* x, double* y, hipStream_t s) {
  bool dia;
  int rc = csr_ensure_dia64(A, &dia);
  if (rc) return rc;
  if (dia) {
    bool sym;
    if ((rc = csr_ensure_dias(A, 0, &sym))) return rc;
    const DiaSymRows<double> ds = sym ? dias_f64(A) : DiaSymRows<double>{};
    return launch_spmv_dia<double>(A, dia_f64(A), sym ? &ds : nullptr, x, y, OFH_PROF_SPMV64, s);
  }
  bool sell;
  if ((rc = csr_ensure_sell(A, false, &sell))) return rc;
  if (sell) {
    const int nb = (A->n + kBlock - 1) / kBlock;
    prof_kernel(A->ctx, OFH_PROF_SPMV64, k_spmv_sell<double, 2>, dim3(nb), dim3(kBlock), s, A->n,
                nb, A->sell_cptr, A->sell_clen, A->sell_col, A->sell_col16, A->sell_val, x, y);
    OFH_HIP(hipGetLastError());
    return OFH_OK;
  }
  return launch<double, kTile64, 0>(A, A->val, x, y, 0.0, s);
}

int launch_spmv32(ofh_csr A, const float* x, float* y, hipStream_t s) {
  bool dia;
  int rc = csr_ensure_dia(A, &dia);
  if (rc) return rc;
  if (dia) {
    bool sym;
    if ((rc = csr_ensure_dias(A, 1, &sym))) return rc;
    const DiaSymRows<float> ds = sym ? dias_f32(A, false) : DiaSymRows<float>{};
    return launch_spmv_dia<float>(A, dia_f32(A), sym ? &ds : nullptr, x, y, OFH_PROF_SPMV32, s);
  }
  bool sell;
  if ((rc = csr_ensure_sell(A, true, &sell))) return rc;
  if (sell) {
    const int nb = (A->n + kBlock - 1) / kBlock;
    prof_kernel(A->ctx, OFH_PROF_SPMV32, k_spmv_sell<float, 4>, dim3(nb), dim3(kBlock), s, A->n,
                nb, A->sell_cptr, A->sell_clen, A->sell_col, A->sell_col16, A->sell_val32, x, y);
    OFH_HIP(hipGetLastError());
    return OFH_OK;
  }
  if ((rc = csr_ensure_val32(A))) return rc;
  return launch<float, kTile32, 0>(A, A->val32, x, y, 0.0f, s);
}

}  // namespace ofh

using namespace ofh;

extern "C" {

int ofh_spmv_f64(ofh_csr A, const double* x, double* y) {
  OFH_CHECK(A && x && y, OFH_E_INVALID, "ofh_spmv_f64: null argument");
  return launch_spmv64(A, x, y, A->ctx->stream);
}

int ofh_spmv_add_f64(ofh_csr A, double a, const double* x, double* y) {
  OFH_CHECK(A && x && y, OFH_E_INVALID, "ofh_spmv_add_f64: null argument");
  // MultAdd(x,y) (SpMatrix_impl.h:737-745) accumulates a_k*x onto y;
  // MultAdd(a,x,y) (:748-757) accumulates a*a_k*x.
  if (a == 1.0) return launch<double, kTile64, 1>(A, A->val, x, y, 1.0, A->ctx->stream);
  return launch<double, kTile64, 2>(A, A->val, x, y, a, A->ctx->stream);
}

int ofh_spmv_f32(ofh_csr A, const float* x, float* y) {
  OFH_CHECK(A && x && y, OFH_E_INVALID, "ofh_spmv_f32: null argument");
  return launch_spmv32(A, x, y, A->ctx->stream);
}

}  // extern "C"
