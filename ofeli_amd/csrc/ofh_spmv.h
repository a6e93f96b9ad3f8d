// ofh_spmv.h -- device-side CSR-stream row sums shared by the SpMV and the
// fused Krylov kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "ofh_internal.h"

namespace ofh {

// Products of one nonzero: a_k * x_c, or (s*a_k) * x_c when SCALED (the
// operation order of SpMatrix::MultAdd(a,x,y), SpMatrix_impl.h:748-757).
template <typename T, bool SCALED>
__device__ __forceinline__ T nz_product(T a, T xv, T s) {
  if (SCALED) return (s * a) * xv;
  return a * xv;
}

// Column of nonzero k: int32 column, or (compressed form) the block's first
// row plus a 16-bit offset -- FE orderings keep every column of a 256-row
// block within +-32767 of its first row, which saves 2 of the 12 (FP64) or
// 8 (FP32) bytes streamed per nonzero.
struct ColIndex {
  const int* __restrict__ col;
  const int16_t* __restrict__ rel;  // non-null: compressed form
};

__device__ __forceinline__ int col_at(const ColIndex& ci, int r0, int k) {
  return ci.rel ? r0 + (int)ci.rel[k] : ci.col[k];  // uniform branch, unswitched
}

// Gather of the multiplied vector: x[c].
template <typename T>
struct GatherVec {
  const T* __restrict__ x;
  __device__ __forceinline__ T operator()(int c) const { return x[c]; }
};

// Row sum of row r0+threadIdx.x over the rows [r0,r1) of one workgroup.
// The sum starts at `init` and adds the row's products left to right.
// Rows whose run exceeds TILE nonzeros come as a single-row block; that
// row's sum is returned in thread 0.  Must be called by all kBlock threads.
// X is the gather functor of the multiplied vector (a plain vector, or a
// vector formed on the fly from others).
template <typename T, int TILE, bool SCALED = false, class X = GatherVec<T>>
__device__ __forceinline__ T csr_stream_rowsum_g(int r0, int r1, const int* __restrict__ rp,
                                                 ColIndex ci, const T* __restrict__ val,
                                                 const X& x, T* lds, T init,
                                                 T scale = T(1)) {
  const int k0 = rp[r0], k1 = rp[r1];
  const int t = threadIdx.x;
  if (k1 - k0 <= TILE) {
    int k = k0 + t;
    // 4 independent val/col loads in flight per thread (measured faster on
    // MI355X than 16-byte pair loads: 35 vs 47 us on the 96^3 system)
    for (; k + 3 * kBlock < k1; k += 4 * kBlock) {
      int c0 = col_at(ci, r0, k), c1 = col_at(ci, r0, k + kBlock),
          c2 = col_at(ci, r0, k + 2 * kBlock), c3 = col_at(ci, r0, k + 3 * kBlock);
      T a0 = val[k], a1 = val[k + kBlock], a2 = val[k + 2 * kBlock], a3 = val[k + 3 * kBlock];
      T x0 = x(c0), x1 = x(c1), x2 = x(c2), x3 = x(c3);
      lds[k - k0] = nz_product<T, SCALED>(a0, x0, scale);
      lds[k + kBlock - k0] = nz_product<T, SCALED>(a1, x1, scale);
      lds[k + 2 * kBlock - k0] = nz_product<T, SCALED>(a2, x2, scale);
      lds[k + 3 * kBlock - k0] = nz_product<T, SCALED>(a3, x3, scale);
    }
    for (; k < k1; k += kBlock)
      lds[k - k0] = nz_product<T, SCALED>(val[k], x(col_at(ci, r0, k)), scale);
    __syncthreads();
    T s = init;
    const int r = r0 + t;
    if (r < r1) {
      const int a = rp[r] - k0, e = rp[r + 1] - k0;
      for (int j = a; j < e; ++j) s += lds[j];
    }
    return s;
  }
  // a single row longer than the tile: chunk through LDS, thread 0 sums
  T s = init;
  for (int c0 = k0; c0 < k1; c0 += TILE) {
    const int c1 = min(c0 + TILE, k1);
    __syncthreads();
    for (int k = c0 + t; k < c1; k += kBlock)
      lds[k - c0] = nz_product<T, SCALED>(val[k], x(col_at(ci, r0, k)), scale);
    __syncthreads();
    if (t == 0)
      for (int j = 0; j < c1 - c0; ++j) s += lds[j];
  }
  return s;
}

// ---------------------------------------------------------------------------
// SELL-64: one row per lane.  A chunk of 64 consecutive rows (one wave) is
// stored position-major -- entry j of every row of the chunk in 64
// consecutive slots -- and padded to its longest row with the value -0.0 and
// a column whose byte offset lies past the end of x, which the buffer load
// returns as +0.0: the padding adds -0.0, which leaves every sum unchanged.
// Each lane adds its row's products left to right, the reference's order.
// ---------------------------------------------------------------------------
constexpr int kSellChunk = 64;
constexpr uint32_t kSellPadCol = 0x3FFFFFFEu;  // x byte offset >= 2^31 for 4- and 8-byte x
// 16-bit form of the columns: column - first row of the chunk, in
// [-32767, 32767]; the padding slots hold kSellPad16 and read as kSellPadCol.
constexpr int kSellPad16 = -32768;

using XRsrc = __amdgpu_buffer_rsrc_t;

template <typename T>
__device__ __forceinline__ XRsrc x_rsrc(const T* x, int n) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(x), (short)0, n * (int)sizeof(T),
                                           0x00020000);
}

template <typename T>
__device__ __forceinline__ T x_load(XRsrc r, uint32_t c) {
  if constexpr (sizeof(T) == 8)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)(c * 8u), 0, 0));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)(c * 4u), 0, 0));
}

// Gather of x[c] through a buffer descriptor (padding columns read +0.0).
template <typename T>
struct SellGather {
  XRsrc x;
  __device__ __forceinline__ T operator()(uint32_t c) const { return x_load<T>(x, c); }
};

// Column of stored entry e, from the 32-bit or the 16-bit array.
struct SellCol32 {
  const uint32_t* __restrict__ c;
  __device__ __forceinline__ uint32_t operator()(int e) const { return c[e]; }
};
struct SellCol16 {
  const int16_t* __restrict__ c;
  int row0;  // first row of the chunk
  __device__ __forceinline__ uint32_t operator()(int e) const {
    const int o = c[e];
    return o == kSellPad16 ? kSellPadCol : (uint32_t)(row0 + o);
  }
};

// Sum of row `row` (chunk q = row / 64) of the SELL copy, from 0, with U
// entries' loads in flight; C reads the columns, G gathers the multiplied
// vector and must return +0.0 for the padding column.
template <typename T, int U, class C, class G>
__device__ __forceinline__ T sell_rowsum_cg(int row, const int* __restrict__ cptr,
                                            const int* __restrict__ clen, const C& col,
                                            const T* __restrict__ val, const G& g) {
  const int q = row / kSellChunk;
  const int base = cptr[q] + (row & (kSellChunk - 1));
  const int len = clen[q];
  T s = T(0);
  int j = 0;
  for (; j + U <= len; j += U) {
    uint32_t c[U];
    T a[U], v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      c[u] = col(base + (j + u) * kSellChunk);
      a[u] = val[base + (j + u) * kSellChunk];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = g(c[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) s += a[u] * v[u];
  }
  for (; j < len; ++j) s += val[base + j * kSellChunk] * g(col(base + j * kSellChunk));
  return s;
}

// The same from the 16-bit columns when the copy has them (col16 non-null,
// the same for every lane of the launch), else from the 32-bit ones.
template <typename T, int U, class G>
__device__ __forceinline__ T sell_rowsum_g(int row, const int* __restrict__ cptr,
                                           const int* __restrict__ clen,
                                           const uint32_t* __restrict__ col,
                                           const int16_t* __restrict__ col16,
                                           const T* __restrict__ val, const G& g) {
  if (col16)
    return sell_rowsum_cg<T, U>(row, cptr, clen, SellCol16{col16, row & ~(kSellChunk - 1)}, val,
                                g);
  return sell_rowsum_cg<T, U>(row, cptr, clen, SellCol32{col}, val, g);
}

template <typename T, int U>
__device__ __forceinline__ T sell_rowsum(int row, const int* __restrict__ cptr,
                                         const int* __restrict__ clen,
                                         const uint32_t* __restrict__ col,
                                         const int16_t* __restrict__ col16,
                                         const T* __restrict__ val, XRsrc xr) {
  return sell_rowsum_g<T, U>(row, cptr, clen, col, col16, val, SellGather<T>{xr});
}

// ---------------------------------------------------------------------------
// Diagonal copy (FP32 and FP64): row i has W slots -- slot d the value of the
// entry at column i + off[d] (offsets ascending, so slot order is column
// order), slot W-1 the bit mask of the slots the row really has.  Rows are
// stored by chunks of 64 (one wave), slot-major inside a chunk: slot d of
// the chunk's 64 rows occupies 64 consecutive elements, so every load of a
// wave reads one contiguous run.  A lane loads its row's W slots and every
// x[i + off[d]] at once: the addresses need no column stream, so all loads of
// a row are in flight together, and across a wave each x load reads 64
// consecutive entries.  The sum adds the present entries left to right from
// +0.0, which is the CSR row sum bit for bit; x is read through a buffer
// descriptor, so a column outside [0, n) of an absent slot reads +0.0 and is
// masked out anyway.
// ---------------------------------------------------------------------------
constexpr int kDiaMaxDiag = 15;
constexpr int kDiaChunk = 64;

template <typename T>
struct DiaRows {
  int n;
  const T* __restrict__ v;  // ceil(n / 64) * W * 64
  int off[kDiaMaxDiag];     // unused slots 0
};
using DiaF32 = DiaRows<float>;
using DiaF64 = DiaRows<double>;

// The diagonal copies of A as the kernels take them (host side).
DiaF32 dia_f32(ofh_csr A);
DiaF64 dia_f64(ofh_csr A);

template <typename T>
struct DiaMaskBits;
template <>
struct DiaMaskBits<float> {
  using U = uint32_t;
};
template <>
struct DiaMaskBits<double> {
  using U = uint64_t;
};

template <int W, typename T, class G>
__device__ __forceinline__ T dia_rowsum_g(int row, const DiaRows<T>& da, const G& g) {
  static_assert(W == 4 || W == 8 || W == 16, "4, 8 or 16 slots per row");
  const T* __restrict__ base =
      da.v + (size_t)(row / kDiaChunk) * (W * kDiaChunk) + (row % kDiaChunk);
  T a[W];
#pragma unroll
  for (int d = 0; d < W; ++d) a[d] = base[d * kDiaChunk];
  T xv[W - 1];
#pragma unroll
  for (int d = 0; d < W - 1; ++d) xv[d] = g((uint32_t)(row + da.off[d]));
  using U = typename DiaMaskBits<T>::U;
  const U m = __builtin_bit_cast(U, a[W - 1]);
  T s = T(0);
#pragma unroll
  for (int d = 0; d < W - 1; ++d) {
    const T t = s + a[d] * xv[d];
    s = (m >> d & U(1)) ? t : s;
  }
  return s;
}

// Symmetric diagonal copy, for a matrix whose diagonals come in pairs +-o and
// whose values are symmetric bit for bit (a_ij == a_ji, checked when the copy
// is made): row i keeps K = W / 2 slots -- slot 0 the diagonal, slot k the
// value at column i + up[k] (up ascending; unused slots 0) -- in the same
// chunked, slot-major layout, and the lower value a_{i, i-up[k]} is read as
// a_{i-up[k], i}, slot k of row i - up[k]: across a wave one more contiguous
// run per diagonal, read at about the time the partner rows' own workgroup
// reads it.  A 16-bit mask per row holds the present entries in summation
// order: bit K-1-k the lower entry k, bit K-1 the diagonal, bit K-1+k the
// upper entry k.  The sum runs over the lower entries (k from K-1 down to
// 1), the diagonal and the upper entries in ascending column order from
// +0.0: the CSR row sum bit for bit, with half the value bytes.
template <typename T>
struct DiaSymRows {
  int n;
  const T* __restrict__ u;         // ceil(n / 64) * K * 64
  const uint16_t* __restrict__ m;  // n
  int up[8];                       // up[0] = 0 (the diagonal)
};

template <int K, typename T, class G>
__device__ __forceinline__ T dias_rowsum_g(int row, const DiaSymRows<T>& ds, const G& g) {
  static_assert(K == 2 || K == 4 || K == 8, "2, 4 or 8 slots per row");
  constexpr int cs = K * kDiaChunk;
  const T* __restrict__ base = ds.u + (size_t)(row / kDiaChunk) * cs + (row % kDiaChunk);
  T au[K], al[K], xu[K], xl[K];
#pragma unroll
  for (int k = 0; k < K; ++k) au[k] = base[k * kDiaChunk];
#pragma unroll
  for (int k = 1; k < K; ++k) {
    // the partner row; below row 0 the entry is absent (masked): any valid row
    const int pr = max(row - ds.up[k], 0);
    al[k] = ds.u[(size_t)(pr / kDiaChunk) * cs + k * kDiaChunk + (pr % kDiaChunk)];
    xl[k] = g((uint32_t)(row - ds.up[k]));
    xu[k] = g((uint32_t)(row + ds.up[k]));
  }
  xu[0] = g((uint32_t)row);
  const uint32_t m = ds.m[row];
  T s = T(0);
#pragma unroll
  for (int k = K - 1; k >= 1; --k) {
    const T t = s + al[k] * xl[k];
    s = (m >> (K - 1 - k) & 1u) ? t : s;
  }
  {
    const T t = s + au[0] * xu[0];
    s = (m >> (K - 1) & 1u) ? t : s;
  }
#pragma unroll
  for (int k = 1; k < K; ++k) {
    const T t = s + au[k] * xu[k];
    s = (m >> (K - 1 + k) & 1u) ? t : s;
  }
  return s;
}

// The row sum from the diagonal copy of W slots per row, or (SYM) from its
// symmetric form of W / 2 slots.
template <int W, bool SYM, typename T, class G>
__device__ __forceinline__ T dia_row(int row, const DiaRows<T>& da, const DiaSymRows<T>& ds,
                                     const G& g) {
  if constexpr (SYM)
    return dias_rowsum_g<W / 2>(row, ds, g);
  else
    return dia_rowsum_g<W>(row, da, g);
}

// The symmetric copies as the kernels take them (host side): FP64, FP32,
// FP32 of D^-1/2 A D^-1/2.
DiaSymRows<double> dias_f64(ofh_csr A);
DiaSymRows<float> dias_f32(ofh_csr A, bool scaled);

template <typename T, int TILE, bool SCALED = false>
__device__ __forceinline__ T csr_stream_rowsum(int r0, int r1, const int* __restrict__ rp,
                                               ColIndex ci, const T* __restrict__ val,
                                               const T* __restrict__ x, T* lds, T init,
                                               T scale = T(1)) {
  return csr_stream_rowsum_g<T, TILE, SCALED>(r0, r1, rp, ci, val, GatherVec<T>{x}, lds, init,
                                              scale);
}

// The FP64 matrix as a product kernel reads it: the CSR-stream partition
// (nblocks row blocks) and, when sell.cptr is set, its SELL-64 copy, or when
// dia.v is set its FP64 diagonal copy of dia_w slots per row (either: 256
// rows per workgroup, nblocks = ceil(n / 256)).
struct Sell64 {
  int n;
  const int* __restrict__ cptr;
  const int* __restrict__ clen;
  const uint32_t* __restrict__ col;
  const int16_t* __restrict__ col16;
  const double* __restrict__ val;
};
struct MatRows64 {
  int nblocks;
  const int* __restrict__ rb;
  const int* __restrict__ rp;
  ColIndex ci;
  const double* __restrict__ val;
  Sell64 sell;
  DiaF64 dia;
  int dia_w;
};

// (A x)_i for this lane's row i of the workgroup; `valid` is false for lanes
// past the last row.  Every lane of the workgroup must call it (the CSR-stream
// form synchronizes the workgroup).  Both forms give the CSR row sums bitwise.
template <bool SELL>
__device__ __forceinline__ double mat_rowsum64(const MatRows64& m, const double* __restrict__ x,
                                               int& i, bool& valid) {
  if constexpr (SELL) {
    i = xcd_block(blockIdx.x, m.nblocks) * kBlock + threadIdx.x;
    valid = i < m.sell.n;
    if (i / kSellChunk * kSellChunk >= m.sell.n) return 0.0;  // a whole wave past the end
    if (m.dia.v) {  // uniform: the diagonal copy
      if (!valid) return 0.0;
      const SellGather<double> g{x_rsrc(x, m.dia.n)};
      if (m.dia_w == 4) return dia_rowsum_g<4>(i, m.dia, g);
      if (m.dia_w == 8) return dia_rowsum_g<8>(i, m.dia, g);
      return dia_rowsum_g<16>(i, m.dia, g);
    }
    return sell_rowsum<double, 2>(i, m.sell.cptr, m.sell.clen, m.sell.col, m.sell.col16, m.sell.val,
                                  x_rsrc(x, m.sell.n));
  } else {
    __shared__ double lds[kTile64];
    const int blk = xcd_block(blockIdx.x, m.nblocks);
    const int r0 = m.rb[blk], r1 = m.rb[blk + 1];
    const double s = csr_stream_rowsum<double, kTile64>(r0, r1, m.rp, m.ci, m.val, x, lds, 0.0);
    i = r0 + threadIdx.x;
    valid = i < r1;
    return s;
  }
}

// The FP64 matrix for mat_rowsum64<*sell>: the FP64 diagonal copy, else the
// SELL-64 copy, when it is worth using and one partial per workgroup fits
// the reduction slots (*sell then selects the 256-rows-per-workgroup form).
inline int mat_rows64(ofh_csr A, MatRows64* m, bool* sell) {
  *sell = false;
  bool dia = false;
  const int nbs = (A->n + kBlock - 1) / kBlock;
  if (nbs <= kMaxPartials) {
    int rc = csr_ensure_dia64(A, &dia);
    if (!rc && !dia) rc = csr_ensure_sell(A, false, sell);
    if (rc) return rc;
  }
  *m = MatRows64{*sell || dia ? nbs : A->nblocks, A->row_blocks, A->row_ptr,
                 ColIndex{A->col, A->col16}, A->val,
                 Sell64{A->n, A->sell_cptr, A->sell_clen, A->sell_col, A->sell_col16,
                        A->sell_val},
                 dia ? dia_f64(A) : DiaF64{}, dia ? A->dia_w : 0};
  if (dia) *sell = true;
  return OFH_OK;
}

}  // namespace ofh
