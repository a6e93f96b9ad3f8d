// ofh_mpnn.hip -- learned initial guess for a sequence of solves A x = b.
//
// No reference counterpart: the reference starts every solve from zero
// (Equa::run, Vect_impl.h:547) or from the previous step's solution
// (TimeStepping.cpp:207).  Here a message-passing network on the graph of A
// predicts the start vector of the next solve from the last two solutions
// and the new right-hand side.  The definition (features, layers, loss) is
// written out in include/ofeli_hip.h next to the entry points.
//
// Device design:
//  * activations are FP32, row-major with the row count padded to a
//    multiple of 16 (pad rows are kept at exactly zero by every kernel);
//  * the six dense products per layer (forward Z = H Wab, U = [H,M] Wu;
//    backward g Wu^T, [H,M]^T g, H^T dZ, dZ Wab^T) run on the matrix cores:
//    v_mfma_f32_16x16x16_bf16 with BF16 operands and FP32 accumulation.  A
//    wavefront keeps the whole BF16 weight matrix in registers and streams
//    16-row tiles of activations;
//  * message aggregation, encoder and decoder are FP32 vector code, one
//    wavefront per graph node with one lane per hidden channel;
//  * every reduction over graph nodes (weight and bias gradients, the loss)
//    writes per-workgroup partials that a following kernel folds in a fixed
//    order: no atomics, results are bitwise reproducible;
//  * Adam updates the FP32 master weights and refreshes their BF16 copies.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ofh_internal.h"
#include "ofh_spmv.h"

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace ofh {
namespace mp {

constexpr int kH = 64;        // hidden channels
constexpr int kH2 = 128;
constexpr int kFin = 8;       // input features (4 used, 4 zero)
constexpr int kLayer = kH * kH2 + kH + kH + kH2 * kH + kH;  // parameters per layer
constexpr int kNodeGrid = 2048;  // workgroups of the node-reduction kernels (8 waves per SIMD)
constexpr int kGemmGrid = 512;   // workgroups of the weight-gradient products
constexpr int kMaxLayers = 8;

// parameter offsets in the flat FP32 vector
__host__ __device__ constexpr int off_enc_W() { return 0; }
__host__ __device__ constexpr int off_enc_b() { return kFin * kH; }
__host__ __device__ inline int off_layer(int l) { return kFin * kH + kH + l * kLayer; }
__host__ __device__ inline int off_Wab(int l) { return off_layer(l); }
__host__ __device__ inline int off_we(int l) { return off_layer(l) + kH * kH2; }
__host__ __device__ inline int off_bm(int l) { return off_layer(l) + kH * kH2 + kH; }
__host__ __device__ inline int off_Wu(int l) { return off_layer(l) + kH * kH2 + 2 * kH; }
__host__ __device__ inline int off_bu(int l) { return off_layer(l) + 2 * kH * kH2 + 2 * kH; }
__host__ __device__ inline int off_dec(int L) { return kFin * kH + kH + L * kLayer; }
inline int64_t n_params(int L) { return off_dec(L) + kH + 1; }

__device__ __forceinline__ short f2bf(float f) {
  // round to nearest even (the activations here are finite)
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

// ---------------------------------------------------------------------------
// features
// ---------------------------------------------------------------------------
// The base x_b = x_n + ext (x_n - x_p) (ext = 0: x_n exactly).
__device__ __forceinline__ double xbase(const double* __restrict__ xn,
                                        const double* __restrict__ xp, double ext, int i) {
  return xn[i] + ext * (xn[i] - xp[i]);
}

// r = b - A x_b, z = D^-1 r; partial sums of z^2 per workgroup.
// x_b gathered at column c (+0.0 past either end, where the entry is absent)
struct XbGather {
  XRsrc xn, xp;
  double ext;
  __device__ __forceinline__ double operator()(uint32_t c) const {
    const double a = x_load<double>(xn, c), p = x_load<double>(xp, c);
    return a + ext * (a - p);
  }
};

// The row sums of the next three kernels: with W > 0 from the FP64 diagonal
// copy of W slots per row (a row per lane, workgroups remapped across the
// XCDs, each writing the partial of its own 256 rows), else a thread per row
// over its CSR entries; both are the CSR row sum bit for bit.
template <int W>
__device__ __forceinline__ int mp_row_block() {
  return W ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x;
}

template <int W>
__global__ __launch_bounds__(256) void k_mp_resid(int n, const int* __restrict__ rp,
                                                  const int* __restrict__ col,
                                                  const double* __restrict__ val, DiaF64 dia,
                                                  const double* __restrict__ dinv,
                                                  const double* __restrict__ xn,
                                                  const double* __restrict__ xp, double ext,
                                                  const double* __restrict__ b, double* r,
                                                  double* z, double* part) {
  __shared__ double red[4];
  const int blk = mp_row_block<W>();
  const int i = blk * 256 + threadIdx.x;
  double zz = 0.0;
  if (i < n) {
    double s = 0.0;
    if constexpr (W > 0)
      s = dia_rowsum_g<W>(i, dia, XbGather{x_rsrc(xn, n), x_rsrc(xp, n), ext});
    else
      for (int k = rp[i]; k < rp[i + 1]; ++k) s += val[k] * xbase(xn, xp, ext, col[k]);
    const double ri = b[i] - s;
    const double zi = dinv[i] * ri;
    r[i] = ri;
    z[i] = zi;
    zz = zi * zi;
  }
  zz = block_sum<256>(zz, red);
  if (threadIdx.x == 0) part[blk] = zz;
}

// s = rms(z) (folded by every workgroup), z2 = D^-1 (r - A z), features.
template <int W>
__global__ __launch_bounds__(256) void k_mp_feat(int n, int n_pad, const int* __restrict__ rp,
                                                 const int* __restrict__ col,
                                                 const double* __restrict__ val, DiaF64 dia,
                                                 const double* __restrict__ dinv,
                                                 const double* __restrict__ r,
                                                 const double* __restrict__ z,
                                                 const double* __restrict__ xn,
                                                 const double* __restrict__ xp,
                                                 const float* __restrict__ ld,
                                                 const double* __restrict__ part, int nparts,
                                                 float* F, double* scal) {
  __shared__ double red[4];
  double t = 0.0;
  for (int p = threadIdx.x; p < nparts; p += 256) t += part[p];
  t = block_sum<256>(t, red);
  double s = sqrt(t / (double)n);
  if (!(s > 0.0)) s = 1.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) scal[0] = s;
  const int i = mp_row_block<W>() * 256 + threadIdx.x;
  if (i >= n_pad) return;
  float f[kFin] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (i < n) {
    double az = 0.0;
    if constexpr (W > 0)
      az = dia_rowsum_g<W>(i, dia, SellGather<double>{x_rsrc(z, n)});
    else
      for (int k = rp[i]; k < rp[i + 1]; ++k) az += val[k] * z[col[k]];
    const double z2 = dinv[i] * (r[i] - az);
    f[0] = (float)(z[i] / s);
    f[1] = (float)(z2 / s);
    f[2] = (float)((xn[i] - xp[i]) / s);
    f[3] = ld[i];
  }
  float4* o = (float4*)(F + (size_t)i * kFin);
  o[0] = make_float4(f[0], f[1], f[2], f[3]);
  o[1] = make_float4(f[4], f[5], f[6], f[7]);
}

// H0 = relu(F W_enc + b_enc); one thread per (row, channel).
__global__ __launch_bounds__(256) void k_mp_enc(int n, int n_pad, const float* __restrict__ F,
                                                const float* __restrict__ P, float* H0) {
  // a thread per row and 4 channels (16-byte stores)
  const int idx = blockIdx.x * 256 + threadIdx.x;
  const int row = idx / (kH / 4), c = 4 * (idx % (kH / 4));
  if (row >= n_pad) return;
  float h[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  if (row < n) {
#pragma unroll
    for (int q = 0; q < 4; ++q) h[q] = P[off_enc_b() + c + q];
#pragma unroll
    for (int f = 0; f < kFin; ++f) {
      const float ff = F[(size_t)row * kFin + f];
#pragma unroll
      for (int q = 0; q < 4; ++q) h[q] += ff * P[off_enc_W() + f * kH + c + q];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) h[q] = fmaxf(h[q], 0.0f);
  }
  *reinterpret_cast<float4*>(H0 + (size_t)row * kH + c) = make_float4(h[0], h[1], h[2], h[3]);
}

// ---------------------------------------------------------------------------
// dense products on the matrix cores
// ---------------------------------------------------------------------------
// Y = act(A W + bias) + R1 + R2 over 16-row tiles.  A[r][k] is X1[r][k] for
// k < K1 and X2[r][k-K1] beyond (a column concatenation), optionally masked
// by Mk[r][k] > 0.  W is BF16, [K][O] row-major, or [O][K] read transposed.
// Y2 (optional) receives act(A W + bias) alone.
struct RowGemmArgs {
  int n, n_pad;
  const float* X1;
  int ld1;
  const float* X2;
  int ld2;
  int K1;
  const float* Mk;
  int ldm;
  const uint16_t* W;
  const float* bias;
  const float* R1;
  int ldr1;
  const float* R2;
  int ldr2;
  float* Y;
  float* Y2;
};

// MFMA 16x16x16 BF16 operand layout (lane l, item t = 0..3):
//   A fragment: A[l%16][4*(l/16)+t];  B fragment: B[4*(l/16)+t][l%16];
//   accumulator: D[4*(l/16)+t][l%16].
template <int K, int O, bool TRANS, bool RELU, bool MASK>
__global__ __launch_bounds__(256, 3) void k_mp_rowgemm(RowGemmArgs a) {
  constexpr int KB = K / 16, JB = O / 16;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lg = lane >> 4;
  bf16x4 bw[KB][JB];
#pragma unroll
  for (int kb = 0; kb < KB; ++kb)
#pragma unroll
    for (int jb = 0; jb < JB; ++jb)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int k = kb * 16 + 4 * lg + t, j = jb * 16 + li;
        bw[kb][jb][t] = (short)(TRANS ? a.W[j * K + k] : a.W[k * O + j]);
      }
  const int ntiles = a.n_pad / 16;
  for (int tile = blockIdx.x * 4 + wave; tile < ntiles; tile += gridDim.x * 4) {
    const int r = tile * 16 + li;
    f32x4 acc[JB];
#pragma unroll
    for (int jb = 0; jb < JB; ++jb) acc[jb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
      const int k0 = kb * 16 + 4 * lg;
      float4 v = k0 < a.K1 ? *(const float4*)(a.X1 + (size_t)r * a.ld1 + k0)
                           : *(const float4*)(a.X2 + (size_t)r * a.ld2 + (k0 - a.K1));
      if (MASK) {
        const float4 m = *(const float4*)(a.Mk + (size_t)r * a.ldm + k0);
        v.x = m.x > 0.f ? v.x : 0.f;
        v.y = m.y > 0.f ? v.y : 0.f;
        v.z = m.z > 0.f ? v.z : 0.f;
        v.w = m.w > 0.f ? v.w : 0.f;
      }
      bf16x4 av;
      av[0] = f2bf(v.x);
      av[1] = f2bf(v.y);
      av[2] = f2bf(v.z);
      av[3] = f2bf(v.w);
#pragma unroll
      for (int jb = 0; jb < JB; ++jb)
        acc[jb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bw[kb][jb], acc[jb], 0, 0, 0);
    }
#pragma unroll
    for (int jb = 0; jb < JB; ++jb)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int row = tile * 16 + 4 * lg + t, c = jb * 16 + li;
        float v = acc[jb][t];
        if (a.bias) v += a.bias[c];
        if (RELU) v = fmaxf(v, 0.0f);
        float y = v;
        if (a.R1) y += a.R1[(size_t)row * a.ldr1 + c];
        if (a.R2) y += a.R2[(size_t)row * a.ldr2 + c];
        if (row >= a.n) v = y = 0.0f;
        a.Y[(size_t)row * O + c] = y;
        if (a.Y2) a.Y2[(size_t)row * O + c] = v;
      }
  }
}

// Weight gradient: part[wg] = sum over this workgroup's rows of A^T G (K x O)
// and, with COLSUM, the column sums of G (bias gradient).  A as in the row
// product (X1 | X2); G optionally masked by Mg > 0.  Rows are split into
// contiguous tile ranges per wavefront; the four wavefronts of a workgroup
// add into LDS in wave order.
struct WgradArgs {
  int n_pad;
  const float* X1;
  int ld1;
  const float* X2;
  int ld2;
  int K1;
  const float* G;
  int ldg;
  const float* Mg;
  int ldm;
  float* part;
};

template <int K, int O, bool MASK, bool COLSUM>
__global__ __launch_bounds__(256) void k_mp_wgrad(WgradArgs a) {
  constexpr int MB = K / 16, JB = O / 16;
  constexpr int W = K * O + (COLSUM ? O : 0);
  __shared__ float red[K * O + O];
  OFH_LDS_BUDGET(sizeof(float) * (K * O + O));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lg = lane >> 4;
  f32x4 acc[MB][JB];
#pragma unroll
  for (int mb = 0; mb < MB; ++mb)
#pragma unroll
    for (int jb = 0; jb < JB; ++jb) acc[mb][jb] = f32x4{0.f, 0.f, 0.f, 0.f};
  float cs[JB];
#pragma unroll
  for (int jb = 0; jb < JB; ++jb) cs[jb] = 0.f;
  const int ntiles = a.n_pad / 16;
  const int nw = gridDim.x * 4, w = blockIdx.x * 4 + wave;
  const int t0 = (int)((int64_t)ntiles * w / nw), t1 = (int)((int64_t)ntiles * (w + 1) / nw);
#pragma unroll 2
  for (int tile = t0; tile < t1; ++tile) {
    bf16x4 av[MB], bv[JB];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const size_t row = (size_t)tile * 16 + 4 * lg + t;
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) {
        const int k = mb * 16 + li;
        const float x = k < a.K1 ? a.X1[row * a.ld1 + k] : a.X2[row * a.ld2 + (k - a.K1)];
        av[mb][t] = f2bf(x);
      }
#pragma unroll
      for (int jb = 0; jb < JB; ++jb) {
        const int c = jb * 16 + li;
        float g = a.G[row * a.ldg + c];
        if (MASK) g = a.Mg[row * a.ldm + c] > 0.f ? g : 0.f;
        bv[jb][t] = f2bf(g);
        if (COLSUM) cs[jb] += g;
      }
    }
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
      for (int jb = 0; jb < JB; ++jb)
        acc[mb][jb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[mb], bv[jb], acc[mb][jb], 0, 0, 0);
  }
  if (COLSUM) {
#pragma unroll
    for (int jb = 0; jb < JB; ++jb) {
      cs[jb] += __shfl_xor(cs[jb], 16, 64);
      cs[jb] += __shfl_xor(cs[jb], 32, 64);
    }
  }
  for (int wv = 0; wv < 4; ++wv) {
    if (wave == wv) {
#pragma unroll
      for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int jb = 0; jb < JB; ++jb)
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const int idx = (mb * 16 + 4 * lg + t) * O + jb * 16 + li;
            red[idx] = (wv == 0 ? 0.f : red[idx]) + acc[mb][jb][t];
          }
      if (COLSUM && lg == 0) {
#pragma unroll
        for (int jb = 0; jb < JB; ++jb) {
          const int idx = K * O + jb * 16 + li;
          red[idx] = (wv == 0 ? 0.f : red[idx]) + cs[jb];
        }
      }
    }
    __syncthreads();
  }
  float* out = a.part + (size_t)blockIdx.x * W;
  for (int i = threadIdx.x; i < W; i += 256) out[i] = red[i];
}

// out[i] = sum_p part[p][i], i < width: a workgroup per 64 columns, four
// groups of 64 threads summing every fourth partial, then the four group
// sums added in group order (fixed order: bitwise reproducible).
__global__ __launch_bounds__(256) void k_mp_fold(const float* __restrict__ part, int nparts,
                                                 int width, float* __restrict__ out) {
  __shared__ float red[4][64];
  const int c = blockIdx.x * 64 + (threadIdx.x & 63), g = threadIdx.x >> 6;
  float s = 0.f;
  if (c < width) {
#pragma unroll 8
    for (int p = g; p < nparts; p += 4) s += part[(size_t)p * width + c];
  }
  red[g][threadIdx.x & 63] = s;
  __syncthreads();
  if (g == 0 && c < width)
    out[c] = ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
}

// ---------------------------------------------------------------------------
// message aggregation
// ---------------------------------------------------------------------------
// M_i = sum_{j != i} relu(Z_i[c] + Z_j[64+c] + e_ij we[c] + bm[c]).
__global__ __launch_bounds__(256) void k_mp_edge_fwd(int n, int n_pad, const int* __restrict__ rp,
                                                     const int* __restrict__ col,
                                                     const float* __restrict__ e,
                                                     const float* __restrict__ Z,
                                                     const float* __restrict__ we,
                                                     const float* __restrict__ bm, float* M) {
  // 16 lanes per node, 4 channels per lane (16-byte loads and stores); the
  // row's columns and edge weights are loaded lane-parallel, 16 at a time,
  // and broadcast within the node's lanes, and the gathers of a group of 8
  // neighbour rows are issued before any is summed.  Per channel the sum
  // runs over the row's entries in order, as one lane per channel would.
  constexpr int G = 16;
  const int l = threadIdx.x & (G - 1);
  const int i = blockIdx.x * (256 / G) + (threadIdx.x / G);
  if (i >= n_pad) return;
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  if (i < n) {
    const float4 za = *reinterpret_cast<const float4*>(Z + (size_t)i * kH2 + 4 * l);
    const float4 w = *reinterpret_cast<const float4*>(we + 4 * l);
    const float4 bb = *reinterpret_cast<const float4*>(bm + 4 * l);
    for (int k0 = rp[i], k1 = rp[i + 1]; k0 < k1; k0 += G) {
      const int m = min(G, k1 - k0);
      const int jl = l < m ? col[k0 + l] : i;
      const float el = l < m ? e[k0 + l] : 0.f;
      for (int t0 = 0; t0 < m; t0 += 8) {
        float4 zv[8];
        int jt[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          jt[t] = __shfl(jl, t0 + t, G);  // lanes past m hold i: an in-range row
          zv[t] = *reinterpret_cast<const float4*>(Z + (size_t)jt[t] * kH2 + kH + 4 * l);
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          const float et = __shfl(el, t0 + t, G);
          if (t0 + t < m && jt[t] != i) {
            acc.x += fmaxf(za.x + zv[t].x + et * w.x + bb.x, 0.f);
            acc.y += fmaxf(za.y + zv[t].y + et * w.y + bb.y, 0.f);
            acc.z += fmaxf(za.z + zv[t].z + et * w.z + bb.z, 0.f);
            acc.w += fmaxf(za.w + zv[t].w + et * w.w + bb.w, 0.f);
          }
        }
      }
    }
  }
  *reinterpret_cast<float4*>(M + (size_t)i * kH + 4 * l) = acc;
}

// Backward of the aggregation.  dM = T[:, 64:].  With s_ij = dM_i [pre_ij > 0]:
// dZ_v[c] = sum_u s_vu,  dZ_v[64+c] = sum_u s_uv (pre_uv recomputed from row v,
// e_uv = e_vu by construction), and per-workgroup partials of
// sum s_ij e_ij (we) and sum s_ij (bm).
__global__ __launch_bounds__(256) void k_mp_edge_bwd(int n, int n_pad, const int* __restrict__ rp,
                                                     const int* __restrict__ col,
                                                     const float* __restrict__ e,
                                                     const float* __restrict__ Z,
                                                     const float* __restrict__ we,
                                                     const float* __restrict__ bm,
                                                     const float* __restrict__ T, float* dZ,
                                                     float* part) {
  __shared__ float red[2 * kH];
  const int c = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const float w = we[c], bb = bm[c];
  float gwe = 0.f, gbm = 0.f;
  for (int v = blockIdx.x * 4 + wave; v < n_pad; v += gridDim.x * 4) {
    float dza = 0.f, dzc = 0.f;
    if (v < n) {
      const float za = Z[(size_t)v * kH2 + c], zc = Z[(size_t)v * kH2 + kH + c];
      const float dm = T[(size_t)v * kH2 + kH + c];
      for (int k0 = rp[v], k1 = rp[v + 1]; k0 < k1; k0 += 64) {
        const int m = min(64, k1 - k0);
        const int ul = c < m ? col[k0 + c] : v;
        const float el = c < m ? e[k0 + c] : 0.f;
#pragma unroll 4
        for (int t = 0; t < m; ++t) {
          const int u = __shfl(ul, t, 64);
          const float ek = __shfl(el, t, 64);
          if (u == v) continue;
          const float zuc = Z[(size_t)u * kH2 + kH + c], zua = Z[(size_t)u * kH2 + c];
          const float tu = T[(size_t)u * kH2 + kH + c];
          if (za + zuc + ek * w + bb > 0.f) {
            dza += dm;
            gwe += dm * ek;
            gbm += dm;
          }
          if (zua + zc + ek * w + bb > 0.f) dzc += tu;
        }
      }
    }
    dZ[(size_t)v * kH2 + c] = dza;
    dZ[(size_t)v * kH2 + kH + c] = dzc;
  }
  for (int wv = 0; wv < 4; ++wv) {
    if (wave == wv) {
      red[c] = (wv == 0 ? 0.f : red[c]) + gwe;
      red[kH + c] = (wv == 0 ? 0.f : red[kH + c]) + gbm;
    }
    __syncthreads();
  }
  if (threadIdx.x < 2 * kH) part[(size_t)blockIdx.x * 2 * kH + threadIdx.x] = red[threadIdx.x];
}

// ---------------------------------------------------------------------------
// decoder, loss, encoder backward
// ---------------------------------------------------------------------------
// d = s (H w_dec + b_dec), the learned direction.
__global__ __launch_bounds__(256) void k_mp_dec(int n, const float* __restrict__ H,
                                                const float* __restrict__ P, int off,
                                                const double* __restrict__ scal, double* dv) {
  // 16 lanes per row, 4 channels per lane; the sum is the butterfly over the
  // 64 channels (partners 32, 16, 8, 4 across lanes, then 2 and 1 within a
  // lane), the order one lane per channel with xor shuffles adds in
  const int l = threadIdx.x & 15;
  const int i = blockIdx.x * 16 + (threadIdx.x >> 4);
  if (i >= n) return;
  const float4 h = *reinterpret_cast<const float4*>(H + (size_t)i * kH + 4 * l);
  float v[4] = {h.x * P[off + 4 * l], h.y * P[off + 4 * l + 1], h.z * P[off + 4 * l + 2],
                h.w * P[off + 4 * l + 3]};
#pragma unroll
  for (int o = 8; o > 0; o >>= 1)
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] += __shfl_xor(v[q], o, 16);
  const float d = ((v[0] + v[2]) + (v[1] + v[3])) + P[off + kH];
  if (l == 0) dv[i] = scal[0] * (double)d;
}

// q = A d; line-search partials.  A-norm objective (obj 0): d.q -> slot 0,
// d.r -> slot 1; minimum residual (obj 1): q.q -> 0, q.r -> 1, r.r -> 2.
template <int W>
__global__ __launch_bounds__(256) void k_mp_ls(int n, const int* __restrict__ rp,
                                               const int* __restrict__ col,
                                               const double* __restrict__ val, DiaF64 dia,
                                               const double* __restrict__ dv,
                                               const double* __restrict__ r, double* q,
                                               double* part, int nparts, int obj) {
  __shared__ double red[12];
  const int blk = mp_row_block<W>();
  const int i = blk * 256 + threadIdx.x;
  double v[3] = {0.0, 0.0, 0.0};
  if (i < n) {
    double a = 0.0;
    if constexpr (W > 0)
      a = dia_rowsum_g<W>(i, dia, SellGather<double>{x_rsrc(dv, n)});
    else
      for (int k = rp[i]; k < rp[i + 1]; ++k) a += val[k] * dv[col[k]];
    q[i] = a;
    const double w = obj ? a : dv[i];
    v[0] = w * a;
    v[1] = w * r[i];
    v[2] = r[i] * r[i];
  }
  block_sum_k<256, 3>(v, red);
  if (threadIdx.x == 0) {
    part[blk] = v[0];
    part[nparts + blk] = v[1];
    if (obj) part[2 * nparts + blk] = v[2];
  }
}

// alpha = slot 1 / slot 0 (0 if slot 0 <= 0): d.r / d.Ad (A-norm) or
// Ad.r / Ad.Ad (minimum residual); x0 = x_b + alpha d.
__global__ __launch_bounds__(256) void k_mp_apply(int n, const double* __restrict__ part,
                                                  int nparts, const double* __restrict__ xn,
                                                  const double* __restrict__ xp, double ext,
                                                  const double* __restrict__ dv, double* scal,
                                                  double* x0) {
  __shared__ double red[8];
  double v[2] = {0.0, 0.0};
  for (int p = threadIdx.x; p < nparts; p += 256) {
    v[0] += part[p];
    v[1] += part[nparts + p];
  }
  block_sum_k<256, 2>(v, red);
  const double dAd = v[0], dr = v[1];
  const double alpha = dAd > 0.0 ? dr / dAd : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    scal[4] = alpha;
    scal[5] = dAd;
  }
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) x0[i] = xbase(xn, xp, ext, i) + alpha * dv[i];
}

// Training: g = A (x* - x_b); partials (x* - x_b).g -> slot 2 (E0), d.g -> slot 3.
__global__ __launch_bounds__(256) void k_mp_loss(int n, const int* __restrict__ rp,
                                                 const int* __restrict__ col,
                                                 const double* __restrict__ val,
                                                 const double* __restrict__ xn,
                                                 const double* __restrict__ xp, double ext,
                                                 const double* __restrict__ xs,
                                                 const double* __restrict__ dv, double* gv,
                                                 double* part, int nparts) {
  __shared__ double red[8];
  const int i = blockIdx.x * 256 + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (i < n) {
    double a = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      const int j = col[k];
      a += val[k] * (xs[j] - xbase(xn, xp, ext, j));
    }
    gv[i] = a;
    v[0] = (xs[i] - xbase(xn, xp, ext, i)) * a;
    v[1] = dv[i] * a;
  }
  block_sum_k<256, 2>(v, red);
  if (threadIdx.x == 0) {
    part[2 * nparts + blockIdx.x] = v[0];
    part[3 * nparts + blockIdx.x] = v[1];
  }
}

// Training, minimum-residual objective: folds q.q, q.r, r.r of k_mp_ls
// (every workgroup); L = 1 - (q.r)^2 / (q.q r.r) and dL/dd = A w with
// w = -2 (q.r) / (q.q r.r) r + 2 (q.r)^2 / ((q.q)^2 r.r) q, formed at the
// gather (A symmetric), into gv.
__global__ __launch_bounds__(256) void k_mp_loss_res(int n, const int* __restrict__ rp,
                                                     const int* __restrict__ col,
                                                     const double* __restrict__ val,
                                                     const double* __restrict__ r,
                                                     const double* __restrict__ q,
                                                     const double* __restrict__ part, int nparts,
                                                     double* gv, double* scal) {
  __shared__ double red[12];
  double v[3] = {0.0, 0.0, 0.0};
  for (int p = threadIdx.x; p < nparts; p += 256) {
    v[0] += part[p];
    v[1] += part[nparts + p];
    v[2] += part[2 * nparts + p];
  }
  block_sum_k<256, 3>(v, red);
  const double qq = v[0], c = v[1], R = v[2];
  double L, a1 = 0.0, a2 = 0.0;
  if (!(R > 0.0)) {
    L = 0.0;  // the base solves the system
  } else if (!(qq > 0.0)) {
    L = 1.0;  // no direction
  } else {
    L = 1.0 - c * c / (qq * R);
    a1 = -2.0 * c / (qq * R);
    a2 = 2.0 * c * c / (qq * qq * R);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    scal[1] = R;
    scal[2] = c;
    scal[3] = L;
  }
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) {
    double a = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      const int j = col[k];
      a += val[k] * (a1 * r[j] + a2 * q[j]);
    }
    gv[i] = a;
  }
}

// Folds d.Ad, E0, d.g (every workgroup); L = 1 - (d.g)^2 / (d.Ad E0) and
// dL/dd = -2 (d.g) / (d.Ad E0) (g - (d.g / d.Ad) A d); decoder backward:
// dH = (s dL/dd) w_dec^T and partials of the decoder gradient.
__global__ __launch_bounds__(256) void k_mp_dec_bwd(int n, int n_pad, const float* __restrict__ H,
                                                    const float* __restrict__ P, int off,
                                                    const double* __restrict__ q,
                                                    const double* __restrict__ gv,
                                                    const double* __restrict__ lpart, int nparts,
                                                    double* scal, float* dH, float* part,
                                                    int obj) {
  __shared__ double dred[12];
  __shared__ float red[kH + 1];
  double c1 = 1.0, c2 = 0.0;  // obj 1: gv is dL/dd already (k_mp_loss_res)
  if (!obj) {
    double v[3] = {0.0, 0.0, 0.0};
    for (int p = threadIdx.x; p < nparts; p += 256) {
      v[0] += lpart[p];
      v[1] += lpart[2 * nparts + p];
      v[2] += lpart[3 * nparts + p];
    }
    block_sum_k<256, 3>(v, dred);
    const double dAd = v[0], E0 = v[1], dg = v[2];
    double L;
    c1 = 0.0;
    if (!(E0 > 0.0)) {
      L = 0.0;  // the base is already the solution
    } else if (!(dAd > 0.0)) {
      L = 1.0;  // no direction
    } else {
      L = 1.0 - dg * dg / (dAd * E0);
      c1 = -2.0 * dg / (dAd * E0);
      c2 = dg / dAd;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      scal[1] = E0;
      scal[2] = dg;
      scal[3] = L;
    }
  }
  const double sc = scal[0];
  const int c = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const float w = P[off + c];
  float gw = 0.f, gb = 0.f;
  for (int i = blockIdx.x * 4 + wave; i < n_pad; i += gridDim.x * 4) {
    float dd = 0.f;
    if (i < n) dd = (float)(sc * c1 * (gv[i] - c2 * q[i]));
    dH[(size_t)i * kH + c] = dd * w;
    gw += H[(size_t)i * kH + c] * dd;
    gb += dd;
  }
  for (int wv = 0; wv < 4; ++wv) {
    if (wave == wv) {
      red[c] = (wv == 0 ? 0.f : red[c]) + gw;
      if (c == 0) red[kH] = (wv == 0 ? 0.f : red[kH]) + gb;
    }
    __syncthreads();
  }
  if (threadIdx.x <= kH) part[(size_t)blockIdx.x * (kH + 1) + threadIdx.x] = red[threadIdx.x];
}

// Encoder backward: g0 = dH [H0 > 0]; partials of F^T g0 (8 x 64) and sum g0.
__global__ __launch_bounds__(256) void k_mp_enc_bwd(int n_pad, const float* __restrict__ F,
                                                    const float* __restrict__ H0,
                                                    const float* __restrict__ dH, float* part) {
  constexpr int W = kFin * kH + kH;
  __shared__ float red[W];
  OFH_LDS_BUDGET(sizeof(float) * W);
  const int c = threadIdx.x & 63, wave = threadIdx.x >> 6;
  float gw[kFin] = {0, 0, 0, 0, 0, 0, 0, 0};
  float gb = 0.f;
  for (int i = blockIdx.x * 4 + wave; i < n_pad; i += gridDim.x * 4) {
    const float g = H0[(size_t)i * kH + c] > 0.f ? dH[(size_t)i * kH + c] : 0.f;
#pragma unroll
    for (int f = 0; f < kFin; ++f) gw[f] += F[(size_t)i * kFin + f] * g;
    gb += g;
  }
  for (int wv = 0; wv < 4; ++wv) {
    if (wave == wv) {
#pragma unroll
      for (int f = 0; f < kFin; ++f) red[f * kH + c] = (wv == 0 ? 0.f : red[f * kH + c]) + gw[f];
      red[kFin * kH + c] = (wv == 0 ? 0.f : red[kFin * kH + c]) + gb;
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < W; i += 256) part[(size_t)blockIdx.x * W + i] = red[i];
}

// ---------------------------------------------------------------------------
// optimiser
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mp_adam(int64_t np, float* __restrict__ p,
                                                 const float* __restrict__ g, float* m, float* v,
                                                 float lr, float b1, float b2, float eps,
                                                 float bc1, float bc2) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= np) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// BF16 copies of every layer's Wab (64 x 128) and Wu (128 x 64).
__global__ __launch_bounds__(256) void k_mp_to_bf16(int L, const float* __restrict__ p,
                                                    uint16_t* wbf) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int per = 2 * kH * kH2;
  if (i >= L * per) return;
  const int l = i / per, k = i % per;
  const int src = k < kH * kH2 ? off_Wab(l) + k : off_Wu(l) + (k - kH * kH2);
  wbf[i] = (uint16_t)f2bf(p[src]);
}

}  // namespace mp
}  // namespace ofh

// ---------------------------------------------------------------------------
// object
// ---------------------------------------------------------------------------
struct ofh_mpnn_s {
  ofh_csr A = nullptr;
  int L = 0;
  double ext = 0.0;  // base x_b = x_n + ext (x_n - x_p)
  int n = 0, n_pad = 0;
  int64_t np = 0;
  float* params = nullptr;
  float* grads = nullptr;
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  int adam_t = 0;
  uint16_t* wbf = nullptr;  // per layer: Wab then Wu, BF16
  float* e = nullptr;       // nnz edge weights a_ij / sqrt(a_ii a_jj)
  int* tpos = nullptr;      // nnz: position of a_ji
  float* ld = nullptr;      // n: ln a_ii - mean
  float* F = nullptr;       // n_pad x 8
  float* Hs = nullptr;      // (L+1) x n_pad x 64
  float* Zs = nullptr;      // L x n_pad x 128
  float* Ms = nullptr;      // L x n_pad x 64
  float* Us = nullptr;      // L x n_pad x 64
  float* dH0 = nullptr;     // n_pad x 64
  float* dH1 = nullptr;     // n_pad x 64
  float* T = nullptr;       // n_pad x 128
  float* dZ = nullptr;      // n_pad x 128
  double* r = nullptr;      // n
  double* z = nullptr;      // n
  double* q = nullptr;      // n: A d
  double* dv = nullptr;     // n: the learned direction d
  double* scal = nullptr;   // [s, E0 | r.r, d.g | q.r, loss, alpha, slot 0 of the line search]
  int obj = 0;              // objective: 0 A-norm error, 1 residual norm
  double* dpart = nullptr;  // 4 x nparts (row reductions)
  float* part = nullptr;    // per-workgroup partials of the node reductions
  double* h_pinned = nullptr;
  int nparts = 0;
  hipStream_t cap = nullptr;  // capture-only stream for the inference graph
  hipGraphExec_t gexec = nullptr;
  const void* gkey[4] = {nullptr, nullptr, nullptr, nullptr};
  bool use_graph = true;
};

namespace ofh {
namespace mp {

static void free_all(ofh_mpnn M) {
  void* ptrs[] = {M->params, M->grads, M->adam_m, M->adam_v, M->wbf, M->e, M->tpos, M->ld, M->F,
                  M->Hs, M->Zs, M->Ms, M->Us, M->dH0, M->dH1, M->T, M->dZ, M->r, M->z, M->q,
                  M->dv, M->scal, M->dpart, M->part};
  for (void* p : ptrs)
    if (p) hipFree(p);
  if (M->h_pinned) hipHostFree(M->h_pinned);
  if (M->gexec) hipGraphExecDestroy(M->gexec);
  if (M->cap) hipStreamDestroy(M->cap);
}

static inline float* Hl(ofh_mpnn M, int l) { return M->Hs + (size_t)l * M->n_pad * kH; }
static inline float* Zl(ofh_mpnn M, int l) { return M->Zs + (size_t)l * M->n_pad * kH2; }
static inline float* Ml(ofh_mpnn M, int l) { return M->Ms + (size_t)l * M->n_pad * kH; }
static inline float* Ul(ofh_mpnn M, int l) { return M->Us + (size_t)l * M->n_pad * kH; }
static inline const uint16_t* Wab_bf(ofh_mpnn M, int l) { return M->wbf + (size_t)l * 2 * kH * kH2; }
static inline const uint16_t* Wu_bf(ofh_mpnn M, int l) { return M->wbf + (size_t)l * 2 * kH * kH2 + kH * kH2; }

static int gemm_grid(int n_pad) { return std::max(1, std::min(2048, (n_pad / 16 + 3) / 4)); }

// Forward pass; training keeps U for the backward pass.
static int forward(ofh_mpnn M, const double* xn, const double* xp, const double* b, double* x0,
                   bool training, hipStream_t s) {
  ofh_csr A = M->A;
  const int n = M->n, n_pad = M->n_pad;
  const int g1 = (n + 255) / 256, gp = (n_pad + 255) / 256;
  // the FP64 diagonal copy when the entry point made it current (never here:
  // this may run under stream capture)
  const bool dia = A->dia_use64 && A->dia_val64 && A->dia_version64 == A->version;
  const DiaF64 da = dia ? dia_f64(A) : DiaF64{};
  const int dw = dia ? A->dia_w : 0;
#define OFH_MP_ROWS(KERNEL, GRID, ...)                                                        \
  do {                                                                                        \
    if (dw == 4)                                                                              \
      hipLaunchKernelGGL(KERNEL<4>, dim3(GRID), dim3(256), 0, s, __VA_ARGS__);                \
    else if (dw == 8)                                                                         \
      hipLaunchKernelGGL(KERNEL<8>, dim3(GRID), dim3(256), 0, s, __VA_ARGS__);                \
    else if (dw == 16)                                                                        \
      hipLaunchKernelGGL(KERNEL<16>, dim3(GRID), dim3(256), 0, s, __VA_ARGS__);               \
    else                                                                                      \
      hipLaunchKernelGGL(KERNEL<0>, dim3(GRID), dim3(256), 0, s, __VA_ARGS__);                \
  } while (0)
  OFH_MP_ROWS(k_mp_resid, g1, n, A->row_ptr, A->col, A->val, da, A->dinv, xn, xp, M->ext, b,
              M->r, M->z, M->dpart);
  OFH_MP_ROWS(k_mp_feat, gp, n, n_pad, A->row_ptr, A->col, A->val, da, A->dinv, M->r, M->z, xn,
              xp, M->ld, M->dpart, g1, M->F, M->scal);
  hipLaunchKernelGGL(k_mp_enc, dim3((n_pad * (kH / 4) + 255) / 256), dim3(256), 0, s, n, n_pad, M->F,
                     M->params, Hl(M, 0));
  const int gg = gemm_grid(n_pad);
  for (int l = 0; l < M->L; ++l) {
    RowGemmArgs a{};
    a.n = n;
    a.n_pad = n_pad;
    a.X1 = Hl(M, l);
    a.ld1 = kH;
    a.K1 = kH;
    a.W = Wab_bf(M, l);
    a.Y = Zl(M, l);
    hipLaunchKernelGGL((k_mp_rowgemm<kH, kH2, false, false, false>), dim3(gg), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_mp_edge_fwd, dim3((n_pad + 15) / 16), dim3(256), 0, s, n, n_pad,
                       A->row_ptr, A->col, M->e, Zl(M, l), M->params + off_we(l),
                       M->params + off_bm(l), Ml(M, l));
    RowGemmArgs u{};
    u.n = n;
    u.n_pad = n_pad;
    u.X1 = Hl(M, l);
    u.ld1 = kH;
    u.X2 = Ml(M, l);
    u.ld2 = kH;
    u.K1 = kH;
    u.W = Wu_bf(M, l);
    u.bias = M->params + off_bu(l);
    u.R1 = Hl(M, l);
    u.ldr1 = kH;
    u.Y = Hl(M, l + 1);
    u.Y2 = training ? Ul(M, l) : nullptr;
    hipLaunchKernelGGL((k_mp_rowgemm<kH2, kH, false, true, false>), dim3(gg), dim3(256), 0, s, u);
  }
  hipLaunchKernelGGL(k_mp_dec, dim3((n + 15) / 16), dim3(256), 0, s, n, Hl(M, M->L), M->params,
                     off_dec(M->L), M->scal, M->dv);
  OFH_MP_ROWS(k_mp_ls, g1, n, A->row_ptr, A->col, A->val, da, M->dv, M->r, M->q, M->dpart, g1,
              M->obj);
#undef OFH_MP_ROWS
  hipLaunchKernelGGL(k_mp_apply, dim3(g1), dim3(256), 0, s, n, M->dpart, g1, xn, xp, M->ext,
                     M->dv, M->scal, x0);
  OFH_HIP(hipGetLastError());
  return OFH_OK;
}

static int fold(const float* part, int nparts, int width, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_mp_fold, dim3((width + 63) / 64), dim3(256), 0, s, part, nparts, width, out);
  return OFH_OK;
}

// Loss and gradient into M->grads (forward must have run with training = true).
static int backward(ofh_mpnn M, const double* xn, const double* xp, const double* xs,
                    hipStream_t s) {
  ofh_csr A = M->A;
  const int n = M->n, n_pad = M->n_pad, L = M->L;
  const int g1 = (n + 255) / 256;
  const int gg = gemm_grid(n_pad);
  double* gv;
  if (M->obj == 1) {
    // dL/dd = A w into z (free after the forward pass); r is still b - A x_b
    gv = M->z;
    hipLaunchKernelGGL(k_mp_loss_res, dim3(g1), dim3(256), 0, s, n, A->row_ptr, A->col, A->val,
                       M->r, M->q, M->dpart, g1, gv, M->scal);
  } else {
    // g = A (x* - x_b) goes to r (free after the forward pass)
    gv = M->r;
    hipLaunchKernelGGL(k_mp_loss, dim3(g1), dim3(256), 0, s, n, A->row_ptr, A->col, A->val, xn,
                       xp, M->ext, xs, M->dv, gv, M->dpart, g1);
  }
  float* dH = M->dH0;
  float* dHn = M->dH1;
  hipLaunchKernelGGL(k_mp_dec_bwd, dim3(kNodeGrid), dim3(256), 0, s, n, n_pad, Hl(M, L), M->params,
                     off_dec(L), M->q, gv, M->dpart, g1, M->scal, dH, M->part, M->obj);
  fold(M->part, kNodeGrid, kH + 1, M->grads + off_dec(L), s);
  for (int l = L - 1; l >= 0; --l) {
    // T = (dH [U > 0]) Wu^T  (n x 128)
    RowGemmArgs t{};
    t.n = n;
    t.n_pad = n_pad;
    t.X1 = dH;
    t.ld1 = kH;
    t.K1 = kH;
    t.Mk = Ul(M, l);
    t.ldm = kH;
    t.W = Wu_bf(M, l);
    t.Y = M->T;
    hipLaunchKernelGGL((k_mp_rowgemm<kH, kH2, true, false, true>), dim3(gg), dim3(256), 0, s, t);
    // dWu = [H, M]^T (dH [U > 0]), dbu = column sums
    WgradArgs wu{};
    wu.n_pad = n_pad;
    wu.X1 = Hl(M, l);
    wu.ld1 = kH;
    wu.X2 = Ml(M, l);
    wu.ld2 = kH;
    wu.K1 = kH;
    wu.G = dH;
    wu.ldg = kH;
    wu.Mg = Ul(M, l);
    wu.ldm = kH;
    wu.part = M->part;
    hipLaunchKernelGGL((k_mp_wgrad<kH2, kH, true, true>), dim3(kGemmGrid), dim3(256), 0, s, wu);
    // Wu (128 x 64) and bu (64) are adjacent in the parameter vector
    fold(M->part, kGemmGrid, kH2 * kH + kH, M->grads + off_Wu(l), s);
    // aggregation backward -> dZ, we/bm gradients (adjacent)
    hipLaunchKernelGGL(k_mp_edge_bwd, dim3(kNodeGrid), dim3(256), 0, s, n, n_pad, A->row_ptr,
                       A->col, M->e, Zl(M, l), M->params + off_we(l), M->params + off_bm(l), M->T,
                       M->dZ, M->part);
    fold(M->part, kNodeGrid, 2 * kH, M->grads + off_we(l), s);
    // dWab = H^T dZ
    WgradArgs wa{};
    wa.n_pad = n_pad;
    wa.X1 = Hl(M, l);
    wa.ld1 = kH;
    wa.K1 = kH;
    wa.G = M->dZ;
    wa.ldg = kH2;
    wa.part = M->part;
    hipLaunchKernelGGL((k_mp_wgrad<kH, kH2, false, false>), dim3(kGemmGrid), dim3(256), 0, s, wa);
    fold(M->part, kGemmGrid, kH * kH2, M->grads + off_Wab(l), s);
    // dH_prev = dH + T[:, :64] + dZ Wab^T
    RowGemmArgs h{};
    h.n = n;
    h.n_pad = n_pad;
    h.X1 = M->dZ;
    h.ld1 = kH2;
    h.K1 = kH2;
    h.W = Wab_bf(M, l);
    h.R1 = dH;
    h.ldr1 = kH;
    h.R2 = M->T;
    h.ldr2 = kH2;
    h.Y = dHn;
    hipLaunchKernelGGL((k_mp_rowgemm<kH2, kH, true, false, false>), dim3(gg), dim3(256), 0, s, h);
    std::swap(dH, dHn);
  }
  hipLaunchKernelGGL(k_mp_enc_bwd, dim3(kNodeGrid), dim3(256), 0, s, n_pad, M->F, Hl(M, 0), dH,
                     M->part);
  fold(M->part, kNodeGrid, kFin * kH + kH, M->grads + off_enc_W(), s);
  OFH_HIP(hipGetLastError());
  return OFH_OK;
}

// The checks the network needs on A (host, once per pattern): positive
// diagonal, symmetric pattern, values symmetric to 1e-12 relative; returns
// the position of a_ji for every entry a_ij.
static int graph_check(ofh_csr A, std::vector<int>& tpos) {
  const int n = A->n;
  std::vector<int> rp(n + 1), col(A->nnz);
  std::vector<double> val(A->nnz);
  int rc = ofh_csr_get_pattern(A, rp.data(), col.data());
  if (rc) return rc;
  if ((rc = ofh_csr_get_values(A, val.data()))) return rc;
  std::vector<double> diag(n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int k = rp[i]; k < rp[i + 1]; ++k)
      if (col[k] == i) diag[i] = val[k];
  for (int i = 0; i < n; ++i)
    OFH_CHECK(diag[i] > 0.0, OFH_E_INVALID,
              "ofh_mpnn: the diagonal must be positive (row " + std::to_string(i + 1) + ")");
  tpos.assign(A->nnz, 0);
  for (int i = 0; i < n; ++i)
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      const int j = col[k];
      const int* p = std::lower_bound(col.data() + rp[j], col.data() + rp[j + 1], i);
      OFH_CHECK(p != col.data() + rp[j + 1] && *p == i, OFH_E_INVALID,
                "ofh_mpnn: the matrix pattern must be symmetric");
      const double aji = val[p - col.data()];
      OFH_CHECK(std::fabs(aji - val[k]) <= 1e-12 * std::sqrt(diag[i] * diag[j]), OFH_E_INVALID,
                "ofh_mpnn: the matrix must be symmetric");
      tpos[k] = (int)(p - col.data());
    }
  return OFH_OK;
}

// e_ij = (a_ij + a_ji) / (2 sqrt(a_ii a_jj)) (e_ij = e_ji exactly), one row
// per thread, and the partial sums of ln a_ii per workgroup
__global__ __launch_bounds__(256) void k_mp_edge_weights(int n, const int* __restrict__ rp,
                                                         const int* __restrict__ col,
                                                         const double* __restrict__ val,
                                                         const int* __restrict__ tpos,
                                                         const int* __restrict__ dpos,
                                                         float* __restrict__ e,
                                                         double* __restrict__ part) {
  __shared__ double red[4];
  const int i = blockIdx.x * 256 + threadIdx.x;
  double l = 0.0;
  if (i < n) {
    const double aii = val[dpos[i]];
    for (int k = rp[i]; k < rp[i + 1]; ++k)
      e[k] = (float)((val[k] + val[tpos[k]]) / (2.0 * sqrt(aii * val[dpos[col[k]]])));
    l = log(aii);
  }
  l = block_sum<256>(l, red);
  if (threadIdx.x == 0) part[blockIdx.x] = l;
}

// ld_i = ln a_ii - mean_j ln a_jj (every workgroup folds the partials alike)
__global__ __launch_bounds__(256) void k_mp_logdiag(int n, const double* __restrict__ val,
                                                    const int* __restrict__ dpos,
                                                    const double* __restrict__ part, int nparts,
                                                    float* __restrict__ ld) {
  __shared__ double red[4];
  double t = 0.0;
  for (int p = threadIdx.x; p < nparts; p += 256) t += part[p];
  t = block_sum<256>(t, red);
  const double mean = t / (double)n;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) ld[i] = (float)(log(val[dpos[i]]) - mean);
}

// The graph inputs of the network from A's current values, on the GPU.
static int graph_weights(ofh_mpnn M, hipStream_t s) {
  ofh_csr A = M->A;
  int rc = csr_ensure_dinv(A);  // diagonal positions and 1/a_ii of the current values
  if (rc) return rc;
  const int nb = (A->n + 255) / 256;
  hipLaunchKernelGGL(k_mp_edge_weights, dim3(nb), dim3(256), 0, s, A->n, A->row_ptr, A->col,
                     A->val, M->tpos, A->diag_pos, M->e, M->dpart);
  hipLaunchKernelGGL(k_mp_logdiag, dim3(nb), dim3(256), 0, s, A->n, A->val, A->diag_pos,
                     M->dpart, nb, M->ld);
  OFH_HIP(hipGetLastError());
  return OFH_OK;
}

static int refresh_bf16(ofh_mpnn M, hipStream_t s) {
  const int tot = M->L * 2 * kH * kH2;
  hipLaunchKernelGGL(k_mp_to_bf16, dim3((tot + 255) / 256), dim3(256), 0, s, M->L, M->params, M->wbf);
  OFH_HIP(hipGetLastError());
  return OFH_OK;
}

struct SplitMix {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform(double lo, double hi) { return lo + (hi - lo) * ((next() >> 11) * (1.0 / 9007199254740992.0)); }
};

}  // namespace mp
}  // namespace ofh

using namespace ofh;
using namespace ofh::mp;

extern "C" {

int ofh_mpnn_create(ofh_csr A, int layers, int hidden, int extrapolate, uint64_t seed,
                    ofh_mpnn* out) {
  OFH_CHECK(A && out, OFH_E_INVALID, "ofh_mpnn_create: bad argument");
  OFH_CHECK(hidden == kH, OFH_E_UNSUPPORTED, "ofh_mpnn_create: hidden width must be 64");
  OFH_CHECK(layers >= 1 && layers <= kMaxLayers, OFH_E_INVALID, "ofh_mpnn_create: layers in 1..8");
  OFH_CHECK(extrapolate == 0 || extrapolate == 1, OFH_E_INVALID, "ofh_mpnn_create: extrapolate is 0 or 1");
  OFH_CHECK(A->n > 0 && A->n <= kMaxRows, OFH_E_INVALID, "ofh_mpnn_create: matrix size");
  const int n = A->n;
  std::vector<int> tpos;
  int rc = graph_check(A, tpos);
  if (rc) return rc;
  ofh_mpnn M = new ofh_mpnn_s();
  M->A = A;
  M->L = layers;
  M->ext = extrapolate;
  M->n = n;
  M->n_pad = (n + 15) / 16 * 16;
  M->np = n_params(layers);
  M->nparts = (n + 255) / 256;
  const size_t np = (size_t)M->n_pad;
  auto alloc = [&](void** p, size_t bytes) { return hipMalloc(p, std::max<size_t>(bytes, 256)); };
  hipError_t he = hipSuccess;
  auto A_ = [&](void** p, size_t bytes) {
    if (he == hipSuccess) he = alloc(p, bytes);
  };
  A_((void**)&M->params, sizeof(float) * M->np);
  A_((void**)&M->grads, sizeof(float) * M->np);
  A_((void**)&M->adam_m, sizeof(float) * M->np);
  A_((void**)&M->adam_v, sizeof(float) * M->np);
  A_((void**)&M->wbf, sizeof(uint16_t) * layers * 2 * kH * kH2);
  A_((void**)&M->e, sizeof(float) * A->nnz);
  A_((void**)&M->tpos, sizeof(int) * A->nnz);
  A_((void**)&M->ld, sizeof(float) * n);
  A_((void**)&M->F, sizeof(float) * np * kFin);
  A_((void**)&M->Hs, sizeof(float) * (layers + 1) * np * kH);
  A_((void**)&M->Zs, sizeof(float) * layers * np * kH2);
  A_((void**)&M->Ms, sizeof(float) * layers * np * kH);
  A_((void**)&M->Us, sizeof(float) * layers * np * kH);
  A_((void**)&M->dH0, sizeof(float) * np * kH);
  A_((void**)&M->dH1, sizeof(float) * np * kH);
  A_((void**)&M->T, sizeof(float) * np * kH2);
  A_((void**)&M->dZ, sizeof(float) * np * kH2);
  A_((void**)&M->r, sizeof(double) * n);
  A_((void**)&M->z, sizeof(double) * n);
  A_((void**)&M->q, sizeof(double) * n);
  A_((void**)&M->dv, sizeof(double) * n);
  A_((void**)&M->scal, sizeof(double) * 8);
  A_((void**)&M->dpart, sizeof(double) * 4 * M->nparts);
  A_((void**)&M->part, sizeof(float) * std::max<size_t>((size_t)kGemmGrid * (kH2 * kH + kH2),
                                                         (size_t)kNodeGrid * (kFin * kH + kH)));
  if (he == hipSuccess) he = hipHostMalloc((void**)&M->h_pinned, 64, hipHostMallocDefault);
  if (he != hipSuccess) {
    free_all(M);
    delete M;
    return fail(OFH_E_NOMEM, std::string("ofh_mpnn_create: ") + hipGetErrorString(he));
  }
  // initial parameters: Glorot-uniform weights (decoder included), zero
  // biases; the line search keeps even the untrained guess no worse than x_b
  std::vector<float> p(M->np, 0.0f);
  SplitMix rng{seed};
  auto glorot = [&](int off, int fin, int fout) {
    const double a = std::sqrt(6.0 / (fin + fout));
    for (int i = 0; i < fin * fout; ++i) p[off + i] = (float)rng.uniform(-a, a);
  };
  glorot(off_enc_W(), kFin, kH);
  for (int l = 0; l < layers; ++l) {
    glorot(off_Wab(l), kH, kH2);
    for (int c = 0; c < kH; ++c) p[off_we(l) + c] = (float)rng.uniform(-1.0, 1.0);
    glorot(off_Wu(l), kH2, kH);
  }
  glorot(off_dec(layers), kH, 1);
  hipStream_t s = A->ctx->stream;
  int ok = OFH_OK;
  auto H_ = [&](hipError_t x) {
    if (x != hipSuccess && ok == OFH_OK) ok = fail(OFH_E_HIP, hipGetErrorString(x));
  };
  H_(hipMemcpyAsync(M->params, p.data(), sizeof(float) * M->np, hipMemcpyHostToDevice, s));
  H_(hipMemcpyAsync(M->tpos, tpos.data(), sizeof(int) * A->nnz, hipMemcpyHostToDevice, s));
  if (ok == OFH_OK) ok = graph_weights(M, s);
  H_(hipMemsetAsync(M->adam_m, 0, sizeof(float) * M->np, s));
  H_(hipMemsetAsync(M->adam_v, 0, sizeof(float) * M->np, s));
  H_(hipMemsetAsync(M->grads, 0, sizeof(float) * M->np, s));
  if (ok == OFH_OK) ok = refresh_bf16(M, s);
  H_(hipStreamSynchronize(s));
  if (ok != OFH_OK) {
    free_all(M);
    delete M;
    return ok;
  }
  *out = M;
  return OFH_OK;
}

int ofh_mpnn_destroy(ofh_mpnn M) {
  if (!M) return OFH_OK;
  hipDeviceSynchronize();
  free_all(M);
  delete M;
  return OFH_OK;
}

int ofh_mpnn_refresh(ofh_mpnn M) {
  OFH_CHECK(M, OFH_E_INVALID, "ofh_mpnn_refresh: null network");
  return graph_weights(M, M->A->ctx->stream);
}

int ofh_mpnn_num_params(ofh_mpnn M, int64_t* count) {
  OFH_CHECK(M && count, OFH_E_INVALID, "ofh_mpnn_num_params: bad argument");
  *count = M->np;
  return OFH_OK;
}

int ofh_mpnn_get_params(ofh_mpnn M, float* h_params) {
  OFH_CHECK(M && h_params, OFH_E_INVALID, "ofh_mpnn_get_params: bad argument");
  hipStream_t s = M->A->ctx->stream;
  OFH_HIP(hipMemcpyAsync(h_params, M->params, sizeof(float) * M->np, hipMemcpyDeviceToHost, s));
  OFH_HIP(hipStreamSynchronize(s));
  return OFH_OK;
}

int ofh_mpnn_set_params(ofh_mpnn M, const float* h_params) {
  OFH_CHECK(M && h_params, OFH_E_INVALID, "ofh_mpnn_set_params: bad argument");
  hipStream_t s = M->A->ctx->stream;
  OFH_HIP(hipMemcpyAsync(M->params, h_params, sizeof(float) * M->np, hipMemcpyHostToDevice, s));
  OFH_HIP(hipMemsetAsync(M->adam_m, 0, sizeof(float) * M->np, s));
  OFH_HIP(hipMemsetAsync(M->adam_v, 0, sizeof(float) * M->np, s));
  M->adam_t = 0;
  int rc = refresh_bf16(M, s);
  if (rc) return rc;
  OFH_HIP(hipStreamSynchronize(s));
  return OFH_OK;
}

int ofh_mpnn_use_graph(ofh_mpnn M, int enable) {
  OFH_CHECK(M, OFH_E_INVALID, "null network");
  M->use_graph = enable != 0;
  return OFH_OK;
}

int ofh_mpnn_guess(ofh_mpnn M, const double* d_xn, const double* d_xp, const double* d_b,
                   double* d_x0) {
  OFH_CHECK(M && d_xn && d_xp && d_b && d_x0, OFH_E_INVALID, "ofh_mpnn_guess: bad argument");
  hipStream_t s = M->A->ctx->stream;
  bool dia = false;  // the FP64 diagonal copy the row sums read, made current outside any capture
  if (int rc = csr_ensure_dia64(M->A, &dia)) return rc;
  if (!M->use_graph) return forward(M, d_xn, d_xp, d_b, d_x0, false, s);
  // the ~16 launches of an inference are captured once per argument set and
  // replayed (graph of kernels only)
  const void* key[4] = {d_xn, d_xp, d_b, d_x0};
  if (!M->gexec || !std::equal(key, key + 4, M->gkey)) {
    if (M->gexec) {
      OFH_HIP(hipGraphExecDestroy(M->gexec));
      M->gexec = nullptr;
    }
    if (!M->cap) OFH_HIP(hipStreamCreateWithFlags(&M->cap, hipStreamNonBlocking));
    OFH_HIP(hipStreamBeginCapture(M->cap, hipStreamCaptureModeThreadLocal));
    int rc = forward(M, d_xn, d_xp, d_b, d_x0, false, M->cap);
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(M->cap, &g);
    if (rc) {
      if (g) hipGraphDestroy(g);
      return rc;
    }
    OFH_HIP(ce);
    hipError_t ie = hipGraphInstantiate(&M->gexec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    OFH_HIP(ie);
    std::copy(key, key + 4, M->gkey);
  }
  OFH_HIP(hipGraphLaunch(M->gexec, s));
  return OFH_OK;
}

int ofh_mpnn_set_objective(ofh_mpnn M, int objective) {
  OFH_CHECK(M && (objective == 0 || objective == 1), OFH_E_INVALID,
            "ofh_mpnn_set_objective: 0 (A-norm error) or 1 (residual norm)");
  if (objective != M->obj && M->gexec) {  // the captured inference holds the old one
    OFH_HIP(hipGraphExecDestroy(M->gexec));
    M->gexec = nullptr;
  }
  M->obj = objective;
  return OFH_OK;
}

int ofh_mpnn_loss_grad(ofh_mpnn M, const double* d_xn, const double* d_xp, const double* d_b,
                       const double* d_xstar, double* d_x0, double* h_loss, float* h_grad) {
  OFH_CHECK(M && d_xn && d_xp && d_b && (d_xstar || M->obj == 1) && d_x0, OFH_E_INVALID,
            "ofh_mpnn_loss_grad: bad argument");
  hipStream_t s = M->A->ctx->stream;
  bool dia = false;  // the FP64 diagonal copy the forward row sums read
  int rc = csr_ensure_dia64(M->A, &dia);
  if (!rc) rc = forward(M, d_xn, d_xp, d_b, d_x0, true, s);
  if (rc) return rc;
  if ((rc = backward(M, d_xn, d_xp, d_xstar, s))) return rc;
  OFH_HIP(hipMemcpyAsync(M->h_pinned, M->scal, sizeof(double) * 4, hipMemcpyDeviceToHost, s));
  if (h_grad) OFH_HIP(hipMemcpyAsync(h_grad, M->grads, sizeof(float) * M->np, hipMemcpyDeviceToHost, s));
  OFH_HIP(hipStreamSynchronize(s));
  if (h_loss) *h_loss = M->h_pinned[3];
  return OFH_OK;
}

int ofh_mpnn_train_step(ofh_mpnn M, const double* d_xn, const double* d_xp, const double* d_b,
                        const double* d_xstar, double* d_x0, float lr, double* h_loss) {
  OFH_CHECK(M && d_xn && d_xp && d_b && (d_xstar || M->obj == 1) && d_x0 && lr > 0.f,
            OFH_E_INVALID,
            "ofh_mpnn_train_step: bad argument");
  hipStream_t s = M->A->ctx->stream;
  bool dia = false;  // the FP64 diagonal copy the forward row sums read
  int rc = csr_ensure_dia64(M->A, &dia);
  if (!rc) rc = forward(M, d_xn, d_xp, d_b, d_x0, true, s);
  if (rc) return rc;
  if ((rc = backward(M, d_xn, d_xp, d_xstar, s))) return rc;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  M->adam_t += 1;
  const float bc1 = 1.f - std::pow(b1, (float)M->adam_t), bc2 = 1.f - std::pow(b2, (float)M->adam_t);
  hipLaunchKernelGGL(k_mp_adam, dim3((unsigned)((M->np + 255) / 256)), dim3(256), 0, s, M->np,
                     M->params, M->grads, M->adam_m, M->adam_v, lr, b1, b2, eps, bc1, bc2);
  if ((rc = refresh_bf16(M, s))) return rc;
  if (h_loss) {
    OFH_HIP(hipMemcpyAsync(M->h_pinned, M->scal, sizeof(double) * 4, hipMemcpyDeviceToHost, s));
    OFH_HIP(hipStreamSynchronize(s));
    *h_loss = M->h_pinned[3];
  }
  return OFH_OK;
}

}  // extern "C"
