// ofh_assembly.hip -- sparse pattern and finite-element assembly on the GPU.
//
// Pattern (NodeGraph/FinalizeGraph/StoreGraph, GraphOfMatrix.cpp:52-67,
// 347-375, 619-634): the columns of the row of a free DOF are the free DOFs of
// every node that shares an element with its node.  Instead of sorting the
// full list of element DOF pairs, the GPU builds the node-to-element
// adjacency (elements in ascending order) and lets one thread per row merge
// the DOFs of its neighbours into a sorted, duplicate-free list.
//
// Assembly (Matrix<T_>::Assembly, Matrix_impl.h:172-201; Vect::Assembly,
// Vect_impl.h:1522-1534): one thread per element computes the element matrix
// and right-hand side into a buffer; a precomputed gather map lists, for every
// nonzero, the element-matrix entries that land on it in ascending element
// order, and one thread per nonzero sums them from 0.0 in that order.  This is
// the reference's summation order, without atomics, so A and b are bitwise
// the reference's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "ofh_elem.h"
#include "ofh_internal.h"

struct ofh_asm_s {
  ofh_ctx ctx = nullptr;
  int dim = 0, nb_nodes = 0, nb_el = 0, nen = 0, nb_dof = 0, nee = 0, nb_eq = 0;
  double* coord = nullptr;   // nb_nodes x dim
  int* elem = nullptr;       // nb_el x nen, 0-based node ids
  int* dof = nullptr;        // nb_nodes x nb_dof, 1-based labels, 0 = eliminated
  int* code = nullptr;       // nb_nodes x nb_dof
  int* adj_ptr = nullptr;    // nb_nodes + 1
  int* adj_el = nullptr;     // elements around each node, ascending
  int* adj_loc = nullptr;    // local index of the node in that element
  int* row_node = nullptr;   // nb_eq: node of each equation
  int* row_comp = nullptr;   // nb_eq: DOF component of each equation
  int* err = nullptr;          // device error slots
  ofh_csr A = nullptr;
  std::vector<int> h_dof;
  uint8_t* lmap = nullptr;     // per (adjacency entry, DOF component): each local column's
                               // position in the row (0xFF: eliminated DOF)
  int* trac = nullptr;         // side tractions: side nodes + per-DOF contribution lists
  size_t trac_cap = 0;         // ints held by trac (grown on demand)
};

namespace ofh {

constexpr int kMaxRowCand = 96;  // distinct columns per row the merge can hold

__global__ void k_adj_count(int nb_el, int nen, const int* __restrict__ elem, int* cnt) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nb_el * nen) return;
  atomicAdd(&cnt[elem[t]], 1);
}

__global__ void k_adj_fill(int nb_el, int nen, const int* __restrict__ elem,
                           const int* __restrict__ ptr, int* cursor, int* adj_el, int* adj_loc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nb_el * nen) return;
  const int node = elem[t];
  const int slot = ptr[node] + atomicAdd(&cursor[node], 1);
  adj_el[slot] = t / nen;
  adj_loc[slot] = t % nen;
}

// the slots above are taken in arbitrary order: sort each node's list
__global__ void k_adj_sort(int nb_nodes, const int* __restrict__ ptr, int* adj_el, int* adj_loc) {
  const int nd = blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= nb_nodes) return;
  const int a = ptr[nd], e = ptr[nd + 1];
  for (int i = a + 1; i < e; ++i) {
    const int ke = adj_el[i], kl = adj_loc[i];
    int j = i - 1;
    while (j >= a && adj_el[j] > ke) {
      adj_el[j + 1] = adj_el[j];
      adj_loc[j + 1] = adj_loc[j];
      --j;
    }
    adj_el[j + 1] = ke;
    adj_loc[j + 1] = kl;
  }
}

// Sorted distinct columns of row r into cols[] (returns count, -1 on overflow).
__device__ int row_columns(int node, int nen, int nb_dof, const int* __restrict__ elem,
                           const int* __restrict__ dof, const int* __restrict__ adj_ptr,
                           const int* __restrict__ adj_el, int* cols) {
  int n = 0;
  for (int t = adj_ptr[node]; t < adj_ptr[node + 1]; ++t) {
    const int* en = elem + (size_t)adj_el[t] * nen;
    for (int j = 0; j < nen; ++j)
      for (int l = 0; l < nb_dof; ++l) {
        const int d = dof[en[j] * nb_dof + l];
        if (!d) continue;
        const int c = d - 1;
        int pos = n;
        while (pos > 0 && cols[pos - 1] > c) --pos;
        if (pos > 0 && cols[pos - 1] == c) continue;
        if (n == kMaxRowCand) return -1;
        for (int q = n; q > pos; --q) cols[q] = cols[q - 1];
        cols[pos] = c;
        ++n;
      }
  }
  return n;
}

__global__ void k_row_count(int nb_eq, int nen, int nb_dof, const int* __restrict__ elem,
                            const int* __restrict__ dof, const int* __restrict__ adj_ptr,
                            const int* __restrict__ adj_el, const int* __restrict__ row_node,
                            int* row_len, int* err) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nb_eq) return;
  int cols[kMaxRowCand];
  const int n = row_columns(row_node[r], nen, nb_dof, elem, dof, adj_ptr, adj_el, cols);
  if (n < 0) atomicMin(err, r);
  row_len[r] = n < 0 ? 0 : n;
}

__global__ void k_row_fill(int nb_eq, int nen, int nb_dof, const int* __restrict__ elem,
                           const int* __restrict__ dof, const int* __restrict__ adj_ptr,
                           const int* __restrict__ adj_el, const int* __restrict__ row_node,
                           const int* __restrict__ row_ptr, int* col) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nb_eq) return;
  int cols[kMaxRowCand];
  const int n = row_columns(row_node[r], nen, nb_dof, elem, dof, adj_ptr, adj_el, cols);
  const int base = row_ptr[r];
  for (int q = 0; q < n; ++q) col[base + q] = cols[q];
}

__device__ __forceinline__ int find_col(const int* __restrict__ col, int a, int e, int c) {
  int lo = a, hi = e - 1;
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1;
    const int v = col[mid];
    if (v == c) return mid;
    if (v < c) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

// Positions within its row of the columns an element adds to: for every
// adjacency entry t of a node (element e, local node i) and DOF component k
// of that node, lmap[(t nb_dof + k) nee + j nb_dof + l] = position of the
// column of local DOF (j, l) in the row of DOF (node, k), or 0xFF when that
// column (or the row) is an eliminated DOF.  Rows hold at most kMaxRowCand
// (< 255) columns.
__global__ void k_local_map(int nb_nodes, int nen, int nb_dof, int nee,
                            const int* __restrict__ elem, const int* __restrict__ dof,
                            const int* __restrict__ adj_ptr, const int* __restrict__ adj_el,
                            const int* __restrict__ row_ptr, const int* __restrict__ col,
                            uint8_t* __restrict__ lmap) {
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= nb_nodes) return;
  for (int k = 0; k < nb_dof; ++k) {
    const int dr = dof[node * nb_dof + k];
    const int ra = dr ? row_ptr[dr - 1] : 0, re = dr ? row_ptr[dr] : 0;
    for (int t = adj_ptr[node]; t < adj_ptr[node + 1]; ++t) {
      const int* en = elem + (size_t)adj_el[t] * nen;
      uint8_t* out = lmap + ((size_t)t * nb_dof + k) * nee;
      for (int j = 0; j < nen; ++j)
        for (int l = 0; l < nb_dof; ++l) {
          const int d = dof[en[j] * nb_dof + l];
          const int p = (dr && d) ? find_col(col, ra, re, d - 1) : -1;
          out[j * nb_dof + l] = p < 0 ? (uint8_t)0xFF : (uint8_t)(p - ra);
        }
    }
  }
}

// ---------------------------------------------------------------------------
// element routines and the per-equation assembly
// ---------------------------------------------------------------------------
struct ElemArgs {
  int kind, nen, nb_dof, nee, dim;
  const double* coord;
  const int* elem;
  const int* code;
  const double* coef;
  const double* f;
  const double* bc;
  const double* uprev;
  double p0, p1, p2;  // kind parameters
  double g0, g1, w0, w1;  // 2-point Gauss rule (Q4)
  int* err;
};

__device__ __forceinline__ void report(int* err, int e, int rc) {
  atomicMin(&err[rc == -1 ? 1 : 2], e);
}

// Row `in` of element e's matrix (em_row, NEE entries, when want_row) and
// its right-hand-side entry (*rhs, when want_rhs), computed with the
// floating-point operations, in the order, of the equation classes' element
// routines: Laplace2DT3::LHS / BodyRHS, DC2DT3 Diffusion / LCapacity /
// BodyRHS and the backward-Euler combination, Laplace3DT4 / DC3DT4, then
// updateBC (Equation_impl.h:160-176) and the time term (Equa_Therm.h:203).
// (Elas2DQ4: elem_rows_q4.)
// Returns 0, or -1 / -2 for a negative / null Jacobian.
template <int KIND>
__device__ __forceinline__ int elem_row(const ElemArgs& a, int e, int in, bool want_row,
                                        bool want_rhs, double* em_row, double* rhs) {
  static_assert(KIND != OFH_ELAS2DQ4, "Elas2DQ4 rows come in pairs: elem_rows_q4");
  constexpr int NEN = KIND == OFH_LAPLACE3DT4 ? 4 : 3;
  constexpr int NB = 1;
  constexpr int NEE = NEN;
  constexpr int DIM = KIND == OFH_LAPLACE3DT4 ? 3 : 2;
  int nd[NEN];
#pragma unroll
  for (int i = 0; i < NEN; ++i) nd[i] = a.elem[(size_t)e * NEN + i];
  double x[NEN * DIM];
#pragma unroll
  for (int i = 0; i < NEN; ++i)
#pragma unroll
    for (int c = 0; c < DIM; ++c) x[DIM * i + c] = a.coord[(size_t)nd[i] * DIM + c];
  const bool need_row = want_row || (want_rhs && a.bc);
  double row[NEE];
  double r = 0.0;

  if constexpr (KIND == OFH_LAPLACE3DT4) {
    double d[12], meas;
    const int rc = tet_shape(reinterpret_cast<const double(&)[12]>(x), meas, d);
    if (rc) return rc;
    const bool var = a.coef != nullptr;
    const double c = var ? 1.0 * a.coef[e] * meas : 0.0;
    if (need_row) {
      double gi[3];  // d of node `in`, picked by node rather than by a 12-way register index
#pragma unroll
      for (int m = 0; m < 3; ++m)
        gi[m] = in == 0 ? d[m] : in == 1 ? d[3 + m] : in == 2 ? d[6 + m] : d[9 + m];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double dd = gi[0] * d[3 * j] + gi[1] * d[3 * j + 1] + gi[2] * d[3 * j + 2];
        row[j] = var ? 0.0 + (0.0 + c * dd) : meas * dd;
      }
    }
    if (want_rhs && a.f) {
      const double fi = a.f[nd[in]];
      r += var ? fi * 0.25 * meas : 0.25 * meas * fi;
    }
  } else {
    double d[6], meas;
    const int rc = tri_shape(reinterpret_cast<const double(&)[6]>(x), meas, d);
    if (rc) return rc;
    const double k = KIND == OFH_LAPLACE2DT3 ? 0.0 : a.coef[e];
    double ea0[3];
    const double gi0 = in == 0 ? d[0] : in == 1 ? d[2] : d[4];  // d of node `in`
    const double gi1 = in == 0 ? d[1] : in == 1 ? d[3] : d[5];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double dd = gi0 * d[2 * j] + gi1 * d[2 * j + 1] + 0.0 * 0.0;
      ea0[j] = KIND == OFH_LAPLACE2DT3 ? 0.0 + meas * dd : 0.0 + 1.0 * k * meas * dd;
    }
    double ea1[3] = {0.0, 0.0, 0.0};
    double dti = 0.0;
    if constexpr (KIND == OFH_DC2DT3_BE) {
      dti = 1. / a.p0;
      const double c = kThird * meas * a.p1 * a.p2 * 1.0;
      ea1[in] += c;  // lumped capacity: the diagonal entry only
#pragma unroll
      for (int j = 0; j < 3; ++j) row[j] = dti * ea1[j] + ea0[j];
    } else {
#pragma unroll
      for (int j = 0; j < 3; ++j) row[j] = ea0[j];
    }
    if (want_rhs && a.f) {
      if (KIND == OFH_LAPLACE2DT3) {
        r += a.f[nd[in]] * meas * kThird;
      } else {
        const double c = kThird * meas;
        r += c * a.f[nd[in]];
      }
    }
    if (want_rhs && a.bc) {  // updateBC, Equation_impl.h:160-176
#pragma unroll
      for (int j = 0; j < 3; ++j)
        if (a.code[nd[j]] > 0) r -= row[j] * a.bc[nd[j]];
    }
    if constexpr (KIND == OFH_DC2DT3_BE) {  // eRHS += dti*eA1*u^n, Equa_Therm.h:203
      if (want_rhs && a.uprev) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j) s += dti * ea1[j] * a.uprev[nd[j]];
        r = r + s;
      }
    }
    if (want_row)
#pragma unroll
      for (int j = 0; j < NEE; ++j) em_row[j] = row[j];
    if (want_rhs) *rhs = r;
    return 0;
  }
  if (want_rhs && a.bc) {  // updateBC, Equation_impl.h:160-176
#pragma unroll
    for (int j = 0; j < NEN; ++j)
#pragma unroll
      for (int l = 0; l < NB; ++l)
        if (a.code[nd[j] * NB + l] > 0) r -= row[j * NB + l] * a.bc[nd[j] * NB + l];
  }
  if (want_row)
#pragma unroll
    for (int j = 0; j < NEE; ++j) em_row[j] = row[j];
  if (want_rhs) *rhs = r;
  return 0;
}

// Both rows (2 il and 2 il + 1: the node's two DOFs) of an Elas2DQ4 element
// matrix and right-hand side, with the element's geometry computed once:
// each row by the operations of elem_row<OFH_ELAS2DQ4>, in the same order.
__device__ __forceinline__ int elem_rows_q4(const ElemArgs& a, int e, int il, bool want_row,
                                            bool want_rhs, double (&row0)[8], double (&row1)[8],
                                            double& rhs0, double& rhs1) {
  int nd[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) nd[i] = a.elem[(size_t)e * 4 + i];
  double x[8];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    x[2 * i] = a.coord[(size_t)nd[i] * 2];
    x[2 * i + 1] = a.coord[(size_t)nd[i] * 2 + 1];
  }
  const bool need_row = want_row || (want_rhs && a.bc);
  double dx[4][4], dy[4][4], w[4];
  const int rc = quad_shape(x, a.g0, a.g1, a.w0, a.w1, dx, dy, w);
  if (rc) return rc;
  double lambda, G;
  const int mode = (int)a.p1;
  elas_moduli(a.coef[e], a.p0, mode, lambda, G);
  double gx[4], gy[4];  // the node's gradients, picked by node rather than by a register index
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    gx[k] = il == 0 ? dx[0][k] : il == 1 ? dx[1][k] : il == 2 ? dx[2][k] : dx[3][k];
    gy[k] = il == 0 ? dy[0][k] : il == 1 ? dy[1][k] : il == 2 ? dy[2][k] : dy[3][k];
  }
  if (need_row) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int dj = 0; dj < 2; ++dj) {
        row0[2 * j + dj] = q4_entry_g<0>(gx, gy, j, dj, dx, dy, w, G, lambda, mode);
        row1[2 * j + dj] = q4_entry_g<1>(gx, gy, j, dj, dx, dy, w, G, lambda, mode);
      }
  }
  double r0 = 0.0, r1 = 0.0;
  if (want_rhs && a.f) {  // body force (fx, fy per node) at the Gauss points
    double sh[4][4];
    quad_values(a.g0, a.g1, sh);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double si = il == 0 ? sh[0][k] : il == 1 ? sh[1][k] : il == 2 ? sh[2][k] : sh[3][k];
      double fx = 0., fy = 0.;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        fx += sh[j][k] * a.f[nd[j] * 2];
        fy += sh[j][k] * a.f[nd[j] * 2 + 1];
      }
      r0 += w[k] * fx * si;
      r1 += w[k] * fy * si;
    }
  }
  if (want_rhs && a.bc) {  // updateBC, Equation_impl.h:160-176
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int l = 0; l < 2; ++l)
        if (a.code[nd[j] * 2 + l] > 0) {
          r0 -= row0[j * 2 + l] * a.bc[nd[j] * 2 + l];
          r1 -= row1[j * 2 + l] * a.bc[nd[j] * 2 + l];
        }
  }
  rhs0 = r0;
  rhs1 = r1;
  return 0;
}

// The element loop of build() and Matrix::Assembly / Vect::Assembly
// (Matrix_impl.h:172-201, Vect_impl.h:1522-1534) per equation.  A
// workgroup owns R consecutive rows, L lanes per row (L >= the usual row
// length).  The rows' adjacency entries (element, local node) form one
// contiguous range, which the workgroup walks in passes: every thread
// computes element rows and right-hand-side entries of the pass's entries
// into LDS (adjacency reads coalesced, geometry gathered from a small
// neighbourhood), then lane l of row r sums, over the entries of r's node
// in ascending element order, the contributions to nonzeros l and l + L
// in registers (longer rows: to nonzeros beyond 2L in place), lane 0 the
// right-hand side.  The LDS reads of that phase are broadcasts within a
// row's lanes.  Each nonzero and b_r is 0 + its element contributions in
// ascending element order, the reference's sums, with no element-matrix
// buffer and no atomics.
constexpr int kAsmStage = 512;  // element rows staged in LDS per pass
// The entry of a staged element row (row[j * kAsmStage], positions in
// words pos[0..W)) that lands on position q of its matrix row, or 0.0 when
// none does (positions of one element row are distinct; 0xFF never equals
// q).  Adding that 0.0 leaves a sum that started at +0.0 unchanged, bit for
// bit.
template <int W>
__device__ __forceinline__ double pick(const uint32_t* pos, int q, const double* row) {
  const uint32_t rep = 0x01010101u * (uint32_t)q;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    const uint32_t x = pos[w] ^ rep;
    const uint32_t z = (x - 0x01010101u) & ~x & 0x80808080u;  // lowest flag: a zero byte
    if (z) return row[(4 * w + (__builtin_ctz(z) >> 3)) * kAsmStage];
  }
  return 0.0;
}
constexpr int kAsmThreads = 256;  // threads per workgroup of k_asm_rows
static_assert(kAsmStage % kAsmThreads == 0, "a pass stages whole rounds of threads");
template <int KIND>
struct AsmShape {
  static constexpr int NEN = (KIND == OFH_LAPLACE3DT4 || KIND == OFH_ELAS2DQ4) ? 4 : 3;
  static constexpr int NB = KIND == OFH_ELAS2DQ4 ? 2 : 1;
  static constexpr int NEE = NEN * NB;
  static constexpr int L = NB == 2 ? 32 : (NEN == 4 ? 16 : 8);  // lanes per row
  static constexpr int R = kAsmThreads / L;                       // rows per workgroup
  static constexpr int C = kAsmStage / NB;                        // adjacency entries per pass
  static constexpr int WAVES = NB == 2 ? 1 : 5;  // waves per SIMD asked of the register allocator
};
template <int KIND>
__global__ __launch_bounds__(kAsmThreads)
__attribute__((amdgpu_waves_per_eu(AsmShape<KIND>::WAVES))) void k_asm_rows(int nb_eq, ElemArgs a,
                                                          const int* __restrict__ adj_ptr,
                                                          const int* __restrict__ adj_el,
                                                          const int* __restrict__ adj_loc,
                                                          const int* __restrict__ row_node,
                                                          const int* __restrict__ row_comp,
                                                          const uint8_t* __restrict__ lmap,
                                                          const int* __restrict__ row_ptr,
                                                          double* __restrict__ val,
                                                          double* __restrict__ b) {
  using S = AsmShape<KIND>;
  constexpr int NB = S::NB, NEE = S::NEE, L = S::L, R = S::R, C = S::C;
  constexpr int W = (NEE + 3) / 4;              // position words per element row
  __shared__ double s_row[NEE * kAsmStage];      // element rows of the pass, [j][entry, comp]
  __shared__ double s_rhs[kAsmStage];
  __shared__ uint32_t s_pos[kAsmStage * W];      // their column positions (lmap), 4 per word,
                                                 // padded with 0xFF
  OFH_LDS_BUDGET(sizeof(double) * kAsmStage * (NEE + 1) + sizeof(uint32_t) * kAsmStage * W);
  const bool want_mat = val != nullptr, want_rhs = b != nullptr;
  // consecutive equations have ascending nodes: the workgroup's entries are [t0, t1)
  const int r0 = blockIdx.x * R;
  const int r = r0 + threadIdx.x / L, lane = threadIdx.x % L;
  const int rl = min(r0 + R, nb_eq) - 1;
  const int t0 = adj_ptr[row_node[r0]], t1 = adj_ptr[row_node[rl] + 1];
  const bool live = r < nb_eq;
  int comp = 0, ra = 0, len = 0, ta = 0, tb = 0;
  if (live) {
    const int node = row_node[r];
    comp = row_comp[r];
    ra = row_ptr[r];
    len = row_ptr[r + 1] - ra;
    ta = adj_ptr[node];
    tb = adj_ptr[node + 1];
  }
  if (want_mat)  // Matrix::clear (Matrix_impl.h:164-169) of the in-place nonzeros
    for (int q = lane + 2 * L; q < len; q += L) val[ra + q] = 0.0;
  double acc0 = 0.0, acc1 = 0.0, s = 0.0;
  for (int c0 = t0; c0 < t1; c0 += C) {
    const int cn = min(C, t1 - c0);
    __syncthreads();  // the previous pass has been consumed
    if constexpr (NB == 2) {  // one thread per entry: both DOF rows from one geometry
      for (int q = threadIdx.x; q < cn; q += kAsmThreads) {
        const int t = c0 + q;
        const int e = adj_el[t];
        double row0[NEE], row1[NEE], rhs0 = 0.0, rhs1 = 0.0;
        const int rc = elem_rows_q4(a, e, adj_loc[t], want_mat, want_rhs, row0, row1, rhs0, rhs1);
        if (rc) {
          report(a.err, e, rc);  // the build fails; the values are not used
#pragma unroll
          for (int j = 0; j < NEE; ++j) row0[j] = row1[j] = 0.0;
          rhs0 = rhs1 = 0.0;
        }
        if (want_mat) {
#pragma unroll
          for (int j = 0; j < NEE; ++j) {
            s_row[j * kAsmStage + 2 * q] = row0[j];
            s_row[j * kAsmStage + 2 * q + 1] = row1[j];
          }
        }
        s_rhs[2 * q] = rhs0;
        s_rhs[2 * q + 1] = rhs1;
      }
    } else {
#pragma unroll
      for (int k = 0; k < kAsmStage / kAsmThreads; ++k) {  // unrolled: the loads overlap
        const int p = threadIdx.x + k * kAsmThreads;
        if (p >= cn) break;
        const int t = c0 + p;
        const int e = adj_el[t];
        double row[NEE], rhs = 0.0;
        const int rc = elem_row<KIND>(a, e, adj_loc[t], want_mat, want_rhs, row, &rhs);
        if (rc) {
          report(a.err, e, rc);  // the build fails; the values are not used
#pragma unroll
          for (int j = 0; j < NEE; ++j) row[j] = 0.0;
          rhs = 0.0;
        }
        if (want_mat) {
#pragma unroll
          for (int j = 0; j < NEE; ++j) s_row[j * kAsmStage + p] = row[j];
        }
        s_rhs[p] = rhs;
      }
    }
    if (want_mat) {
      const uint8_t* src = lmap + (size_t)c0 * NB * NEE;
      if constexpr (NEE % 4 == 0) {  // word-aligned: the plan's map starts on 256 bytes
        const uint32_t* src4 = reinterpret_cast<const uint32_t*>(src);
        for (int q = threadIdx.x; q < cn * NB * W; q += kAsmThreads) s_pos[q] = src4[q];
      } else {
        for (int q = threadIdx.x; q < cn * NB * W; q += kAsmThreads) {
          const int pe = q / W, j0 = 4 * (q % W);
          uint32_t x = 0xFFFFFFFFu;
#pragma unroll
          for (int k = 0; k < 4; ++k)
            if (j0 + k < NEE)
              x = (x & ~(0xFFu << (8 * k))) | ((uint32_t)src[pe * NEE + j0 + k] << (8 * k));
          s_pos[q] = x;
        }
      }
    }
    __syncthreads();
    const int lo = max(ta, c0), hi = min(tb, c0 + cn);
    for (int t = lo; t < hi; ++t) {
      const int p = (t - c0) * NB + comp;
      if (want_mat) {
        acc0 += pick<W>(s_pos + p * W, lane, s_row + p);
        if (len > L) acc1 += pick<W>(s_pos + p * W, lane + L, s_row + p);
        if (len > 2 * L)
          for (int q = lane + 2 * L; q < len; q += L) val[ra + q] += pick<W>(s_pos + p * W, q, s_row + p);
      }
      s += s_rhs[p];
    }
  }
  if (!live) return;
  if (want_mat) {
    if (lane < len) val[ra + lane] = acc0;
    if (lane + L < len) val[ra + lane + L] = acc1;
  }
  if (want_rhs && lane == 0) b[r] = s;
}

// Elas2DQ4 side tractions (Equa_Solid.h:194-202, Elas2DQ4::BoundaryRHS
// Elas2DQ4.cpp:256-263).  The reference adds side by side into b; here one
// thread per touched equation adds that equation's contributions in side
// order (the host lists them, stably sorted by equation), which is the same
// sequence of additions for every entry of b.
//   sides: nb_sides x 2 1-based nodes; lst_ptr/lst_dof: touched equations
//   (0-based) and their contribution ranges; lst_sk: side * 2 + component.
__global__ void k_traction(int nlist, const int* __restrict__ sides,
                           const int* __restrict__ lst_ptr, const int* __restrict__ lst_dof,
                           const int* __restrict__ lst_sk, const double* __restrict__ coord,
                           const double* __restrict__ tr, double* __restrict__ b, int* err) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nlist) return;
  const int d = lst_dof[t];
  double acc = b[d];
  for (int q = lst_ptr[t]; q < lst_ptr[t + 1]; ++q) {
    const int s = lst_sk[q] >> 1, k = lst_sk[q] & 1;
    const int n0 = sides[2 * s] - 1, n1 = sides[2 * s + 1] - 1;
    const double dlx = 0.5 * (coord[2 * n1] - coord[2 * n0]);
    const double dly = 0.5 * (coord[2 * n1 + 1] - coord[2 * n0 + 1]);
    const double det = sqrt(dlx * dlx + dly * dly + 0.0 * 0.0);  // Line2.cpp:85-86
    if (det == 0.0) {  // the side is skipped and reported
      atomicMin(&err[2], s);
      continue;
    }
    const double c = 0.5 * (2 * det);
    acc += 0.0 + c * tr[2 * s + k];
  }
  b[d] = acc;
}

// Vect::insertBC: u_i = v_j for free DOFs (in order), bc_i otherwise
__global__ void k_insert_bc(int m, const int* __restrict__ dof, const double* __restrict__ v,
                            const double* __restrict__ bc, double* __restrict__ u) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  const int d = dof[t];
  u[t] = d ? v[d - 1] : (bc ? bc[t] : 0.0);
}

static int host_exclusive_scan(hipStream_t s, const int* d_cnt, int n, std::vector<int64_t>& out) {
  std::vector<int> c(n);
  OFH_HIP(hipMemcpyAsync(c.data(), d_cnt, sizeof(int) * n, hipMemcpyDeviceToHost, s));
  OFH_HIP(hipStreamSynchronize(s));
  out.assign(n + 1, 0);
  for (int i = 0; i < n; ++i) out[i + 1] = out[i] + c[i];
  return OFH_OK;
}

static int check_err(ofh_asm P, const char* what) {
  int h[4];
  OFH_HIP(hipMemcpyAsync(h, P->err, sizeof(h), hipMemcpyDeviceToHost, P->ctx->stream));
  OFH_HIP(hipStreamSynchronize(P->ctx->stream));
  const int big = 0x7fffffff;
  if (h[0] != big)
    return fail(OFH_E_PATTERN, std::string(what) + ": row " + std::to_string(h[0] + 1) +
                                   " has more than 96 distinct columns");
  if (h[1] != big)
    return fail(OFH_E_NEGDET, std::string(what) + ": negative determinant of jacobian in element " +
                                  std::to_string(h[1] + 1));
  if (h[2] != big)
    return fail(OFH_E_ZERODET, std::string(what) + ": determinant of jacobian is null (element/side " +
                                   std::to_string(h[2] + 1) + ")");
  return OFH_OK;
}

static int reset_err(ofh_asm P) {
  const int big[4] = {0x7fffffff, 0x7fffffff, 0x7fffffff, 0x7fffffff};
  OFH_HIP(hipMemcpyAsync(P->err, big, sizeof(big), hipMemcpyHostToDevice, P->ctx->stream));
  return OFH_OK;
}

}  // namespace ofh

using namespace ofh;

extern "C" {

int ofh_asm_create(ofh_ctx ctx, int dim, int nb_nodes, const double* h_coord, int nb_el, int nen,
                   const int32_t* h_elem, int nb_dof, const int32_t* h_code, ofh_asm* out) {
  OFH_CHECK(ctx && h_coord && h_elem && h_code && out, OFH_E_INVALID, "ofh_asm_create: null argument");
  OFH_CHECK((dim == 2 || dim == 3) && nb_nodes > 0 && nb_el > 0 && (nen == 3 || nen == 4) &&
                (nb_dof == 1 || nb_dof == 2),
            OFH_E_INVALID, "ofh_asm_create: unsupported mesh shape");
  OFH_CHECK((int64_t)nb_el * nen < (int64_t)0x7fffffff, OFH_E_INVALID,
            "ofh_asm_create: element-node table exceeds 2^31 entries");
  for (int64_t t = 0; t < (int64_t)nb_el * nen; ++t)
    OFH_CHECK(h_elem[t] >= 1 && h_elem[t] <= nb_nodes, OFH_E_INVALID,
              "ofh_asm_create: element node id out of range");
  hipStream_t s = ctx->stream;
  ofh_asm P = new ofh_asm_s();
  P->ctx = ctx;
  P->dim = dim;
  P->nb_nodes = nb_nodes;
  P->nb_el = nb_el;
  P->nen = nen;
  P->nb_dof = nb_dof;
  P->nee = nen * nb_dof;
  // equation numbering (Mesh.cpp:1499-1509)
  P->h_dof.resize((size_t)nb_nodes * nb_dof);
  int neq = 0;
  ofh_number_equations(nb_nodes, nb_dof, h_code, P->h_dof.data(), &neq);
  P->nb_eq = neq;
  OFH_CHECK(neq > 0, OFH_E_INVALID, "ofh_asm_create: every DOF is imposed");
  OFH_CHECK(neq <= kMaxRows, OFH_E_INVALID, "ofh_asm_create: too many equations");
  std::vector<int> rnode(neq), rcomp(neq);
  for (int64_t t = 0; t < (int64_t)nb_nodes * nb_dof; ++t)
    if (P->h_dof[t]) {
      rnode[P->h_dof[t] - 1] = (int)(t / nb_dof);
      rcomp[P->h_dof[t] - 1] = (int)(t % nb_dof);
    }
  std::vector<int> elem0((size_t)nb_el * nen);
  for (size_t t = 0; t < elem0.size(); ++t) elem0[t] = h_elem[t] - 1;
  OFH_HIP(hipMalloc(&P->coord, sizeof(double) * (size_t)nb_nodes * dim));
  OFH_HIP(hipMalloc(&P->elem, sizeof(int) * elem0.size()));
  OFH_HIP(hipMalloc(&P->dof, sizeof(int) * P->h_dof.size()));
  OFH_HIP(hipMalloc(&P->code, sizeof(int) * P->h_dof.size()));
  OFH_HIP(hipMalloc(&P->row_node, sizeof(int) * neq));
  OFH_HIP(hipMalloc(&P->row_comp, sizeof(int) * neq));
  OFH_HIP(hipMalloc(&P->err, sizeof(int) * 4));
  OFH_HIP(hipMemcpyAsync(P->coord, h_coord, sizeof(double) * (size_t)nb_nodes * dim, hipMemcpyHostToDevice, s));
  OFH_HIP(hipMemcpyAsync(P->elem, elem0.data(), sizeof(int) * elem0.size(), hipMemcpyHostToDevice, s));
  OFH_HIP(hipMemcpyAsync(P->dof, P->h_dof.data(), sizeof(int) * P->h_dof.size(), hipMemcpyHostToDevice, s));
  OFH_HIP(hipMemcpyAsync(P->code, h_code, sizeof(int) * P->h_dof.size(), hipMemcpyHostToDevice, s));
  OFH_HIP(hipMemcpyAsync(P->row_node, rnode.data(), sizeof(int) * neq, hipMemcpyHostToDevice, s));
  OFH_HIP(hipMemcpyAsync(P->row_comp, rcomp.data(), sizeof(int) * neq, hipMemcpyHostToDevice, s));
  int rc = reset_err(P);
  if (rc) return rc;

  // node -> element adjacency
  const int tot = nb_el * nen;
  int* cnt;
  OFH_HIP(hipMalloc(&cnt, sizeof(int) * (size_t)std::max(nb_nodes, 1)));
  OFH_HIP(hipMemsetAsync(cnt, 0, sizeof(int) * nb_nodes, s));
  hipLaunchKernelGGL(k_adj_count, dim3((tot + 255) / 256), dim3(256), 0, s, nb_el, nen, P->elem, cnt);
  std::vector<int64_t> aptr;
  if ((rc = host_exclusive_scan(s, cnt, nb_nodes, aptr))) return rc;
  std::vector<int> aptr32(aptr.begin(), aptr.end());
  OFH_HIP(hipMalloc(&P->adj_ptr, sizeof(int) * (nb_nodes + 1)));
  OFH_HIP(hipMalloc(&P->adj_el, sizeof(int) * tot));
  OFH_HIP(hipMalloc(&P->adj_loc, sizeof(int) * tot));
  OFH_HIP(hipMemcpyAsync(P->adj_ptr, aptr32.data(), sizeof(int) * (nb_nodes + 1), hipMemcpyHostToDevice, s));
  OFH_HIP(hipMemsetAsync(cnt, 0, sizeof(int) * nb_nodes, s));
  hipLaunchKernelGGL(k_adj_fill, dim3((tot + 255) / 256), dim3(256), 0, s, nb_el, nen, P->elem,
                     P->adj_ptr, cnt, P->adj_el, P->adj_loc);
  hipLaunchKernelGGL(k_adj_sort, dim3((nb_nodes + 255) / 256), dim3(256), 0, s, nb_nodes,
                     P->adj_ptr, P->adj_el, P->adj_loc);
  OFH_HIP(hipFree(cnt));

  // CSR pattern
  int* rlen;
  OFH_HIP(hipMalloc(&rlen, sizeof(int) * neq));
  hipLaunchKernelGGL(k_row_count, dim3((neq + 255) / 256), dim3(256), 0, s, neq, nen, nb_dof,
                     P->elem, P->dof, P->adj_ptr, P->adj_el, P->row_node, rlen, P->err);
  if ((rc = check_err(P, "NodeGraph"))) return rc;
  std::vector<int64_t> rptr;
  if ((rc = host_exclusive_scan(s, rlen, neq, rptr))) return rc;
  OFH_HIP(hipFree(rlen));
  const int64_t nnz = rptr[neq];
  OFH_CHECK(nnz < (int64_t)0x7fffffff, OFH_E_INVALID, "ofh_asm_create: nnz exceeds 2^31");
  std::vector<int> rp32(rptr.begin(), rptr.end());
  ofh_csr A = new ofh_csr_s();
  A->ctx = ctx;
  A->serial = next_csr_serial();
  A->n = neq;
  A->nnz = nnz;
  OFH_HIP(hipMalloc(&A->row_ptr, sizeof(int) * (neq + 1)));
  OFH_HIP(hipMalloc(&A->col, sizeof(int) * (std::max<int64_t>(nnz, 1) + kPad)));
  OFH_HIP(hipMalloc(&A->val, sizeof(double) * (std::max<int64_t>(nnz, 1) + kPad)));
  OFH_HIP(hipMemsetAsync(A->col + nnz, 0, sizeof(int) * kPad, ctx->stream));
  OFH_HIP(hipMemsetAsync(A->val + nnz, 0, sizeof(double) * kPad, ctx->stream));
  OFH_HIP(hipMemsetAsync(A->val, 0, sizeof(double) * std::max<int64_t>(nnz, 1), s));
  OFH_HIP(hipMemcpyAsync(A->row_ptr, rp32.data(), sizeof(int) * (neq + 1), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_row_fill, dim3((neq + 255) / 256), dim3(256), 0, s, neq, nen, nb_dof,
                     P->elem, P->dof, P->adj_ptr, P->adj_el, P->row_node, A->row_ptr, A->col);
  // row blocks of the CSR-stream SpMV (both tiles)
  {
    int maxr = 0;
    for (int i = 0; i < neq; ++i) maxr = std::max(maxr, rp32[i + 1] - rp32[i]);
    A->max_row_nnz = maxr;
    const int rc = csr_set_row_blocks(A, rp32.data());
    if (rc) return rc;
  }
  P->A = A;
  // column positions of every element contribution within its row
  OFH_HIP(hipMalloc(&P->lmap, std::max<size_t>((size_t)tot * nb_dof * P->nee, 1)));
  hipLaunchKernelGGL(k_local_map, dim3((nb_nodes + 255) / 256), dim3(256), 0, s, nb_nodes, nen,
                     nb_dof, P->nee, P->elem, P->dof, P->adj_ptr, P->adj_el, A->row_ptr, A->col,
                     P->lmap);
  OFH_HIP(hipGetLastError());

  OFH_HIP(hipStreamSynchronize(s));
  *out = P;
  return OFH_OK;
}

int ofh_asm_destroy(ofh_asm P) {
  if (!P) return OFH_OK;
  hipDeviceSynchronize();
  void* ptrs[] = {P->coord, P->elem, P->dof, P->code, P->adj_ptr, P->adj_el, P->adj_loc,
                  P->row_node, P->row_comp, P->err, P->lmap, P->trac};
  for (void* p : ptrs) hipFree(p);
  ofh_csr_destroy(P->A);
  delete P;
  return OFH_OK;
}

int ofh_asm_info(ofh_asm P, int* nb_eq, int64_t* nnz) {
  OFH_CHECK(P, OFH_E_INVALID, "null plan");
  if (nb_eq) *nb_eq = P->nb_eq;
  if (nnz) *nnz = P->A->nnz;
  return OFH_OK;
}

int ofh_asm_matrix(ofh_asm P, ofh_csr* A) {
  OFH_CHECK(P && A, OFH_E_INVALID, "null argument");
  *A = P->A;
  return OFH_OK;
}

int ofh_asm_get_dof(ofh_asm P, int32_t* h_dof) {
  OFH_CHECK(P && h_dof, OFH_E_INVALID, "null argument");
  std::memcpy(h_dof, P->h_dof.data(), sizeof(int) * P->h_dof.size());
  return OFH_OK;
}

int ofh_asm_build(ofh_asm P, int kind, const double* d_coef, const double* d_f,
                  const double* d_bc, const double* d_uprev, const double* h_prm, int nprm,
                  int want_matrix, double* d_b) {
  OFH_CHECK(P, OFH_E_INVALID, "ofh_asm_build: null plan");
  const bool tri = kind == OFH_LAPLACE2DT3 || kind == OFH_DC2DT3 || kind == OFH_DC2DT3_BE;
  OFH_CHECK((tri && P->nen == 3 && P->dim == 2 && P->nb_dof == 1) ||
                (kind == OFH_LAPLACE3DT4 && P->nen == 4 && P->dim == 3 && P->nb_dof == 1) ||
                (kind == OFH_ELAS2DQ4 && P->nen == 4 && P->dim == 2 && P->nb_dof == 2),
            OFH_E_INVALID, "ofh_asm_build: element kind does not match the mesh");
  OFH_CHECK(kind == OFH_LAPLACE2DT3 || kind == OFH_LAPLACE3DT4 || d_coef, OFH_E_INVALID,
            "ofh_asm_build: this kind needs a per-element coefficient");
  OFH_CHECK(kind != OFH_DC2DT3_BE || (h_prm && nprm >= 3), OFH_E_INVALID,
            "ofh_asm_build: DC2DT3_BE needs {dt, rho, cp}");
  OFH_CHECK(kind != OFH_ELAS2DQ4 || (h_prm && nprm >= 2), OFH_E_INVALID,
            "ofh_asm_build: ELAS2DQ4 needs {nu, mode}");
  hipStream_t s = P->ctx->stream;
  int rc = reset_err(P);
  if (rc) return rc;
  ElemArgs a{};
  a.kind = kind;
  a.nen = P->nen;
  a.nb_dof = P->nb_dof;
  a.nee = P->nee;
  a.dim = P->dim;
  a.coord = P->coord;
  a.elem = P->elem;
  a.code = P->code;
  a.coef = d_coef;
  a.f = d_f;
  a.bc = d_bc;
  a.uprev = d_uprev;
  a.p0 = nprm > 0 ? h_prm[0] : 0.0;
  a.p1 = nprm > 1 ? h_prm[1] : 0.0;
  a.p2 = nprm > 2 ? h_prm[2] : 0.0;
  double gx[2], gw[2];
  ofh_gauss(2, gx, gw);
  a.g0 = gx[0];
  a.g1 = gx[1];
  a.w0 = gw[0];
  a.w1 = gw[1];
  a.err = P->err;
  ofh_csr A = P->A;
  double* val = want_matrix ? A->val : nullptr;
  if (!val && !d_b) return OFH_OK;
#define OFH_ASM_ROWS(K)                                                                       \
  hipLaunchKernelGGL(k_asm_rows<K>, dim3((P->nb_eq + AsmShape<K>::R - 1) / AsmShape<K>::R),     \
                     dim3(kAsmThreads), 0, s, P->nb_eq, a, P->adj_ptr, P->adj_el, P->adj_loc,   \
                     P->row_node, P->row_comp, P->lmap, A->row_ptr, val, d_b)
  switch (kind) {
    case OFH_LAPLACE2DT3: OFH_ASM_ROWS(OFH_LAPLACE2DT3); break;
    case OFH_DC2DT3: OFH_ASM_ROWS(OFH_DC2DT3); break;
    case OFH_DC2DT3_BE: OFH_ASM_ROWS(OFH_DC2DT3_BE); break;
    case OFH_LAPLACE3DT4: OFH_ASM_ROWS(OFH_LAPLACE3DT4); break;
    case OFH_ELAS2DQ4: OFH_ASM_ROWS(OFH_ELAS2DQ4); break;
    default: return fail(OFH_E_INVALID, "ofh_asm_build: unknown kind");
  }
#undef OFH_ASM_ROWS
  OFH_HIP(hipGetLastError());
  if (want_matrix) csr_values_changed(A);
  return check_err(P, "element loop");
}

int ofh_asm_traction(ofh_asm P, int nb_sides, const int32_t* h_side_nodes,
                     const double* d_traction, double* d_b) {
  OFH_CHECK(P && h_side_nodes && d_traction && d_b && nb_sides >= 0, OFH_E_INVALID,
            "ofh_asm_traction: bad argument");
  OFH_CHECK(P->dim == 2 && P->nb_dof == 2, OFH_E_INVALID, "ofh_asm_traction: 2D, 2 DOF meshes");
  for (int t = 0; t < 2 * nb_sides; ++t)
    OFH_CHECK(h_side_nodes[t] >= 1 && h_side_nodes[t] <= P->nb_nodes, OFH_E_INVALID,
              "ofh_asm_traction: side node out of range");
  if (nb_sides == 0) return OFH_OK;
  hipStream_t s = P->ctx->stream;
  // contributions (equation, side, component) in the reference's order of
  // additions, stably grouped by equation: cnt[d] per equation, then the
  // device buffer [sides | side*2+k per contribution | ranges | equations]
  std::vector<int> cnt(P->nb_eq, 0);
  for (int sd = 0; sd < nb_sides; ++sd)
    for (int i = 0; i < 2; ++i)
      for (int k = 0; k < 2; ++k) {
        const int d = P->h_dof[(h_side_nodes[2 * sd + i] - 1) * 2 + k];
        if (d) ++cnt[d - 1];
      }
  std::vector<int> start(P->nb_eq + 1, 0);
  for (int d = 0; d < P->nb_eq; ++d) start[d + 1] = start[d] + cnt[d];
  const int ncontrib = start[P->nb_eq];
  std::vector<int> sk(ncontrib);
  std::vector<int> cursor(start.begin(), start.end() - 1);
  for (int sd = 0; sd < nb_sides; ++sd)
    for (int i = 0; i < 2; ++i)
      for (int k = 0; k < 2; ++k) {
        const int d = P->h_dof[(h_side_nodes[2 * sd + i] - 1) * 2 + k];
        if (d) sk[cursor[d - 1]++] = sd * 2 + k;
      }
  std::vector<int> ptr(1, 0), eqs;
  for (int d = 0; d < P->nb_eq; ++d)
    if (cnt[d]) {
      eqs.push_back(d);
      ptr.push_back(start[d + 1]);
    }
  const int nlist = (int)eqs.size();
  std::vector<int> host(h_side_nodes, h_side_nodes + 2 * nb_sides);
  host.insert(host.end(), sk.begin(), sk.end());
  host.insert(host.end(), ptr.begin(), ptr.end());
  host.insert(host.end(), eqs.begin(), eqs.end());
  const size_t need = host.size();
  if (P->trac_cap < need) {  // grown on demand, kept by the plan
    OFH_HIP(hipStreamSynchronize(s));
    if (P->trac) OFH_HIP(hipFree(P->trac));
    P->trac = nullptr;
    P->trac_cap = 0;
    OFH_HIP(hipMalloc(&P->trac, sizeof(int) * need));
    P->trac_cap = need;
  }
  OFH_HIP(hipMemcpyAsync(P->trac, host.data(), sizeof(int) * need, hipMemcpyHostToDevice, s));
  OFH_HIP(hipStreamSynchronize(s));  // `host` goes out of scope on every return path
  int rc = reset_err(P);
  if (rc) return rc;
  if (nlist > 0) {
    const int* d_sides = P->trac;
    const int* d_sk = d_sides + 2 * nb_sides;
    const int* d_ptr = d_sk + ncontrib;
    const int* d_eq = d_ptr + ptr.size();
    hipLaunchKernelGGL(k_traction, dim3((nlist + 255) / 256), dim3(256), 0, s, nlist, d_sides,
                       d_ptr, d_eq, d_sk, P->coord, d_traction, d_b, P->err);
    OFH_HIP(hipGetLastError());
  }
  return check_err(P, "Line2");
}

int ofh_asm_insert_bc(ofh_asm P, const double* d_v, const double* d_bc, double* d_u) {
  OFH_CHECK(P && d_v && d_u, OFH_E_INVALID, "ofh_asm_insert_bc: null argument");
  const int m = P->nb_nodes * P->nb_dof;
  hipLaunchKernelGGL(k_insert_bc, dim3((m + 255) / 256), dim3(256), 0, P->ctx->stream, m, P->dof,
                     d_v, d_bc, d_u);
  OFH_HIP(hipGetLastError());
  return OFH_OK;
}

}  // extern "C"
