// ofh_elem.h -- element geometry and element matrices on the device.
//
// Each routine follows the reference's floating-point operation order
// exactly (the library is compiled without FP contraction), so element
// matrices are bitwise those of the reference's host code:
//   Triang3::set        Triang3.cpp:68-93
//   Tetra4::CalculateShape Tetra4.cpp:76-114 (volume = det/6, Tetra4.h:87)
//   Quad4::setLocal / atGauss  Quad4.cpp:124-180
//   Laplace2DT3::LHS/BodyRHS   Laplace2DT3.cpp:131-143
//   DC2DT3 Diffusion/LCapacity/BodyRHS  DC2DT3.cpp:121-147, 248-253
//   Laplace3DT4::LHS/BodyRHS   Laplace3DT4.cpp:118-129
//   DC3DT4::Diffusion/BodyRHS  DC3DT4.cpp:128-135, 207-211
//   Elas2DQ4 PlaneStrain/PlaneStress/Deviator/Dilatation  Elas2DQ4.cpp:121-221
#pragma once

#include <hip/hip_runtime.h>

namespace ofh {

constexpr double kThird = 0.33333333333333333333333333333;  // OFELI_THIRD
constexpr double kSixth = 0.16666666666666666666666666667;  // OFELI_SIXTH

// returns 0, -1 (negative determinant) or -2 (null determinant)
__device__ __forceinline__ int tri_shape(const double (&x)[6], double& area, double (&d)[6]) {
  const double det = (x[2] - x[0]) * (x[5] - x[1]) - (x[3] - x[1]) * (x[4] - x[0]);
  if (det < 0.0) return -1;
  if (det == 0.0) return -2;
  area = 0.5 * det;
  d[0] = (x[3] - x[5]) / det;
  d[1] = (x[4] - x[2]) / det;
  d[2] = (x[5] - x[1]) / det;
  d[3] = (x[0] - x[4]) / det;
  d[4] = (x[1] - x[3]) / det;
  d[5] = (x[2] - x[0]) / det;
  return 0;
}

__device__ __forceinline__ int tet_shape(const double (&x)[12], double& vol, double (&d)[12]) {
  // Jacobian J(r,c) = 0 + sum_k dshl_k.c * x_k.r over k = 0..3 with
  // reference-element derivatives (-1,-1,-1), (1,0,0), (0,1,0), (0,0,1), as
  // the reference accumulates it.  For finite coordinates the products by
  // 0 are +-0 and leave the sum unchanged (a sum that starts at +0 is never
  // -0 in round-to-nearest, and adding +-0 to it is exact), so
  // J(r,c) = (0 - x_0.r) + x_{c+1}.r bit for bit.
  double J[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) J[r][c] = (0.0 - x[r]) + x[3 * (c + 1) + r];
  double IJ[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {  // cofactors, Tetra4.cpp:97-102 (0-based)
    const int j = (i + 1) % 3, k = (j + 1) % 3;
    IJ[i][i] = J[j][j] * J[k][k] - J[j][k] * J[k][j];
    IJ[j][i] = J[j][k] * J[k][i] - J[j][i] * J[k][k];
    IJ[i][j] = J[k][j] * J[i][k] - J[i][j] * J[k][k];
  }
  const double det = J[0][0] * IJ[0][0] + J[1][0] * IJ[0][1] + J[2][0] * IJ[0][2];
  if (det < 0.0) return -1;
  if (det == 0.0) return -2;
  const double c = 1. / det;
  // d_i.c = c * sum_m IJ(m,c) * dshl_i.m.  Node 0: c * ((-IJ0c - IJ1c) - IJ2c).
  // Nodes 1..3 pick one cofactor; the dropped products by 0 can change only
  // the sign of a zero d, and every element-matrix and right-hand-side entry
  // built from d enters sums that start at +0, where that sign cannot show.
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    d[m] = c * ((-IJ[0][m] + -IJ[1][m]) + -IJ[2][m]);
    d[3 * 1 + m] = c * IJ[0][m];
    d[3 * 2 + m] = c * IJ[1][m];
    d[3 * 3 + m] = c * IJ[2][m];
  }
  vol = kSixth * det;
  return 0;
}

// Q4 shape derivatives at the 2x2 Gauss points: dx/dy[node][gp], w[gp];
// gp = 2*i + j for the point (g_i, g_j).
__device__ __forceinline__ int quad_shape(const double (&x)[8], double g0, double g1,
                                          double w0, double w1, double (&dx)[4][4],
                                          double (&dy)[4][4], double (&w)[4]) {
  const double g[2] = {g0, g1}, gw[2] = {w0, w1};
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int gp = 2 * i + j;
      const double s = g[i], t = g[j];
      double lx[4], ly[4];
      lx[0] = -0.25 * (1. - t);
      lx[1] = 0.25 * (1. - t);
      lx[2] = 0.25 * (1. + t);
      lx[3] = -0.25 * (1. + t);
      ly[0] = -0.25 * (1. - s);
      ly[3] = 0.25 * (1. - s);
      ly[1] = -0.25 * (1. + s);
      ly[2] = 0.25 * (1. + s);
      double dxds = 0., dyds = 0., dxdt = 0., dydt = 0.;
#pragma unroll
      for (int a = 0; a < 4; a++) {
        dxds += lx[a] * x[2 * a];
        dxdt += ly[a] * x[2 * a];
        dyds += lx[a] * x[2 * a + 1];
        dydt += ly[a] * x[2 * a + 1];
      }
      const double det = dxds * dydt - dxdt * dyds;
      if (det < 0.0) return -1;
      if (det == 0.0) return -2;
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        dx[a][gp] = (dydt * lx[a] - dyds * ly[a]) / det;
        dy[a][gp] = (dxds * ly[a] - dxdt * lx[a]) / det;
      }
      w[gp] = det * gw[i] * gw[j];
    }
  return 0;
}

// Q4 shape function values at the 2x2 Gauss points, sh[node][gp] with
// gp = 2*i + j for (s, t) = (g_i, g_j) (Quad4::setLocal, Quad4.cpp:127-130).
__device__ __forceinline__ void quad_values(double g0, double g1, double (&sh)[4][4]) {
  const double g[2] = {g0, g1};
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int gp = 2 * i + j;
      const double s = g[i], t = g[j];
      sh[0][gp] = 0.25 * (1. - s) * (1. - t);
      sh[1][gp] = 0.25 * (1. + s) * (1. - t);
      sh[2][gp] = 0.25 * (1. + s) * (1. + t);
      sh[3][gp] = 0.25 * (1. - s) * (1. + t);
    }
}

// One entry (row 2i+di, column 2j+dj) of the Elas2DQ4 element matrix.
// eA0 accumulates the Deviator terms over the Gauss points, then (mode 0,
// as the reference) keeps accumulating the Dilatation terms on top, and eMat
// receives eA0 after each routine: eMat = Dev + (Dev + Dil).  Mode 1 (plane
// stress, separate terms) gives Dev + Dil.
__device__ __forceinline__ double q4_entry(int i, int di, int j, int dj, const double (&dx)[4][4],
                                           const double (&dy)[4][4], const double (&w)[4],
                                           double G, double lambda, int mode) {
  double dev = 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double c = G * w[k] * 1.0;
    const double ax = c * dx[i][k], ay = c * dy[i][k], bx = dx[j][k], by = dy[j][k];
    double t;
    if (di == 0 && dj == 0) t = 2 * ax * bx + ay * by;
    else if (di == 0) t = ay * bx;
    else if (dj == 0) t = ax * by;
    else t = 2 * ay * by + ax * bx;
    dev += t;
  }
  double acc = mode == 0 ? dev : 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double c = lambda * w[k] * 1.0;
    const double ax = c * dx[i][k], ay = c * dy[i][k], bx = dx[j][k], by = dy[j][k];
    double t;
    if (di == 0 && dj == 0) t = ax * bx;
    else if (di == 0) t = ax * by;
    else if (dj == 0) t = ay * bx;
    else t = ay * by;
    acc += t;
  }
  return (0.0 + dev) + acc;
}

// q4_entry with the row's node gradients already picked (gx[k] = dx[i][k],
// gy[k] = dy[i][k]) and its DOF component DI fixed: the same operations.
template <int DI>
__device__ __forceinline__ double q4_entry_g(const double (&gx)[4], const double (&gy)[4], int j,
                                             int dj, const double (&dx)[4][4],
                                             const double (&dy)[4][4], const double (&w)[4],
                                             double G, double lambda, int mode) {
  double dev = 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double c = G * w[k] * 1.0;
    const double ax = c * gx[k], ay = c * gy[k], bx = dx[j][k], by = dy[j][k];
    double t;
    if (DI == 0 && dj == 0) t = 2 * ax * bx + ay * by;
    else if (DI == 0) t = ay * bx;
    else if (dj == 0) t = ax * by;
    else t = 2 * ay * by + ax * bx;
    dev += t;
  }
  double acc = mode == 0 ? dev : 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double c = lambda * w[k] * 1.0;
    const double ax = c * gx[k], ay = c * gy[k], bx = dx[j][k], by = dy[j][k];
    double t;
    if (DI == 0 && dj == 0) t = ax * bx;
    else if (DI == 0) t = ax * by;
    else if (dj == 0) t = ay * bx;
    else t = ay * by;
    acc += t;
  }
  return (0.0 + dev) + acc;
}

__device__ __forceinline__ void elas_moduli(double E, double nu, int mode, double& lambda,
                                            double& G) {
  double l = nu * E / ((1 + nu) * (1 - 2 * nu));
  const double g = 0.5 * E / (1 + nu);
  if (mode == 1) l = 2 * l * g / (l + 2 * g);
  lambda = l;
  G = g;
}

}  // namespace ofh
