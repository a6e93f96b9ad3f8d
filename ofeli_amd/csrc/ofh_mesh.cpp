// ofh_mesh.cpp -- host-side mesh data of the path: the structured mesh
// constructors, equation numbering and the Gauss rule the Q4 kernels need.
//
// Node coordinates are produced the way Mesh::Mesh builds them (a running
// sum of hx / hy / hz, Mesh.cpp:568-607 and 703-726), so coordinates, and
// with them every element matrix, are bitwise the reference's.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../../include/ofeli_hip.h"

namespace {

constexpr double kPi = 3.14159265358979323846264338328;  // OFELI_PI, constants.h:62

// Legendre polynomial P_np and its derivative at y (Gauss.cpp:57-80).
void legendre_p(int np, double y, double& p, double& dp) {
  if (np == 0) { p = 1.0; dp = 0.0; return; }
  if (np == 1) { p = y; dp = 1.0; return; }
  double p0 = 1.0, p1 = y, p2 = 0.0;
  const int nn = np * (np + 1) / 2;
  for (int j = 2; j <= np; j++, p0 = p1, p1 = p2) p2 = ((2 * j - 1) * y * p1 - (j - 1) * p0) / j;
  p = p2;
  if (y == 1.0) dp = nn;
  else if (y == -1.0) dp = np % 2 == 1 ? nn : -nn;
  else dp = np * (y * p1 - p0) / (y * y - 1);
}

}  // namespace

extern "C" {

int ofh_mesh2d_structured(double xmin, double xmax, double ymin, double ymax, int nx, int ny,
                          int cx0, int cxN, int cy0, int cyN, int quad, int nb_dof,
                          double* coord, int32_t* code, int32_t* elem) {
  if (nx < 1 || ny < 1 || nb_dof < 1 || !coord || !code || !elem) return OFH_E_INVALID;
  const double hx = (xmax - xmin) / nx, hy = (ymax - ymin) / ny;
  const int ly = ny + 1;  // nodes per x-line
  int64_t ne = 0;
  double px = xmin;
  for (int i = 0; i <= nx; ++i, px += hx) {
    double py = ymin;
    for (int j = 0; j <= ny; ++j, py += hy) {
      const int64_t nd = (int64_t)i * ly + j;  // 0-based node
      coord[2 * nd] = px;
      coord[2 * nd + 1] = py;
      int c = 0;
      if (j == 0 && cy0 > 0) c = cy0;
      if (j == ny && cyN > 0) c = cyN;
      if (i == 0 && cx0 > 0) c = cx0;
      if (i == nx && cxN > 0) c = cxN;
      code[nd * nb_dof] = c;
      for (int k = 1; k < nb_dof; ++k) code[nd * nb_dof + k] = 0;
      if (i == 0 || j == 0) continue;
      // corners of cell (i,j), 1-based ids: (i-1,j-1) (i,j-1) (i,j) (i-1,j)
      const int32_t sw = (int32_t)((i - 1) * ly + (j - 1) + 1), se = (int32_t)(i * ly + (j - 1) + 1),
                    ne_ = (int32_t)(i * ly + j + 1), nw = (int32_t)((i - 1) * ly + j + 1);
      if (quad) {
        int32_t* e = elem + 4 * ne++;
        e[0] = sw; e[1] = se; e[2] = ne_; e[3] = nw;
      } else {
        int32_t* e = elem + 3 * ne++;
        e[0] = sw; e[1] = se; e[2] = ne_;
        e = elem + 3 * ne++;
        e[0] = ne_; e[1] = nw; e[2] = sw;
      }
    }
  }
  return OFH_OK;
}

int ofh_mesh3d_kuhn(double xmin, double xmax, double ymin, double ymax, double zmin, double zmax,
                    int nx, int ny, int nz, int cx0, int cxN, int cy0, int cyN, int cz0, int czN,
                    double* coord, int32_t* code, int32_t* elem) {
  if (nx < 1 || ny < 1 || nz < 1 || !coord || !code || !elem) return OFH_E_INVALID;
  const double hx = (xmax - xmin) / nx, hy = (ymax - ymin) / ny, hz = (zmax - zmin) / nz;
  int64_t nd = 0;
  double pz = zmin;
  for (int k = 0; k <= nz; ++k, pz += hz) {
    double py = ymin;
    for (int j = 0; j <= ny; ++j, py += hy) {
      double px = xmin;
      for (int i = 0; i <= nx; ++i, px += hx, ++nd) {
        coord[3 * nd] = px;
        coord[3 * nd + 1] = py;
        coord[3 * nd + 2] = pz;
        int c = 0;
        if (j == 0 && cy0 > 0) c = cy0;
        if (j == ny && cyN > 0) c = cyN;
        if (i == 0 && cx0 > 0) c = cx0;
        if (i == nx && cxN > 0) c = cxN;
        if (k == 0 && cz0 > 0) c = cz0;
        if (k == nz && czN > 0) c = czN;
        code[nd] = c;
      }
    }
  }
  // Kuhn split around the diagonal c0-c6: one tet per monotone path
  // c0 -> c0+e_a -> c0+e_a+e_b -> c6, listed with positive orientation.
  static const int path[6][4] = {{0, 1, 2, 6}, {0, 1, 6, 5}, {0, 3, 6, 2},
                                 {0, 3, 7, 6}, {0, 4, 5, 6}, {0, 4, 6, 7}};
  const int64_t sx = 1, sy = nx + 1, sz = (int64_t)(nx + 1) * (ny + 1);
  int64_t ne = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const int64_t o = k * sz + j * sy + i * sx;
        const int64_t c[8] = {o, o + sx, o + sx + sy, o + sy,
                              o + sz, o + sz + sx, o + sz + sx + sy, o + sz + sy};
        for (int t = 0; t < 6; ++t, ++ne)
          for (int a = 0; a < 4; ++a) elem[4 * ne + a] = (int32_t)(c[path[t][a]] + 1);
      }
  return OFH_OK;
}

int ofh_number_equations(int nb_nodes, int nb_dof, const int32_t* code, int32_t* dof,
                         int* nb_eq) {
  if (nb_nodes < 0 || nb_dof < 1 || !code || !dof || !nb_eq) return OFH_E_INVALID;
  int n = 0;
  const int64_t m = (int64_t)nb_nodes * nb_dof;
  for (int64_t t = 0; t < m; ++t) dof[t] = code[t] == 0 ? ++n : 0;
  *nb_eq = n;
  return OFH_OK;
}

int ofh_gauss(int np, double* x, double* w) {
  if (np < 1 || np > 10 || !x || !w) return OFH_E_INVALID;
  for (int i = 0; i <= (np - 1) / 2; i++) {
    double r = std::cos(kPi * (i + 0.75) / (np + 0.5)), p, dp;
    for (;;) {
      legendre_p(np, r, p, dp);
      if (std::fabs(p) <= 2.0 * 2.2204460492503131e-16 * std::fabs(r * dp)) break;
      r -= p / dp;
    }
    x[i] = -r;
    x[np - i - 1] = r;
    w[np - i - 1] = w[i] = 2.0 / ((1.0 - r * r) * dp * dp);
  }
  return OFH_OK;
}

}  // extern "C"
