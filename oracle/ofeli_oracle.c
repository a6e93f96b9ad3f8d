/*
 * ofeli_oracle.c -- CPU restatement of OFELI 5.3.3's sparse linear-solve path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  It is loaded
 * by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by
 * the product library.  Each function cites the reference file:line whose
 * algorithm and floating-point operation order it restates.  Build strictly
 * (gcc -O2 -ffp-contract=off) for parity; the timed CPU baseline may use the
 * reference's own -Ofast (CMakeLists.txt:78-80).
 */
#include "ofeli_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OFELI_THIRD 0.33333333333333333333333333333 /* constants.h:70 */
#define OFELI_SIXTH 0.16666666666666666666666666667 /* constants.h:78 */
#define OFELI_PI 3.14159265358979323846264338328    /* constants.h:62 */
#define DBL_EPS 2.2204460492503131e-16

void orc_free(void* p) { free(p); }

/* ======================================================================= */
/* Meshes                                                                  */
/* ======================================================================= */

oi orc_mesh2d_nb_nodes(oi nx, oi ny) { return (nx + 1) * (ny + 1); }
oi orc_mesh2d_nb_elements(oi nx, oi ny, int quad) { return quad ? nx * ny : 2 * nx * ny; }

/* Structured 2D mesh, Mesh.cpp:547-615.  Nodes run x-outer / y-inner with
 * coordinates accumulated by repeated addition of hx, hy (Mesh.cpp:568-607).
 * Only DOF 1 of each node receives the boundary code (Mesh.cpp:577-588);
 * later assignments win (cy0, cyN, cx0, cxN).  elem holds 1-based node ids:
 * triangles (i-1,j-1),(i,j-1),(i,j) and (i,j),(i-1,j),(i-1,j-1)
 * (Mesh.cpp:600-611); quads (i-1,j-1),(i,j-1),(i,j),(i-1,j) (Mesh.cpp:592-598). */
void orc_mesh2d(double xmin, double xmax, double ymin, double ymax, oi nx, oi ny,
                int cx0, int cxN, int cy0, int cyN, int quad, int nb_dof,
                double* coord, int* code, oi* elem) {
  oi n = 0, ne = 0;
  double px = xmin, py;
  double hx = (xmax - xmin) / (double)nx, hy = (ymax - ymin) / (double)ny;
  for (oi i = 0; i <= nx; ++i) {
    py = ymin;
    for (oi j = 0; j <= ny; ++j) {
      n++;
      coord[2 * (n - 1)] = px;
      coord[2 * (n - 1) + 1] = py;
      int c = 0;
      if (j == 0 && cy0 > 0) c = cy0;
      if (j == ny && cyN > 0) c = cyN;
      if (i == 0 && cx0 > 0) c = cx0;
      if (i == nx && cxN > 0) c = cxN;
      for (int k = 0; k < nb_dof; ++k) code[(n - 1) * nb_dof + k] = 0;
      code[(n - 1) * nb_dof] = c;
      py += hy;
      if (i > 0 && j > 0) {
        /* theNodes[n-ny-3] .. theNodes[n-1] are 0-based; store 1-based ids */
        oi a = n - ny - 2, b = n - 1, cc = n, d = n - ny - 1;
        if (quad) {
          oi* e = elem + 4 * ne++;
          e[0] = a; e[1] = b; e[2] = cc; e[3] = d;
        } else {
          oi* e = elem + 3 * ne++;
          e[0] = a; e[1] = b; e[2] = cc;
          e = elem + 3 * ne++;
          e[0] = cc; e[1] = d; e[2] = a;
        }
      }
    }
    px += hx;
  }
}

/* Boundary sides of the structured 2D mesh with negative codes
 * (Mesh.cpp:617-669): x=xmin, x=xmax, y=ymin, y=ymax in that order.  Side
 * nodes are 1-based node ids; side_code receives -c. */
void orc_mesh2d_sides(oi nx, oi ny, int cx0, int cxN, int cy0, int cyN,
                      oi* side_nodes, int* side_code, oi* nb_sides) {
  oi is = 0, n;
  n = 1;
  if (cx0 < 0)
    for (oi j = 1; j <= ny; ++j) {
      side_nodes[2 * is] = n; side_nodes[2 * is + 1] = n + 1; side_code[is++] = -cx0; n++;
    }
  n = nx * (ny + 1) + 1;
  if (cxN < 0)
    for (oi j = 1; j <= ny; ++j) {
      side_nodes[2 * is] = n; side_nodes[2 * is + 1] = n + 1; side_code[is++] = -cxN; n++;
    }
  n = 1;
  if (cy0 < 0)
    for (oi i = 1; i <= nx; ++i) {
      side_nodes[2 * is] = n; side_nodes[2 * is + 1] = n + ny + 1; side_code[is++] = -cy0;
      n += ny + 1;
    }
  n = ny + 1;
  if (cyN < 0)
    for (oi i = 1; i <= nx; ++i) {
      side_nodes[2 * is] = n; side_nodes[2 * is + 1] = n + ny + 1; side_code[is++] = -cyN;
      n += ny + 1;
    }
  *nb_sides = is;
}

oi orc_mesh3d_nb_nodes(oi nx, oi ny, oi nz) { return (nx + 1) * (ny + 1) * (nz + 1); }

/* Structured 3D mesh.  Nodes and codes follow Mesh.cpp:703-726 (x inner,
 * then y, then z; accumulated coordinates; priority y0,yN,x0,xN,z0,zN).
 * The reference's TETRAHEDRON branch (Mesh.cpp:747-771) does not advance its
 * node counter and produces negatively oriented tets, so the cell split is
 * the project's own: the positively oriented Kuhn split of each cube into 6
 * tets around the diagonal c0-c6 (corner numbering as Mesh.cpp:733-735). */
void orc_mesh3d_kuhn(double xmin, double xmax, double ymin, double ymax,
                     double zmin, double zmax, oi nx, oi ny, oi nz,
                     int cx0, int cxN, int cy0, int cyN, int cz0, int czN,
                     double* coord, int* code, oi* elem) {
  double hx = (xmax - xmin) / (double)nx, hy = (ymax - ymin) / (double)ny,
         hz = (zmax - zmin) / (double)nz;
  double px, py, pz = zmin;
  oi n = 0;
  for (oi k = 0; k <= nz; k++) {
    py = ymin;
    for (oi j = 0; j <= ny; j++) {
      px = xmin;
      for (oi i = 0; i <= nx; i++) {
        coord[3 * n] = px; coord[3 * n + 1] = py; coord[3 * n + 2] = pz;
        int c = 0;
        if (j == 0 && cy0 > 0) c = cy0;
        if (j == ny && cyN > 0) c = cyN;
        if (i == 0 && cx0 > 0) c = cx0;
        if (i == nx && cxN > 0) c = cxN;
        if (k == 0 && cz0 > 0) c = cz0;
        if (k == nz && czN > 0) c = czN;
        code[n] = c;
        n++;
        px += hx;
      }
      py += hy;
    }
    pz += hz;
  }
  /* Kuhn tets (corner indices into c[0..7]) */
  static const int T[6][4] = {{0, 1, 2, 6}, {0, 1, 6, 5}, {0, 3, 6, 2},
                              {0, 3, 7, 6}, {0, 4, 5, 6}, {0, 4, 6, 7}};
  oi nn = (nx + 1) * (ny + 1), ne = 0;
  for (oi k = 1; k <= nz; ++k)
    for (oi j = 1; j <= ny; ++j)
      for (oi i = 1; i <= nx; ++i) {
        oi b = (k - 1) * nn + (j - 1) * (nx + 1) + (i - 1); /* 0-based c0 */
        oi c[8] = {b, b + 1, b + nx + 2, b + nx + 1,
                   b + nn, b + nn + 1, b + nn + nx + 2, b + nn + nx + 1};
        for (int t = 0; t < 6; ++t)
          for (int a = 0; a < 4; ++a) elem[4 * ne + 4 * t + a] = c[T[t][a]] + 1;
        ne += 6;
      }
}

/* DOF numbering with imposed-DOF elimination, Mesh.cpp:1499-1509:
 * code==0 -> next equation number in node order, else 0. */
oi orc_number_equations(oi nb_nodes, int nb_dof, const int* code, oi* dof) {
  oi nb_eq = 0;
  for (oi n = 0; n < nb_nodes; ++n)
    for (int i = 0; i < nb_dof; ++i)
      dof[n * nb_dof + i] = (code[n * nb_dof + i] == 0) ? ++nb_eq : 0;
  return nb_eq;
}

/* ======================================================================= */
/* Sparse pattern: NodeGraph + FinalizeGraph + StoreGraph                  */
/* ======================================================================= */

static void radix_sort_u64(uint64_t* a, uint64_t* tmp, oi n, int bits) {
  /* LSD radix sort, 16-bit digits; sorting is order-only so the result is
   * identical to std::sort of the pair list (GraphOfMatrix.cpp:56). */
  oi* cnt = (oi*)malloc(sizeof(oi) * 65536);
  uint64_t *src = a, *dst = tmp;
  for (int shift = 0; shift < bits; shift += 16) {
    memset(cnt, 0, sizeof(oi) * 65536);
    for (oi i = 0; i < n; ++i) cnt[(src[i] >> shift) & 0xFFFF]++;
    oi s = 0;
    for (int d = 0; d < 65536; ++d) { oi c = cnt[d]; cnt[d] = s; s += c; }
    for (oi i = 0; i < n; ++i) dst[cnt[(src[i] >> shift) & 0xFFFF]++] = src[i];
    uint64_t* t = src; src = dst; dst = t;
  }
  if (src != a) memcpy(a, src, sizeof(uint64_t) * n);
  free(cnt);
}

/* NodeGraph (GraphOfMatrix.cpp:347-375): every element DOF pair (nk,nl) with
 * both labels nonzero is pushed, plus its transpose when nk!=nl; then
 * sort + unique (FinalizeGraph, :52-67) and StoreGraph (:619-634).  Returns
 * nnz and fills row_ptr[nb_eq+1] (0-based) and *col_ind (1-based, malloc'd).
 * Returns ORC_E_PATTERN if some row is empty (StoreGraph would misalign). */
oi orc_node_graph(oi nb_el, int nen, int nb_dof, const oi* elem, const oi* dof,
                  oi nb_eq, oi* row_ptr, oi** col_ind) {
  oi ned = (oi)nen * nb_dof;
  oi cap = nb_el * ned * ned * 2;
  uint64_t* ij = (uint64_t*)malloc(sizeof(uint64_t) * (cap > 0 ? cap : 1));
  if (!ij) return ORC_E_NOMEM;
  oi np = 0;
  for (oi e = 0; e < nb_el; ++e) {
    const oi* en = elem + e * nen;
    for (int in = 0; in < nen; ++in)
      for (int k = 0; k < nb_dof; ++k) {
        oi nk = dof[(en[in] - 1) * nb_dof + k];
        if (!nk) continue;
        for (int jn = 0; jn < nen; ++jn)
          for (int l = 0; l < nb_dof; ++l) {
            oi nl = dof[(en[jn] - 1) * nb_dof + l];
            if (!nl) continue;
            ij[np++] = ((uint64_t)(nk - 1) << 32) | (uint64_t)(nl - 1);
            if (nk != nl) ij[np++] = ((uint64_t)(nl - 1) << 32) | (uint64_t)(nk - 1);
          }
      }
  }
  uint64_t* tmp = (uint64_t*)malloc(sizeof(uint64_t) * (np > 0 ? np : 1));
  if (!tmp) { free(ij); return ORC_E_NOMEM; }
  int bits = 64;
  radix_sort_u64(ij, tmp, np, bits);
  free(tmp);
  oi nu = 0;
  for (oi k = 0; k < np; ++k)
    if (nu == 0 || ij[k] != ij[nu - 1]) ij[nu++] = ij[k];
  oi* ci = (oi*)malloc(sizeof(oi) * (nu > 0 ? nu : 1));
  /* StoreGraph: col_ind 1-based, row_ptr pushed where the row changes */
  oi nr = 0;
  row_ptr[nr++] = 0;
  for (oi k = 0; k < nu; ++k) {
    ci[k] = (oi)(ij[k] & 0xFFFFFFFFu) + 1;
    if (k > 0 && (ij[k] >> 32) > (ij[k - 1] >> 32)) {
      if ((oi)(ij[k] >> 32) != (oi)(ij[k - 1] >> 32) + 1) { free(ij); free(ci); return ORC_E_PATTERN; }
      row_ptr[nr++] = k;
    }
  }
  free(ij);
  if (nr != nb_eq) { free(ci); return ORC_E_PATTERN; }
  row_ptr[nr] = nu;
  *col_ind = ci;
  return nu;
}

/* NodeGraphScal, GraphOfMatrix.cpp:532-546: the overload SpMatrix::setMesh
 * calls for dof != 0 (SpMatrix_impl.h:399-402).  Every element node pair
 * (in, jn), both orders, as node numbers: one row per node, no equation
 * numbering and no imposed-DOF elimination; the matrix has nb_nodes rows
 * (Matrix_impl.h:148-150).  FinalizeGraph/StoreGraph as above. */
oi orc_node_graph_scal(oi nb_el, int nen, const oi* elem, oi nb_nodes, oi* row_ptr,
                       oi** col_ind) {
  oi cap = nb_el * nen * nen * 2;
  uint64_t* ij = (uint64_t*)malloc(sizeof(uint64_t) * (cap > 0 ? cap : 1));
  if (!ij) return ORC_E_NOMEM;
  oi np = 0;
  for (oi e = 0; e < nb_el; ++e) {
    const oi* en = elem + e * nen;
    for (int in = 0; in < nen; ++in)
      for (int jn = 0; jn < nen; ++jn) {
        uint64_t a = (uint64_t)(en[in] - 1), b = (uint64_t)(en[jn] - 1);
        ij[np++] = (a << 32) | b;
        if (en[in] != en[jn]) ij[np++] = (b << 32) | a;
      }
  }
  uint64_t* tmp = (uint64_t*)malloc(sizeof(uint64_t) * (np > 0 ? np : 1));
  if (!tmp) { free(ij); return ORC_E_NOMEM; }
  radix_sort_u64(ij, tmp, np, 64);
  free(tmp);
  oi nu = 0;
  for (oi k = 0; k < np; ++k)
    if (nu == 0 || ij[k] != ij[nu - 1]) ij[nu++] = ij[k];
  oi* ci = (oi*)malloc(sizeof(oi) * (nu > 0 ? nu : 1));
  oi nr = 0;
  row_ptr[nr++] = 0;
  for (oi k = 0; k < nu; ++k) {
    ci[k] = (oi)(ij[k] & 0xFFFFFFFFu) + 1;
    if (k > 0 && (ij[k] >> 32) > (ij[k - 1] >> 32)) {
      /* a node in no element leaves an empty row, which StoreGraph cannot
       * represent (it pushes row_ptr only where the row index changes) */
      if ((oi)(ij[k] >> 32) != (oi)(ij[k - 1] >> 32) + 1) { free(ij); free(ci); return ORC_E_PATTERN; }
      row_ptr[nr++] = k;
    }
  }
  free(ij);
  if (nr != nb_nodes) { free(ci); return ORC_E_PATTERN; }
  row_ptr[nr] = nu;
  *col_ind = ci;
  return nu;
}

/* ======================================================================= */
/* Shape functions                                                         */
/* ======================================================================= */

/* Triang3::set, Triang3.cpp:68-93.  x = {x0,y0,x1,y1,x2,y2};
 * dsh = {dx0,dy0,dx1,dy1,dx2,dy2}. */
int orc_triang3(const double* x, double* area, double* dsh) {
  double det = (x[2] - x[0]) * (x[5] - x[1]) - (x[3] - x[1]) * (x[4] - x[0]);
  if (det < 0.0) return ORC_E_NEGDET;
  if (det == 0.0) return ORC_E_ZERODET;
  *area = 0.5 * det;
  dsh[0] = (x[3] - x[5]) / det;
  dsh[1] = (x[4] - x[2]) / det;
  dsh[2] = (x[5] - x[1]) / det;
  dsh[3] = (x[0] - x[4]) / det;
  dsh[4] = (x[1] - x[3]) / det;
  dsh[5] = (x[2] - x[0]) / det;
  return ORC_OK;
}

/* Tetra4::CalculateShape, Tetra4.cpp:76-114, volume = det/6 (Tetra4.h:87).
 * x = 4 nodes x (x,y,z); dsh = 4 x (dx,dy,dz). */
int orc_tetra4(const double* x, double* vol, double* dsh) {
  static const double L[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  double J[4][4], IJ[4][4]; /* 1-based */
  memset(J, 0, sizeof J);
  memset(IJ, 0, sizeof IJ);
  for (int k = 0; k < 4; k++) {
    const double* p = x + 3 * k;
    J[1][1] += L[k][0] * p[0]; J[1][2] += L[k][1] * p[0]; J[1][3] += L[k][2] * p[0];
    J[2][1] += L[k][0] * p[1]; J[2][2] += L[k][1] * p[1]; J[2][3] += L[k][2] * p[1];
    J[3][1] += L[k][0] * p[2]; J[3][2] += L[k][1] * p[2]; J[3][3] += L[k][2] * p[2];
  }
  for (int i = 1; i <= 3; ++i) {
    int j = i % 3 + 1, k = j % 3 + 1;
    IJ[i][i] = J[j][j] * J[k][k] - J[j][k] * J[k][j];
    IJ[j][i] = J[j][k] * J[k][i] - J[j][i] * J[k][k];
    IJ[i][j] = J[k][j] * J[i][k] - J[i][j] * J[k][k];
  }
  double det = J[1][1] * IJ[1][1] + J[2][1] * IJ[1][2] + J[3][1] * IJ[1][3];
  if (det < 0.0) return ORC_E_NEGDET;
  if (det == 0.0) return ORC_E_ZERODET;
  double c = 1. / det;
  for (int i = 0; i < 4; ++i) {
    dsh[3 * i + 0] = c * (IJ[1][1] * L[i][0] + IJ[2][1] * L[i][1] + IJ[3][1] * L[i][2]);
    dsh[3 * i + 1] = c * (IJ[1][2] * L[i][0] + IJ[2][2] * L[i][1] + IJ[3][2] * L[i][2]);
    dsh[3 * i + 2] = c * (IJ[1][3] * L[i][0] + IJ[2][3] * L[i][1] + IJ[3][3] * L[i][2]);
  }
  *vol = OFELI_SIXTH * det;
  return ORC_OK;
}

/* Gauss-Legendre points by Newton iteration, Gauss.cpp:37-80. */
static void legendre(int np, double y, double* p, double* dp) {
  double p0, p1, p2 = 0;
  if (np == 0) { *p = 1.0; *dp = 0.0; return; }
  if (np == 1) { *p = y; *dp = 1.0; return; }
  p0 = 1.0;
  p1 = y;
  int nn = np * (np + 1) / 2;
  for (int j = 2; j <= np; j++, p0 = p1, p1 = p2)
    p2 = ((2 * j - 1) * y * p1 - (j - 1) * p0) / j;
  *p = p2;
  if (y == 1.0) *dp = nn;
  else if (y == -1.0) *dp = np % 2 == 1 ? nn : -nn;
  else *dp = np * (y * p1 - p0) / (y * y - 1);
}

void orc_gauss(int np, double* x, double* w) {
  double p, dp, r;
  for (int i = 0; i <= (np - 1) / 2; i++) {
    r = cos(OFELI_PI * (i + 0.75) / (np + 0.5));
    while (1) {
      legendre(np, r, &p, &dp);
      if (fabs(p) <= 2.0 * DBL_EPS * fabs(r * dp)) break;
      r -= p / dp;
    }
    x[i] = -r;
    x[np - i - 1] = r;
    w[np - i - 1] = w[i] = 2.0 / ((1.0 - r * r) * dp * dp);
  }
}

/* Quad4::setLocal (Quad4.cpp:124-157) evaluated at the 2x2 Gauss points as
 * in Quad4::atGauss (Quad4.cpp:160-180): sh[4*node+gp], dsh[2*(4*node+gp)],
 * w[gp] = det*wi*wj, gp = i*2+j with (s,t) = (g_i, g_j). */
int orc_quad4_atgauss(const double* x, double* sh, double* dsh, double* w) {
  double gx[2], gw[2];
  orc_gauss(2, gx, gw);
  for (int k = 0; k < 4; ++k) {
    int ij = 0;
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) {
        double s = gx[i], t = gx[j];
        double lsh[4], dx[4], dy[4];
        lsh[0] = 0.25 * (1. - s) * (1. - t);
        lsh[1] = 0.25 * (1. + s) * (1. - t);
        lsh[2] = 0.25 * (1. + s) * (1. + t);
        lsh[3] = 0.25 * (1. - s) * (1. + t);
        dx[0] = -0.25 * (1. - t); dx[1] = 0.25 * (1. - t);
        dx[2] = 0.25 * (1. + t);  dx[3] = -0.25 * (1. + t);
        dy[0] = -0.25 * (1. - s); dy[3] = 0.25 * (1. - s);
        dy[1] = -0.25 * (1. + s); dy[2] = 0.25 * (1. + s);
        double dxds = 0., dyds = 0., dxdt = 0., dydt = 0.;
        for (int a = 0; a < 4; a++) {
          dxds += dx[a] * x[2 * a];
          dxdt += dy[a] * x[2 * a];
          dyds += dx[a] * x[2 * a + 1];
          dydt += dy[a] * x[2 * a + 1];
        }
        double det = dxds * dydt - dxdt * dyds;
        if (det < 0.0) return ORC_E_NEGDET;
        if (det == 0.0) return ORC_E_ZERODET;
        double ax = dx[k], ay = dy[k];
        sh[4 * k + ij] = lsh[k];
        dsh[2 * (4 * k + ij)] = (dydt * ax - dyds * ay) / det;
        dsh[2 * (4 * k + ij) + 1] = (dxds * ay - dxdt * ax) / det;
        w[ij++] = det * gw[i] * gw[j];
      }
  }
  return ORC_OK;
}

/* ======================================================================= */
/* Element matrices                                                        */
/* ======================================================================= */

static inline double dot3(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; /* Point operator, Point.h:217-220 */
}
static inline double dot2(const double* a, const double* b) {
  /* 2D Point: z components are 0 */
  return a[0] * b[0] + a[1] * b[1] + 0.0 * 0.0;
}

/* Elastic moduli, Elas2DQ4.cpp:121-152.  mode 0 = PlaneStrain (what
 * Elas2DQ4::set always ends with, :89); mode 1 = PlaneStress. */
static void elas_moduli(double E, double nu, int mode, double* lambda, double* G) {
  double l = nu * E / ((1 + nu) * (1 - 2 * nu));
  double g = 0.5 * E / (1 + nu);
  if (mode == 1) l = 2 * l * g / (l + 2 * g);
  *lambda = l;
  *G = g;
}

/* Computes the element matrix of one element.
 *   kind ORC_LAPLACE2DT3: Laplace2DT3::LHS, Laplace2DT3.cpp:131-136
 *   kind ORC_DC2DT3:      DC2DT3::Diffusion, DC2DT3.cpp:141-147 (eMat := eA0)
 *   kind ORC_DC2DT3_BE:   LCapacity (:121-128) + Diffusion, then
 *                         eMat = dti*eA1 + eA0 (Equa_Therm.h:192);
 *                         prm = {dt, rho, cp}; ea1 receives eA1
 *   kind ORC_LAPLACE3DT4: Laplace3DT4::LHS (Laplace3DT4.cpp:118-123) when
 *                         coef < 0, else DC3DT4::Diffusion (DC3DT4.cpp:128-135)
 *   kind ORC_ELAS2DQ4:    Elas2DQ4 Deviator + Dilatation (Elas2DQ4.cpp:186-221);
 *                         prm = {nu, mode}: mode 0 = reference parity
 *                         (plane strain, eMat = 2*Dev + Dil: both routines add
 *                         the cumulative eA0), mode 1 = plane stress, Dev + Dil,
 *                         mode 2 = plane strain, Dev + Dil.
 * x holds the element's node coordinates (dim per node). */
int orc_element(int kind, const double* x, double coef, const double* prm,
                double* emat, double* ea1) {
  int rc;
  if (kind == ORC_LAPLACE2DT3 || kind == ORC_DC2DT3 || kind == ORC_DC2DT3_BE) {
    double area, d[6];
    if ((rc = orc_triang3(x, &area, d)) != ORC_OK) return rc;
    double ea0[9];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        double dd = dot2(d + 2 * i, d + 2 * j);
        if (kind == ORC_LAPLACE2DT3) ea0[3 * i + j] = 0.0 + area * dd;
        else ea0[3 * i + j] = 0.0 + 1.0 * coef * area * dd;
      }
    if (kind == ORC_DC2DT3_BE) {
      double dti = 1. / prm[0];
      double c = OFELI_THIRD * area * prm[1] * prm[2] * 1.0;
      double a1[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      a1[0] += c; a1[4] += c; a1[8] += c;
      for (int k = 0; k < 9; ++k) emat[k] = dti * a1[k] + ea0[k];
      if (ea1) memcpy(ea1, a1, sizeof a1);
    } else {
      memcpy(emat, ea0, sizeof ea0);
    }
    return ORC_OK;
  }
  if (kind == ORC_LAPLACE3DT4) {
    double vol, d[12];
    if ((rc = orc_tetra4(x, &vol, d)) != ORC_OK) return rc;
    if (coef < 0) {
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) emat[4 * i + j] = vol * dot3(d + 3 * i, d + 3 * j);
    } else {
      double c = 1.0 * coef * vol;
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) emat[4 * i + j] = 0.0 + (0.0 + c * dot3(d + 3 * i, d + 3 * j));
    }
    return ORC_OK;
  }
  if (kind == ORC_ELAS2DQ4) {
    double sh[16], dsh[32], w[4];
    if ((rc = orc_quad4_atgauss(x, sh, dsh, w)) != ORC_OK) return rc;
    double lambda, G;
    int mode = (int)prm[1];
    elas_moduli(coef, prm[0], mode, &lambda, &G);
    double ea0[64], em[64];
    memset(ea0, 0, sizeof ea0);
    memset(em, 0, sizeof em);
    /* Deviator */
    for (int k = 0; k < 4; ++k) {
      double c = G * w[k] * 1.0;
      for (int i = 0; i < 4; ++i) {
        double ax = c * dsh[2 * (4 * i + k)], ay = c * dsh[2 * (4 * i + k) + 1];
        for (int j = 0; j < 4; ++j) {
          double bx = dsh[2 * (4 * j + k)], by = dsh[2 * (4 * j + k) + 1];
          ea0[(2 * i) * 8 + 2 * j] += 2 * ax * bx + ay * by;
          ea0[(2 * i) * 8 + 2 * j + 1] += ay * bx;
          ea0[(2 * i + 1) * 8 + 2 * j] += ax * by;
          ea0[(2 * i + 1) * 8 + 2 * j + 1] += 2 * ay * by + ax * bx;
        }
      }
    }
    for (int k = 0; k < 64; ++k) em[k] += ea0[k];
    if (mode != 0) memset(ea0, 0, sizeof ea0); /* textbook modes: separate terms */
    /* Dilatation */
    for (int k = 0; k < 4; ++k) {
      double c = lambda * w[k] * 1.0;
      for (int i = 0; i < 4; ++i) {
        double ax = c * dsh[2 * (4 * i + k)], ay = c * dsh[2 * (4 * i + k) + 1];
        for (int j = 0; j < 4; ++j) {
          double bx = dsh[2 * (4 * j + k)], by = dsh[2 * (4 * j + k) + 1];
          ea0[(2 * i) * 8 + 2 * j] += ax * bx;
          ea0[(2 * i) * 8 + 2 * j + 1] += ax * by;
          ea0[(2 * i + 1) * 8 + 2 * j] += ay * bx;
          ea0[(2 * i + 1) * 8 + 2 * j + 1] += ay * by;
        }
      }
    }
    for (int k = 0; k < 64; ++k) em[k] += ea0[k];
    memcpy(emat, em, sizeof em);
    return ORC_OK;
  }
  return ORC_E_ARG;
}

/* ======================================================================= */
/* Assembly                                                                */
/* ======================================================================= */

static oi col_pos(const oi* row_ptr, const oi* col_ind, oi i1, oi j1) {
  /* SpMatrix::_col_index (SpMatrix_impl.h:1008-1015) searches row i for
   * column j; the columns are sorted, so a binary search finds the same k. */
  oi lo = row_ptr[i1 - 1], hi = row_ptr[i1] - 1;
  while (lo <= hi) {
    oi mid = (lo + hi) / 2;
    if (col_ind[mid] == j1) return mid;
    if (col_ind[mid] < j1) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

/* Element loop of the equation classes' build():
 *   Equa_Laplace::build (Equa_Laplace.h:112-133): LHS, BodyRHS, updateBC,
 *     Matrix::Assembly, Vect::Assembly;
 *   Equa_Therm::build (Equa_Therm.h:162-218): terms, eMat, Matrix::Assembly,
 *     BodyRHS, updateBC, time term, Vect::Assembly.
 * a[] and b[] are cleared first (Matrix_impl.h:164-169, Equa.cpp:389) and
 * every element adds in ascending element order (Matrix_impl.h:172-201,
 * Vect_impl.h:1522-1534); updateBC is Equation_impl.h:160-176.
 * coord: dim per node; coef: per-element coefficient (k or E) or NULL;
 * f: nodal source (Laplace/DC kinds) or NULL; bc: nodal Dirichlet values
 * (nb_dof per node) or NULL; u_prev: nodal u^n for ORC_DC2DT3_BE. */
int orc_assemble(int kind, int dim, const double* coord, oi nb_el, const oi* elem,
                 int nb_dof, const oi* dof, const int* code, const double* bc,
                 const double* f, const double* coef, const double* prm,
                 const double* u_prev, oi nb_eq, const oi* row_ptr,
                 const oi* col_ind, double* a, double* b) {
  int nen = (kind == ORC_LAPLACE3DT4 || kind == ORC_ELAS2DQ4) ? 4 : 3;
  int nee = nen * nb_dof;
  oi nnz = row_ptr[nb_eq];
  if (a) for (oi k = 0; k < nnz; ++k) a[k] = 0.0;
  if (b) for (oi k = 0; k < nb_eq; ++k) b[k] = 0.0;
  double x[12], emat[64], ea1[9], erhs[8];
  for (oi e = 0; e < nb_el; ++e) {
    const oi* en = elem + e * nen;
    for (int i = 0; i < nen; ++i)
      for (int c = 0; c < dim; ++c) x[dim * i + c] = coord[(en[i] - 1) * dim + c];
    double ce = coef ? coef[e] : -1.0;
    int rc = orc_element(kind, x, ce, prm, emat, ea1);
    if (rc != ORC_OK) return rc;
    /* BodyRHS */
    for (int k = 0; k < nee; ++k) erhs[k] = 0.0;
    if (f) {
      if (kind == ORC_LAPLACE2DT3) {
        double area, d[6];
        orc_triang3(x, &area, d);
        for (int i = 0; i < 3; i++) erhs[i] += f[en[i] - 1] * area * OFELI_THIRD; /* :139-143 */
      } else if (kind == ORC_DC2DT3 || kind == ORC_DC2DT3_BE) {
        double area, d[6];
        orc_triang3(x, &area, d);
        double c = OFELI_THIRD * area; /* DC2DT3.cpp:248-253 */
        for (int i = 0; i < 3; i++) erhs[i] += c * f[en[i] - 1];
      } else if (kind == ORC_LAPLACE3DT4) {
        double vol, d[12];
        orc_tetra4(x, &vol, d);
        for (int i = 0; i < 4; ++i) {
          if (coef) erhs[i] += f[en[i] - 1] * 0.25 * vol; /* DC3DT4.cpp:207-211 */
          else erhs[i] += 0.25 * vol * f[en[i] - 1];     /* Laplace3DT4.cpp:125-129 */
        }
      } else if (kind == ORC_ELAS2DQ4) {
        /* Body force at the 2x2 Gauss points, f = 2 values per node (fx, fy).
         * Equa_Solid::build (Equa_Solid.h:188-189) calls Elas2DQ4::BodyRHS()
         * (Elas2DQ4.cpp:224-237), whose loop reads _ebf(2j-1), _ebf(2j) for
         * j = 0..3: the pair of node j-1, and for j = 0 two entries before the
         * local vector (undefined behaviour).  Restated here is the intended
         * sum, as the one-argument BodyRHS(f) (Elas2DQ4.cpp:240-253) writes it:
         * node j's force with node j's shape function. */
        double sh[16], dsh[32], w[4];
        orc_quad4_atgauss(x, sh, dsh, w);
        for (int k = 0; k < 4; ++k) {
          double fx = 0., fy = 0.;
          for (int j = 0; j < 4; ++j) {
            fx += sh[4 * j + k] * f[(en[j] - 1) * 2];
            fy += sh[4 * j + k] * f[(en[j] - 1) * 2 + 1];
          }
          for (int i = 0; i < 4; ++i) {
            erhs[2 * i] += w[k] * fx * sh[4 * i + k];
            erhs[2 * i + 1] += w[k] * fy * sh[4 * i + k];
          }
        }
      }
    }
    /* updateBC, Equation_impl.h:160-176 */
    if (bc) {
      for (int in = 0; in < nee; ++in)
        for (int j = 0; j < nen; ++j)
          for (int l = 0; l < nb_dof; ++l) {
            oi nd = en[j] - 1;
            if (code[nd * nb_dof + l] > 0)
              erhs[in] -= emat[in * nee + j * nb_dof + l] * bc[nd * nb_dof + l];
          }
    }
    /* backward-Euler time term, Equa_Therm.h:203 */
    if (kind == ORC_DC2DT3_BE && u_prev) {
      double dti = 1. / prm[0], eu[3], m1[9];
      for (int i = 0; i < 3; ++i) eu[i] = u_prev[en[i] - 1];
      for (int k = 0; k < 9; ++k) m1[k] = dti * ea1[k];
      for (int i = 0; i < 3; ++i) {
        double s = 0.0;
        for (int j = 0; j < 3; ++j) s += m1[3 * i + j] * eu[j];
        erhs[i] = erhs[i] + s;
      }
    }
    /* Matrix::Assembly, Matrix_impl.h:187-200 */
    if (a) {
      int kk = 0;
      for (int i = 0; i < nen; ++i)
        for (int k = 0; k < nb_dof; ++k)
          for (int j = 0; j < nen; ++j)
            for (int l = 0; l < nb_dof; ++l, ++kk) {
              oi di = dof[(en[i] - 1) * nb_dof + k], dj = dof[(en[j] - 1) * nb_dof + l];
              if (di && dj) {
                oi p = col_pos(row_ptr, col_ind, di, dj);
                if (p < 0) return ORC_E_PATTERN;
                a[p] += emat[kk];
              }
            }
    }
    /* Vect::Assembly, Vect_impl.h:1522-1534 */
    if (b) {
      int ii = 0;
      for (int n = 0; n < nen; ++n)
        for (int k = 0; k < nb_dof; ++k, ++ii) {
          oi di = dof[(en[n] - 1) * nb_dof + k];
          if (di) b[di - 1] += erhs[ii];
        }
    }
  }
  return ORC_OK;
}

/* Constant traction per boundary side for Elas2DQ4: Line2 length
 * (Line2.cpp:82-93) and Elas2DQ4::BoundaryRHS() (Elas2DQ4.cpp:256-263),
 * assembled after the element loop (Equa_Solid.h:194-202).  traction holds
 * (tx,ty) per side; side_nodes are 1-based node ids; coord is 2D. */
int orc_traction_rhs(const double* coord, oi nb_sides, const oi* side_nodes,
                     const double* traction, const oi* dof, double* b) {
  for (oi s = 0; s < nb_sides; ++s) {
    oi n0 = side_nodes[2 * s] - 1, n1 = side_nodes[2 * s + 1] - 1;
    double dlx = 0.5 * (coord[2 * n1] - coord[2 * n0]);
    double dly = 0.5 * (coord[2 * n1 + 1] - coord[2 * n0 + 1]);
    double det = sqrt(dlx * dlx + dly * dly + 0.0 * 0.0);
    if (det == 0.0) return ORC_E_ZERODET;
    double len = 2 * det;
    double c = 0.5 * len;
    double srhs[4] = {0, 0, 0, 0};
    for (int i = 0; i < 2; ++i) {
      srhs[2 * i] += c * traction[2 * s];
      srhs[2 * i + 1] += c * traction[2 * s + 1];
    }
    oi nd[2] = {n0, n1};
    for (int i = 0; i < 2; ++i)
      for (int k = 0; k < 2; ++k) {
        oi di = dof[nd[i] * 2 + k];
        if (di) b[di - 1] += srhs[2 * i + k];
      }
  }
  return ORC_OK;
}

/* ======================================================================= */
/* Algebra                                                                 */
/* ======================================================================= */

/* SpMatrix::Mult/MultAdd, SpMatrix_impl.h:728-745: y = 0, then the row sums
 * accumulate left to right. */
void orc_spmv(oi n, const oi* row_ptr, const oi* col_ind, const double* a,
              const double* x, double* y) {
  oi l = 0;
  for (oi i = 0; i < n; ++i) {
    double yi = 0.0;
    for (oi j = 0; j < row_ptr[i + 1] - row_ptr[i]; ++j) {
      yi += a[row_ptr[i] + j] * x[col_ind[l++] - 1];
    }
    y[i] = yi;
  }
}

/* Vect::getNorm2, Vect_impl.h:660-668 */
double orc_norm2(oi n, const double* v) {
  double s = 0;
  for (oi i = 0; i < n; ++i) s += v[i] * v[i];
  return sqrt(s);
}

/* Vect::operator, (dot), Vect_impl.h:2218-2225 */
double orc_dot(oi n, const double* u, const double* v) {
  double p = 0;
  for (oi i = 0; i < n; i++) p += u[i] * v[i];
  return p;
}

/* Prec::setMatrix + inv_diag, Prec_impl.h:164-195, 306-316. */
/* 1-based position of the diagonal entry of each row (0: none, or a zero
 * diagonal value -> ORC_E_ZERODIAG), as SpMatrix::DILUFactorize's id[]. */
static int diag_ids(oi n, const oi* row_ptr, const oi* col_ind, const double* a, oi* id) {
  oi k = 0;
  for (oi i = 0; i < n; i++) {
    id[i] = 0;
    for (oi j = row_ptr[i]; j < row_ptr[i + 1]; j++, k++)
      if (col_ind[j] == i + 1) {
        id[i] = k + 1;
        if (a[k] == 0.0) return ORC_E_ZERODIAG;
      }
    if (id[i] == 0) return ORC_E_ZERODIAG;
  }
  return ORC_OK;
}

/* SpMatrix::DILUFactorize, SpMatrix_impl.h:834-870 (ILUFactorize is the same
 * routine, :873-878): with pivot_i = a_ii, for i = 1..n in turn,
 *   pivot_i <- 1 / pivot_i, then for every upper entry a_il of row i whose
 *   transposed entry a_li lies in the lower part of row l:
 *   pivot_l <- pivot_l - (a_li pivot_i) a_il,
 * so pivot_i = 1/D_i with D_l = a_ll - sum_{i<l} a_li a_il / D_i (D-ILU). */
int orc_dilu_factorize(oi n, const oi* row_ptr, const oi* col_ind, const double* a,
                       double* pivot) {
  oi* id = (oi*)malloc(sizeof(oi) * (n > 0 ? n : 1));
  if (!id) return ORC_E_NOMEM;
  int rc = diag_ids(n, row_ptr, col_ind, a, id);
  if (rc != ORC_OK) { free(id); return rc; }
  for (oi i = 0; i < n; i++) pivot[i] = a[id[i] - 1];
  for (oi i = 0; i < n; i++) {
    pivot[i] = 1.0 / pivot[i];
    for (oi j = id[i]; j < row_ptr[i + 1]; j++) {
      int found = 0;
      double c = 0.0;
      const oi l = col_ind[j] - 1;
      for (oi k = row_ptr[l]; k < id[l] - 1; k++)
        if (col_ind[k] == i + 1) found = 1, c = a[k];
      if (found) pivot[l] -= c * pivot[i] * a[j];
    }
  }
  free(id);
  return ORC_OK;
}

/* SpMatrix::DILUSolve, SpMatrix_impl.h:881-900: forward
 * z_i = pivot_i (b_i - sum_{j<=i} a_ij z_j) (the diagonal term meets z_i = 0),
 * then backward x_i = z_i - pivot_i sum_{j>i} a_ij x_j, i = n..1; that is
 * x = M^-1 b with M = (D+L) D^-1 (D+U). */
int orc_dilu_solve(oi n, const oi* row_ptr, const oi* col_ind, const double* a,
                   const double* pivot, const double* b, double* x) {
  oi* id = (oi*)malloc(sizeof(oi) * (n > 0 ? n : 1));
  double* z = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
  if (!id || !z) { free(id); free(z); return ORC_E_NOMEM; }
  int rc = diag_ids(n, row_ptr, col_ind, a, id);
  if (rc != ORC_OK) { free(id); free(z); return rc; }
  for (oi i = 0; i < n; i++) {
    double s = 0;
    for (oi j = row_ptr[i]; j < id[i]; j++) s += a[j] * z[col_ind[j] - 1];
    z[i] = pivot[i] * (b[i] - s);
  }
  for (oi i = n - 1; i >= 0; i--) {
    double s = 0;
    for (oi j = id[i]; j < row_ptr[i + 1]; j++) s += a[j] * x[col_ind[j] - 1];
    x[i] = z[i] - pivot[i] * s;
  }
  free(id);
  free(z);
  return ORC_OK;
}

int orc_prec_setup(oi n, const oi* row_ptr, const oi* col_ind, const double* a,
                   int prec, double* pivot) {
  if (prec == ORC_DILU || prec == ORC_ILU) return orc_dilu_factorize(n, row_ptr, col_ind, a, pivot);
  if (prec == ORC_DIAG) {
    for (oi i = 1; i <= n; i++) {
      oi p = col_pos(row_ptr, col_ind, i, i);
      double aii = p < 0 ? 0.0 : a[p];
      if (aii == 0.0) return ORC_E_ZERODIAG;
      pivot[i - 1] = 1.0 / aii;
    }
  } else if (prec != ORC_IDENT && prec != ORC_SSOR) {
    /* SSOR_PREC needs no setup; its zero-pivot check runs at every
     * application (SpMatrix_impl.h:917-928). */
    return ORC_E_ARG;
  }
  return ORC_OK;
}

/* SpMatrix::SSORSolve, SpMatrix_impl.h:913-942: forward (D+L)z = b, then
 * backward x_i = z_i - (sum_{j>i} a_ij x_j)/a_ii, i.e. symmetric
 * Gauss-Seidel with M = (D+L) D^-1 (D+U). */
int orc_ssor_solve(oi n, const oi* row_ptr, const oi* col_ind, const double* a,
                   const double* b, double* x) {
  oi* id = (oi*)malloc(sizeof(oi) * (n > 0 ? n : 1));
  double* z = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
  oi k = 0;
  for (oi i = 0; i < n; i++) {
    id[i] = 0;
    for (oi j = row_ptr[i]; j < row_ptr[i + 1]; j++, k++)
      if (col_ind[j] == i + 1) {
        id[i] = k + 1;
        if (a[k] == 0.0) { free(id); free(z); return ORC_E_ZERODIAG; }
      }
    if (id[i] == 0) { free(id); free(z); return ORC_E_ZERODIAG; }
  }
  for (oi i = 0; i < n; i++) {
    double s = 0;
    for (oi j = row_ptr[i]; j < id[i] - 1; j++) s += a[j] * z[col_ind[j] - 1];
    z[i] = (b[i] - s) / a[id[i] - 1];
  }
  for (oi i = n - 1; i >= 0; i--) {
    double s = 0;
    for (oi j = id[i]; j < row_ptr[i + 1]; j++) s += a[j] * x[col_ind[j] - 1];
    x[i] = z[i] - s / a[id[i] - 1];
  }
  free(id);
  free(z);
  return ORC_OK;
}

/* Prec::solve(b,x), Prec_impl.h:226-255. */
int orc_prec_apply(oi n, const oi* row_ptr, const oi* col_ind, const double* a,
                   int prec, const double* pivot, const double* b, double* x) {
  if (prec == ORC_IDENT) {
    for (oi i = 0; i < n; i++) x[i] = b[i];
  } else if (prec == ORC_DIAG) {
    for (oi i = 0; i < n; i++) x[i] = b[i] * pivot[i];
  } else if (prec == ORC_SSOR) {
    return orc_ssor_solve(n, row_ptr, col_ind, a, b, x);
  } else if (prec == ORC_DILU || prec == ORC_ILU) {
    return orc_dilu_solve(n, row_ptr, col_ind, a, pivot, b, x);
  } else {
    return ORC_E_ARG;
  }
  return ORC_OK;
}

#define ALLOCV(name) double* name = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double))

/* Preconditioned CG, CG.h:99-157 (with the wrapper CG.h:189-200 building the
 * preconditioner first).  Returns the iteration count, 0 if the initial
 * guess already satisfies the test, or -(max_it+1) without convergence (the
 * loop counter's value after the loop).  Stop test: ||r||/||b|| <= toler on
 * the recursive residual (||b||==0 -> 1).  Negative ORC_E_* values below
 * -1e6 flag setup errors. */
int orc_cg(oi n, const oi* row_ptr, const oi* col_ind, const double* a, int prec,
           const double* b, double* x, int max_it, double toler, double* res_out) {
  ALLOCV(pivot); ALLOCV(r); ALLOCV(p); ALLOCV(q); ALLOCV(z);
  int ret;
  int rc = orc_prec_setup(n, row_ptr, col_ind, a, prec, pivot);
  if (rc != ORC_OK) { ret = -1000000 + rc; goto done; }
  double res, nrm = orc_norm2(n, b);
  double rho = 0, rho_1 = 1.0, beta = 0;
  if (nrm == 0) nrm = 1;
  orc_spmv(n, row_ptr, col_ind, a, x, q);
  for (oi i = 0; i < n; i++) r[i] = b[i] - q[i];
  if ((res = orc_norm2(n, r) / nrm) <= toler) { ret = 0; if (res_out) *res_out = res; goto done; }
  int it;
  for (it = 1; it <= max_it; it++) {
    rc = orc_prec_apply(n, row_ptr, col_ind, a, prec, pivot, r, z);
    if (rc != ORC_OK) { ret = -1000000 + rc; goto done; }
    rho = orc_dot(n, r, z);
    if (it == 1) {
      for (oi i = 0; i < n; i++) p[i] = z[i];
    } else {
      beta = rho / rho_1;
      for (oi i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
    }
    orc_spmv(n, row_ptr, col_ind, a, p, q);
    double alpha = rho / orc_dot(n, p, q);
    for (oi i = 0; i < n; i++) x[i] += alpha * p[i];
    for (oi i = 0; i < n; i++) r[i] -= alpha * q[i];
    res = orc_norm2(n, r) / nrm;
    if (res <= toler) { ret = it; if (res_out) *res_out = res; goto done; }
    rho_1 = rho;
  }
  if (res_out) *res_out = res;
  ret = -it;
done:
  free(pivot); free(r); free(p); free(q); free(z);
  return ret;
}

/* Preconditioned BiCGStab, BiCGStab.h:87-162 (right preconditioning of p and
 * s; breakdown returns 2 for rho==0 and 3 for omega==0; strict '<' tests). */
int orc_bicgstab(oi n, const oi* row_ptr, const oi* col_ind, const double* a,
                 int prec, const double* b, double* x, int max_it, double toler,
                 double* res_out) {
  ALLOCV(pivot); ALLOCV(p); ALLOCV(pp); ALLOCV(s); ALLOCV(ss);
  ALLOCV(t); ALLOCV(v); ALLOCV(r); ALLOCV(rr);
  int ret, it;
  double rho_1 = 0, rho_2 = 0, alpha = 0, beta = 0, omega = 0, res;
  int rc = orc_prec_setup(n, row_ptr, col_ind, a, prec, pivot);
  if (rc != ORC_OK) { ret = -1000000 + rc; goto done; }
  double nrm = orc_norm2(n, b);
  orc_spmv(n, row_ptr, col_ind, a, x, t);
  for (oi i = 0; i < n; i++) r[i] = b[i] - t[i];
  for (oi i = 0; i < n; i++) rr[i] = r[i];
  if (nrm == 0.0) nrm = 1;
  if ((res = orc_norm2(n, r) / nrm) <= toler) { ret = 0; goto out; }
  for (it = 1; it <= max_it; it++) {
    rho_1 = orc_dot(n, rr, r);
    if (rho_1 == 0.0) { res = orc_norm2(n, r) / nrm; ret = 2; goto out; }
    if (it == 1) {
      for (oi i = 0; i < n; i++) p[i] = r[i];
    } else {
      beta = (rho_1 / rho_2) * (alpha / omega);
      for (oi i = 0; i < n; i++) p[i] = r[i] + beta * (p[i] - omega * v[i]);
    }
    rc = orc_prec_apply(n, row_ptr, col_ind, a, prec, pivot, p, pp);
    if (rc != ORC_OK) { ret = -1000000 + rc; goto done; }
    orc_spmv(n, row_ptr, col_ind, a, pp, v);
    alpha = rho_1 / orc_dot(n, rr, v);
    for (oi i = 0; i < n; i++) s[i] = r[i] - alpha * v[i];
    if ((res = orc_norm2(n, s) / nrm) < toler) {
      for (oi i = 0; i < n; i++) x[i] += alpha * pp[i];
      ret = it;
      goto out;
    }
    rc = orc_prec_apply(n, row_ptr, col_ind, a, prec, pivot, s, ss);
    if (rc != ORC_OK) { ret = -1000000 + rc; goto done; }
    orc_spmv(n, row_ptr, col_ind, a, ss, t);
    omega = orc_dot(n, t, s) / orc_dot(n, t, t);
    for (oi i = 0; i < n; i++) x[i] += alpha * pp[i] + omega * ss[i];
    for (oi i = 0; i < n; i++) r[i] = s[i] - omega * t[i];
    rho_2 = rho_1;
    if ((res = orc_norm2(n, r) / nrm) < toler) { ret = it; goto out; }
    if (omega == 0.0) { res = orc_norm2(n, r) / nrm; ret = 3; goto out; }
  }
  ret = -it;
out:
  if (res_out) *res_out = res;
done:
  free(pivot); free(p); free(pp); free(s); free(ss); free(t); free(v); free(r); free(rr);
  return ret;
}

/* SpMatrix::TMult, SpMatrix_impl.h:761-767, as BiCG.h:148 uses it: y = A^T x,
 * accumulated row by row (y_j += a_ij x_i for i = 1..n) into a cleared y.
 * The reference's loop indexes _col_ind[k-1] for entry k (the previous
 * entry's 1-based column used as a 0-based index, and _col_ind[-1] for the
 * first entry) and adds into y without clearing it; neither can be
 * reproduced, so the product its callers need is computed. */
void orc_spmv_t(oi n, const oi* row_ptr, const oi* col_ind, const double* a,
                const double* x, double* y) {
  for (oi j = 0; j < n; j++) y[j] = 0.0;
  for (oi i = 0; i < n; i++)
    for (oi k = row_ptr[i]; k < row_ptr[i + 1]; k++) y[col_ind[k] - 1] += a[k] * x[i];
}

/* Prec::TransSolve for ILU_PREC, Prec_impl.h:269-284: x = M^-T b with
 * M = (D+L) D^-1 (D+U) and pivot = 1/D, by the reference's column sweeps:
 *   i = 1..n:  z_i = t_i, then t_j -= (pivot_i z_i) a_ij for the upper a_ij;
 *   i = n..1:  x_i = pivot_i z_i, then z_j -= a_ij x_i for the lower a_ij,
 * with t = b.  The reference indexes Prec's own _row_ptr/_col_ind, which are
 * never filled (Prec.h:173); the matrix's arrays are what it means. */
int orc_dilu_trans_solve(oi n, const oi* row_ptr, const oi* col_ind, const double* a,
                         const double* pivot, const double* b, double* x) {
  oi* id = (oi*)malloc(sizeof(oi) * (n > 0 ? n : 1));
  double* z = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
  double* t = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
  if (!id || !z || !t) { free(id); free(z); free(t); return ORC_E_NOMEM; }
  int rc = diag_ids(n, row_ptr, col_ind, a, id);
  if (rc != ORC_OK) { free(id); free(z); free(t); return rc; }
  for (oi i = 0; i < n; i++) t[i] = b[i];
  for (oi i = 0; i < n; i++) {
    z[i] = t[i];
    const double tmp = pivot[i] * z[i];
    for (oi j = id[i]; j < row_ptr[i + 1]; j++) t[col_ind[j] - 1] -= tmp * a[j];
  }
  for (oi i = n - 1; i >= 0; i--) {
    x[i] = pivot[i] * z[i];
    for (oi j = row_ptr[i]; j < id[i] - 1; j++) z[col_ind[j] - 1] -= a[j] * x[i];
  }
  free(id);
  free(z);
  free(t);
  return ORC_OK;
}

/* Prec::TransSolve(b,x), Prec_impl.h:259-296: x = b, then DIAG_PREC scales
 * by the pivots and ILU_PREC solves with M^T; IDENT, DILU and SSOR have no
 * case in the switch and leave x = b. */
int orc_prec_trans_apply(oi n, const oi* row_ptr, const oi* col_ind, const double* a,
                         int prec, const double* pivot, const double* b, double* x) {
  if (prec == ORC_DIAG) {
    for (oi i = 0; i < n; i++) x[i] = b[i] * pivot[i];
  } else if (prec == ORC_ILU) {
    return orc_dilu_trans_solve(n, row_ptr, col_ind, a, pivot, b, x);
  } else if (prec == ORC_IDENT || prec == ORC_DILU || prec == ORC_SSOR) {
    for (oi i = 0; i < n; i++) x[i] = b[i];
  } else {
    return ORC_E_ARG;
  }
  return ORC_OK;
}

/* CGS, CGS.h:91-171, as written: the vector q is declared but never
 * assigned, so it stays zero: from the second iteration u = r + beta*0 and
 * p = u + beta*(0 + beta*p).  Returns 0 (converged at start), it
 * (||r||/||b|| < toler after iteration it), 2 when (s,r) = 0 (whatever the
 * iteration), or -max_it. */
int orc_cgs(oi n, const oi* row_ptr, const oi* col_ind, const double* a, int prec,
            const double* b, double* x, int max_it, double toler, double* res_out) {
  ALLOCV(pivot); ALLOCV(r); ALLOCV(p); ALLOCV(y); ALLOCV(q); ALLOCV(z); ALLOCV(v);
  ALLOCV(u); ALLOCV(w); ALLOCV(s); ALLOCV(tmp);
  int ret, it;
  double rho_1 = 0, rho_2 = 0, alpha = 0, beta = 0, res;
  int rc = orc_prec_setup(n, row_ptr, col_ind, a, prec, pivot);
  if (rc != ORC_OK) { ret = -1000000 + rc; goto done; }
  double nrm = orc_norm2(n, b);
  orc_spmv(n, row_ptr, col_ind, a, x, tmp);
  for (oi i = 0; i < n; i++) s[i] = b[i] - tmp[i];
  for (oi i = 0; i < n; i++) r[i] = s[i];
  if (nrm == 0) nrm = 1;
  if ((res = orc_norm2(n, r) / nrm) <= toler) { ret = 0; goto out; }
  for (it = 1; it <= max_it; it++) {
    rho_1 = orc_dot(n, s, r);
    if (rho_1 == 0.0) { res = orc_norm2(n, r) / nrm; ret = 2; goto out; }
    if (it == 1) {
      for (oi i = 0; i < n; i++) u[i] = r[i];
      for (oi i = 0; i < n; i++) p[i] = u[i];
    } else {
      beta = rho_1 / rho_2;
      for (oi i = 0; i < n; i++) u[i] = r[i] + beta * q[i];
      for (oi i = 0; i < n; i++) p[i] = u[i] + beta * (q[i] + beta * p[i]);
    }
    rc = orc_prec_apply(n, row_ptr, col_ind, a, prec, pivot, p, y);  /* y = p; P.solve(y) */
    if (rc != ORC_OK) { ret = -1000000 + rc; goto done; }
    orc_spmv(n, row_ptr, col_ind, a, y, v);
    alpha = rho_1 / orc_dot(n, s, v);
    for (oi i = 0; i < n; i++) tmp[i] = u[i] - alpha * v[i];
    for (oi i = 0; i < n; i++) tmp[i] += u[i];
    rc = orc_prec_apply(n, row_ptr, col_ind, a, prec, pivot, tmp, w);  /* P.solve(w) */
    if (rc != ORC_OK) { ret = -1000000 + rc; goto done; }
    for (oi i = 0; i < n; i++) x[i] += alpha * w[i];
    orc_spmv(n, row_ptr, col_ind, a, w, z);
    for (oi i = 0; i < n; i++) r[i] += -alpha * z[i];
    rho_2 = rho_1;
    if ((res = orc_norm2(n, r) / nrm) < toler) { ret = it; goto out; }
  }
  ret = -max_it;
out:
  if (res_out) *res_out = res;
done:
  free(pivot); free(r); free(p); free(y); free(q); free(z); free(v); free(u); free(w);
  free(s); free(tmp);
  return ret;
}

/* BiCG, BiCG.h:94-170, with the transposed products it means (orc_spmv_t,
 * orc_prec_trans_apply): z = P r, zz = P^T rr, rho = (z,rr); p = z + beta p,
 * pp = zz + beta pp; q = A p, qq = A^T pp; alpha = rho/(pp,q);
 * x += alpha p, r -= alpha q, rr -= alpha qq.  The in-loop stop test divides
 * by ||b|| itself (BiCG.h:158), the start test by ||b|| or 1.  Returns 0, it
 * (converged, or rho = 0 at iteration it), or -(max_it+1). */
int orc_bicg(oi n, const oi* row_ptr, const oi* col_ind, const double* a, int prec,
             const double* b, double* x, int max_it, double toler, double* res_out) {
  ALLOCV(pivot); ALLOCV(r); ALLOCV(z); ALLOCV(zz); ALLOCV(p); ALLOCV(pp); ALLOCV(q);
  ALLOCV(qq); ALLOCV(rr);
  int ret, it;
  double rho_1 = 0, rho_2 = 0, alpha = 0, beta = 0, res;
  int rc = orc_prec_setup(n, row_ptr, col_ind, a, prec, pivot);
  if (rc != ORC_OK) { ret = -1000000 + rc; goto done; }
  const double bn = orc_norm2(n, b);
  double nrm = bn;
  orc_spmv(n, row_ptr, col_ind, a, x, r);
  for (oi i = 0; i < n; i++) r[i] = b[i] - r[i];
  for (oi i = 0; i < n; i++) rr[i] = r[i];
  if (nrm == 0.0) nrm = 1;
  if ((res = orc_norm2(n, r) / nrm) <= toler) { ret = 0; goto out; }
  for (it = 1; it <= max_it; it++) {
    rc = orc_prec_apply(n, row_ptr, col_ind, a, prec, pivot, r, z);
    if (rc == ORC_OK) rc = orc_prec_trans_apply(n, row_ptr, col_ind, a, prec, pivot, rr, zz);
    if (rc != ORC_OK) { ret = -1000000 + rc; goto done; }
    rho_1 = orc_dot(n, z, rr);
    if (rho_1 == 0.0) { res = orc_norm2(n, r) / nrm; ret = it; goto out; }
    if (it == 1) {
      for (oi i = 0; i < n; i++) p[i] = z[i];
      for (oi i = 0; i < n; i++) pp[i] = zz[i];
    } else {
      beta = rho_1 / rho_2;
      for (oi i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
      for (oi i = 0; i < n; i++) pp[i] = zz[i] + beta * pp[i];
    }
    orc_spmv(n, row_ptr, col_ind, a, p, q);
    orc_spmv_t(n, row_ptr, col_ind, a, pp, qq);
    alpha = rho_1 / orc_dot(n, pp, q);
    for (oi i = 0; i < n; i++) x[i] += alpha * p[i];
    for (oi i = 0; i < n; i++) r[i] += -alpha * q[i];
    for (oi i = 0; i < n; i++) rr[i] += -alpha * qq[i];
    rho_2 = rho_1;
    if ((res = orc_norm2(n, r) / bn) < toler) { ret = it; goto out; }
  }
  ret = -it;
out:
  if (res_out) *res_out = res;
done:
  free(pivot); free(r); free(z); free(zz); free(p); free(pp); free(q); free(qq); free(rr);
  return ret;
}

/* Givens helpers, GMRes.h:67-115. */
static void g_rotate(double* dx, double* dy, double cs, double sn) {
  double temp = cs * *dx + sn * *dy;
  *dy = cs * *dy - sn * *dx;
  *dx = temp;
}
static void g_set_rotation(double dx, double dy, double* cs, double* sn) {
  if (dy == 0.0) { *cs = 1.0; *sn = 0.0; }
  else if (fabs(dy) > fabs(dx)) {
    double temp = dx / dy;
    *sn = 1.0 / sqrt(1.0 + temp * temp);
    *cs = temp * *sn;
  } else {
    double temp = dy / dx;
    *cs = 1.0 / sqrt(1.0 + temp * temp);
    *sn = temp * *cs;
  }
}

/* Left-preconditioned restarted GMRES(m), GMRes.h:148-241, with update()
 * (:78-94).  H is (m+1)x(m+1), H(i,j) 1-based stored at H[(i-1)*(m+1)+(j-1)].
 * As in the reference, a run that is already converged returns 1, and the
 * end-of-cycle update always uses k = m-1. */
int orc_gmres(oi n, const oi* row_ptr, const oi* col_ind, const double* a,
              int prec, const double* b, double* x, int m, int max_it,
              double toler, double* res_out) {
  int M1 = m + 1;
  double* H = (double*)calloc((size_t)M1 * M1, sizeof(double));
  double* s = (double*)calloc(M1, sizeof(double));
  double* cs = (double*)calloc(M1, sizeof(double));
  double* sn = (double*)calloc(M1, sizeof(double));
  double* y = (double*)calloc(M1, sizeof(double));
  double* V = (double*)calloc((size_t)M1 * (n > 0 ? n : 1), sizeof(double));
  ALLOCV(pivot); ALLOCV(w); ALLOCV(r); ALLOCV(tmp);
#define HH(i, j) H[((i)-1) * M1 + ((j)-1)]
  int ret, it = 1;
  double res = 0;
  int rc = orc_prec_setup(n, row_ptr, col_ind, a, prec, pivot);
  if (rc != ORC_OK) { ret = -1000000 + rc; goto done; }
  double normb = orc_norm2(n, b);
  orc_spmv(n, row_ptr, col_ind, a, x, tmp);
  for (oi i = 0; i < n; i++) r[i] = b[i] - tmp[i];
  memcpy(tmp, r, sizeof(double) * n);
  rc = orc_prec_apply(n, row_ptr, col_ind, a, prec, pivot, tmp, r);
  if (rc != ORC_OK) { ret = -1000000 + rc; goto done; }
  double beta = orc_norm2(n, r);
  if (normb == 0.0) normb = 1;
  if ((res = orc_norm2(n, r) / normb) <= toler) { ret = it; goto out; }
  while (it <= max_it) {
    double sc = 1.0 / beta;
    for (oi l = 0; l < n; l++) V[l] = sc * r[l];
    for (int k = 0; k < M1; ++k) s[k] = 0.0;
    s[0] = beta;
    for (int i = 0; i < m && it <= max_it; i++, it++) {
      orc_spmv(n, row_ptr, col_ind, a, V + (oi)i * n, tmp);
      orc_prec_apply(n, row_ptr, col_ind, a, prec, pivot, tmp, w);
      for (int k = 0; k <= i; k++) {
        HH(k + 1, i + 1) = orc_dot(n, w, V + (oi)k * n);
        for (oi j = 0; j < n; j++) w[j] -= HH(k + 1, i + 1) * V[(oi)k * n + j];
      }
      HH(i + 2, i + 1) = orc_norm2(n, w);
      double sc2 = 1.0 / HH(i + 2, i + 1);
      for (oi l = 0; l < n; l++) V[(oi)(i + 1) * n + l] = sc2 * w[l];
      for (int k = 0; k < i; k++) g_rotate(&HH(k + 1, i + 1), &HH(k + 2, i + 1), cs[k], sn[k]);
      g_set_rotation(HH(i + 1, i + 1), HH(i + 2, i + 1), &cs[i], &sn[i]);
      g_rotate(&HH(i + 1, i + 1), &HH(i + 2, i + 1), cs[i], sn[i]);
      g_rotate(&s[i], &s[i + 1], cs[i], sn[i]);
      if ((res = fabs(s[i + 1]) / normb) < toler) {
        /* update(x, i, H, s, v) */
        for (int q = 0; q < M1; ++q) y[q] = s[q];
        for (int ii = i; ii >= 0; ii--) {
          y[ii] /= HH(ii + 1, ii + 1);
          for (int jj = ii - 1; jj >= 0; jj--) y[jj] -= HH(jj + 1, ii + 1) * y[ii];
        }
        for (int jj = 0; jj <= i; jj++)
          for (oi l = 0; l < n; l++) x[l] += V[(oi)jj * n + l] * y[jj];
        ret = it;
        goto out;
      }
    }
    /* update(x, m-1, H, s, v) */
    for (int q = 0; q < M1; ++q) y[q] = s[q];
    for (int ii = m - 1; ii >= 0; ii--) {
      y[ii] /= HH(ii + 1, ii + 1);
      for (int jj = ii - 1; jj >= 0; jj--) y[jj] -= HH(jj + 1, ii + 1) * y[ii];
    }
    for (int jj = 0; jj <= m - 1; jj++)
      for (oi l = 0; l < n; l++) x[l] += V[(oi)jj * n + l] * y[jj];
    orc_spmv(n, row_ptr, col_ind, a, x, tmp);
    for (oi i = 0; i < n; i++) tmp[i] = b[i] - tmp[i];
    orc_prec_apply(n, row_ptr, col_ind, a, prec, pivot, tmp, r);
    beta = orc_norm2(n, r);
    if ((res = beta / normb) < toler) { ret = it; goto out; }
  }
  ret = -it;
out:
  if (res_out) *res_out = res;
done:
#undef HH
  free(H); free(s); free(cs); free(sn); free(y); free(V);
  free(pivot); free(w); free(r); free(tmp);
  return ret;
}

/* Vect::insertBC(mesh, v, bc, dof=0), Vect_impl.h:1097-1116: free DOFs take
 * successive entries of v, imposed ones the bc value. */
void orc_insert_bc(oi nb_nodes, int nb_dof, const int* code, const double* v,
                   const double* bc, double* u) {
  oi i = 0, j = 0;
  for (oi n = 0; n < nb_nodes; ++n)
    for (int k = 0; k < nb_dof; ++k, i++) {
      if (code[n * nb_dof + k] == 0) u[i] = v[j++];
      else u[i] = bc ? bc[i] : 0.0;
    }
}
