/*
 * ofeli_oracle.h -- CPU restatement of OFELI 5.3.3's sparse linear-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product library (ofeli_amd/libofeli_hip.so) never links, loads or calls it.
 *
 * Conventions follow the reference (file:line citations are relative to the
 * OFELI source tree):
 *   - node ids and DOF labels are 1-based, DOF label 0 = eliminated
 *     (Mesh.cpp:1499-1509);
 *   - CSR row_ptr is 0-based, col_ind is 1-based, sorted ascending per row
 *     (GraphOfMatrix.cpp:619-634);
 *   - all arithmetic is FP64 in the reference's operation order.
 * Index type is int64 (the reference uses size_t).
 */
#ifndef OFELI_ORACLE_H
#define OFELI_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int64_t oi;

/* Element kinds for orc_assemble. */
enum {
  ORC_LAPLACE2DT3 = 1,  /* Laplace2DT3: LHS + BodyRHS(f) + updateBC          */
  ORC_DC2DT3      = 2,  /* DC2DT3 steady Diffusion(k_e) + BodyRHS(f)         */
  ORC_DC2DT3_BE   = 3,  /* DC2DT3 backward Euler: LCapacity + Diffusion      */
  ORC_LAPLACE3DT4 = 4,  /* Laplace3DT4 LHS (k==NULL) / DC3DT4 Diffusion(k_e) */
  ORC_ELAS2DQ4    = 5   /* Elas2DQ4 Deviator + Dilatation                    */
};

/* Preconditioners, numbered as the reference's Preconditioner enum
 * (Matrix.h:84-93). */
enum { ORC_IDENT = 1, ORC_DIAG = 2, ORC_DILU = 3, ORC_ILU = 4, ORC_SSOR = 5 };

/* Error codes (the reference throws OFELIException in these cases). */
enum {
  ORC_OK = 0,
  ORC_E_NEGDET = -1,     /* Triang3/Tetra4/Quad4: negative determinant      */
  ORC_E_ZERODET = -2,    /* null determinant                                */
  ORC_E_ZERODIAG = -3,   /* Prec::inv_diag / SSORSolve zero pivot           */
  ORC_E_PATTERN = -4,    /* entry not in the sparse pattern                 */
  ORC_E_ARG = -5,
  ORC_E_NOMEM = -6
};

/* ---- meshes (L1) ---- */
oi  orc_mesh2d_nb_nodes(oi nx, oi ny);
oi  orc_mesh2d_nb_elements(oi nx, oi ny, int quad);
void orc_mesh2d(double xmin, double xmax, double ymin, double ymax, oi nx, oi ny,
                int cx0, int cxN, int cy0, int cyN, int quad, int nb_dof,
                double* coord, int* code, oi* elem);
void orc_mesh2d_sides(oi nx, oi ny, int cx0, int cxN, int cy0, int cyN,
                      oi* side_nodes, int* side_code, oi* nb_sides);
oi  orc_mesh3d_nb_nodes(oi nx, oi ny, oi nz);
void orc_mesh3d_kuhn(double xmin, double xmax, double ymin, double ymax,
                     double zmin, double zmax, oi nx, oi ny, oi nz,
                     int cx0, int cxN, int cy0, int cyN, int cz0, int czN,
                     double* coord, int* code, oi* elem);
oi  orc_number_equations(oi nb_nodes, int nb_dof, const int* code, oi* dof);

/* ---- sparse pattern (L3) ---- */
oi  orc_node_graph(oi nb_el, int nen, int nb_dof, const oi* elem, const oi* dof,
                   oi nb_eq, oi* row_ptr, oi** col_ind);
oi  orc_node_graph_scal(oi nb_el, int nen, const oi* elem, oi nb_nodes, oi* row_ptr,
                        oi** col_ind);
void orc_free(void* p);

/* ---- shape functions (L2) ---- */
int orc_triang3(const double* x, double* area, double* dsh);
int orc_tetra4(const double* x, double* vol, double* dsh);
void orc_gauss(int np, double* x, double* w);
int orc_quad4_atgauss(const double* x, double* sh, double* dsh, double* w);

/* ---- element matrices (L5) ---- */
int orc_element(int kind, const double* x, double coef, const double* prm,
                double* emat, double* ea1);

/* ---- assembly (L3/L5) ---- */
int orc_assemble(int kind, int dim, const double* coord, oi nb_el, const oi* elem,
                 int nb_dof, const oi* dof, const int* code, const double* bc,
                 const double* f, const double* coef, const double* prm,
                 const double* u_prev, oi nb_eq, const oi* row_ptr,
                 const oi* col_ind, double* a, double* b);
int orc_traction_rhs(const double* coord, oi nb_sides, const oi* side_nodes,
                     const double* traction, const oi* dof, double* b);

/* ---- algebra (L3/L4) ---- */
void orc_spmv(oi n, const oi* row_ptr, const oi* col_ind, const double* a,
              const double* x, double* y);
double orc_norm2(oi n, const double* v);
double orc_dot(oi n, const double* u, const double* v);
int orc_prec_setup(oi n, const oi* row_ptr, const oi* col_ind, const double* a,
                   int prec, double* pivot);
int orc_prec_apply(oi n, const oi* row_ptr, const oi* col_ind, const double* a,
                   int prec, const double* pivot, const double* b, double* x);
/* DILU_PREC / ILU_PREC (the reference's ILU routines are its DILU ones). */
int orc_dilu_factorize(oi n, const oi* row_ptr, const oi* col_ind, const double* a,
                       double* pivot);
int orc_dilu_solve(oi n, const oi* row_ptr, const oi* col_ind, const double* a,
                   const double* pivot, const double* b, double* x);
int orc_ssor_solve(oi n, const oi* row_ptr, const oi* col_ind, const double* a,
                   const double* b, double* x);
/* y = A^T x (SpMatrix::TMult as its callers intend) */
void orc_spmv_t(oi n, const oi* row_ptr, const oi* col_ind, const double* a,
                const double* x, double* y);
/* Prec::TransSolve(b,x): DIAG scales, ILU solves M^T x = b, others copy */
int orc_dilu_trans_solve(oi n, const oi* row_ptr, const oi* col_ind, const double* a,
                         const double* pivot, const double* b, double* x);
int orc_prec_trans_apply(oi n, const oi* row_ptr, const oi* col_ind, const double* a,
                         int prec, const double* pivot, const double* b, double* x);
int orc_cgs(oi n, const oi* row_ptr, const oi* col_ind, const double* a, int prec,
            const double* b, double* x, int max_it, double toler, double* res_out);
int orc_bicg(oi n, const oi* row_ptr, const oi* col_ind, const double* a, int prec,
             const double* b, double* x, int max_it, double toler, double* res_out);
int orc_cg(oi n, const oi* row_ptr, const oi* col_ind, const double* a, int prec,
           const double* b, double* x, int max_it, double toler, double* res_out);
int orc_bicgstab(oi n, const oi* row_ptr, const oi* col_ind, const double* a,
                 int prec, const double* b, double* x, int max_it, double toler,
                 double* res_out);
int orc_gmres(oi n, const oi* row_ptr, const oi* col_ind, const double* a,
              int prec, const double* b, double* x, int m, int max_it,
              double toler, double* res_out);
void orc_insert_bc(oi nb_nodes, int nb_dof, const int* code, const double* v,
                   const double* bc, double* u);

#ifdef __cplusplus
}
#endif
#endif
