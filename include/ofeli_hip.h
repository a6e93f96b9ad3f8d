/*
 * ofeli_hip.h -- C-ABI of libofeli_hip, the MI355X (gfx950) implementation of
 * OFELI's sparse linear-solve path.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes.  The
 * C++ drop-in classes (OFELI::SpMatrix, Vect, Prec, LinearSolver, the equation
 * classes) sit on top of these entry points; each entry point names the
 * reference function it replaces (paths relative to the OFELI 5.3.3 tree).
 *
 * Memory: arguments named d_* are device pointers (hipMalloc'd, or any device
 * allocation on the context's device); h_* are host pointers.  All device
 * work is ordered on the context's stream.  Entry points that return host
 * results (iteration counts, norms) synchronise that stream before returning.
 *
 * Status: every function returns OFH_OK (0) or a negative OFH_E_* code; the
 * message of the last failure on the calling thread is ofh_last_error().
 * Solvers report the reference's return convention separately in
 * ofh_solve_info.nb_it (CG.h:154-156, BiCGStab.h:117-157, GMRes.h:236-240).
 */
#ifndef OFELI_HIP_H
#define OFELI_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define OFH_OK 0
#define OFH_E_INVALID (-1)     /* bad argument / size / handle               */
#define OFH_E_HIP (-2)         /* HIP runtime failure                        */
#define OFH_E_ZERO_PIVOT (-3)  /* Prec::inv_diag / SSORSolve zero diagonal   */
#define OFH_E_NEGDET (-4)      /* Triang3/Tetra4/Quad4 negative determinant  */
#define OFH_E_PATTERN (-5)     /* entry outside the sparse pattern / empty row */
#define OFH_E_NOMEM (-6)
#define OFH_E_UNSUPPORTED (-7) /* DIRECT_SOLVER, not provided by this path */
#define OFH_E_ZERODET (-8)     /* null determinant                           */

/* Iteration and Preconditioner enums: same values as Matrix.h:75-93. */
#define OFH_DIRECT_SOLVER 1
#define OFH_CG_SOLVER 2
#define OFH_CGS_SOLVER 3
#define OFH_BICG_SOLVER 4
#define OFH_BICG_STAB_SOLVER 5
#define OFH_GMRES_SOLVER 6
#define OFH_IDENT_PREC 1
#define OFH_DIAG_PREC 2
#define OFH_DILU_PREC 3
#define OFH_ILU_PREC 4
#define OFH_SSOR_PREC 5

/* Element kinds of the assembly plan (equation classes' build()). */
#define OFH_LAPLACE2DT3 1 /* Laplace2DT3.cpp:93-143                          */
#define OFH_DC2DT3 2      /* DC2DT3 steady diffusion, per-element k          */
#define OFH_DC2DT3_BE 3   /* DC2DT3 backward Euler (Equa_Therm.h:162-218)    */
#define OFH_LAPLACE3DT4 4 /* Laplace3DT4.cpp:76-129 / DC3DT4 Diffusion(k_e)  */
#define OFH_ELAS2DQ4 5    /* Elas2DQ4.cpp:76-221                             */

typedef struct ofh_ctx_s* ofh_ctx;
typedef struct ofh_csr_s* ofh_csr;

typedef struct ofh_solve_info {
  int nb_it;          /* reference convention: it>0 converged after it steps,
                         0 converged at start, -(max_it+1) not converged,
                         2/3 BiCGStab breakdowns                             */
  int iterations;     /* iterations actually performed                       */
  double relres;      /* final ||r||/||b|| as the reference measures it      */
  double true_relres; /* ||b-Ax||/||b|| recomputed in FP64 at the end        */
  int outer;          /* mixed precision: FP64 residual replacements          */
  int fallback;       /* mixed precision: 1 if the FP64 fallback ran          */
  double seconds;     /* wall time of the solve (host clock, synchronised)    */
} ofh_solve_info;

/* ---------------------------------------------------------------- context */
const char* ofh_last_error(void);
int ofh_ctx_create(int device, ofh_ctx* out);
int ofh_ctx_destroy(ofh_ctx ctx);
/* Order all further work on an existing hipStream_t (NULL is the null /
 * default stream); the context's own stream is released. */
int ofh_ctx_set_stream(ofh_ctx ctx, void* hip_stream);
int ofh_ctx_synchronize(ofh_ctx ctx);
/* Solvers replay chunks of iterations from HIP graphs captured per argument
 * set and cached in the context (enable = 1), or launch every kernel
 * directly (enable = 0, the default: measured faster on MI355X).  Results
 * are identical either way. */
int ofh_ctx_use_graphs(ofh_ctx ctx, int enable);
/* Form of ofh_cg_mixed's FP32 inner iteration (no reference counterpart; a
 * tuning and test control): 0 (default) chooses by size, 1 takes two kernels
 * per inner iteration (direction + product, update), 2 takes one kernel (the
 * single-reduction iteration) wherever it applies: IDENT_PREC or DIAG_PREC
 * on a matrix with an FP32 diagonal copy whose values are symmetric bit for
 * bit (its (p, Ap) uses the symmetry).  By default the one-kernel form is
 * taken there up to 300,000 rows with 16 slots per row of the copy and
 * 650,000 rows with 8 or 4 (measured crossovers on MI355X).  Both
 * forms meet the same FP64-verified tolerance; their iterates differ in
 * rounding. */
int ofh_ctx_mixed_form(ofh_ctx ctx, int form);
/* Per-kernel timing (no reference counterpart; the reference's Timer
 * measures whole calls with clock(), Timer.h:93-129).  enable = 1 clears
 * earlier records and, until disabled, gives each launch of the kernels
 * below (made by ofh_cg, ofh_cg_mixed, ofh_spmv_f64 and ofh_spmv_f32) a
 * start and a stop event attached to its own dispatch, so each pair times
 * the kernel's execution; launches replayed from captured graphs are not
 * timed.  ofh_ctx_profile_read copies the durations (ms) of up to cap
 * recorded launches of one kind to h_ms and the number recorded to *count;
 * launches the host enqueued past the solver's stop point are recorded too
 * (they return at once). */
#define OFH_PROF_MX_DIRSPMV 0 /* mixed CG: direction + FP32 product        */
#define OFH_PROF_MX_UPD 1     /* mixed CG: x, r, z update + dot products   */
#define OFH_PROF_CG_DIRSPMV 2 /* FP64 CG, two-kernel form: direction + A p */
#define OFH_PROF_CG_UPD 3     /* FP64 CG: x, r, z update + dot products    */
#define OFH_PROF_CG_SPMV 4    /* FP64 CG, three-kernel form: q = A p       */
#define OFH_PROF_CG_DIR 5     /* FP64 CG, three-kernel form: direction     */
#define OFH_PROF_SPMV64 6     /* ofh_spmv_f64 from the SELL-64 copy        */
#define OFH_PROF_SPMV32 7     /* ofh_spmv_f32 from the SELL-64 copy        */
#define OFH_PROF_MX_ITER 8    /* mixed CG, one-kernel inner iteration      */
int ofh_ctx_profile(ofh_ctx ctx, int enable);
int ofh_ctx_profile_read(ofh_ctx ctx, int kind, double* h_ms, int cap, int* count);
int ofh_malloc(ofh_ctx ctx, size_t bytes, void** d_ptr);
int ofh_free(ofh_ctx ctx, void* d_ptr);
int ofh_memcpy_h2d(ofh_ctx ctx, void* d_dst, const void* h_src, size_t bytes);
int ofh_memcpy_d2h(ofh_ctx ctx, void* h_dst, const void* d_src, size_t bytes);
int ofh_memcpy_d2d(ofh_ctx ctx, void* d_dst, const void* d_src, size_t bytes);

/* ------------------------------------------------------------ SpMatrix CSR */
/* SpMatrix(row_ptr,col_ind,a) (SpMatrix.h:176-210): host arrays, row_ptr
 * 0-based (n+1), col 0-based (the reference stores col_ind 1-based; the C++
 * layer converts), sorted ascending in each row.  h_val may be NULL. */
int ofh_csr_create(ofh_ctx ctx, int n, int64_t nnz, const int32_t* h_row_ptr,
                   const int32_t* h_col, const double* h_val, ofh_csr* out);
int ofh_csr_destroy(ofh_csr A);
int ofh_csr_info(ofh_csr A, int* n, int64_t* nnz, int* max_row_nnz);
int ofh_csr_get_pattern(ofh_csr A, int32_t* h_row_ptr, int32_t* h_col);
/* Values: SpMatrix::get()/set (SpMatrix_impl.h:980-1005) as bulk copies. */
int ofh_csr_set_values(ofh_csr A, const double* h_val);
int ofh_csr_set_values_device(ofh_csr A, const double* d_val);
int ofh_csr_get_values(ofh_csr A, double* h_val);
int ofh_csr_values_device(ofh_csr A, double** d_val);

/* Column-index storage of the SpMV stream (no reference counterpart: an
 * internal format choice).  enable = 1 streams 16-bit offsets from the
 * first row of each 256-row block instead of 32-bit columns when every
 * column fits (otherwise the call keeps 32-bit columns); 0 (the default)
 * streams 32-bit columns.  Results are identical either way. */
int ofh_csr_compress_columns(ofh_csr A, int enable);

/* The SELL-64 copy that y = A x and the Krylov products stream (no reference
 * counterpart: an internal format, reported for measurement).  Builds the
 * copy if the matrix has none yet; *in_use = 0 when its padding would exceed
 * a quarter of nnz (the CSR-stream kernels run), else 1 with *stored = the
 * stored entries (nnz + padding) and *col_bytes = 2 (16-bit offsets from the
 * chunk's first row) or 4 (32-bit columns). */
int ofh_csr_sell_info(ofh_csr A, int* in_use, int64_t* stored, int* col_bytes);

/* The diagonal copies of the values (no reference counterpart: an internal
 * format, reported for measurement): where every entry lies on at most 15
 * diagonals (column - row offsets) and each row's columns ascend, row i is
 * stored as 4, 8 or 16 slots (the values of the diagonals and a presence
 * mask) and no column index.  The FP32 copy (read by ofh_spmv_f32 and the
 * mixed CG's inner product) is used when it takes no more bytes than the
 * SELL-64 FP32 copy with 16-bit columns, the FP64 copy (ofh_spmv_f64, the
 * FP64 CG products) when it takes no more than the SELL-64 FP64 copy.
 * Builds what the matrix has not yet; *in_use32 / *in_use64 = 1 for a copy in
 * use, *diagonals and *row_slots describe the layout (0 when the pattern does
 * not fit); otherwise the SELL-64 or CSR-stream kernels run.  *symmetric = 1
 * when the diagonals pair up (+-o) and the values are symmetric bit for bit
 * (a_ij == a_ji, checked on the GPU for each set of values): the products
 * then read the symmetric form, row_slots / 2 slots per row (the diagonal
 * and the upper values; a lower value is read from its mirror) and a 16-bit
 * presence mask per row, with the same row sums bit for bit. */
int ofh_csr_dia_info(ofh_csr A, int* in_use32, int* in_use64, int* diagonals, int* row_slots,
                     int* symmetric);

/* y = A x: SpMatrix::Mult (SpMatrix_impl.h:728-734). */
int ofh_spmv_f64(ofh_csr A, const double* d_x, double* d_y);
/* y += a A x: SpMatrix::MultAdd (SpMatrix_impl.h:737-757). */
int ofh_spmv_add_f64(ofh_csr A, double a, const double* d_x, double* d_y);
/* y = A^T x: SpMatrix::TMult (SpMatrix_impl.h:761-767) as BiCG uses it.
 * y is overwritten (the reference adds into y and indexes the previous
 * entry's column, which is not reproduced).  A CSR copy of A^T is built on
 * the first call and its values re-gathered after every value change; each
 * y_j sums a_ij x_i in ascending i. */
int ofh_spmv_t_f64(ofh_csr A, const double* d_x, double* d_y);
/* FP32 SpMV on the FP32 mirror of the values (mixed-precision inner loop). */
int ofh_spmv_f32(ofh_csr A, const float* d_x, float* d_y);

/* ------------------------------------------------------------ Vect BLAS-1 */
/* Vect::operator, (Vect_impl.h:2218-2225) and getNorm2 (:660-668). */
int ofh_dot_f64(ofh_ctx ctx, int n, const double* d_x, const double* d_y, double* h_out);
int ofh_norm2_f64(ofh_ctx ctx, int n, const double* d_x, double* h_out);
/* y = a*x + y (Vect operator+= of a*x, Vect_impl.h:2143-2173). */
int ofh_axpy_f64(ofh_ctx ctx, int n, double a, const double* d_x, double* d_y);

/* ------------------------------------------------------------ Prec / SSOR */
/* Prec<T_>::solve(b,x) (Prec_impl.h:226-255) for IDENT, DIAG, DILU, ILU and
 * SSOR.  DILU_PREC and ILU_PREC are the reference's D-ILU
 * (SpMatrix::DILUFactorize / DILUSolve, SpMatrix_impl.h:834-900; its ILU
 * routines call these): pivots 1/D_l with D_l = a_ll - sum_{i<l} a_li a_il / D_i
 * (factorized on the GPU by level-scheduled rows, again whenever the
 * values change), applied as M = (D+L) D^-1 (D+U) by exact triangular
 * sweeps.  ofh_cg, ofh_bicgstab, ofh_gmres and ofh_linear_solve accept them. */
int ofh_prec_apply(ofh_csr A, int prec, const double* d_b, double* d_x);
/* Prec::TransSolve(b,x) (Prec_impl.h:259-296): DIAG_PREC scales by 1/a_ii,
 * ILU_PREC solves M^T x = b (two exact sweeps over A^T, bitwise the
 * reference's column-oriented loops), IDENT/DILU/SSOR_PREC copy b (the
 * reference's switch has no case for them). */
int ofh_prec_trans_apply(ofh_csr A, int prec, const double* d_b, double* d_x);
/* SpMatrix::SSORSolve (SpMatrix_impl.h:913-942): symmetric Gauss-Seidel. */
int ofh_ssor_solve(ofh_csr A, const double* d_b, double* d_x);

/* ---------------------------------------------------------------- solvers */
/* CG (CG.h:99-157 via CG.h:189-200): x holds the initial guess. */
int ofh_cg(ofh_csr A, int prec, const double* d_b, double* d_x, int max_it,
           double toler, ofh_solve_info* info);
/* CG that allocates nothing (no reference counterpart: for callers that
 * cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The ofh_cg iteration for
 * IDENT_PREC / DIAG_PREC (CG.h:99-157; DIAG pivots 1/a_ii found per call),
 * results within rounding of ofh_cg's; its work vectors, partial sums and
 * solver state live in a static device array of the library (16.8 MB per
 * GPU, held until the process ends) instead of the context's workspace, and
 * no SELL copy or other matrix companion is built.  At most 262,144 rows
 * (OFH_E_INVALID beyond).  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU race (serialise them, e.g. one stream). */
int ofh_cg_noalloc(ofh_csr A, int prec, const double* d_b, double* d_x, int max_it,
                   double toler, ofh_solve_info* info);
/* BiCGStab (BiCGStab.h:87-162 via :194-216). */
int ofh_bicgstab(ofh_csr A, int prec, const double* d_b, double* d_x, int max_it,
                 double toler, ofh_solve_info* info);
/* GMRes(m) (GMRes.h:148-241 via :274-283). */
int ofh_gmres(ofh_csr A, int prec, const double* d_b, double* d_x, int m, int max_it,
              double toler, ofh_solve_info* info);
/* CGS (CGS.h:91-171 via :203-212), as written: its q is never assigned and
 * stays 0, so it is not the textbook CGS and often does not converge.
 * Returns 0, it, 2 when (s,r) = 0, or -max_it. */
int ofh_cgs(ofh_csr A, int prec, const double* d_b, double* d_x, int max_it,
            double toler, ofh_solve_info* info);
/* BiCG (BiCG.h:94-170 via :202-211) with the true transposed products
 * (ofh_spmv_t_f64, ofh_prec_trans_apply).  Returns 0, it (also when
 * (z,rr) = 0), or -(max_it+1); the in-loop test divides by ||b|| itself. */
int ofh_bicg(ofh_csr A, int prec, const double* d_b, double* d_x, int max_it,
             double toler, ofh_solve_info* info);
/* LinearSolver::solve (LinearSolver.cpp:194-244) dispatch; GMRES uses
 * m = min(n/5, 50) instead of the reference's n/5 (LinearSolver.cpp:238). */
int ofh_linear_solve(ofh_csr A, int solver, int prec, const double* d_b, double* d_x,
                     int max_it, double toler, ofh_solve_info* info);
/* Mixed precision CG (no reference counterpart; judged against ofh_cg):
 * FP32 SpMV and vector updates, FP64 Krylov dot products, FP64 solution
 * accumulation and FP64 residual replacement b-Ax; stops on the true FP64
 * residual ||b-Ax||/||b|| <= toler; falls back to FP64 CG from the current
 * iterate if the FP32 iteration stagnates.  prec: IDENT, DIAG, or SSOR /
 * DILU / ILU (z = M^-1 r by the FP64 level-scheduled sweeps of the FP32
 * residual, rounded back to FP32). */
int ofh_cg_mixed(ofh_csr A, int prec, const double* d_b, double* d_x, int max_it,
                 double toler, ofh_solve_info* info);

/* ------------------------------------------------------- batched systems */
/* Many independent systems in one persistent launch (no reference routine:
 * the reference solves such a set with one LinearSolver::solve per system,
 * LinearSolver.cpp:194-244 -> CG.h:99-157).  A is block-diagonal: system s
 * owns rows [h_sys_rows[s], h_sys_rows[s+1]) and its columns stay inside
 * that range (checked).  Typically A is the plan matrix of several disjoint
 * meshes assembled together. */
typedef struct ofh_batch_s* ofh_batch;
int ofh_batch_create(ofh_csr A, int nsys, const int32_t* h_sys_rows, ofh_batch* out);
int ofh_batch_destroy(ofh_batch B);
/* CG + DIAG_PREC/IDENT_PREC on every system (x holds the initial guesses).
 * mixed = 0: FP64, the reference iteration per system; mixed = 1: FP32
 * inner iteration with FP64 residual replacement, stopping on the true FP64
 * residual.  Per-system results (reference convention in h_status) may be
 * NULL; total->iterations is the largest count. */
int ofh_batch_cg(ofh_batch B, int prec, int mixed, const double* d_b, double* d_x, int max_it,
                 double toler, int32_t* h_iters, int32_t* h_status, double* h_relres,
                 ofh_solve_info* total);
/* ofh_batch_cg replays its launch (queue reset + persistent kernel) from a
 * HIP graph captured for the current arguments (enable = 1, the default), or
 * launches it directly (enable = 0).  Results are identical either way. */
int ofh_batch_use_graph(ofh_batch B, int enable);
/* Launch shape: workgroup size {256,512,1024} x threads per row {1,2,4,8},
 * or (1024, 0): the resident form, one workgroup per system with the
 * direction in LDS, r in registers and the matrix read from an ELL copy
 * made on its first use (every system at most 12288 rows).  Both
 * forms give bitwise the same iterates at workgroup size 1024. */
int ofh_batch_set_config(ofh_batch B, int block_threads, int threads_per_row);
int ofh_batch_get_config(ofh_batch B, int* block_threads, int* threads_per_row);
/* Times every launch shape on this batch (x is restored; the resident form
 * only when every system fits) and keeps the fastest for later ofh_batch_cg
 * calls.  The choice is cached for the process, keyed by prec, mixed and
 * every system's rows and nonzeros: a later batch of the same shape takes
 * it without timing (*best_ms then reports the cached time). */
int ofh_batch_autotune(ofh_batch B, int prec, int mixed, const double* d_b, double* d_x,
                       int max_it, double toler, int* best_block, int* best_tpr,
                       double* best_ms);

/* ------------------------------------------------------ meshes (host side) */
/* Structured rectangle, Mesh::Mesh(xmin,xmax,ymin,ymax,nx,ny,cx0,cxN,cy0,cyN,
 * opt,nb_dof) (Mesh.cpp:547-676): h_coord (nn x 2), h_code (nn x nb_dof; the
 * boundary code goes to DOF 1 as in the reference), h_elem (ne x nen,
 * 1-based node ids).  quad = 0: 2 triangles per cell, 1: quadrilaterals. */
int ofh_mesh2d_structured(double xmin, double xmax, double ymin, double ymax, int nx, int ny,
                          int cx0, int cxN, int cy0, int cyN, int quad, int nb_dof,
                          double* h_coord, int32_t* h_code, int32_t* h_elem);
/* Structured box, node order and codes of Mesh.cpp:703-726; cells split into
 * 6 positively oriented Kuhn tetrahedra (the reference's TETRAHEDRON split,
 * Mesh.cpp:747-771, yields negative Jacobians). */
int ofh_mesh3d_kuhn(double xmin, double xmax, double ymin, double ymax, double zmin,
                    double zmax, int nx, int ny, int nz, int cx0, int cxN, int cy0, int cyN,
                    int cz0, int czN, double* h_coord, int32_t* h_code, int32_t* h_elem);
/* OFELI XML mesh files (.m).  ofh_mesh_read is Mesh::get(file)
 * (Mesh.cpp:2639-2670) as XMLParser::read_mesh / read_mesh_data parse it
 * (XMLParser.cpp:1175-1220, 2210-2293): coordinates are rounded to single
 * precision (stof), node and side marks become nb_dof codes by DOFCode
 * (MeshUtil.cpp:101-116; a negative mark yields its digits, positive).
 * Shapes use the reference's ElementShape values (OFELI_Config.h:272-283:
 * LINE 2, TRIANGLE 3, QUADRILATERAL 4, TETRAHEDRON 5, ...).  Labels are
 * 1-based; codes are nb_nodes x nb_dof, side codes nb_sides x nb_dof.
 * Any output pointer of ofh_mesh_get may be NULL. */
typedef struct ofh_mesh_s* ofh_mesh;
int ofh_mesh_read(const char* path, ofh_mesh* out);
int ofh_mesh_info(ofh_mesh m, int* dim, int* nb_dof, int* nb_nodes, int* nb_elements,
                  int* el_shape, int* nen, int* nb_sides, int* sd_shape, int* nsn,
                  int* nb_materials);
/* Each non-NULL buffer comes with its capacity in elements; a buffer shorter
 * than its array (sizes from ofh_mesh_info) fails with OFH_E_INVALID before
 * anything is copied.  The reader rejects files whose records disagree in
 * width (a second <Mesh>, a later <Elements>/<Sides> changing shape or
 * nodes), so the sizes ofh_mesh_info reports are exact. */
int ofh_mesh_get(ofh_mesh m, double* h_coord, int64_t coord_len, int32_t* h_code,
                 int64_t code_len, int32_t* h_elem, int64_t elem_len, int32_t* h_elem_code,
                 int64_t elem_code_len, int32_t* h_side, int64_t side_len,
                 int32_t* h_side_code, int64_t side_code_len);
int ofh_mesh_material(ofh_mesh m, int i, int* code, char* name, int name_len);
int ofh_mesh_destroy(ofh_mesh m);
/* Mesh::put(file) (Mesh.cpp:2734-2800), byte for byte; the <Material>
 * section is written when there are several materials or the only one is
 * not named "Generic". */
int ofh_mesh_write(const char* path, int dim, int nb_dof, int nb_nodes, const double* h_coord,
                   const int32_t* h_code, int nb_elements, int el_shape, int nen,
                   const int32_t* h_elem, const int32_t* h_elem_code, int nb_sides,
                   int sd_shape, int nsn, const int32_t* h_side, const int32_t* h_side_code,
                   int nb_materials, const int32_t* mat_code, const char* const* mat_name);
/* DOF numbering with imposed-DOF elimination, Mesh::NumberEquations with
 * removeImposedDOF (Mesh.cpp:1499-1509): returns nb_eq, labels 1-based. */
int ofh_number_equations(int nb_nodes, int nb_dof, const int32_t* h_code, int32_t* h_dof,
                         int* nb_eq);
/* Gauss-Legendre points/weights, Gauss::Gauss (Gauss.cpp:37-80). */
int ofh_gauss(int np, double* h_x, double* h_w);

/* ------------------------------------------------------- assembly (device) */
typedef struct ofh_asm_s* ofh_asm;
/* Uploads the mesh, numbers the equations (imposed DOFs eliminated) and builds
 * on the GPU the CSR pattern of NodeGraph/FinalizeGraph/StoreGraph
 * (GraphOfMatrix.cpp:52-67, 347-375, 619-634) plus the element-to-nonzero
 * gather map used by ofh_asm_build.  h_elem holds 1-based node ids. */
int ofh_asm_create(ofh_ctx ctx, int dim, int nb_nodes, const double* h_coord, int nb_el,
                   int nen, const int32_t* h_elem, int nb_dof, const int32_t* h_code,
                   ofh_asm* out);
int ofh_asm_destroy(ofh_asm P);
int ofh_asm_info(ofh_asm P, int* nb_eq, int64_t* nnz);
/* The plan's matrix (pattern + values); owned by the plan. */
int ofh_asm_matrix(ofh_asm P, ofh_csr* A);
int ofh_asm_get_dof(ofh_asm P, int32_t* h_dof);
/* The equation classes' element loop (Equa_Laplace.h:112-133,
 * Equa_Therm.h:162-218, Equa_Solid.h:168-204): element matrices and
 * right-hand sides (BodyRHS, updateBC lifting, backward-Euler term) computed
 * per element, then each matrix entry and RHS entry summed over its elements
 * in ascending element order (Matrix_impl.h:172-201, Vect_impl.h:1522-1534)
 * with no atomics: bitwise the reference's sums.
 *   d_coef: per-element coefficient (k for DC/3D kinds, E for ELAS2DQ4) or NULL
 *   d_f:    nodal source (LAPLACE/DC kinds) or NULL
 *   d_bc:   nodal Dirichlet values (nb_dof per node) or NULL
 *   d_uprev: nodal u^n (DC2DT3_BE)
 *   h_prm:  DC2DT3_BE {dt, rho, cp}; ELAS2DQ4 {nu, mode}: mode 0 reference
 *           parity (plane strain, 2*Dev+Dil, as Elas2DQ4::set always ends
 *           with PlaneStrain() and both routines add the cumulative eA0),
 *           1 plane stress Dev+Dil, 2 plane strain Dev+Dil
 * want_matrix = 0 assembles the right-hand side only. d_b may be NULL. */
int ofh_asm_build(ofh_asm P, int kind, const double* d_coef, const double* d_f,
                  const double* d_bc, const double* d_uprev, const double* h_prm, int nprm,
                  int want_matrix, double* d_b);
/* Elas2DQ4 constant side tractions (Line2 length, Elas2DQ4::BoundaryRHS,
 * Elas2DQ4.cpp:256-263), added to d_b after the element loop in side order.
 * h_side_nodes: nb_sides x 2 1-based node ids; d_traction: nb_sides x 2. */
int ofh_asm_traction(ofh_asm P, int nb_sides, const int32_t* h_side_nodes,
                     const double* d_traction, double* d_b);
/* Vect::insertBC (Vect_impl.h:1097-1116): nodal vector from the free-DOF
 * solution and the Dirichlet values (d_bc may be NULL: zero). */
int ofh_asm_insert_bc(ofh_asm P, const double* d_v, const double* d_bc, double* d_u);

/* --------------------------------- start vectors from previous solutions */
/* No reference routine (the reference starts from zero or from the previous
 * solution, TimeStepping.cpp:207).  For solves A x_k = b_k with one SPD
 * matrix, keeps a basis V of the span of recent solutions and returns the
 * start vector of that span closest to A^-1 b (Fischer 1998):
 *   mode 0: V A-orthonormal, x0 = V V^T b minimises the A-norm error;
 *   mode 1: AV orthonormal, x0 = V (AV)^T b minimises ||b - A x0||_2, the
 *           quantity of the solvers' stop test.
 * ofh_proj_update keeps the solution in a ring of the last max_vectors
 * solutions; when V is full it is rebuilt from the last max_vectors/2 of
 * them, otherwise the solution is orthogonalised against V (classical
 * Gram-Schmidt, two passes), normalised and appended.  A direction below
 * 1e-6 of the solution's norm is stored as zero (it only carries residual
 * noise).  The span always holds the newest solution, so the guess is never
 * worse than the warm start in the mode's norm.  An append costs one SpMV and
 * about 6 x max_vectors vector passes; a guess 1 + 2 x max_vectors passes.
 * max_vectors: 8, 16 or 32. */
typedef struct ofh_proj_s* ofh_proj;
int ofh_proj_create(ofh_csr A, int max_vectors, int mode, ofh_proj* out);
int ofh_proj_destroy(ofh_proj P);
int ofh_proj_reset(ofh_proj P);
int ofh_proj_size(ofh_proj P, int* count);
int ofh_proj_guess(ofh_proj P, const double* d_b, double* d_x0);
int ofh_proj_update(ofh_proj P, const double* d_x);

/* ------------------------------------------ learned initial guess (device) */
/* No reference routine: the reference starts a solve from zero (Equa::run,
 * Vect_impl.h:547) or from the previous step's solution (TimeStepping.cpp:207).
 * For a sequence of solves A x_{k+1} = b_{k+1} with A symmetric (pattern
 * exactly, values to 1e-12 relative), positive diagonal, a message-passing
 * network on the graph of A predicts a search
 * direction d from the base x_b and the new b; the start vector is the base
 * plus the exact A-norm line-search step along d, so it is never worse than
 * the base:
 *   x_b = x_n (extrapolate = 0) or x_n + (x_n - x_p) (extrapolate = 1)
 *   r = b - A x_b, z = D^-1 r, s = rms(z), z2 = D^-1 (r - A z)
 *   F = [z/s, z2/s, (x_n - x_p)/s, ln a_ii - mean, 0 x 4]      (FP32)
 *   H = relu(F W_enc + b_enc)                                    (8 -> 64)
 *   per layer: Z = H W_ab (64 -> 128, BF16 matrix cores),
 *              M_i = sum_{j != i} relu(Z_i[:64] + Z_j[64:] + e_ij w_e + b_m),
 *              e_ij = (a_ij + a_ji) / (2 sqrt(a_ii a_jj)),
 *              H += relu([H, M] W_u + b_u) (128 -> 64, BF16 matrix cores)
 *   d = s (H w_dec + b_dec),  x0 = x_b + alpha d,  alpha = d.r / d.Ad.
 * Training minimises L = 1 - (d.g)^2 / (d.Ad E0) with g = A (x* - x_b),
 * E0 = (x* - x_b).g and x* the solver's solution: the A-norm error left
 * after the line search relative to the base's (scale-invariant in d); Adam
 * (0.9, 0.999, 1e-8) on FP32 master weights.
 * Parameters are one FP32 vector: W_enc[8][64], b_enc[64], per layer
 * W_ab[64][128], w_e[64], b_m[64], W_u[128][64], b_u[64]; w_dec[64], b_dec.
 * Objective 1 (ofh_mpnn_set_objective) aligns the start with CG's stop test
 * on ||r||: alpha = Ad.r / Ad.Ad (minimum residual along d) and
 * L = 1 - (Ad.r)^2 / (Ad.Ad r.r), the squared residual norm left relative to
 * the base's; it needs no solution (d_xstar may be NULL). */
typedef struct ofh_mpnn_s* ofh_mpnn;
int ofh_mpnn_create(ofh_csr A, int layers, int hidden, int extrapolate, uint64_t seed,
                    ofh_mpnn* out);
int ofh_mpnn_destroy(ofh_mpnn M);
int ofh_mpnn_num_params(ofh_mpnn M, int64_t* count);
int ofh_mpnn_get_params(ofh_mpnn M, float* h_params);
/* Re-reads the edge weights and diagonal features from the matrix's current
 * values (same pattern), e.g. after a new assembly with other coefficients;
 * the parameters are kept, so one network serves a family of matrices. */
int ofh_mpnn_refresh(ofh_mpnn M);
/* Sets the parameters and resets the optimiser state. */
int ofh_mpnn_set_params(ofh_mpnn M, const float* h_params);
/* ofh_mpnn_guess replays its kernels from a HIP graph captured per argument
 * set (enable = 1, the default) or launches them directly. */
int ofh_mpnn_use_graph(ofh_mpnn M, int enable);
/* objective: 0 A-norm error (the default), 1 residual norm. */
int ofh_mpnn_set_objective(ofh_mpnn M, int objective);
/* x0 = x_b + alpha d; stream-ordered, no host synchronisation. */
int ofh_mpnn_guess(ofh_mpnn M, const double* d_xn, const double* d_xp, const double* d_b,
                   double* d_x0);
/* Forward + loss + backward: d_x0 receives the guess, h_loss the loss L and
 * h_grad (may be NULL) the gradient of the parameter vector. */
int ofh_mpnn_loss_grad(ofh_mpnn M, const double* d_xn, const double* d_xp, const double* d_b,
                       const double* d_xstar, double* d_x0, double* h_loss, float* h_grad);
/* One Adam step on one sample (h_loss may be NULL: then no host sync). */
int ofh_mpnn_train_step(ofh_mpnn M, const double* d_xn, const double* d_xp, const double* d_b,
                        const double* d_xstar, double* d_x0, float lr, double* h_loss);

#ifdef __cplusplus
}
#endif
#endif /* OFELI_HIP_H */
