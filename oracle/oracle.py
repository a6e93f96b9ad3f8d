"""ctypes front end of the CPU oracle (ofeli_oracle.c).

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg import this module; the product package
(ofeli_amd) never does.  All routines restate OFELI 5.3.3's algorithms in
their reference operation order; the C source cites file:line for each.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_STRICT = os.path.join(HERE, "liborc.so")       # -O2 -ffp-contract=off
LIB_FAST = os.path.join(HERE, "liborc_fast.so")    # the reference's -Ofast

LAPLACE2DT3, DC2DT3, DC2DT3_BE, LAPLACE3DT4, ELAS2DQ4 = 1, 2, 3, 4, 5
IDENT, DIAG, DILU, ILU, SSOR = 1, 2, 3, 4, 5

_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_oi = C.c_int64
_libs = {}


def build(fast=False):
    """Compile the oracle with gcc (strict, and -Ofast for timing)."""
    import subprocess
    src = os.path.join(HERE, "ofeli_oracle.c")
    subprocess.check_call(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared",
                           "-o", LIB_STRICT, src, "-lm"])
    subprocess.check_call(["gcc", "-Ofast", "-fPIC", "-shared",
                           "-o", LIB_FAST, src, "-lm"])


def lib(fast=False):
    key = "fast" if fast else "strict"
    if key in _libs:
        return _libs[key]
    path = LIB_FAST if fast else LIB_STRICT
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    L.orc_mesh2d.argtypes = [C.c_double] * 4 + [_oi, _oi] + [C.c_int] * 6 + [_f64p, _i32p, _i64p]
    L.orc_mesh2d_sides.argtypes = [_oi, _oi] + [C.c_int] * 4 + [_i64p, _i32p, C.POINTER(_oi)]
    L.orc_mesh3d_kuhn.argtypes = [C.c_double] * 6 + [_oi] * 3 + [C.c_int] * 6 + [_f64p, _i32p, _i64p]
    L.orc_number_equations.argtypes = [_oi, C.c_int, _i32p, _i64p]
    L.orc_number_equations.restype = _oi
    L.orc_node_graph.argtypes = [_oi, C.c_int, C.c_int, _i64p, _i64p, _oi, _i64p,
                                 C.POINTER(C.POINTER(_oi))]
    L.orc_node_graph.restype = _oi
    L.orc_node_graph_scal.argtypes = [_oi, C.c_int, _i64p, _oi, _i64p,
                                      C.POINTER(C.POINTER(_oi))]
    L.orc_node_graph_scal.restype = _oi
    L.orc_free.argtypes = [C.c_void_p]
    L.orc_triang3.argtypes = [_f64p, C.POINTER(C.c_double), _f64p]
    L.orc_tetra4.argtypes = [_f64p, C.POINTER(C.c_double), _f64p]
    L.orc_gauss.argtypes = [C.c_int, _f64p, _f64p]
    L.orc_quad4_atgauss.argtypes = [_f64p, _f64p, _f64p, _f64p]
    L.orc_element.argtypes = [C.c_int, _f64p, C.c_double, C.c_void_p, _f64p, C.c_void_p]
    L.orc_assemble.argtypes = [C.c_int, C.c_int, _f64p, _oi, _i64p, C.c_int, _i64p, _i32p,
                               C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                               _oi, _i64p, _i64p, _f64p, _f64p]
    L.orc_traction_rhs.argtypes = [_f64p, _oi, _i64p, _f64p, _i64p, _f64p]
    L.orc_spmv.argtypes = [_oi, _i64p, _i64p, _f64p, _f64p, _f64p]
    L.orc_norm2.argtypes = [_oi, _f64p]
    L.orc_norm2.restype = C.c_double
    L.orc_dot.argtypes = [_oi, _f64p, _f64p]
    L.orc_dot.restype = C.c_double
    L.orc_ssor_solve.argtypes = [_oi, _i64p, _i64p, _f64p, _f64p, _f64p]
    L.orc_dilu_factorize.argtypes = [_oi, _i64p, _i64p, _f64p, _f64p]
    L.orc_dilu_solve.argtypes = [_oi, _i64p, _i64p, _f64p, _f64p, _f64p, _f64p]
    L.orc_prec_setup.argtypes = [_oi, _i64p, _i64p, _f64p, C.c_int, _f64p]
    L.orc_spmv_t.argtypes = [_oi, _i64p, _i64p, _f64p, _f64p, _f64p]
    L.orc_dilu_trans_solve.argtypes = [_oi, _i64p, _i64p, _f64p, _f64p, _f64p, _f64p]
    L.orc_prec_trans_apply.argtypes = [_oi, _i64p, _i64p, _f64p, C.c_int, _f64p, _f64p, _f64p]
    for fn in ("orc_cg", "orc_bicgstab", "orc_cgs", "orc_bicg"):
        getattr(L, fn).argtypes = [_oi, _i64p, _i64p, _f64p, C.c_int, _f64p, _f64p,
                                   C.c_int, C.c_double, C.POINTER(C.c_double)]
    L.orc_gmres.argtypes = [_oi, _i64p, _i64p, _f64p, C.c_int, _f64p, _f64p, C.c_int,
                            C.c_int, C.c_double, C.POINTER(C.c_double)]
    L.orc_insert_bc.argtypes = [_oi, C.c_int, _i32p, _f64p, C.c_void_p, _f64p]
    _libs[key] = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


class Mesh:
    """Plain arrays of a mesh: coord (nn,dim), code (nn,nb_dof), elem (ne,nen) 1-based."""

    def __init__(self, coord, code, elem, nb_dof, sides=None, side_code=None):
        self.coord, self.code, self.elem, self.nb_dof = coord, code, elem, nb_dof
        self.sides, self.side_code = sides, side_code

    @property
    def dim(self):
        return self.coord.shape[1]

    @property
    def nb_nodes(self):
        return self.coord.shape[0]

    @property
    def nb_elements(self):
        return self.elem.shape[0]


def mesh2d(nx, ny, quad=False, nb_dof=1, codes=(1, 1, 1, 1), box=(0.0, 1.0, 0.0, 1.0)):
    """Mesh.cpp:547-676 structured rectangle (codes = cx0,cxN,cy0,cyN)."""
    L = lib()
    nn = (nx + 1) * (ny + 1)
    ne = nx * ny if quad else 2 * nx * ny
    nen = 4 if quad else 3
    coord = np.zeros((nn, 2))
    code = np.zeros((nn, nb_dof), np.int32)
    elem = np.zeros((ne, nen), np.int64)
    L.orc_mesh2d(*box, nx, ny, *codes, int(quad), nb_dof, coord, code, elem)
    sides = np.zeros((2 * (nx + ny), 2), np.int64)
    scode = np.zeros(2 * (nx + ny), np.int32)
    ns = _oi(0)
    L.orc_mesh2d_sides(nx, ny, *codes, sides, scode, C.byref(ns))
    return Mesh(coord, code, elem, nb_dof, sides[: ns.value].copy(), scode[: ns.value].copy())


def mesh3d_kuhn(nx, ny, nz, codes=(1, 1, 1, 1, 1, 1), box=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0)):
    """Structured cube (node order Mesh.cpp:703-726) with the Kuhn 6-tet split."""
    L = lib()
    nn = (nx + 1) * (ny + 1) * (nz + 1)
    ne = 6 * nx * ny * nz
    coord = np.zeros((nn, 3))
    code = np.zeros((nn, 1), np.int32)
    elem = np.zeros((ne, 4), np.int64)
    L.orc_mesh3d_kuhn(*box, nx, ny, nz, *codes, coord, code, elem)
    return Mesh(coord, code, elem, 1)


def number_equations(code):
    """Mesh.cpp:1499-1509; returns (dof (nn,nb_dof) int64, nb_eq)."""
    code = np.ascontiguousarray(code, np.int32)
    dof = np.zeros(code.shape, np.int64)
    neq = lib().orc_number_equations(code.shape[0], code.shape[1], code, dof)
    return dof, int(neq)


def node_graph(elem, nb_dof, dof, nb_eq):
    """NodeGraph/FinalizeGraph/StoreGraph: row_ptr 0-based, col_ind 1-based."""
    L = lib()
    elem = np.ascontiguousarray(elem, np.int64)
    dof = np.ascontiguousarray(dof, np.int64)
    row_ptr = np.zeros(nb_eq + 1, np.int64)
    cp = C.POINTER(_oi)()
    nnz = L.orc_node_graph(elem.shape[0], elem.shape[1], nb_dof, elem, dof, nb_eq, row_ptr,
                           C.byref(cp))
    if nnz < 0:
        raise RuntimeError("orc_node_graph failed: %d" % nnz)
    col = np.ctypeslib.as_array(cp, shape=(nnz,)).copy()
    L.orc_free(cp)
    return row_ptr, col


def node_graph_scal(elem, nb_nodes):
    """NodeGraphScal (GraphOfMatrix.cpp:532-546), the pattern of
    SpMatrix(mesh, dof != 0): node numbers, no elimination; row_ptr 0-based,
    col_ind 1-based."""
    L = lib()
    elem = np.ascontiguousarray(elem, np.int64)
    row_ptr = np.zeros(nb_nodes + 1, np.int64)
    cp = C.POINTER(_oi)()
    nnz = L.orc_node_graph_scal(elem.shape[0], elem.shape[1], elem, nb_nodes, row_ptr,
                                C.byref(cp))
    if nnz < 0:
        raise RuntimeError("orc_node_graph_scal failed: %d" % nnz)
    col = np.ctypeslib.as_array(cp, shape=(nnz,)).copy()
    L.orc_free(cp)
    return row_ptr, col


def triang3(x):
    area = C.c_double()
    d = np.zeros(6)
    rc = lib().orc_triang3(_f64(np.ravel(x)), C.byref(area), d)
    return rc, area.value, d.reshape(3, 2)


def tetra4(x):
    vol = C.c_double()
    d = np.zeros(12)
    rc = lib().orc_tetra4(_f64(np.ravel(x)), C.byref(vol), d)
    return rc, vol.value, d.reshape(4, 3)


def gauss(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gauss(n, x, w)
    return x, w


def quad4_atgauss(x):
    sh = np.zeros(16)
    dsh = np.zeros(32)
    w = np.zeros(4)
    rc = lib().orc_quad4_atgauss(_f64(np.ravel(x)), sh, dsh, w)
    return rc, sh.reshape(4, 4), dsh.reshape(4, 4, 2), w


def element(kind, x, coef=-1.0, prm=None):
    nee = {LAPLACE2DT3: 3, DC2DT3: 3, DC2DT3_BE: 3, LAPLACE3DT4: 4, ELAS2DQ4: 8}[kind]
    emat = np.zeros(nee * nee)
    ea1 = np.zeros(9)
    p = _f64(prm)
    rc = lib().orc_element(kind, _f64(np.ravel(x)), coef, _ptr(p), emat, _ptr(ea1))
    return rc, emat.reshape(nee, nee)


def assemble(kind, mesh, dof, nb_eq, row_ptr, col_ind, f=None, coef=None, bc=None,
             prm=None, u_prev=None, matrix=True, rhs=True):
    """Element loop of the equation classes' build(); returns (a, b)."""
    L = lib()
    nnz = int(row_ptr[-1])
    a = np.zeros(max(nnz, 1))
    b = np.zeros(max(nb_eq, 1))
    f, coef, bc, prm, u_prev = map(_f64, (f, coef, bc, prm, u_prev))
    rc = L.orc_assemble(kind, mesh.dim, np.ascontiguousarray(mesh.coord), mesh.nb_elements,
                        np.ascontiguousarray(mesh.elem), mesh.nb_dof,
                        np.ascontiguousarray(dof, np.int64),
                        np.ascontiguousarray(mesh.code, np.int32), _ptr(bc), _ptr(f),
                        _ptr(coef), _ptr(prm), _ptr(u_prev), nb_eq, row_ptr, col_ind, a, b)
    if rc != 0:
        raise RuntimeError("orc_assemble failed: %d" % rc)
    return a[:nnz], b[:nb_eq]


def traction_rhs(mesh, sides, traction, dof, b):
    b = np.ascontiguousarray(b, np.float64).copy()
    rc = lib().orc_traction_rhs(np.ascontiguousarray(mesh.coord), sides.shape[0],
                                np.ascontiguousarray(sides, np.int64), _f64(traction),
                                np.ascontiguousarray(dof, np.int64), b)
    if rc != 0:
        raise RuntimeError("orc_traction_rhs failed: %d" % rc)
    return b


def spmv(row_ptr, col_ind, a, x):
    n = row_ptr.shape[0] - 1
    y = np.zeros(n)
    lib().orc_spmv(n, row_ptr, col_ind, _f64(a), _f64(x), y)
    return y


def spmv_t(row_ptr, col_ind, a, x):
    """y = A^T x (SpMatrix::TMult as BiCG uses it)."""
    n = row_ptr.shape[0] - 1
    y = np.zeros(n)
    lib().orc_spmv_t(n, row_ptr, col_ind, _f64(a), _f64(x), y)
    return y


def prec_trans_apply(row_ptr, col_ind, a, prec, b):
    """x of Prec::TransSolve(b,x): DIAG scales, ILU solves M^T x = b, IDENT,
    DILU and SSOR return b."""
    n = row_ptr.shape[0] - 1
    pivot = np.zeros(n)
    if prec in (DIAG, DILU, ILU):
        if lib().orc_prec_setup(n, row_ptr, col_ind, _f64(a), prec, pivot) != 0:
            raise RuntimeError("zero pivot")
    x = np.zeros(n)
    if lib().orc_prec_trans_apply(n, row_ptr, col_ind, _f64(a), prec, pivot, _f64(b), x) != 0:
        raise RuntimeError("preconditioner error")
    return x


def ssor_solve(row_ptr, col_ind, a, b):
    n = row_ptr.shape[0] - 1
    x = np.zeros(n)
    rc = lib().orc_ssor_solve(n, row_ptr, col_ind, _f64(a), _f64(b), x)
    if rc != 0:
        raise RuntimeError("zero pivot")
    return x


def dilu_factorize(row_ptr, col_ind, a):
    """pivot = 1/D of SpMatrix::DILUFactorize (DILU_PREC and ILU_PREC)."""
    n = row_ptr.shape[0] - 1
    pivot = np.zeros(n)
    if lib().orc_dilu_factorize(n, row_ptr, col_ind, _f64(a), pivot) != 0:
        raise RuntimeError("zero pivot")
    return pivot


def dilu_solve(row_ptr, col_ind, a, pivot, b):
    """x = M^-1 b of SpMatrix::DILUSolve, M = (D+L) D^-1 (D+U)."""
    n = row_ptr.shape[0] - 1
    x = np.zeros(n)
    if lib().orc_dilu_solve(n, row_ptr, col_ind, _f64(a), _f64(pivot), _f64(b), x) != 0:
        raise RuntimeError("zero pivot")
    return x


def _solve(fn, row_ptr, col_ind, a, b, x0, prec, max_it, toler, fast, extra=()):
    L = lib(fast)
    n = row_ptr.shape[0] - 1
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    res = C.c_double(0.0)
    args = [n, row_ptr, col_ind, _f64(a), prec, _f64(b), x, *extra, max_it, toler,
            C.byref(res)]
    it = getattr(L, fn)(*args)
    return int(it), x, res.value


def cg(row_ptr, col_ind, a, b, x0=None, prec=DIAG, max_it=1000, toler=1.4901161193847656e-08,
       fast=False):
    """CG.h:99-157 -> (ret, x, res); ret = it, 0, or -(max_it+1)."""
    return _solve("orc_cg", row_ptr, col_ind, a, b, x0, prec, max_it, toler, fast)


def bicgstab(row_ptr, col_ind, a, b, x0=None, prec=DIAG, max_it=1000,
             toler=1.4901161193847656e-08, fast=False):
    return _solve("orc_bicgstab", row_ptr, col_ind, a, b, x0, prec, max_it, toler, fast)


def cgs(row_ptr, col_ind, a, b, x0=None, prec=DIAG, max_it=1000,
        toler=1.4901161193847656e-08, fast=False):
    """CGS.h:91-171 as written -> (ret, x, res); ret = it, 0, 2, or -max_it."""
    return _solve("orc_cgs", row_ptr, col_ind, a, b, x0, prec, max_it, toler, fast)


def bicg(row_ptr, col_ind, a, b, x0=None, prec=DIAG, max_it=1000,
         toler=1.4901161193847656e-08, fast=False):
    """BiCG.h:94-170 with true transposed products -> (ret, x, res);
    ret = it, 0, or -(max_it+1)."""
    return _solve("orc_bicg", row_ptr, col_ind, a, b, x0, prec, max_it, toler, fast)


def gmres(row_ptr, col_ind, a, b, m, x0=None, prec=DIAG, max_it=1000,
          toler=1.4901161193847656e-08, fast=False):
    return _solve("orc_gmres", row_ptr, col_ind, a, b, x0, prec, max_it, toler, fast, (m,))


def insert_bc(code, v, bc=None):
    code = np.ascontiguousarray(code, np.int32)
    u = np.zeros(code.size)
    bcp = _f64(bc)
    lib().orc_insert_bc(code.shape[0], code.shape[1], code, _f64(v), _ptr(bcp), u)
    return u


def to_dense(row_ptr, col_ind, a):
    n = row_ptr.shape[0] - 1
    A = np.zeros((n, n))
    for i in range(n):
        for k in range(row_ptr[i], row_ptr[i + 1]):
            A[i, col_ind[k] - 1] += a[k]
    return A
