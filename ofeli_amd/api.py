"""Python front end of libofeli_hip: objects mirroring OFELI's SpMatrix,
the solver functions and the equation classes' assembly.

Device vectors are torch tensors on the GPU (float64 unless noted).  The
context runs its kernels on torch's current stream, so torch work and
library work are ordered without extra synchronisation.
"""
import atexit
import ctypes as C
import os
import weakref

import numpy as np

from . import (DIAG_PREC, SolveInfo, check, lib)


def _ptr(t):
    """Device pointer of a contiguous torch tensor (or None)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a GPU tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return C.c_void_p(t.data_ptr())


def _f64(t, n=None):
    import torch
    if t.dtype != torch.float64:
        raise TypeError("expected float64 tensor, got %s" % t.dtype)
    if n is not None and t.numel() != n:
        raise ValueError("expected %d entries, got %d" % (n, t.numel()))
    return _ptr(t)


def _np(a, dt):
    a = np.ascontiguousarray(a, dtype=dt)
    ct = {np.int32: C.c_int32, np.int64: C.c_int64, np.float64: C.c_double}[dt]
    return a, a.ctypes.data_as(C.POINTER(ct))


_live_contexts = weakref.WeakSet()


@atexit.register
def _close_all():
    # runs before torch releases the HIP runtime at interpreter exit
    for c in list(_live_contexts):
        c.close()


class Context:
    """ofh_ctx: device, stream and workspace of the library.  Objects made
    from a context are closed before it."""

    def __init__(self, device=0, use_torch_stream=True):
        L = lib()
        h = C.c_void_p()
        check(L.ofh_ctx_create(device, C.byref(h)))
        self.h = h
        self.device = device
        self._children = weakref.WeakSet()
        _live_contexts.add(self)
        if use_torch_stream:
            import torch
            s = torch.cuda.current_stream(device)
            check(L.ofh_ctx_set_stream(self.h, C.c_void_p(s.cuda_stream)))

    def synchronize(self):
        check(lib().ofh_ctx_synchronize(self.h))

    def profile(self, enable=True):
        """Per-kernel timing of the solver loops on (clears earlier records) or off."""
        check(lib().ofh_ctx_profile(self.h, int(bool(enable))))

    def profile_read(self, kind):
        """Durations (ms) of the recorded launches of one kind (OFH_PROF_*)."""
        cnt = C.c_int()
        check(lib().ofh_ctx_profile_read(self.h, int(kind), None, 0, C.byref(cnt)))
        ms = np.zeros(max(cnt.value, 1))
        check(lib().ofh_ctx_profile_read(self.h, int(kind), ms.ctypes.data_as(C.POINTER(C.c_double)),
                                         cnt.value, C.byref(cnt)))
        return ms[: cnt.value]

    def use_graphs(self, enable=True):
        """Replay chunks of solver iterations from captured graphs (default off)."""
        check(lib().ofh_ctx_use_graphs(self.h, int(bool(enable))))

    def mixed_form(self, form=0):
        """Form of cg_mixed's inner iteration: 0 by size (default), 1 two kernels,
        2 one kernel wherever z = r on a diagonal copy."""
        check(lib().ofh_ctx_mixed_form(self.h, int(form)))

    def close(self):
        if self.h:
            for ch in list(self._children):
                ch.close()
            lib().ofh_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CsrMatrix:
    """Device CSR (SpMatrix storage): row_ptr 0-based, col 0-based sorted."""

    def __init__(self, ctx, n=None, row_ptr=None, col=None, val=None, handle=None, owner=None):
        self.ctx = ctx
        self._own = handle is None
        self._owner = owner
        if self._own:
            ctx._children.add(self)
        if handle is None:
            rp, rpp = _np(row_ptr, np.int32)
            cl, clp = _np(col, np.int32)
            vp = None
            if val is not None:
                v, vp = _np(val, np.float64)
            h = C.c_void_p()
            check(lib().ofh_csr_create(ctx.h, int(n), int(rp[-1]), rpp, clp, vp, C.byref(h)))
            handle = h
        self.h = handle
        n_ = C.c_int()
        nnz = C.c_int64()
        mr = C.c_int()
        check(lib().ofh_csr_info(self.h, C.byref(n_), C.byref(nnz), C.byref(mr)))
        self.n, self.nnz, self.max_row_nnz = n_.value, nnz.value, mr.value

    def compress_columns(self, enable=True):
        """16-bit block-relative column offsets in the SpMV stream (default off)."""
        check(lib().ofh_csr_compress_columns(self.h, int(bool(enable))))

    def sell_info(self):
        """The SELL-64 copy the products stream: {"in_use", "stored", "col_bytes"}."""
        u, st, cb = C.c_int(), C.c_int64(), C.c_int()
        check(lib().ofh_csr_sell_info(self.h, C.byref(u), C.byref(st), C.byref(cb)))
        return {"in_use": bool(u.value), "stored": st.value, "col_bytes": cb.value}

    def dia_info(self):
        """The diagonal copies of the values: {"fp32", "fp64" (in use), "diagonals",
        "row_slots", "symmetric" (the products read the symmetric form)}."""
        u32, u64, nd, w, sym = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(lib().ofh_csr_dia_info(self.h, C.byref(u32), C.byref(u64), C.byref(nd), C.byref(w),
                                     C.byref(sym)))
        return {"fp32": bool(u32.value), "fp64": bool(u64.value), "diagonals": nd.value,
                "row_slots": w.value, "symmetric": bool(sym.value)}

    def pattern(self):
        rp = np.zeros(self.n + 1, np.int32)
        cl = np.zeros(max(self.nnz, 1), np.int32)
        check(lib().ofh_csr_get_pattern(self.h, rp.ctypes.data_as(C.POINTER(C.c_int32)),
                                        cl.ctypes.data_as(C.POINTER(C.c_int32))))
        return rp, cl[: self.nnz]

    def values(self):
        v = np.zeros(max(self.nnz, 1))
        check(lib().ofh_csr_get_values(self.h, v.ctypes.data_as(C.POINTER(C.c_double))))
        return v[: self.nnz]

    def set_values(self, v):
        import torch
        if isinstance(v, torch.Tensor):
            check(lib().ofh_csr_set_values_device(self.h, _f64(v, self.nnz)))
        else:
            a, ap = _np(v, np.float64)
            assert a.size == self.nnz
            check(lib().ofh_csr_set_values(self.h, ap))

    def close(self):
        if self._own and self.h:
            lib().ofh_csr_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def spmv(A, x, y):
    """y = A x (SpMatrix::Mult)."""
    check(lib().ofh_spmv_f64(A.h, _f64(x, A.n), _f64(y, A.n)))


def spmv_add(A, a, x, y):
    """y += a A x (SpMatrix::MultAdd)."""
    check(lib().ofh_spmv_add_f64(A.h, float(a), _f64(x, A.n), _f64(y, A.n)))


def spmv_t(A, x, y):
    """y = A^T x (SpMatrix::TMult as BiCG uses it; y is overwritten)."""
    check(lib().ofh_spmv_t_f64(A.h, _f64(x, A.n), _f64(y, A.n)))


def spmv_f32(A, x, y):
    check(lib().ofh_spmv_f32(A.h, _ptr(x), _ptr(y)))


def dot(ctx, x, y):
    out = C.c_double()
    check(lib().ofh_dot_f64(ctx.h, x.numel(), _f64(x), _f64(y), C.byref(out)))
    return out.value


def norm2(ctx, x):
    out = C.c_double()
    check(lib().ofh_norm2_f64(ctx.h, x.numel(), _f64(x), C.byref(out)))
    return out.value


def axpy(ctx, a, x, y):
    """y = a x + y (Vect += a*x)."""
    check(lib().ofh_axpy_f64(ctx.h, y.numel(), float(a), _f64(x, y.numel()), _f64(y)))


def prec_apply(A, prec, b, x):
    check(lib().ofh_prec_apply(A.h, prec, _f64(b, A.n), _f64(x, A.n)))


def prec_trans_apply(A, prec, b, x):
    """x of Prec::TransSolve(b,x): DIAG scales, ILU solves M^T x = b,
    IDENT/DILU/SSOR copy b."""
    check(lib().ofh_prec_trans_apply(A.h, prec, _f64(b, A.n), _f64(x, A.n)))


def ssor_solve(A, b, x):
    check(lib().ofh_ssor_solve(A.h, _f64(b, A.n), _f64(x, A.n)))


def _solve(fn, A, b, x, *args):
    info = SolveInfo()
    check(getattr(lib(), fn)(A.h, *args[:1], _f64(b, A.n), _f64(x, A.n), *args[1:],
                             C.byref(info)))
    return info


def cg(A, b, x, prec=DIAG_PREC, max_it=1000, toler=1.4901161193847656e-08):
    """CG (CG.h:99-157); x is the initial guess and receives the solution."""
    return _solve("ofh_cg", A, b, x, prec, int(max_it), float(toler))


def bicgstab(A, b, x, prec=DIAG_PREC, max_it=1000, toler=1.4901161193847656e-08):
    return _solve("ofh_bicgstab", A, b, x, prec, int(max_it), float(toler))


def cgs(A, b, x, prec=DIAG_PREC, max_it=1000, toler=1.4901161193847656e-08):
    """CGS (CGS.h:91-171) as the reference writes it (q stays 0)."""
    return _solve("ofh_cgs", A, b, x, prec, int(max_it), float(toler))


def bicg(A, b, x, prec=DIAG_PREC, max_it=1000, toler=1.4901161193847656e-08):
    """BiCG (BiCG.h:94-170) with the true transposed products."""
    return _solve("ofh_bicg", A, b, x, prec, int(max_it), float(toler))


def gmres(A, b, x, m, prec=DIAG_PREC, max_it=1000, toler=1.4901161193847656e-08):
    return _solve("ofh_gmres", A, b, x, prec, int(m), int(max_it), float(toler))


def cg_mixed(A, b, x, prec=DIAG_PREC, max_it=1000, toler=1.4901161193847656e-08):
    return _solve("ofh_cg_mixed", A, b, x, prec, int(max_it), float(toler))


def linear_solve(A, b, x, solver, prec=DIAG_PREC, max_it=1000, toler=1.4901161193847656e-08):
    info = SolveInfo()
    check(lib().ofh_linear_solve(A.h, int(solver), int(prec), _f64(b, A.n), _f64(x, A.n),
                                 int(max_it), float(toler), C.byref(info)))
    return info


# ---------------------------------------------------------------- meshes
def mesh2d(nx, ny, quad=False, nb_dof=1, codes=(1, 1, 1, 1), box=(0.0, 1.0, 0.0, 1.0)):
    """Structured rectangle (Mesh.cpp:547-676): coord, code, elem (1-based)."""
    nn = (nx + 1) * (ny + 1)
    ne = nx * ny if quad else 2 * nx * ny
    coord = np.zeros((nn, 2))
    code = np.zeros((nn, nb_dof), np.int32)
    elem = np.zeros((ne, 4 if quad else 3), np.int32)
    check(lib().ofh_mesh2d_structured(*map(float, box), nx, ny, *codes, int(quad), nb_dof,
                                      coord.ctypes.data_as(C.POINTER(C.c_double)),
                                      code.ctypes.data_as(C.POINTER(C.c_int32)),
                                      elem.ctypes.data_as(C.POINTER(C.c_int32))))
    return coord, code, elem


class MeshFile:
    """An OFELI XML mesh (.m) as read by Mesh::get(file): coord (nn x dim),
    code (nn x nb_dof), elem (ne x nen, 1-based), elem_code, sides
    (ns x nsn, 1-based), side_code (ns x nb_dof), shapes as ElementShape
    values (LINE 2, TRIANGLE 3, QUADRILATERAL 4, TETRAHEDRON 5, ...) and
    the (code, name) materials."""

    def __init__(self, dim, nb_dof, coord, code, elem, elem_code, el_shape, sides, side_code,
                 sd_shape, materials=()):
        self.dim, self.nb_dof = int(dim), int(nb_dof)
        self.coord, self.code, self.elem, self.elem_code = coord, code, elem, elem_code
        self.sides, self.side_code = sides, side_code
        self.el_shape, self.sd_shape = int(el_shape), int(sd_shape)
        self.materials = list(materials)


def _i32(a):
    return np.ascontiguousarray(a, np.int32)


def read_mesh(path):
    """Mesh::get(file) for the OFELI XML format (C++ reader)."""
    h = C.c_void_p()
    check(lib().ofh_mesh_read(os.fsencode(path), C.byref(h)))
    try:
        v = [C.c_int() for _ in range(10)]
        check(lib().ofh_mesh_info(h, *[C.byref(x) for x in v]))
        dim, nb_dof, nn, ne, el_shape, nen, ns, sd_shape, nsn, nm = [x.value for x in v]
        coord = np.zeros((nn, dim))
        code = np.zeros((nn, nb_dof), np.int32)
        elem = np.zeros((ne, nen), np.int32)
        elem_code = np.zeros(ne, np.int32)
        sides = np.zeros((ns, max(nsn, 1)), np.int32)
        side_code = np.zeros((ns, nb_dof), np.int32)
        ptr = lambda a, t: a.ctypes.data_as(C.POINTER(t))
        check(lib().ofh_mesh_get(h, ptr(coord, C.c_double), coord.size,
                                 ptr(code, C.c_int32), code.size,
                                 ptr(elem, C.c_int32), elem.size,
                                 ptr(elem_code, C.c_int32), elem_code.size,
                                 ptr(sides, C.c_int32), sides.size,
                                 ptr(side_code, C.c_int32), side_code.size))
        mats = []
        for i in range(nm):
            c, name = C.c_int(), C.create_string_buffer(256)
            check(lib().ofh_mesh_material(h, i, C.byref(c), name, 256))
            mats.append((c.value, name.value.decode()))
    finally:
        lib().ofh_mesh_destroy(h)
    return MeshFile(dim, nb_dof, coord, code, elem, elem_code, el_shape, sides[:, :nsn], side_code,
                    sd_shape, mats)


def write_mesh(path, m):
    """Mesh::put(file), byte for byte (C++ writer)."""
    coord = np.ascontiguousarray(m.coord, np.float64)
    code, elem, ec = _i32(m.code), _i32(m.elem), _i32(m.elem_code)
    sides, sc = _i32(m.sides), _i32(m.side_code)
    mc = _i32([c for c, _ in m.materials])
    names = (C.c_char_p * max(len(m.materials), 1))(*[n.encode() for _, n in m.materials])
    ptr = lambda a, t: a.ctypes.data_as(C.POINTER(t))
    nsn = sides.shape[1] if sides.ndim == 2 and sides.shape[0] else 0
    check(lib().ofh_mesh_write(os.fsencode(path), m.dim, m.nb_dof, coord.shape[0],
                               ptr(coord, C.c_double), ptr(code, C.c_int32), elem.shape[0],
                               m.el_shape, elem.shape[1] if elem.ndim == 2 else 0,
                               ptr(elem, C.c_int32), ptr(ec, C.c_int32), sides.shape[0],
                               m.sd_shape, nsn, ptr(sides, C.c_int32), ptr(sc, C.c_int32),
                               len(m.materials), ptr(mc, C.c_int32), names))


def mesh3d_kuhn(nx, ny, nz, codes=(1, 1, 1, 1, 1, 1), box=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0)):
    nn = (nx + 1) * (ny + 1) * (nz + 1)
    coord = np.zeros((nn, 3))
    code = np.zeros((nn, 1), np.int32)
    elem = np.zeros((6 * nx * ny * nz, 4), np.int32)
    check(lib().ofh_mesh3d_kuhn(*map(float, box), nx, ny, nz, *codes,
                                coord.ctypes.data_as(C.POINTER(C.c_double)),
                                code.ctypes.data_as(C.POINTER(C.c_int32)),
                                elem.ctypes.data_as(C.POINTER(C.c_int32))))
    return coord, code, elem


def number_equations(code):
    code = np.ascontiguousarray(code, np.int32)
    dof = np.zeros_like(code)
    n = C.c_int()
    check(lib().ofh_number_equations(code.shape[0], code.shape[1],
                                     code.ctypes.data_as(C.POINTER(C.c_int32)),
                                     dof.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(n)))
    return dof, n.value


def gauss(np_):
    x = np.zeros(np_)
    w = np.zeros(np_)
    check(lib().ofh_gauss(np_, x.ctypes.data_as(C.POINTER(C.c_double)),
                          w.ctypes.data_as(C.POINTER(C.c_double))))
    return x, w


class AssemblyPlan:
    """Mesh on the device + CSR pattern + gather map (equation build)."""

    def __init__(self, ctx, coord, elem, code):
        coord = np.ascontiguousarray(coord, np.float64)
        elem = np.ascontiguousarray(elem, np.int32)
        code = np.ascontiguousarray(code, np.int32)
        self.ctx = ctx
        ctx._children.add(self)
        self.nb_nodes, self.dim = coord.shape
        self.nb_el, self.nen = elem.shape
        self.nb_dof = code.shape[1]
        h = C.c_void_p()
        check(lib().ofh_asm_create(ctx.h, self.dim, self.nb_nodes,
                                   coord.ctypes.data_as(C.POINTER(C.c_double)), self.nb_el,
                                   self.nen, elem.ctypes.data_as(C.POINTER(C.c_int32)),
                                   self.nb_dof, code.ctypes.data_as(C.POINTER(C.c_int32)),
                                   C.byref(h)))
        self.h = h
        neq = C.c_int()
        nnz = C.c_int64()
        check(lib().ofh_asm_info(self.h, C.byref(neq), C.byref(nnz)))
        self.nb_eq, self.nnz = neq.value, nnz.value
        a = C.c_void_p()
        check(lib().ofh_asm_matrix(self.h, C.byref(a)))
        self.A = CsrMatrix(ctx, handle=a, owner=self)

    def dof(self):
        d = np.zeros((self.nb_nodes, self.nb_dof), np.int32)
        check(lib().ofh_asm_get_dof(self.h, d.ctypes.data_as(C.POINTER(C.c_int32))))
        return d

    def build(self, kind, coef=None, f=None, bc=None, u_prev=None, prm=(), matrix=True, b=None):
        p = np.ascontiguousarray(prm if len(prm) else [0.0], np.float64)
        check(lib().ofh_asm_build(self.h, int(kind), _ptr(coef), _ptr(f), _ptr(bc), _ptr(u_prev),
                                  p.ctypes.data_as(C.POINTER(C.c_double)), len(prm),
                                  int(bool(matrix)), _ptr(b)))

    def traction(self, sides, traction, b):
        s = np.ascontiguousarray(sides, np.int32)
        check(lib().ofh_asm_traction(self.h, s.shape[0], s.ctypes.data_as(C.POINTER(C.c_int32)),
                                     _ptr(traction), _ptr(b)))

    def insert_bc(self, v, bc, u):
        check(lib().ofh_asm_insert_bc(self.h, _ptr(v), _ptr(bc), _ptr(u)))

    def close(self):
        if self.h:
            self.A.h = None
            lib().ofh_asm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Independent systems = the diagonal blocks of one CSR, solved together
    by the persistent batched CG (ofh_batch_*)."""

    def __init__(self, A, sys_rows):
        sr = np.ascontiguousarray(sys_rows, np.int32)
        self.A = A
        self.nsys = len(sr) - 1
        self.sys_rows = sr
        h = C.c_void_p()
        check(lib().ofh_batch_create(A.h, self.nsys, sr.ctypes.data_as(C.POINTER(C.c_int32)),
                                     C.byref(h)))
        self.h = h
        A.ctx._children.add(self)

    def cg(self, b, x, prec=DIAG_PREC, mixed=False, max_it=1000, toler=1e-8):
        it = np.zeros(self.nsys, np.int32)
        st = np.zeros(self.nsys, np.int32)
        rr = np.zeros(self.nsys)
        info = SolveInfo()
        check(lib().ofh_batch_cg(self.h, int(prec), int(bool(mixed)), _f64(b, self.A.n),
                                 _f64(x, self.A.n), int(max_it), float(toler),
                                 it.ctypes.data_as(C.POINTER(C.c_int32)),
                                 st.ctypes.data_as(C.POINTER(C.c_int32)),
                                 rr.ctypes.data_as(C.POINTER(C.c_double)), C.byref(info)))
        return it, st, rr, info

    def set_config(self, block, tpr):
        check(lib().ofh_batch_set_config(self.h, int(block), int(tpr)))

    def use_graph(self, enable=True):
        check(lib().ofh_batch_use_graph(self.h, int(bool(enable))))

    def config(self):
        b, t = C.c_int(), C.c_int()
        check(lib().ofh_batch_get_config(self.h, C.byref(b), C.byref(t)))
        return b.value, t.value

    def autotune(self, b, x, prec=DIAG_PREC, mixed=False, max_it=1000, toler=1e-8):
        bb, bt, ms = C.c_int(), C.c_int(), C.c_double()
        check(lib().ofh_batch_autotune(self.h, int(prec), int(bool(mixed)), _f64(b, self.A.n),
                                       _f64(x, self.A.n), int(max_it), float(toler),
                                       C.byref(bb), C.byref(bt), C.byref(ms)))
        return bb.value, bt.value, ms.value

    def close(self):
        if self.h:
            lib().ofh_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ProjectionGuess:
    """Start vectors from the span of previous solutions (ofh_proj_*): the
    A-norm-closest vector (minres=False) or the one of least residual
    (minres=True)."""

    def __init__(self, A, m=16, minres=False):
        self.A = A
        h = C.c_void_p()
        check(lib().ofh_proj_create(A.h, int(m), int(bool(minres)), C.byref(h)))
        self.h = h
        self.m = m
        A.ctx._children.add(self)

    def guess(self, b, x0):
        check(lib().ofh_proj_guess(self.h, _f64(b, self.A.n), _f64(x0, self.A.n)))

    def update(self, x):
        check(lib().ofh_proj_update(self.h, _f64(x, self.A.n)))

    def reset(self):
        check(lib().ofh_proj_reset(self.h))

    def size(self):
        c = C.c_int()
        check(lib().ofh_proj_size(self.h, C.byref(c)))
        return c.value

    def close(self):
        if self.h:
            lib().ofh_proj_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_with_projection(A, b, x, proj, fn=None, prec=DIAG_PREC, toler=1e-8, tighten=1e-2,
                          max_it=200000, guess=True):
    """One solve of a sequence sharing A: start from the projection of b on
    the previous solutions (x holds it on return of proj.guess); if that start
    already meets toler the solve ends at once; otherwise the system is solved
    to toler*tighten and the solution joins the projection basis (stored
    solutions carry residuals far below toler, so later combinations of them
    pass the test without iterations).  Returns the SolveInfo of the last
    solver call."""
    fn = fn or cg
    if guess:
        proj.guess(b, x)
    info = fn(A, b, x, prec=prec, max_it=0, toler=toler)
    if info.nb_it == 0:
        return info
    info = fn(A, b, x, prec=prec, max_it=max_it, toler=toler * tighten)
    proj.update(x)
    return info


class LearnedGuess:
    """Message-passing network predicting the start vector of the next solve
    of a sequence A x_{k+1} = b_{k+1} (ofh_mpnn_*); A symmetric.
    objective "energy": A-norm line search and loss (needs the solution x*);
    "residual": minimum-residual line search and loss (x* not used)."""

    def __init__(self, A, layers=4, hidden=64, seed=0, extrapolate=False, objective="energy"):
        if objective not in ("energy", "residual"):
            raise ValueError("objective is 'energy' or 'residual'")
        self.A = A
        self.layers = layers
        h = C.c_void_p()
        check(lib().ofh_mpnn_create(A.h, int(layers), int(hidden), int(bool(extrapolate)),
                                    C.c_uint64(seed), C.byref(h)))
        self.h = h
        self.objective = objective
        check(lib().ofh_mpnn_set_objective(self.h, int(objective == "residual")))
        cnt = C.c_int64()
        check(lib().ofh_mpnn_num_params(self.h, C.byref(cnt)))
        self.nparams = cnt.value
        A.ctx._children.add(self)

    def refresh(self):
        """Re-read the graph weights from A's current values (same pattern)."""
        check(lib().ofh_mpnn_refresh(self.h))

    def params(self):
        p = np.zeros(self.nparams, np.float32)
        check(lib().ofh_mpnn_get_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.nparams:
            raise ValueError("expected %d parameters" % self.nparams)
        check(lib().ofh_mpnn_set_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))))

    def use_graph(self, enable=True):
        check(lib().ofh_mpnn_use_graph(self.h, int(bool(enable))))

    def guess(self, x_n, x_p, b, x0):
        n = self.A.n
        check(lib().ofh_mpnn_guess(self.h, _f64(x_n, n), _f64(x_p, n), _f64(b, n), _f64(x0, n)))

    def loss_grad(self, x_n, x_p, b, x_star, x0, want_grad=True):
        n = self.A.n
        loss = C.c_double()
        g = np.zeros(self.nparams, np.float32) if want_grad else None
        check(lib().ofh_mpnn_loss_grad(self.h, _f64(x_n, n), _f64(x_p, n), _f64(b, n),
                                       None if x_star is None else _f64(x_star, n), _f64(x0, n),
                                       C.byref(loss),
                                       g.ctypes.data_as(C.POINTER(C.c_float)) if want_grad else None))
        return loss.value, g

    def train_step(self, x_n, x_p, b, x_star, x0, lr=1e-3, want_loss=True):
        n = self.A.n
        loss = C.c_double()
        check(lib().ofh_mpnn_train_step(self.h, _f64(x_n, n), _f64(x_p, n), _f64(b, n),
                                        None if x_star is None else _f64(x_star, n), _f64(x0, n),
                                        float(lr),
                                        C.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def close(self):
        if self.h:
            lib().ofh_mpnn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def concat_meshes(meshes):
    """Disjoint union of (coord, code, elem) meshes, node ids renumbered; the
    free DOFs of mesh s stay contiguous, so its equations are rows
    sys_rows[s]..sys_rows[s+1] of the plan matrix."""
    coords, codes, elems, sys_rows = [], [], [], [0]
    off = 0
    for coord, code, elem in meshes:
        coords.append(coord)
        codes.append(code)
        elems.append(elem + off)
        off += coord.shape[0]
        sys_rows.append(sys_rows[-1] + int((code == 0).sum()))
    return (np.concatenate(coords), np.concatenate(codes), np.concatenate(elems),
            np.array(sys_rows, np.int32))
