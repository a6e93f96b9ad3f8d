"""ofeli_amd -- MI355X-native sparse linear-solve path of OFELI.

This package is a thin ctypes front end of libofeli_hip.so (C-ABI declared in
include/ofeli_hip.h).  Device buffers are torch tensors on the GPU (torch is
only the allocator and stream provider); every computation runs in the HIP
kernels of libofeli_hip.  If the library is missing the import of the
native layer fails loudly: there is no CPU fallback.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libofeli_hip.so")

# enums (Matrix.h:75-93)
DIRECT_SOLVER, CG_SOLVER, CGS_SOLVER, BICG_SOLVER, BICG_STAB_SOLVER, GMRES_SOLVER = 1, 2, 3, 4, 5, 6
IDENT_PREC, DIAG_PREC, DILU_PREC, ILU_PREC, SSOR_PREC = 1, 2, 3, 4, 5
LAPLACE2DT3, DC2DT3, DC2DT3_BE, LAPLACE3DT4, ELAS2DQ4 = 1, 2, 3, 4, 5
# per-kernel timing kinds (ofh_ctx_profile)
(PROF_MX_DIRSPMV, PROF_MX_UPD, PROF_CG_DIRSPMV, PROF_CG_UPD, PROF_CG_SPMV, PROF_CG_DIR,
 PROF_SPMV64, PROF_SPMV32, PROF_MX_ITER) = range(9)

OFH_OK = 0
ERRORS = {-1: "OFH_E_INVALID", -2: "OFH_E_HIP", -3: "OFH_E_ZERO_PIVOT", -4: "OFH_E_NEGDET",
          -5: "OFH_E_PATTERN", -6: "OFH_E_NOMEM", -7: "OFH_E_UNSUPPORTED", -8: "OFH_E_ZERODET"}


class OfhError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "?"), code, msg))
        self.code = code


class SolveInfo(C.Structure):
    _fields_ = [("nb_it", C.c_int), ("iterations", C.c_int), ("relres", C.c_double),
                ("true_relres", C.c_double), ("outer", C.c_int), ("fallback", C.c_int),
                ("seconds", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """Load libofeli_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libofeli_hip.so not built: run `python ofeli_amd/build.py` "
                          "(or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    from . import _abi
    _abi.declare(L)
    _lib = L
    return L


def check(rc):
    if rc != OFH_OK:
        raise OfhError(rc, lib().ofh_last_error().decode())
    return rc


from .api import (Context, CsrMatrix, AssemblyPlan, Batch, LearnedGuess, ProjectionGuess,  # noqa: E402
                  concat_meshes, solve_with_projection, cg, bicgstab, gmres, linear_solve,  # noqa: E402
                  cgs, bicg, cg_mixed, spmv, spmv_add, spmv_t, spmv_f32, prec_apply,  # noqa: E402
                  prec_trans_apply, ssor_solve, dot, norm2, axpy, MeshFile, read_mesh,  # noqa: E402
                  write_mesh)
