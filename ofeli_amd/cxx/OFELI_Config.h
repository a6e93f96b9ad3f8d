// OFELI_Config.h -- types, enums, globals and the exception type of the
// drop-in C++ layer (counterparts: include/OFELI_Config.h:51,374;
// include/OFELIException.h; include/linear_algebra/Matrix.h:75-93).
//
// The drop-in layer keeps OFELI's host API (namespace OFELI, 1-based element
// access, return conventions) and runs every computation of the solve path
// through libofeli_hip (include/ofeli_hip.h) on the GPU.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "ofeli_hip.h"

namespace OFELI {

typedef double real_t;

// Verbosity global of OFELI_Config.h:374 (0: silent)
inline int Verbosity = 0;

class OFELIException : public std::runtime_error {
 public:
  explicit OFELIException(const std::string& s) : std::runtime_error(s) {}
};

enum Iteration {
  DIRECT_SOLVER = OFH_DIRECT_SOLVER,
  CG_SOLVER = OFH_CG_SOLVER,
  CGS_SOLVER = OFH_CGS_SOLVER,
  BICG_SOLVER = OFH_BICG_SOLVER,
  BICG_STAB_SOLVER = OFH_BICG_STAB_SOLVER,
  GMRES_SOLVER = OFH_GMRES_SOLVER
};

enum Preconditioner {
  IDENT_PREC = OFH_IDENT_PREC,
  DIAG_PREC = OFH_DIAG_PREC,
  DILU_PREC = OFH_DILU_PREC,
  ILU_PREC = OFH_ILU_PREC,
  SSOR_PREC = OFH_SSOR_PREC
};

// Mesh element shapes used by the structured constructors (Mesh.cpp:547-836)
enum ElementShape { LINE = 1, TRIANGLE = 3, QUADRILATERAL = 4, TETRAHEDRON = 5 };

enum TimeScheme { NONE = 0, FORWARD_EULER = 1, BACKWARD_EULER = 2, CRANK_NICOLSON = 3 };

// Throws OFELIException with the library's message on a failed C-ABI call.
inline void hip_check(int rc, const char* where) {
  if (rc != OFH_OK)
    throw OFELIException(std::string("In ") + where + ": " + ofh_last_error());
}

// Process-wide device context (device 0), created on first use.  OFELI's
// API has no handle to carry one, exactly as it keeps Verbosity global.
inline ofh_ctx hip_context() {
  static ofh_ctx ctx = [] {
    ofh_ctx c = nullptr;
    hip_check(ofh_ctx_create(0, &c), "hip_context()");
    return c;
  }();
  return ctx;
}

// RAII device buffer of doubles.
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t n) { resize(n); }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  ~DeviceBuffer() { release(); }
  void resize(size_t n) {
    if (n <= n_) return;
    release();
    void* p = nullptr;
    hip_check(ofh_malloc(hip_context(), n * sizeof(double), &p), "DeviceBuffer");
    p_ = static_cast<double*>(p);
    n_ = n;
  }
  void upload(const double* h, size_t n) {
    resize(n);
    if (n) hip_check(ofh_memcpy_h2d(hip_context(), p_, h, n * sizeof(double)), "upload");
  }
  void download(double* h, size_t n) const {
    if (n) hip_check(ofh_memcpy_d2h(hip_context(), h, p_, n * sizeof(double)), "download");
  }
  double* get() const { return p_; }
  size_t size() const { return n_; }

 private:
  void release() {
    if (p_) ofh_free(hip_context(), p_);
    p_ = nullptr;
    n_ = 0;
  }
  double* p_ = nullptr;
  size_t n_ = 0;
};

}  // namespace OFELI
