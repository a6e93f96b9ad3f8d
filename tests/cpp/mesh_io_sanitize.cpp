// mesh_io_sanitize.cpp -- drives the library's host-side mesh code (the .m
// reader and writer, structured generators, equation numbering) under
// AddressSanitizer and UndefinedBehaviorSanitizer.  Built with g++ from
// ofeli_amd/csrc/ofh_mesh_io.cpp and ofh_mesh.cpp only (no GPU code); run by
// tests/test_sanitize_host.py.  Usage: mesh_io_sanitize <file.m>...; every
// file is read, queried with exactly sized buffers, written back and read
// again; files the reader rejects are reported and skipped.  Exit status 0
// unless a sanitizer fires or a round trip differs.
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "ofeli_hip.h"

static int round_trip(const char* path, const std::string& tmp) {
  ofh_mesh m = nullptr;
  if (ofh_mesh_read(path, &m) != OFH_OK) {
    std::printf("rejected %s\n", path);
    return 0;
  }
  int dim, nb_dof, nn, ne, es, nen, ns, ss, nsn, nm;
  ofh_mesh_info(m, &dim, &nb_dof, &nn, &ne, &es, &nen, &ns, &ss, &nsn, &nm);
  std::vector<double> coord((size_t)nn * dim);
  std::vector<int32_t> code((size_t)nn * nb_dof), elem((size_t)ne * nen), ecode(ne),
      side((size_t)ns * nsn), scode((size_t)ns * nb_dof);
  int rc = ofh_mesh_get(m, coord.data(), (int64_t)coord.size(), code.data(), (int64_t)code.size(),
                        elem.data(), (int64_t)elem.size(), ecode.data(), (int64_t)ecode.size(),
                        side.data(), (int64_t)side.size(), scode.data(), (int64_t)scode.size());
  if (rc != OFH_OK) {
    std::printf("get failed on %s\n", path);
    ofh_mesh_destroy(m);
    return 1;
  }
  std::vector<int32_t> mcode(nm);
  std::vector<std::string> names(nm);
  std::vector<const char*> pnames(nm);
  for (int i = 0; i < nm; ++i) {
    char buf[256];
    ofh_mesh_material(m, i, &mcode[i], buf, (int)sizeof(buf));
    names[i] = buf;
    pnames[i] = names[i].c_str();
  }
  ofh_mesh_destroy(m);
  rc = ofh_mesh_write(tmp.c_str(), dim, nb_dof, nn, coord.data(), code.data(), ne, es, nen,
                      elem.data(), ecode.data(), ns, ss, nsn, side.data(), scode.data(), nm,
                      mcode.data(), pnames.data());
  if (rc != OFH_OK) {
    std::printf("write failed on %s\n", path);
    return 1;
  }
  ofh_mesh m2 = nullptr;
  if (ofh_mesh_read(tmp.c_str(), &m2) != OFH_OK) {
    std::printf("re-read failed on %s\n", path);
    return 1;
  }
  int nn2, ne2;
  ofh_mesh_info(m2, nullptr, nullptr, &nn2, &ne2, nullptr, nullptr, nullptr, nullptr, nullptr,
                nullptr);
  ofh_mesh_destroy(m2);
  std::printf("ok %s %d %d\n", path, nn, ne);
  return (nn2 == nn && ne2 == ne) ? 0 : 1;
}

int main(int argc, char** argv) {
  int bad = 0;
  const std::string tmp = std::string(argc > 1 ? argv[1] : "x") + ".sanitize.tmp";
  for (int i = 1; i < argc; ++i) bad |= round_trip(argv[i], tmp);
  // structured generators and numbering at odd sizes
  for (int nx : {1, 3, 7})
    for (int quad : {0, 1}) {
      const int ny = nx + 1, nn = (nx + 1) * (ny + 1), ne = quad ? nx * ny : 2 * nx * ny;
      std::vector<double> c(2 * nn);
      std::vector<int32_t> code(2 * nn), el(ne * (quad ? 4 : 3)), dof(2 * nn);
      bad |= ofh_mesh2d_structured(0, 1, 0, 1, nx, ny, 1, 0, 2, 0, quad, 2, c.data(), code.data(),
                                   el.data()) != OFH_OK;
      int neq = 0;
      bad |= ofh_number_equations(nn, 2, code.data(), dof.data(), &neq) != OFH_OK;
    }
  {
    const int nx = 3, ny = 2, nz = 4, nn = 4 * 3 * 5;
    std::vector<double> c(3 * nn);
    std::vector<int32_t> code(nn), el(6 * nx * ny * nz * 4);
    bad |= ofh_mesh3d_kuhn(0, 1, 0, 1, 0, 1, nx, ny, nz, 1, 1, 1, 1, 1, 1, c.data(), code.data(),
                           el.data()) != OFH_OK;
  }
  std::remove(tmp.c_str());
  return bad;
}
