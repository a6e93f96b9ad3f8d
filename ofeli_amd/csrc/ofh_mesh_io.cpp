// ofh_mesh_io.cpp -- OFELI's XML mesh files (.m): reader and writer.
//
// Reading follows Mesh::get(file) (Mesh.cpp:2639-2670) through XMLParser
// (read_mesh, XMLParser.cpp:1175-1220; read_mesh_data, :2210-2293): the text
// of the <Nodes>, <Elements>, <Sides> and <Material> sections of <Mesh> is
// split at white space and consumed as records of
//   nodes:    dim coordinates (stof: rounded to single precision), a mark
//   elements: `nodes` node labels, a code
//   sides:    `nodes` node labels, a mark
//   material: a code, a name,
// and marks become per-DOF codes by DOFCode (MeshUtil.cpp:101-116): the
// decimal digits of the mark, one per DOF, times the mark's sign applied to
// a quotient of the signed mark -- so a negative mark yields the digits of
// its absolute value, positive.  Without element/side attributes the shapes
// default by dimension: line/2; triangle/3 with line/2 sides; tetrahedron/4
// with triangle/3 sides.
//
// Writing is Mesh::put (Mesh.cpp:2734-2800), byte for byte: coordinates as
// "  %18.8e", the node mark sum_j |code_j| 10^(n-j) (negated, as a size_t, when
// code_1 < 0) as "%10zu", element rows "%8d" x (nodes + code), side rows
// "%8d" x nodes and the mark sum_j code_j 10^(n-j) (size_t) as "%6zu".
#include <cerrno>
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ofeli_hip.h"

struct ofh_mesh_s {
  int dim = 2, nb_dof = 1;
  int el_shape = 0, nen = 0, sd_shape = 0, nsn = 0;
  std::vector<double> coord;
  std::vector<int32_t> code, elem, elem_code, side, side_code, mat_code;
  std::vector<std::string> mat_name;
};

namespace {

const char* const kShapeName[] = {"none", "point", "line", "triangle", "quadrilateral",
                                  "tetrahedron", "hexahedron", "pentahedron"};
constexpr int kNbShapes = 8;

int shape_of(const std::string& s) {
  for (int k = 0; k < kNbShapes; ++k)
    if (s == kShapeName[k]) return k;
  return -1;
}

// DOFCode, MeshUtil.cpp:101-116
void dof_code(int mark, int nb_dof, int32_t* code) {
  int j = mark, sign = 1;
  if (mark < 0) sign = -1;
  for (int k = 0; k < nb_dof; ++k) {
    const int kk = (int)std::pow(10., (double)(nb_dof - k - 1));
    const int m = j / kk;
    code[k] = m * sign;
    j -= m * kk;
  }
}

bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\f' || c == '\v'; }

// next white-space separated token of [p, e) into [tb, te); false at the end
bool next_token(const char*& p, const char* e, const char*& tb, const char*& te) {
  while (p < e && is_space(*p)) ++p;
  if (p == e) return false;
  tb = p;
  while (p < e && !is_space(*p)) ++p;
  te = p;
  return true;
}

// stoi / stof of one token: the longest numeric prefix, an error when there
// is none or the value is out of range.  A token that from_chars reads whole
// gets the same correctly rounded value as strtof; any other token (a leading
// '+', a hexadecimal float, trailing characters) goes through the C library.
bool to_int(const char* b, const char* e, int& v) {
  const auto r = std::from_chars(b, e, v);
  if (r.ec == std::errc() && r.ptr == e) return true;
  if (r.ec == std::errc::result_out_of_range) return false;
  char buf[64];
  const size_t n = (size_t)(e - b);
  if (n == 0 || n >= sizeof(buf)) return false;
  std::memcpy(buf, b, n);
  buf[n] = 0;
  char* q;
  errno = 0;
  const long x = std::strtol(buf, &q, 10);
  if (q == buf || errno == ERANGE || x < INT32_MIN || x > INT32_MAX) return false;
  v = (int)x;
  return true;
}

bool to_float(const char* b, const char* e, double& v) {
  float f;
  const auto r = std::from_chars(b, e, f);
  if (r.ec == std::errc() && r.ptr == e) {
    v = (double)f;
    return true;
  }
  if (r.ec == std::errc::result_out_of_range) return false;
  char buf[64];
  const size_t n = (size_t)(e - b);
  if (n == 0 || n >= sizeof(buf)) return false;
  std::memcpy(buf, b, n);
  buf[n] = 0;
  char* q;
  errno = 0;
  const float x = std::strtof(buf, &q);
  if (q == buf || errno == ERANGE) return false;
  v = (double)x;
  return true;
}

// attribute name="value" pairs of a start tag's text
std::vector<std::pair<std::string, std::string>> attributes(const char* b, const char* e) {
  std::vector<std::pair<std::string, std::string>> out;
  const char* p = b;
  while (p < e && !is_space(*p) && *p != '/' && *p != '>') ++p;  // tag name
  while (p < e) {
    while (p < e && (is_space(*p) || *p == '/')) ++p;
    const char* nb = p;
    while (p < e && *p != '=' && !is_space(*p)) ++p;
    std::string name(nb, p);
    while (p < e && (is_space(*p) || *p == '=')) ++p;
    if (p >= e || (*p != '"' && *p != '\'')) break;
    const char q = *p++;
    const char* vb = p;
    while (p < e && *p != q) ++p;
    out.emplace_back(name, std::string(vb, p));
    if (p < e) ++p;
  }
  return out;
}

enum Section { kNone, kNodes, kElements, kSides, kMaterial };

bool read_section(ofh_mesh_s& m, Section sec, const char* b, const char* e, std::string& err) {
  const char *tb, *te;
  const char* p = b;
  auto bad = [&](const char* what) {
    err = std::string("mesh file: bad or incomplete ") + what + " record";
    return false;
  };
  int32_t code[16];
  while (next_token(p, e, tb, te)) {
    if (sec == kNodes) {
      double x[3] = {0, 0, 0};
      for (int d = 0; d < m.dim; ++d) {
        if (d > 0 && !next_token(p, e, tb, te)) return bad("node");
        if (!to_float(tb, te, x[d])) return bad("node");
      }
      int mark;
      if (!next_token(p, e, tb, te) || !to_int(tb, te, mark)) return bad("node");
      for (int d = 0; d < m.dim; ++d) m.coord.push_back(x[d]);
      dof_code(mark, m.nb_dof, code);
      m.code.insert(m.code.end(), code, code + m.nb_dof);
    } else if (sec == kElements || sec == kSides) {
      const int nn = sec == kElements ? m.nen : m.nsn;
      std::vector<int32_t>& ids = sec == kElements ? m.elem : m.side;
      for (int k = 0; k < nn; ++k) {
        int v;
        if ((k > 0 && !next_token(p, e, tb, te)) || !to_int(tb, te, v))
          return bad(sec == kElements ? "element" : "side");
        ids.push_back(v);
      }
      int c;
      if (!next_token(p, e, tb, te) || !to_int(tb, te, c))
        return bad(sec == kElements ? "element" : "side");
      if (sec == kElements) {
        m.elem_code.push_back(c);
      } else {
        dof_code(c, m.nb_dof, code);
        m.side_code.insert(m.side_code.end(), code, code + m.nb_dof);
      }
    } else if (sec == kMaterial) {
      int c;
      if (!to_int(tb, te, c) || !next_token(p, e, tb, te)) return bad("material");
      m.mat_code.push_back(c);
      m.mat_name.emplace_back(tb, te);
    }
  }
  return true;
}

bool parse(const std::string& text, ofh_mesh_s& m, std::string& err) {
  const char* p = text.data();
  const char* e = p + text.size();
  bool in_mesh = false, seen_mesh = false;
  Section sec = kNone;
  const char* cdata = nullptr;
  while (p < e) {
    const char* lt = (const char*)std::memchr(p, '<', (size_t)(e - p));
    if (!lt) break;
    if (sec != kNone && cdata && !read_section(m, sec, cdata, lt, err)) return false;
    cdata = nullptr;
    if (lt + 1 < e && lt[1] == '?') {  // declaration
      const char* q = std::strstr(lt, "?>");
      if (!q) return (err = "mesh file: unterminated declaration", false);
      p = q + 2;
      continue;
    }
    if (lt + 3 < e && std::strncmp(lt, "<!--", 4) == 0) {
      const char* q = std::strstr(lt, "-->");
      if (!q) return (err = "mesh file: unterminated comment", false);
      p = q + 3;
      if (sec != kNone) cdata = p;  // the section's text goes on after the comment
      continue;
    }
    const char* gt = (const char*)std::memchr(lt, '>', (size_t)(e - lt));
    if (!gt) return (err = "mesh file: unterminated tag", false);
    const bool closing = lt + 1 < gt && lt[1] == '/';
    const char* nb = lt + (closing ? 2 : 1);
    const char* ne = nb;
    while (ne < gt && !is_space(*ne) && *ne != '/') ++ne;
    const std::string name(nb, ne);
    const bool self_closing = !closing && gt > lt && gt[-1] == '/';
    p = gt + 1;
    if (closing) {
      if (name == "Mesh") in_mesh = false;
      if (sec != kNone) sec = kNone;
      continue;
    }
    if (name == "Mesh") {
      // One mesh per file: a second <Mesh> could change dim or nb_dof under
      // records already read with the first one's widths.
      if (seen_mesh) return (err = "mesh file: more than one <Mesh> element", false);
      in_mesh = seen_mesh = true;
      for (auto& a : attributes(ne, gt)) {
        if (a.first == "dim") m.dim = std::atoi(a.second.c_str());
        else if (a.first == "nb_dof") m.nb_dof = std::atoi(a.second.c_str());
        else if (a.first == "file") return (err = "mesh file: <Mesh file=...> is not supported", false);
      }
      if (m.dim < 1 || m.dim > 3) return (err = "mesh file: dim must be 1, 2 or 3", false);
      if (m.nb_dof < 1 || m.nb_dof > 9) return (err = "mesh file: nb_dof must be 1..9", false);
      m.el_shape = m.dim == 1 ? 2 : m.dim == 2 ? 3 : 5;
      m.nen = m.dim + 1;
      m.sd_shape = m.dim == 2 ? 2 : m.dim == 3 ? 3 : 0;
      m.nsn = m.dim == 1 ? 0 : m.dim;
      continue;
    }
    if (!in_mesh) continue;
    Section s = kNone;
    if (name == "Nodes") s = kNodes;
    else if (name == "Elements") s = kElements;
    else if (name == "Sides") s = kSides;
    else if (name == "Material") s = kMaterial;
    if (s == kElements || s == kSides) {
      int& shape = s == kElements ? m.el_shape : m.sd_shape;
      int& nodes = s == kElements ? m.nen : m.nsn;
      const bool have = !(s == kElements ? m.elem_code : m.side_code).empty();
      const int shape0 = shape, nodes0 = nodes;
      for (auto& a : attributes(ne, gt)) {
        if (a.first == "shape") {
          const int sh = shape_of(a.second);
          if (sh < 0) return (err = "mesh file: unknown shape " + a.second, false);
          shape = sh;
        } else if (a.first == "nodes") {
          nodes = std::atoi(a.second.c_str());
        }
      }
      if (nodes < 1 || nodes > 27) return (err = "mesh file: bad nodes attribute", false);
      // Records of one kind share one width: a later section may repeat the
      // shape and node count, not change them.
      if (have && (shape != shape0 || nodes != nodes0))
        return (err = "mesh file: a later section changes the element or side shape", false);
    }
    if (s != kNone && !self_closing) {
      sec = s;
      cdata = p;
    }
  }
  if (!seen_mesh) return (err = "mesh file: no <Mesh> element", false);
  // text after the last tag (a cut file) would otherwise be dropped silently
  if (in_mesh || sec != kNone) return (err = "mesh file: unterminated <Mesh> or section", false);
  const int64_t nn = (int64_t)m.code.size() / m.nb_dof;
  // Every array holds whole records of the widths ofh_mesh_info reports.
  if ((int64_t)m.code.size() != nn * m.nb_dof || (int64_t)m.coord.size() != nn * m.dim ||
      (int64_t)m.elem.size() != (int64_t)m.elem_code.size() * m.nen ||
      (int64_t)m.side_code.size() % m.nb_dof != 0 ||
      (int64_t)m.side.size() != (int64_t)(m.side_code.size() / m.nb_dof) * m.nsn ||
      m.mat_code.size() != m.mat_name.size())
    return (err = "mesh file: inconsistent record widths", false);
  if (nn > INT32_MAX || (int64_t)m.elem_code.size() > INT32_MAX)
    return (err = "mesh file: too many records", false);
  for (int32_t v : m.elem)
    if (v < 1 || v > nn) return (err = "mesh file: element node label out of range", false);
  for (int32_t v : m.side)
    if (v < 1 || v > nn) return (err = "mesh file: side node label out of range", false);
  return true;
}

// right-aligned fields as iostream's setw (the writer's hot loops)
template <class T>
void put_int(std::string& out, T v, int width) {
  char buf[32];
  const auto r = std::to_chars(buf, buf + sizeof(buf), v);
  const int n = (int)(r.ptr - buf);
  if (n < width) out.append((size_t)(width - n), ' ');
  out.append(buf, (size_t)n);
}

// "%18.8e": to_chars gives printf's correctly rounded digits and exponent
void put_sci(std::string& out, double v) {
  char buf[48];
  const auto r = std::to_chars(buf, buf + sizeof(buf), v, std::chars_format::scientific, 8);
  const int n = (int)(r.ptr - buf);
  if (n < 18) out.append((size_t)(18 - n), ' ');
  out.append(buf, (size_t)n);
}

}  // namespace

extern "C" {

int ofh_mesh_read(const char* path, ofh_mesh* out) {
  if (!path || !out) return OFH_E_INVALID;
  *out = nullptr;
  FILE* f = std::fopen(path, "rb");
  if (!f) return OFH_E_INVALID;
  std::string text;
  char buf[1 << 16];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, n);
  std::fclose(f);
  ofh_mesh_s* m = new ofh_mesh_s();
  std::string err;
  if (!parse(text, *m, err)) {
    delete m;
    return OFH_E_INVALID;
  }
  *out = m;
  return OFH_OK;
}

int ofh_mesh_info(ofh_mesh m, int* dim, int* nb_dof, int* nb_nodes, int* nb_elements,
                  int* el_shape, int* nen, int* nb_sides, int* sd_shape, int* nsn,
                  int* nb_materials) {
  if (!m) return OFH_E_INVALID;
  if (dim) *dim = m->dim;
  if (nb_dof) *nb_dof = m->nb_dof;
  if (nb_nodes) *nb_nodes = (int)(m->code.size() / m->nb_dof);
  if (nb_elements) *nb_elements = (int)m->elem_code.size();
  if (el_shape) *el_shape = m->el_shape;
  if (nen) *nen = m->nen;
  if (nb_sides) *nb_sides = (int)(m->side_code.size() / m->nb_dof);
  if (sd_shape) *sd_shape = m->sd_shape;
  if (nsn) *nsn = m->nsn;
  if (nb_materials) *nb_materials = (int)m->mat_code.size();
  return OFH_OK;
}

int ofh_mesh_get(ofh_mesh m, double* coord, int64_t coord_len, int32_t* code, int64_t code_len,
                 int32_t* elem, int64_t elem_len, int32_t* elem_code, int64_t elem_code_len,
                 int32_t* side, int64_t side_len, int32_t* side_code, int64_t side_code_len) {
  if (!m) return OFH_E_INVALID;
  // every requested buffer must hold its whole array: check all before copying any
  auto fits = [](const void* dst, int64_t len, size_t need) {
    return !dst || (len >= 0 && (uint64_t)len >= need);
  };
  if (!fits(coord, coord_len, m->coord.size()) || !fits(code, code_len, m->code.size()) ||
      !fits(elem, elem_len, m->elem.size()) ||
      !fits(elem_code, elem_code_len, m->elem_code.size()) ||
      !fits(side, side_len, m->side.size()) || !fits(side_code, side_code_len, m->side_code.size()))
    return OFH_E_INVALID;
  auto cp = [](auto* dst, const auto& v) {
    if (dst && !v.empty()) std::memcpy(dst, v.data(), sizeof(v[0]) * v.size());
  };
  cp(coord, m->coord);
  cp(code, m->code);
  cp(elem, m->elem);
  cp(elem_code, m->elem_code);
  cp(side, m->side);
  cp(side_code, m->side_code);
  return OFH_OK;
}

int ofh_mesh_material(ofh_mesh m, int i, int* code, char* name, int name_len) {
  if (!m || i < 0 || i >= (int)m->mat_code.size()) return OFH_E_INVALID;
  if (code) *code = m->mat_code[i];
  if (name && name_len > 0) {
    std::snprintf(name, (size_t)name_len, "%s", m->mat_name[i].c_str());
  }
  return OFH_OK;
}

int ofh_mesh_destroy(ofh_mesh m) {
  delete m;
  return OFH_OK;
}

int ofh_mesh_write(const char* path, int dim, int nb_dof, int nb_nodes, const double* coord,
                   const int32_t* code, int nb_elements, int el_shape, int nen,
                   const int32_t* elem, const int32_t* elem_code, int nb_sides, int sd_shape,
                   int nsn, const int32_t* side, const int32_t* side_code, int nb_materials,
                   const int32_t* mat_code, const char* const* mat_name) {
  if (!path || dim < 1 || dim > 3 || nb_dof < 1 || nb_dof > 9 || nb_nodes < 0 ||
      nb_elements < 0 || nb_sides < 0 || nb_materials < 0)
    return OFH_E_INVALID;
  if ((nb_nodes && (!coord || !code)) || (nb_elements && (!elem || !elem_code || nen < 1)) ||
      (nb_sides && (!side || !side_code || nsn < 1)) ||
      (nb_materials && (!mat_code || !mat_name)))
    return OFH_E_INVALID;
  if ((nb_elements && (el_shape < 0 || el_shape >= kNbShapes)) ||
      (nb_sides && (sd_shape < 0 || sd_shape >= kNbShapes)))
    return OFH_E_INVALID;
  std::string out;
  out.reserve((size_t)nb_nodes * (20 * dim + 12) + (size_t)nb_elements * 8 * (nen + 1) + 4096);
  out += "<?xml version=\"1.0\"?>\n<OFELI_File>\n<info>\n   <title></title>\n   <date></date>\n"
         "   <author></author>\n</info>\n";
  char buf[64];
  std::snprintf(buf, sizeof(buf), "<Mesh dim=\"%d\" nb_dof=\"%d\">\n", dim, nb_dof);
  out += buf;
  size_t pw[10];
  for (int j = 0; j < nb_dof; ++j) pw[j] = (size_t)std::pow(10., (double)(nb_dof - 1 - j));
  if (nb_nodes > 0) {
    out += "   <Nodes>\n";
    for (int64_t i = 0; i < nb_nodes; ++i) {
      for (int d = 0; d < dim; ++d) {
        out += "  ";
        put_sci(out, coord[i * dim + d]);
      }
      const int32_t* c = code + i * nb_dof;
      size_t mk = 0;
      for (int j = 0; j < nb_dof; ++j) mk += (size_t)std::abs(c[j]) * pw[j];
      if (c[0] < 0) mk *= (size_t)-1;  // size_t m; m *= sign (Mesh.cpp:2762-2767)
      put_int(out, mk, 10);
      out += '\n';
    }
    out += "   </Nodes>\n";
  }
  if (nb_elements > 0) {
    std::snprintf(buf, sizeof(buf), "   <Elements shape=\"%s\" nodes=\"%d\">\n",
                  kShapeName[el_shape], nen);
    out += buf;
    for (int64_t e = 0; e < nb_elements; ++e) {
      for (int k = 0; k < nen; ++k) put_int(out, elem[e * nen + k], 8);
      put_int(out, elem_code[e], 8);
      out += '\n';
    }
    out += "   </Elements>\n";
  }
  if (nb_sides > 0) {
    std::snprintf(buf, sizeof(buf), "   <Sides shape=\"%s\" nodes=\"%d\">\n", kShapeName[sd_shape],
                  nsn);
    out += buf;
    for (int64_t s = 0; s < nb_sides; ++s) {
      for (int k = 0; k < nsn; ++k) put_int(out, side[s * nsn + k], 8);
      size_t mk = 0;
      for (int j = 0; j < nb_dof; ++j) mk += (size_t)(int64_t)side_code[s * nb_dof + j] * pw[j];
      put_int(out, mk, 6);
      out += '\n';
    }
    out += "   </Sides>\n";
  }
  if (nb_materials > 1 || (nb_materials == 1 && std::strcmp(mat_name[0], "Generic") != 0)) {
    out += "   <Material>\n";
    for (int i = 0; i < nb_materials; ++i) {
      std::snprintf(buf, sizeof(buf), "%9d   ", mat_code[i]);
      out += buf;
      out += mat_name[i];
      out += "\n";
    }
    out += "   </Material>\n";
  }
  out += "</Mesh>\n</OFELI_File>\n";
  FILE* f = std::fopen(path, "wb");
  if (!f) return OFH_E_INVALID;
  const size_t w = std::fwrite(out.data(), 1, out.size(), f);
  const int rc = std::fclose(f);
  return (w == out.size() && rc == 0) ? OFH_OK : OFH_E_INVALID;
}

}  // extern "C"
