"""Oracle (test infrastructure only): OFELI's XML mesh files (.m).

read() restates Mesh::get(file) (Mesh.cpp:2639-2670) as XMLParser parses a
mesh (read_mesh, XMLParser.cpp:1175-1220; read_mesh_data, :2210-2293) and
write() restates Mesh::put (Mesh.cpp:2734-2800).  Only tests/, smoke() and
bench.py's CPU baseline use this module.
"""
import ctypes
import math
import re

import numpy as np

_libc = ctypes.CDLL(None)
_libc.strtof.restype = ctypes.c_float
_libc.strtof.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p)]

SHAPES = ["none", "point", "line", "triangle", "quadrilateral", "tetrahedron", "hexahedron",
          "pentahedron"]


def stof(tok):
    """std::stof: the C library's decimal -> float conversion."""
    return float(_libc.strtof(tok.encode(), None))


def dof_code(mark, nb_dof):
    """DOFCode, MeshUtil.cpp:101-116: j keeps the signed mark, so the
    quotients are negative for a negative mark and the sign factor makes
    them positive again."""
    j, sign = mark, (-1 if mark < 0 else 1)
    code = []
    for k in range(nb_dof):
        kk = int(10.0 ** (nb_dof - k - 1))
        m = int(j / kk)  # C++ integer division truncates toward zero
        code.append(m * sign)
        j -= m * kk
    return code


def _attrs(text):
    return dict(re.findall(r'(\w+)\s*=\s*["\']([^"\']*)["\']', text))


def read(path):
    """Returns dict(dim, nb_dof, coord, code, elem, elem_code, el_shape,
    sides, side_code, sd_shape, materials)."""
    text = open(path, "rb").read().decode("latin-1")
    text = re.sub(r"<!--.*?-->", " ", text, flags=re.S)
    text = re.sub(r"<\?.*?\?>", " ", text, flags=re.S)
    mm = re.search(r"<Mesh\b([^>]*)>(.*?)</Mesh>", text, flags=re.S)
    a = _attrs(mm.group(1))
    dim, nb_dof = int(a.get("dim", 2)), int(a.get("nb_dof", 1))
    el_shape, nen = {1: (2, 2), 2: (3, 3), 3: (5, 4)}[dim]
    sd_shape, nsn = {1: (0, 0), 2: (2, 2), 3: (3, 3)}[dim]
    body = mm.group(2)
    out = dict(dim=dim, nb_dof=nb_dof, materials=[])
    coord, code = [], []
    sec = re.search(r"<Nodes\b[^>]*>(.*?)</Nodes>", body, flags=re.S)
    toks = sec.group(1).split() if sec else []
    rec = dim + 1
    assert len(toks) % rec == 0
    for i in range(0, len(toks), rec):
        coord.append([stof(t) for t in toks[i:i + dim]])
        code.append(dof_code(int(toks[i + dim]), nb_dof))
    out["coord"] = np.array(coord, float).reshape(-1, dim)
    out["code"] = np.array(code, np.int32).reshape(-1, nb_dof)
    for tag, key, shape, nn in (("Elements", "elem", el_shape, nen), ("Sides", "sides", sd_shape, nsn)):
        sec = re.search(r"<%s\b([^>]*)>(.*?)</%s>" % (tag, tag), body, flags=re.S)
        ids, codes = [], []
        if sec:
            at = _attrs(sec.group(1))
            shape = SHAPES.index(at["shape"]) if "shape" in at else shape
            nn = int(at.get("nodes", nn))
            toks = sec.group(2).split()
            assert len(toks) % (nn + 1) == 0
            for i in range(0, len(toks), nn + 1):
                ids.append([int(t) for t in toks[i:i + nn]])
                c = int(toks[i + nn])
                codes.append(c if key == "elem" else dof_code(c, nb_dof))
        out[key] = np.array(ids, np.int32).reshape(-1, nn)
        if key == "elem":
            out["elem_code"], out["el_shape"] = np.array(codes, np.int32), shape
        else:
            out["side_code"] = np.array(codes, np.int32).reshape(-1, nb_dof)
            out["sd_shape"] = shape
    sec = re.search(r"<Material\b[^>]*>(.*?)</Material>", body, flags=re.S)
    if sec:
        toks = sec.group(1).split()
        out["materials"] = [(int(toks[i]), toks[i + 1]) for i in range(0, len(toks), 2)]
    return out


def write(path, m):
    """Mesh::put, Mesh.cpp:2734-2800."""
    n = m["nb_dof"]
    lines = ['<?xml version="1.0"?>', "<OFELI_File>", "<info>", "   <title></title>",
             "   <date></date>", "   <author></author>", "</info>",
             '<Mesh dim="%d" nb_dof="%d">' % (m["dim"], n)]
    if len(m["coord"]):
        lines.append("   <Nodes>")
        for x, c in zip(m["coord"], m["code"]):
            mk = sum(abs(int(c[j])) * int(math.pow(10.0, n - 1 - j)) for j in range(n))
            if c[0] < 0:
                mk = (-mk) % 2 ** 64  # size_t m; m *= sign
            lines.append("".join("  %18.8e" % v for v in x) + "%10d" % mk)
        lines.append("   </Nodes>")
    if len(m["elem"]):
        lines.append('   <Elements shape="%s" nodes="%d">' % (SHAPES[m["el_shape"]],
                                                           m["elem"].shape[1]))
        for e, c in zip(m["elem"], m["elem_code"]):
            lines.append("".join("%8d" % v for v in e) + "%8d" % c)
        lines.append("   </Elements>")
    if len(m["sides"]):
        lines.append('   <Sides shape="%s" nodes="%d">' % (SHAPES[m["sd_shape"]],
                                                        m["sides"].shape[1]))
        for s, c in zip(m["sides"], m["side_code"]):
            mk = sum(int(c[j]) * int(math.pow(10.0, n - 1 - j)) for j in range(n)) % 2 ** 64
            lines.append("".join("%8d" % v for v in s) + "%6d" % mk)
        lines.append("   </Sides>")
    mats = m.get("materials", [])
    if len(mats) > 1 or (len(mats) == 1 and mats[0][1] != "Generic"):
        lines.append("   <Material>")
        for c, name in mats:
            lines.append("%9d   %s" % (c, name))
        lines.append("   </Material>")
    lines += ["</Mesh>", "</OFELI_File>"]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
