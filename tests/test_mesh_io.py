"""OFELI XML mesh files (.m): the oracle (oracle/mesh_io.py) pinned by closed
forms, geometry and the reference's own files, and the C++ reader/writer of
the library (host code, no GPU) against it.  The fixtures under
tests/golden/meshes are mesh files of the reference's demos (laplace/square.m,
solid/lelas2d/beam.m, thermal/stdc2/proj.m, solid/lelas3d/beam.m)."""
import os
import re

import numpy as np
import pytest

from oracle import mesh_io as M

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = ["square", "beam_q4", "proj", "beam_tet"]


def path(name):
    return os.path.join(HERE, "golden", "meshes", name + ".m")


def measure(coord, elem):
    """Total area (2D triangles / quadrilaterals) or volume (tetrahedra)."""
    x = coord[elem - 1]
    if elem.shape[1] == 3 and coord.shape[1] == 2:
        d1, d2 = x[:, 1] - x[:, 0], x[:, 2] - x[:, 0]
        return np.abs(d1[:, 0] * d2[:, 1] - d1[:, 1] * d2[:, 0]).sum() / 2
    if elem.shape[1] == 4 and coord.shape[1] == 2:  # shoelace
        xs, ys = x[..., 0], x[..., 1]
        return np.abs((xs * np.roll(ys, -1, 1) - ys * np.roll(xs, -1, 1)).sum(1)).sum() / 2
    d = x[:, 1:] - x[:, :1]
    return np.abs(np.linalg.det(d)).sum() / 6


def test_dof_code_digits():
    """DOFCode: the nb_dof decimal digits of the mark, most significant
    first; a negative mark gives the digits of its absolute value (the
    reference keeps the signed mark in the quotient and multiplies by the
    sign again)."""
    rng = np.random.default_rng(0)
    for nb in (1, 2, 3):
        for mark in rng.integers(0, 10 ** nb, 40):
            digits = [int(c) for c in str(int(mark)).zfill(nb)]
            assert M.dof_code(int(mark), nb) == digits
            assert M.dof_code(-int(mark), nb) == digits
    assert M.dof_code(12, 2) == [1, 2] and M.dof_code(120, 3) == [1, 2, 0]


@pytest.mark.parametrize("name", FILES)
def test_coordinates_are_single_precision(name):
    """Mesh::get reads coordinates with stof: every value is a float."""
    m = M.read(path(name))
    assert np.array_equal(m["coord"], m["coord"].astype(np.float32).astype(np.float64))
    text = open(path(name), encoding="latin-1").read()
    first = re.search(r"<Nodes>\s*(\S+)", text).group(1)
    assert m["coord"][0, 0] == float(np.float32(float(first)))


@pytest.mark.parametrize("name,expect", [("square", 1.0), ("beam_q4", None), ("proj", None),
                                         ("beam_tet", 10.0)])
def test_elements_tile_the_domain(name, expect):
    """The elements of each demo mesh tile its rectangle / box: total area or
    volume equals the bounding box's (to single-precision coordinates)."""
    m = M.read(path(name))
    box = np.prod(m["coord"].max(0) - m["coord"].min(0))
    total = measure(m["coord"], m["elem"])
    assert abs(total - box) <= 1e-6 * box
    if expect is not None:
        assert abs(box - expect) <= 1e-6 * expect
    assert m["elem"].min() >= 1 and m["elem"].max() == len(m["coord"])


def test_writer_reproduces_the_reference_output():
    """laplace/square.m was written by Mesh::put: the records of its element
    and side sections are reproduced byte for byte (its tag lines carry two
    spaces before nodes=, the current Mesh.cpp:2775,2785 one), its node
    section up to the last digit of single-precision coordinates (the same
    values once read)."""
    m = M.read(path("square"))
    M.write("/tmp/_orc_square.m", m)
    a = open(path("square"), encoding="latin-1").read()
    b = open("/tmp/_orc_square.m").read()
    for tag in ("Elements", "Sides"):
        pa = re.search(r"<%s[^>]*>(.*?)</%s>" % (tag, tag), a, re.S).group(1)
        pb = re.search(r"<%s[^>]*>(.*?)</%s>" % (tag, tag), b, re.S).group(1)
        assert pa == pb
    assert np.array_equal(M.read("/tmp/_orc_square.m")["coord"], m["coord"])


@pytest.mark.parametrize("name", FILES)
def test_read_write_read_is_idempotent(name, tmp_path):
    m = M.read(path(name))
    M.write(str(tmp_path / "a.m"), m)
    m2 = M.read(str(tmp_path / "a.m"))
    for k in ("coord", "code", "elem", "elem_code", "sides", "side_code"):
        assert np.array_equal(m[k], m2[k]), k
    M.write(str(tmp_path / "b.m"), m2)
    assert open(tmp_path / "a.m").read() == open(tmp_path / "b.m").read()


# ---- the library's C++ reader and writer (host code) against the oracle ----

@pytest.mark.parametrize("name", FILES)
def test_library_reader_matches_oracle(name):
    import ofeli_amd as oa
    r = M.read(path(name))
    c = oa.read_mesh(path(name))
    assert (c.dim, c.nb_dof, c.el_shape, c.sd_shape) == (r["dim"], r["nb_dof"], r["el_shape"],
                                                        r["sd_shape"])
    for k, a in (("coord", "coord"), ("code", "code"), ("elem", "elem"),
                 ("elem_code", "elem_code"), ("sides", "sides"), ("side_code", "side_code")):
        assert np.array_equal(r[k], getattr(c, a)), k
    assert c.materials == r["materials"]


@pytest.mark.parametrize("name", FILES)
def test_library_writer_matches_oracle(name, tmp_path):
    import ofeli_amd as oa
    c = oa.read_mesh(path(name))
    oa.write_mesh(str(tmp_path / "c.m"), c)
    M.write(str(tmp_path / "o.m"), M.read(path(name)))
    assert open(tmp_path / "c.m", "rb").read() == open(tmp_path / "o.m", "rb").read()


def test_library_materials_and_negative_codes(tmp_path):
    """Materials round trip; a negative node code is written as the size_t
    the reference computes (2^64 - mark), which the reader then rejects as
    the reference's stoi does."""
    import ofeli_amd as oa
    c = oa.read_mesh(path("proj"))
    c.materials = [(1, "Copper"), (2, "Iron")]
    oa.write_mesh(str(tmp_path / "m.m"), c)
    assert oa.read_mesh(str(tmp_path / "m.m")).materials == [(1, "Copper"), (2, "Iron")]
    c.code = c.code.copy()
    c.code[0, 0] = -1
    oa.write_mesh(str(tmp_path / "n.m"), c)
    assert "18446744073709551615" in open(tmp_path / "n.m").read()
    with pytest.raises(oa.OfhError):
        oa.read_mesh(str(tmp_path / "n.m"))


def test_library_reader_errors(tmp_path):
    import ofeli_amd as oa
    with pytest.raises(oa.OfhError):
        oa.read_mesh(str(tmp_path / "missing.m"))
    text = open(path("proj"), encoding="latin-1").read()
    cut = re.sub(r"(<Nodes>\s*\S+\s+\S+)\s.*?(</Nodes>)", r"\1 \2", text, flags=re.S)
    (tmp_path / "cut.m").write_text(cut, encoding="latin-1")
    with pytest.raises(oa.OfhError):
        oa.read_mesh(str(tmp_path / "cut.m"))
