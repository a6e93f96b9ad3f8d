"""The library's host-side mesh code (ofh_mesh_io.cpp, ofh_mesh.cpp) under
AddressSanitizer + UndefinedBehaviorSanitizer (CPU only, g++, no GPU code):
every golden mesh and a set of malformed variants (mixed record widths, a
second <Mesh>, truncated sections, out-of-range labels, bad attributes) are
read, copied out with exactly sized buffers, written and read back."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
MESHES = os.path.join(HERE, "golden", "meshes")


def variants(text):
    yield "mixed_widths", text.replace(
        "<Elements", '<Elements shape="quadrilateral" nodes="4">\n 1 2 3 4 1\n</Elements>\n'
        "<Elements", 1)
    yield "two_meshes", text.replace("</Mesh>", '</Mesh>\n<Mesh dim="3" nb_dof="2">\n'
                                     "<Nodes> 0 0 0 1 </Nodes>\n</Mesh>", 1)
    yield "truncated_nodes", text.replace("</Nodes>", " 0.5 </Nodes>", 1)
    yield "bad_label", text.replace("</Elements>", " 1 2 999999 1\n</Elements>", 1)
    yield "zero_nodes_attr", text.replace('nodes="3"', 'nodes="0"', 1)
    yield "huge_nodes_attr", text.replace('nodes="3"', 'nodes="99"', 1)
    yield "unterminated", text[: len(text) // 2]
    yield "sides_widen", text.replace('<Sides shape="line"  nodes="2">',
                                      '<Sides shape="line"  nodes="2">\n 1 2 1\n</Sides>\n'
                                      '<Sides shape="triangle" nodes="3">', 1)


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_host_mesh_code_under_sanitizers(tmp_path):
    exe = str(tmp_path / "mesh_io_sanitize")
    csrc = os.path.join(ROOT, "ofeli_amd", "csrc")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-I" + os.path.join(ROOT, "include"),
           os.path.join(HERE, "cpp", "mesh_io_sanitize.cpp"),
           os.path.join(csrc, "ofh_mesh_io.cpp"), os.path.join(csrc, "ofh_mesh.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, timeout=300)
    files = [os.path.join(MESHES, f) for f in sorted(os.listdir(MESHES)) if f.endswith(".m")]
    text = open(os.path.join(MESHES, "square.m"), encoding="latin-1").read()
    for name, t in variants(text):
        p = tmp_path / (name + ".m")
        p.write_text(t, encoding="latin-1")
        files.append(str(p))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([exe] + files, capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
    ok = [l for l in r.stdout.splitlines() if l.startswith("ok ")]
    rejected = [l for l in r.stdout.splitlines() if l.startswith("rejected ")]
    assert len(ok) >= 4  # the golden meshes
    for name in ("mixed_widths", "two_meshes", "truncated_nodes", "bad_label",
                 "zero_nodes_attr", "unterminated", "sides_widen"):
        assert any(name in l for l in rejected), (name, r.stdout)
