"""The tile tables of the fused stage kernel (csrc/host_tiles.cpp), on the
CPU: a small C++ program (tests/tile_check.cpp) compiled here from the
library's own host sources builds the tiles of the shipped quad mesh, the
10k-triangle mesh and the 200k-triangle benchmark mesh at several tile
sizes and runs check_tiles on them (owned cells partition the mesh; every
local index points at the global entity the stage reads), then plants one
mistake at a time in a correct set and requires check_tiles to reject it.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from optimisation2forme_amd import mesher
from tests import helpers as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "optimisation2forme_amd", "csrc")

PLANTED = ["vertex_slot", "edge_slot", "stencil_cell", "owned_twice", "face_cell_dropped", "vertex_stencil",
           "stress_slot", "tile_slot", "gather_cell", "stress_weight", "edge_normal", "concat_member",
           "concat_cell_offset", "concat_stencil_offset"]


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("tiles")
    exe = str(d / "tile_check")
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
           os.path.join(ROOT, "tests", "tile_check.cpp"), os.path.join(CSRC, "host_mesh.cpp"),
           os.path.join(CSRC, "host_tiles.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def run(exe, desc, d, sizes):
    np.asarray(desc["xy"], np.float64).tofile(str(d / "xy.bin"))
    np.asarray(desc["cell_nv"], np.int32).tofile(str(d / "cnv.bin"))
    np.asarray(desc["cell_vtx"], np.int32).tofile(str(d / "cvx.bin"))
    np.asarray(desc["bnd_vtx"], np.int32).tofile(str(d / "bv.bin"))
    np.asarray(desc["bnd_tag"], np.int8).tofile(str(d / "bt.bin"))
    r = subprocess.run([exe, str(d)] + [str(s) for s in sizes], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    tiles, planted, untagged = {}, {}, None
    for line in r.stdout.strip().splitlines():
        w = line.split()
        if w[0] == "tiles":
            tiles[int(w[1])] = w[2:]
        elif w[0] == "planted":
            planted[w[1]] = w[2]
        elif w[0] == "untagged":
            untagged = w[1]
        elif w[0] == "concat":
            assert w[1] == "ok", line
    return tiles, planted, untagged


@pytest.mark.parametrize("name,sizes", [
    ("shipped_quads", [192, 64, 500]),
    ("small_10k", [192, 37, 96, 1000]),
])
def test_tiles_valid_and_checker_catches_planted_mistakes(checker, tmp_path, name, sizes):
    desc = H.shipped_mesh() if name == "shipped_quads" else mesher.small_10k()
    nc = len(desc["cell_nv"])
    tiles, planted, untagged = run(checker, desc, tmp_path, sizes)
    for t in sizes:
        st = tiles[t]
        assert st[0] == "ok", (t, st)
        ntiles, owned, largest, gd_ratio = int(st[1]), int(st[2]), int(st[3]), float(st[4])
        assert owned == nc and largest <= t
        assert ntiles >= -(-nc // t)
        # every cell is a gradient cell at least once (its own tile)
        assert gd_ratio >= 1.0
    assert planted == {k: "detected" for k in PLANTED}
    assert untagged == "refused"


@pytest.mark.slow
def test_tiles_of_the_200k_benchmark_mesh(checker, tmp_path):
    """The transonic 199,680-triangle mesh at the sizes the automatic choice
    lands near on 256 CUs (3 or 4 resident tiles per CU): tables valid, and
    the rim each tile recomputes stays near the measured 1.3-1.5x of the
    gradients."""
    tiles, planted, untagged = run(checker, mesher.transonic_200k(), tmp_path, [196, 264])
    for t, lo, hi in [(196, 1.35, 1.55), (264, 1.3, 1.5)]:
        st = tiles[t]
        assert st[0] == "ok"
        assert int(st[2]) == 199680
        assert lo < float(st[4]) < hi
    assert planted == {k: "detected" for k in PLANTED}
    assert untagged == "refused"
