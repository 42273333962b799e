"""The host geometry tangent (csrc/host_tangent.cpp: derivatives of cell
areas and centres, edge lengths, normals and midpoints and the stencil
weights along a vertex displacement; NACA camber displacements) against
central differences of the host geometry builders, on the CPU: a small C++
program (tests/geometry_tangent_check.cpp) compiled here from the
library's own host sources."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from optimisation2forme_amd import mesher

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "optimisation2forme_amd", "csrc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("geomtan")
    exe = str(d / "geometry_tangent_check")
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
           os.path.join(ROOT, "tests", "geometry_tangent_check.cpp"), os.path.join(CSRC, "host_mesh.cpp"),
           os.path.join(CSRC, "host_tangent.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def run_case(exe, desc, d):
    np.asarray(desc["xy"], np.float64).tofile(str(d / "xy.bin"))
    np.asarray(desc["cell_nv"], np.int32).tofile(str(d / "cnv.bin"))
    np.asarray(desc["cell_vtx"], np.int32).tofile(str(d / "cvx.bin"))
    np.asarray(desc["bnd_vtx"], np.int32).tofile(str(d / "bv.bin"))
    np.asarray(desc["bnd_tag"], np.int8).tofile(str(d / "bt.bin"))
    r = subprocess.run([exe, str(d)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = {}
    for line in r.stdout.strip().splitlines():
        case, name, a, b, *rest = line.split()
        if name == "sparse":  # the sparse form expanded: entries differing, entries listed
            rows[(case, name)] = (int(a), int(b))
            continue
        if name == "perturb":  # the sparse perturbation against the full update
            rows.setdefault((case, name), []).append((int(a), int(b)))
            continue
        rows[(case, name)] = (float(a), int(b), float(rest[0]))
    return rows


def test_geometry_tangent_against_central_differences(checker, tmp_path):
    """Cambered 5.3k-triangle O-mesh: every derived-geometry array's
    tangent within 3e-6 of its central difference (relative to the array's
    largest difference; the differences' own h^2 truncation reaches 1e-6 on
    the stencil weights, whose derivatives are large near the wall), for a
    random displacement of every vertex and for the NACA m and p
    displacements through the reference's partial update."""
    desc = mesher.naca_mesh(m=0.03, p=0.45, t=0.11, n_around=68, n_radial=40, wall_spacing=2e-3, growth=1.15)
    rows = run_case(checker, desc, tmp_path)
    sparse = {k: rows.pop(k) for k in list(rows) if k[1] == "sparse"}
    perturb = {k: rows.pop(k) for k in list(rows) if k[1] == "perturb"}
    assert len(rows) == 27 and len(sparse) == 3
    # the FD gradient's perturbed geometry (base entries overwritten by the
    # listed ones) is the full copy-and-update's, bit for bit, for +-1e-6
    # and a large 1e-2
    for case in ("naca_m", "naca_p"):
        for bad, listed in perturb[(case, "perturb")]:
            print("%-7s perturb: %d entries listed, %d differ" % (case, listed, bad))
            assert bad == 0 and listed > 0
        assert len(perturb[(case, "perturb")]) == 3
    # the sparse form the device path uploads is the dense form, bit for bit
    for case in ("full", "naca_m", "naca_p"):
        bad, listed = sparse[(case, "sparse")]
        print("%-7s sparse: %d entries listed, %d differ" % (case, listed, bad))
        assert bad == 0 and listed > 0
    for (case, name), (err, skipped, scale) in sorted(rows.items()):
        print("%-7s %-5s err %.3g  left out %d  scale %.3g" % (case, name, err, skipped, scale))
        assert err < 3e-6, (case, name)
        assert skipped == 0, (case, name)
    # the reference's update keeps x (midpoints and centres do not move in x)
    for case in ("naca_m", "naca_p"):
        assert rows[(case, "emx")][2] == 0.0 and rows[(case, "ccx")][2] == 0.0
    # and moves what the wall displacement reaches
    for name in ("cvol", "ccy", "elen", "enx", "eny", "emy", "sw"):
        assert rows[("naca_m", name)][2] > 0.0


def test_quad_mesh_tangent(checker, tmp_path):
    """The shipped quad mesh (parallelogram areas, mean-value weights on
    valence-4 vertices), random displacement of every vertex."""
    from tests import helpers as H
    rows = run_case(checker, H.shipped_mesh(), tmp_path)
    assert all(rows[(case, "sparse")][0] == 0 for case in ("full", "naca_m", "naca_p"))
    assert all(bad == 0 for case in ("naca_m", "naca_p") for bad, _ in rows[(case, "perturb")])
    for name in ("cvol", "ccx", "ccy", "elen", "enx", "eny", "sw"):
        err, skipped, scale = rows[("full", name)]
        print("full %-5s err %.3g  left out %d" % (name, err, skipped))
        assert err < 3e-6 and scale > 0.0, name
