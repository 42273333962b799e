"""CPU checks of the C ABI library: it loads, exports every symbol that
include/o2f.h declares, and its host-side parts (defaults, msh2 reader,
error path) behave.  No kernel runs here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from optimisation2forme_amd import capi, mesher
from oracle import msh2
from tests import helpers as H

HEADER = os.path.join(H.GOLDEN, "..", "..", "include", "o2f.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(o2f_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    names = declared_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
    assert sorted(capi.EXPORTS) == names


def test_params_default_is_simulation_params_default():
    p = capi.Params()
    capi.lib().o2f_params_default(C.byref(p))
    assert (p.gamma, p.viscosity, p.Pr, p.cp, p.cfl_number) == (1.4, 1e-6, 0.71, 1005.0, 0.1)
    assert (p.rho_ref, p.p_ref, p.inflow_velocity, p.venkat_k) == (1.225, 101325.0, 1.0, 1e-2)
    assert (p.limiter, p.interp_mode, p.bc_inflow_outflow_x) == (0, 0, 1)


def test_msh2_reader_matches_the_oracle_reader(tmp_path):
    a = capi.read_msh2(H.shipped_mesh_path())
    b = H.shipped_mesh()
    for k in ("xy", "cell_nv", "cell_vtx", "bnd_vtx", "bnd_tag"):
        assert np.array_equal(a[k], b[k]), k
    d = mesher.naca_mesh(m=0.02, n_around=36, n_radial=6, wall_spacing=0.01, radius=5.0)
    path = str(tmp_path / "m.msh")
    mesher.write_msh2(path, d)
    c = capi.read_msh2(path)
    e = msh2.read_msh2(path)
    for k in ("xy", "cell_nv", "cell_vtx", "bnd_vtx", "bnd_tag"):
        assert np.array_equal(c[k], e[k]), k
    for k in ("xy", "cell_nv", "cell_vtx"):
        assert np.array_equal(c[k], d[k]), k
    # boundary lines come back grouped by physical tag (FreeStream 42 before Airfoil 44)
    key = lambda v, t: sorted(zip(map(tuple, v), t))
    assert key(c["bnd_vtx"], c["bnd_tag"]) == key(d["bnd_vtx"], d["bnd_tag"])


def test_msh2_reader_errors():
    with pytest.raises(capi.O2FError) as ei:
        capi.read_msh2("/nonexistent/mesh.msh")
    assert ei.value.code == -4


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_context_creation_fails_loudly_without_a_gpu():
    with pytest.raises(capi.O2FError) as ei:
        capi.Solver(H.tiny_mesh(), capi.make_params())
    assert ei.value.code == -3 and "no HIP device" in str(ei.value)


@pytest.mark.parametrize("which", ["shipped_quads", "small_triangles", "tiny_triangles"])
def test_library_topology_and_geometry_equal_the_oracle(which):
    """The library's host preparation (edge order, left/right cells, normals,
    areas, centres, per-cell edge lists) is bit-identical to the oracle's."""
    from oracle import oracle as O
    desc = {"shipped_quads": H.shipped_mesh, "small_triangles": H.small_mesh,
            "tiny_triangles": H.tiny_mesh}[which]()
    t = capi.mesh_topology(desc)
    m = O.Mesh(desc)
    e, c = m.edges(), m.cells()
    for k in ("va", "vb", "cl", "cr", "bc", "len", "nx", "ny", "mx", "my"):
        assert np.array_equal(t[k], e[k]), k
    for k in ("vol", "cx", "cy", "edges"):
        assert np.array_equal(t[k], c[k]), k
