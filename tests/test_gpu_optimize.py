"""The outer shape-optimisation loop on the GPU: two iterations from the
reference's shipped parameter file on a small triangle mesh."""
import os
import sys

import numpy as np
import pytest

from optimisation2forme_amd import optimize as OPT
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _announce(request):
    print("\n[gpu case] %s" % request.node.nodeid, flush=True)
    sys.stdout.flush()
    yield


def test_two_outer_iterations(tmp_path):
    flow = dict(viscosity=1.8e-5, Pr=0.71, cfl_number=0.5, inflow_velocity=0.3, interp_mode="mvc")
    params = os.path.join(H.GOLDEN, "naca_2D_shape_params_0.txt")
    hist = OPT.run(params, 2, flow, str(tmp_path), mesh=dict(n_around=64, n_radial=20, wall_spacing=4e-3,
                                                              radius=20.0),
                   steps=200, adjoint_tol=1e-6, adjoint_chunk=100, adjoint_max=300, require_adjoint=False)
    assert len(hist) == 2
    (i0, m0, p0, J0, g0, _, _), (i1, m1, p1, J1, g1, _, _) = hist
    assert (m0, p0) == (0.0, 0.4)
    assert (m1, p1) == OPT.sign_step(m0, p0, g0["m"], g0["p"])
    assert np.isfinite(J0) and np.isfinite(J1) and np.isfinite(g0["m"])
    assert g0["p"] == 0.0            # at m = 0 the camber position does not move the wall (F5)
    for k in range(2):
        line = open(tmp_path / ("shape_opt_%d.log" % k)).read()
        assert line.startswith(" J=") and " m=" in line and " p=" in line
    v, marked = OPT.read_params(str(tmp_path / "shape_params_2.txt"))
    assert marked == ["m", "p"]
    assert (v["m"], v["p"]) == OPT.sign_step(m1, p1, g1["m"], g1["p"])


def test_outer_iteration_exact_adjoint_and_tangent_gradient(tmp_path):
    """The same loop with the exact adjoint, its exact right-hand side and
    the forward-mode gradient; the first iteration's tangent gradient
    against the reference-form one on the same state and lambda."""
    flow = dict(viscosity=1.8e-5, Pr=0.71, cfl_number=0.5, inflow_velocity=0.3, interp_mode="mvc")
    params = os.path.join(H.GOLDEN, "naca_2D_shape_params_0.txt")
    mesh = dict(n_around=64, n_radial=20, wall_spacing=4e-3, radius=20.0)
    kw = dict(mesh=mesh, steps=200, adjoint_tol=1e-6, adjoint_chunk=100, adjoint_max=300, require_adjoint=False,
              adjoint="exact", exact_rhs=True)
    h_tan = OPT.run(params, 2, flow, str(tmp_path / "tan"), gradient="tangent", **kw)
    h_ref = OPT.run(params, 1, flow, str(tmp_path / "ref"), gradient="reference", **kw)
    (_, m0, p0, J0, g_tan, _, _), (_, m1, p1, _, _, _, _) = h_tan
    _, _, _, J0r, g_ref, _, _ = h_ref[0]
    assert J0 == J0r
    print("dJ/dm tangent %.10g reference %.10g" % (g_tan["m"], g_ref["m"]), flush=True)
    assert abs(g_tan["m"] - g_ref["m"]) <= 1e-4 * abs(g_ref["m"])
    assert g_tan["p"] == 0.0 and g_ref["p"] == 0.0      # m = 0: p does not move the wall
    assert (m1, p1) == OPT.sign_step(m0, p0, g_tan["m"], g_tan["p"])
