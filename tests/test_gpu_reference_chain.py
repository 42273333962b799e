"""The device time loop against the reference's own shipped runs.

Run k of the reference's optimisation chain loaded naca_2D_flow_k on
naca_2D_shape_k.msh, integrated it with Simulator::time_loop
(src/simulation.cpp:68-126) for final_time = 0.005 s with
mesh_opti/config.txt's parameters (Mach 0.3, mu 1.8e-5, Pr 0.71, CFL 0.5)
and saved naca_2D_flow_{k+1}.  The device, from the same start on the same
mesh, must end on the same state.  The shipped files come from the
author's arm64 build (FMA allowed), so the bar is a tolerance: 5e-12 of
each variable's scale after ~3350 steps (the oracle lands at 4e-14 ..
2.5e-12 on these three runs).  Changing one parameter moves the end state
far outside that bar, so the pin tells the limiter, the viscosity and the
heat flux apart.
"""
import sys

import numpy as np
import pytest

from optimisation2forme_amd import capi
from tests import helpers as H

pytestmark = pytest.mark.gpu

TOL = 5e-12


@pytest.fixture(autouse=True)
def _announce(request):
    print("\n[gpu case] %s" % request.node.nodeid, flush=True)
    sys.stdout.flush()
    yield


def run_chain(k, **change):
    flow = dict(H.SHIPPED)
    flow.update(change)
    s = capi.Solver(H.chain_mesh(k), H.capi_params(flow))
    s.set_state(H.chain_flow(k))
    steps, t, _, _ = s.run_until(H.CHAIN_FINAL_TIME)
    U = s.state()
    s.close()
    return steps, t, U


@pytest.mark.parametrize("k", H.CHAIN)
def test_device_time_loop_reproduces_the_shipped_run(k):
    steps, t, U = run_chain(k)
    err = H.var_rel_err(U, H.chain_flow(k + 1))
    print("run %d: %d steps to t = %.17g, error per variable %s" % (k, steps, t, err), flush=True)
    assert t >= H.CHAIN_FINAL_TIME and steps > 3000
    assert max(err) <= TOL, err


@pytest.mark.parametrize("change,floor", [
    (dict(venkat_k=0.05), 1e-7),         # k_eps of computeVenkatAlpha
    (dict(limiter=1), 1e-4),             # Barth-Jespersen instead of Venkatakrishnan
    (dict(viscosity=1.8e-5 * 1.01), 1e-6),
    (dict(heat_flux=1), 1e-4),           # Fourier's law instead of the reference's R x Fourier
])
def test_the_shipped_run_tells_the_scheme_apart(change, floor):
    _, _, U = run_chain(0, **change)
    err = H.var_rel_err(U, H.chain_flow(1))
    print("%s: error per variable %s" % (change, err), flush=True)
    assert np.all(np.isfinite(U))
    assert max(err) >= floor, err


def test_driver_reproduces_the_shipped_run(tmp_path):
    """The navier_stokes driver on the reference's own config.txt keys
    (mesh_opti/config.txt, with the three paths pointed at the fixtures):
    the saved state is the shipped flow_1."""
    import subprocess
    from oracle import msh2
    from tests.test_driver import DRIVER
    out = tmp_path / "flow_1"
    cfg = tmp_path / "config.txt"
    cfg.write_text("""viscosity    = 1.8e-5
Pr           = 0.71
cfl_number   = 0.2
final_time   = 10000.0
save_interval = 1000
verbosity = 1
reload_config = 1000
initial_condition = one_x
boundary_conditions = inflow_outflow_x
noise_level = 0
inflow_velocity = 0.3
cfl_number   = 0.5
final_time   = 0.01
final_time   = 0.005
state_load_path = %s
state_save_path = %s
output_directory = %s
mesh_name = %s
""" % (H.chain_flow_path(0), out, tmp_path / "results", H.chain_mesh_path(0)))
    r = subprocess.run([DRIVER, str(cfg)], cwd=str(tmp_path), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    err = H.var_rel_err(msh2.read_state(str(out)), H.chain_flow(1))
    print("driver: error per variable %s" % err, flush=True)
    assert max(err) <= TOL, err
