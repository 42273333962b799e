"""optimizeShape end to end through the navier_stokes driver, against the
reference's own optimisation runs and against the oracle chain.

Run k of the shipped chain (run_opti.sh:211-254) integrated naca_2D_flow_k
on naca_2D_shape_k.msh for 0.005 s (mesh_opti/config.txt), then
optimizeShape (src/shape_optimization.cpp:1022-1067) solved the adjoint
(BiCGSTAB to 1e-6 in chunks of 500) and logged J and dJ/d(m, p); the script
then stepped each parameter against the sign of its gradient
(run_opti.sh:91-120: +h when the gradient is negative, -h otherwise) into
naca_2D_shape_params_{k+1}.txt.  The driver repeats runs 0, 15 and 29 with
the reference's settings: it must save flow_{k+1} and log gradients whose
signs reproduce the recorded steps.
"""
import json
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from oracle import msh2
from oracle import oracle as O
from tests import helpers as H
from tests.test_driver import DRIVER

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _announce(request):
    print("\n[gpu case] %s" % request.node.nodeid, flush=True)
    sys.stdout.flush()
    yield


SHIPPED_CONFIG = """viscosity    = 1.8e-5
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
final_time   = %(final_time).17g
state_load_path = %(load)s
state_save_path = %(save)s
output_directory = %(out)s
shape_opt_enable = 1
shape_param_file = %(params)s
shape_log_file = %(log)s
mesh_name = %(mesh)s
%(extra)s
"""


def run_driver(tmp_path, k, final_time, extra="", load=None, timeout=600):
    """The driver on mesh k with shape-parameter file k, starting from flow
    `load` (default k)."""
    cfg = tmp_path / "config.txt"
    cfg.write_text(SHIPPED_CONFIG % dict(
        final_time=final_time, load=H.chain_flow_path(k if load is None else load), save=tmp_path / "flow",
        out=tmp_path / "results",
        params=os.path.join(H.GOLDEN, "naca_2D_shape_params_%d.txt" % k), log=tmp_path / "shape.log",
        mesh=H.chain_mesh_path(k), extra=extra))
    r = subprocess.run([DRIVER, str(cfg)], cwd=str(tmp_path), capture_output=True, text=True, timeout=timeout)
    print(r.stdout[-800:], flush=True)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    log = (tmp_path / "shape.log").read_text()
    m = re.fullmatch(r" J=(\S+) m=(\S+) p=(\S+)\n", log)
    assert m, repr(log)
    preview = dict(re.findall(r"dJ/d(\w+) = (\S+)", r.stdout))
    J_preview = float(re.search(r"\[preview\] J courant = (\S+)", r.stdout).group(1))
    return r.stdout, [float(x) for x in m.groups()], J_preview, {k_: float(v) for k_, v in preview.items()}


@pytest.mark.parametrize("k", H.CHAIN)
def test_driver_repeats_the_reference_run_and_its_gradient_signs(k, tmp_path):
    out, (J, gm, gp), _, _ = run_driver(tmp_path, k, H.CHAIN_FINAL_TIME)
    it = re.findall(r"\[adjoint\] iters=(\d+)\s+relres=(\S+)", out)[-1]
    print("run %d: J %.8f, dJ/dm %.6e, dJ/dp %.6e, BiCGSTAB %s iterations to %s" % (k, J, gm, gp, it[0], it[1]))
    assert float(it[1]) <= 1e-6
    err = H.var_rel_err(msh2.read_state(str(tmp_path / "flow")), H.chain_flow(k + 1))
    assert max(err) <= 5e-12, err
    traj = json.load(open(os.path.join(H.GOLDEN, "trajectory.json")))
    for g, name in ((gm, "m"), (gp, "p")):
        step = traj[k + 1][name] - traj[k][name]
        # run_opti.sh: + step when the gradient is negative, - step otherwise
        assert (step > 0) == (g < 0), (name, g, step)
    if k == 0:
        assert gp == 0.0      # at m = 0 the mean line is flat whatever p (yc_naca4)


def test_driver_shape_gradient_equals_the_oracle_chain(tmp_path):
    """One time step from flow_1 on shape 0 (the device step and dt are
    bitwise the oracle's), then optimizeShape with the adjoint solved to a
    relative residual of 0.5 (a few hundred iterations):
    the logged J and gradient equal the oracle's computeDirectObjectiveTerm,
    solveAdjointSystem (its dot products summed in the device's order, so
    lambda is bitwise the device's) and computeShapeGradient on the same
    face states, to the printed digits."""
    from optimisation2forme_amd import capi
    p = H.oracle_params(H.SHIPPED)
    m = O.Mesh(H.chain_mesh(0))
    m.set_state(H.chain_flow(1))
    # the device's dt of this state (the driver's first step uses it)
    s = capi.Solver(H.chain_mesh(0), H.capi_params(H.SHIPPED))
    s.set_state(H.chain_flow(1))
    dt = s.dt()
    s.close()
    out, (J, gm, gp), J_pre, pre = run_driver(tmp_path, 0, 0.5 * dt, load=1, extra="adjoint_tol = 0.5")
    it_dev, rr_dev = re.findall(r"\[adjoint\] iters=(\d+)\s+relres=(\S+)", out)[-1]
    m.rk3_step(p, dt)
    J_o = m.objective(p, 50.0)
    O.bicgstab_config(device_order=True)
    try:
        st, lam, its, rr = m.bicgstab(p, 50.0, tol=0.5, chunk=500, max_total=100000)
    finally:
        O.bicgstab_config(device_order=False)
    gm_o = m.shape_gradient(p, lam, "m", m0=0.0, p0=0.4, t=0.12, beta=50.0)
    gp_o = m.shape_gradient(p, lam, "p", m0=0.0, p0=0.4, t=0.12, beta=50.0)
    print("driver J %.8f dJ/dm %.8f dJ/dp %.8f | oracle J %.10f dJ/dm %.10f dJ/dp %.10f (%d iterations)"
          % (J_pre, pre["m"], pre["p"], J_o, gm_o, gp_o, its), flush=True)
    assert st == 0 and its == int(it_dev) and its > 0
    assert float("%.8f" % J_o) == J_pre and float("%.8f" % J_o) == J
    assert abs(pre["m"] - gm_o) <= 5e-9 + 1e-12 * abs(gm_o)
    assert abs(gm - gm_o) <= 5e-7 * abs(gm_o)
    assert gp == 0.0 and gp_o == 0.0 and pre["p"] == 0.0


def test_driver_on_the_renumbered_mesh(tmp_path):
    """mesh_reorder = rcm_xcd: the driver renumbers shape 0's mesh for the
    device (o2f_mesh_reorder) and keeps the state files in the mesh file's
    cell order.  Run 0 from flow_0 still lands on the reference's flow_1
    (the sums over each cell's edges run in another order, nothing else
    changes), and optimizeShape logs the same J and gradient signs as the
    run on the file's numbering."""
    a = tmp_path / "file_order"
    b = tmp_path / "renumbered"
    a.mkdir()
    b.mkdir()
    _, (Ja, gma, gpa), _, _ = run_driver(a, 0, H.CHAIN_FINAL_TIME)
    _, (Jb, gmb, gpb), _, _ = run_driver(b, 0, H.CHAIN_FINAL_TIME, extra="mesh_reorder = rcm_xcd")
    Ua, Ub = msh2.read_state(str(a / "flow")), msh2.read_state(str(b / "flow"))
    ref = H.chain_flow(1)
    ea, eb, eab = H.var_rel_err(Ua, ref), H.var_rel_err(Ub, ref), H.var_rel_err(Ub, Ua)
    print("flow_1: file order %s, renumbered %s, between them %s; J %.15g / %.15g; dJ/dm %.8e / %.8e"
          % (max(ea), max(eb), max(eab), Ja, Jb, gma, gmb), flush=True)
    assert max(eb) <= 5e-12 and max(eab) <= 5e-12
    assert abs(Jb - Ja) <= 1e-10 * abs(Ja)
    assert np.sign(gmb) == np.sign(gma) and abs(gmb - gma) <= 1e-3 * abs(gma)
    assert gpb == gpa == 0.0    # m = 0: the mean line is flat whatever p


def test_driver_refuses_an_unknown_mesh_reorder(tmp_path):
    cfg = tmp_path / "config.txt"
    cfg.write_text("mesh_name = %s\nmesh_reorder = metis\nfinal_time = 0\n" % H.chain_mesh_path(0))
    r = subprocess.run([DRIVER, str(cfg)], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "Unknown mesh_reorder: metis" in (r.stdout + r.stderr)
