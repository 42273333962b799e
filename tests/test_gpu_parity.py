"""GPU parity: the gfx950 path (libo2f.so, through the C ABI) against the CPU
oracle on the same inputs.

Bar: bit-identical wherever both sides sum in the same order (the device
gathers every cell sum in ascending edge id, as the oracle does); where a
reduction order differs (dt's hypot, Krylov dot products, the shape
gradient's lambda . dR) the tolerance is stated in the test.
"""
import os
import sys

import numpy as np
import pytest

from optimisation2forme_amd import capi
from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _announce(request):
    print("\n[gpu case] %s" % request.node.nodeid, flush=True)
    sys.stdout.flush()
    yield


def pair(desc, flow, U=None):
    """(device solver, oracle mesh) on the same mesh, parameters and state."""
    s = capi.Solver(desc, H.capi_params(flow))
    m = O.Mesh(desc)
    p = H.oracle_params(flow)
    if U is None:
        m.init_freestream(p)
        U = m.state()
    s.set_state(U)
    m.set_state(U)
    return s, m, p


def same(a, b):
    return np.array_equal(np.asarray(a), np.asarray(b))


# ------------------------------------------------------------------ HLLC ---

def test_hllc_batch_bit_identical_all_branches():
    rng = np.random.default_rng(7)
    n = 20000
    rho = rng.uniform(0.3, 3.0, (2, n))
    p = rng.uniform(2e4, 3e5, (2, n))
    c = np.sqrt(1.4 * p / rho)
    M = rng.uniform(0.0, 2.5, (2, n))          # sub- and supersonic: all four HLLC branches
    th = rng.uniform(0, 2 * np.pi, (2, n))
    u, v = M * c * np.cos(th), M * c * np.sin(th)
    E = p / 0.4 + 0.5 * rho * (u * u + v * v)
    qL = np.array([rho[0], rho[0] * u[0], rho[0] * v[0], E[0]])
    qR = np.array([rho[1], rho[1] * u[1], rho[1] * v[1], E[1]])
    a = rng.uniform(0, 2 * np.pi, n)
    nrm = np.array([np.cos(a), np.sin(a)])
    Fg = capi.hllc_flux_batch(qL, qR, nrm)
    Fo = O.hllc_batch(qL, qR, nrm)
    assert same(Fg, Fo)


# -------------------------------------------------------------- residual ---

def test_residual_shipped_quad_mesh_bit_identical():
    s, m, p = pair(H.shipped_mesh(), H.SHIPPED, H.shipped_flow())
    Rg = s.residual()
    Ro = m.residual(p)
    assert same(Rg, Ro)
    assert same(s.gradients(), m.gradients())
    fg, fo = s.faces(), m.faces()
    assert same(fg[0], fo[0]) and same(fg[1], fo[1])


def test_advection_then_diffusion_separately():
    s, m, p = pair(H.shipped_mesh(), H.SHIPPED, H.shipped_flow())
    assert same(s.advection(), m.advection(p))
    assert same(s.diffusion(), m.diffusion(p))


@pytest.mark.parametrize("limiter", [0, 1])
@pytest.mark.parametrize("interp_mode", [0, 1])
def test_residual_triangle_mesh_bit_identical(limiter, interp_mode):
    flow = H.subsonic_small(limiter=limiter, interp_mode=interp_mode)
    U = H.developed_state(H.small_mesh(), H.subsonic_small(limiter=limiter, interp_mode=1), steps=15, seed=3)
    s, m, p = pair(H.small_mesh(), flow, U)
    Rg, Ro = s.residual(), m.residual(p)
    assert same(Rg, Ro) or (interp_mode == 0 and np.array_equal(np.isnan(Rg), np.isnan(Ro))
                           and same(np.nan_to_num(Rg), np.nan_to_num(Ro)))
    assert same(np.nan_to_num(s.gradients()), np.nan_to_num(m.gradients()))


def test_residual_transonic_200k_bit_identical_and_deterministic():
    flow = H.transonic()
    desc = H.big_mesh()
    m0 = O.Mesh(desc)
    p = H.oracle_params(flow)
    m0.init_freestream(p)
    m0.run_steps(p, 2)
    U = m0.state()
    s, m, p = pair(desc, flow, U)
    Rg = s.residual()
    assert same(Rg, m.residual(p))
    assert same(Rg, s.residual())   # run to run
    # conservation: interior fluxes cancel, so sum_c V_c R_c is a boundary sum
    # that the device reproduces to rounding
    vol = m.cells()["vol"]
    assert np.allclose((Rg * vol).sum(1), (m.residual_array() * vol).sum(1), rtol=1e-12, atol=1e-9)


# ----------------------------------------------------- time integration ---

def test_dt_matches():
    s, m, p = pair(H.shipped_mesh(), H.SHIPPED, H.shipped_flow())
    dg, do = s.dt(), m.dt(p)
    # hypot(u, v) of the device math library may differ from glibc by 1 ulp
    assert abs(dg - do) <= 4e-16 * do


def test_rk3_steps_per_step_parity_shipped():
    s, m, p = pair(H.shipped_mesh(), H.SHIPPED, H.shipped_flow())
    for k in range(3):
        dt = m.dt(p)
        s.rk3_step(dt)
        m.rk3_step(p, dt)
        assert same(s.state(), m.state()), "step %d" % k


def test_rk3_substeps_are_the_step_decomposition():
    """step_SSP_RK3 as the reference decomposes it (U_old <- U, then
    rk3_substep (0, 1), (3/4, 1/4), (1/3, 2/3)): every substep's state and
    residual bitwise the oracle's, the end state bitwise the fused step."""
    s, m, p = pair(H.shipped_mesh(), H.SHIPPED, H.shipped_flow())
    fused = capi.Solver(H.shipped_mesh(), H.capi_params(H.SHIPPED))
    fused.set_state(H.shipped_flow())
    for k in range(2):
        dt = m.dt(p)
        s.rk3_save_old()
        m.save_old()
        for c0, c1 in ((0.0, 1.0), (3.0 / 4.0, 1.0 / 4.0), (1.0 / 3.0, 2.0 / 3.0)):
            s.rk3_substep(c0, c1, dt)
            m.rk3_substep(p, c0, c1, dt)
            assert same(s.state(), m.state()), (k, c0)
            assert same(s.gradients(), m.gradients()), (k, c0)
        fused.rk3_step(dt)
        assert same(fused.state(), s.state()), k


def test_run_steps_transonic_triangles_within_1e_10():
    flow = H.transonic(limiter=1)
    desc = H.small_mesh()
    s, m, p = pair(desc, flow)
    dts_g, t = s.run_steps(6)
    dts_o = m.run_steps(p, 6)
    assert np.allclose(dts_g, dts_o, rtol=1e-14, atol=0)
    assert abs(t - dts_g.sum()) <= 1e-15 * t
    Ug, Uo = s.state(), m.state()
    assert H.rel_err(Ug, Uo) <= 1e-10          # per-step state tolerance of the workload


def test_time_loop_stops_like_the_reference():
    s, m, p = pair(H.shipped_mesh(), H.SHIPPED, H.shipped_flow())
    dt0 = m.dt(p)
    final = 2.5 * dt0
    n, t = s.time_loop(final)
    steps, tt = 0, 0.0
    while tt < final:                     # src/simulation.cpp:75-124
        dt = m.dt(p)
        m.rk3_step(p, dt)
        tt += dt
        steps += 1
    assert n == steps == 3
    assert abs(t - tt) <= 1e-15 * tt
    assert H.rel_err(s.state(), m.state()) <= 1e-10


def test_residual_norm_and_forces():
    s, m, p = pair(H.shipped_mesh(), H.SHIPPED, H.shipped_flow())
    s.residual()
    m.residual(p)
    assert abs(s.residual_norm() - m.residual_norm(p)) <= 1e-13 * m.residual_norm(p)
    assert s.lift_drag() == m.lift_drag(p)
    assert s.objective(50.0) == m.objective(p, 50.0)


# --------------------------------------------------------------- adjoint ---

def test_adjoint_rhs_and_jt_bit_identical():
    s, m, p = pair(H.shipped_mesh(), H.SHIPPED, H.shipped_flow())
    s.residual()
    m.residual(p)
    assert same(s.adjoint_rhs(50.0), m.adjoint_rhs(p, 50.0))
    x = np.random.default_rng(11).standard_normal((4, m.nc))
    assert same(s.apply_jt(x), m.apply_jt(p, x))


@pytest.fixture
def device_order_dots():
    """The oracle's BiCGSTAB summing its dot products in the device's
    reduction order (512-block grid-stride partials, 256-wide halving tree,
    partials in block order then the same tree), so iterates compare bit for
    bit; reset afterwards."""
    O.bicgstab_config(device_order=True)
    yield
    O.bicgstab_config(device_order=False)


@pytest.mark.parametrize("k", [1, 3, 5, 20])
def test_bicgstab_iterates_match(k, device_order_dots):
    s, m, p = pair(H.shipped_mesh(), H.SHIPPED, H.shipped_flow())
    s.residual()
    m.residual(p)
    lg, itg, rrg, stg = s.bicgstab(50.0, tol=1e-300, chunk=k, max_total=k, check=False)
    sto, lo, ito, rro = m.bicgstab(p, 50.0, tol=1e-300, chunk=k, max_total=k)
    assert itg == ito == k and stg == -6 and sto == -3   # not converged: iterate returned
    assert same(lg, lo) and rrg == rro


def test_bicgstab_bitwise_over_chunk_restarts(device_order_dots):
    """solveAdjointSystem's chunked loop (src/shape_optimization.cpp:694-
    717): four chunks of 100 iterations, each restarted from the last x by
    solveWithGuess, the true residual after each; on the device the loop
    test runs on the device, groups of eight iterations are graph-replayed
    and the body is rotated -- every iterate bitwise the oracle's."""
    s, m, p = pair(H.shipped_mesh(), H.SHIPPED, H.shipped_flow())
    s.residual()
    m.residual(p)
    lg, itg, rrg, stg = s.bicgstab(50.0, tol=1e-300, chunk=100, max_total=350, check=False)
    sto, lo, ito, rro = m.bicgstab(p, 50.0, tol=1e-300, chunk=100, max_total=350)
    assert itg == ito == 400 and stg == -6 and sto == -3
    assert np.all(np.isfinite(lg)) and np.isfinite(rrg)
    assert same(lg, lo) and rrg == rro


def test_bicgstab_restart_branch_bitwise(device_order_dots):
    """The orthogonality restart of Eigen's bicgstab (r0 = r = b - A x when
    |r0.r| < eps^2 |r0|^2; the first restart also resets the iteration
    count).  With eps = 0.3 the branch is taken every few iterations, so a
    chunk of 40 runs past 40 iterations; the device follows it bit for
    bit."""
    s, m, p = pair(H.shipped_mesh(), H.SHIPPED, H.shipped_flow())
    s.residual()
    m.residual(p)
    s.set_restart_eps(0.3)
    O.bicgstab_config(device_order=True, restart_eps=0.3)
    lg, itg, rrg, stg = s.bicgstab(50.0, tol=1e-300, chunk=40, max_total=100, check=False)
    sto, lo, ito, rro = m.bicgstab(p, 50.0, tol=1e-300, chunk=40, max_total=100)
    restarts = O.bicgstab_restarts()
    print("iterations %d, restarts %d, relres %.3e" % (ito, restarts, rro), flush=True)
    assert restarts >= 2 and ito > 100          # restarted, and the first restart reset the count
    assert itg == ito and stg == {0: 0, -2: -5, -3: -6}[sto] and np.isfinite(rrg)
    assert same(lg, lo) and rrg == rro


def test_bicgstab_converges_bitwise_through_many_chunks(device_order_dots):
    """A solve that converges: the 160-triangle mesh's adjoint to a relative
    residual of 1e-4 takes ~17,000 iterations, i.e. ~35 chunks of 500
    restarted from the last x.  The device's lambda, iteration count and
    residual are the oracle's bit for bit (on this operator the two
    summation orders alone change the iteration count by thousands)."""
    flow = H.subsonic_small()
    U = H.developed_state(H.tiny_mesh(), flow, steps=30)
    s, m, p = pair(H.tiny_mesh(), flow, U)
    s.residual()
    m.residual(p)
    lg, itg, rrg, stg = s.bicgstab(0.1, tol=1e-4, chunk=500, max_total=20000, check=False)
    sto, lo, ito, rro = m.bicgstab(p, 0.1, tol=1e-4, chunk=500, max_total=20000)
    print("tiny mesh: status %d / %d, %d iterations, relres %.3e" % (stg, sto, ito, rro), flush=True)
    assert sto == 0 and stg == 0 and rro <= 1e-4 and ito > 10000
    assert itg == ito and rrg == rro and same(lg, lo)


def test_bicgstab_small_mesh_bitwise(device_order_dots):
    flow = H.subsonic_small(interp_mode=1)
    U = H.developed_state(H.small_mesh(), flow, steps=30)
    s, m, p = pair(H.small_mesh(), flow, U)
    s.residual()
    m.residual(p)
    lg, itg, rrg, stg = s.bicgstab(0.1, tol=1e-3, chunk=250, max_total=1000, check=False)
    sto, lo, ito, rro = m.bicgstab(p, 0.1, tol=1e-3, chunk=250, max_total=1000)
    print("small mesh: status %d / %d, %d iterations, relres %.3e" % (stg, sto, ito, rro), flush=True)
    assert itg == ito and rrg == rro and same(lg, lo)
    assert stg == {0: 0, -2: -5, -3: -6}[sto]


def test_bicgstab_graph_replays_are_stable():
    # Each call captures the eight-iteration group once and replays it
    # (here 12 times per 100-iteration chunk, with the true-residual check
    # and the next chunk's set-up run eagerly between chunks).  Three calls,
    # with other device work in between, must agree bit for bit, in every
    # operator form that the loop drives.
    flow = H.subsonic_small(interp_mode=1)
    U = H.developed_state(H.small_mesh(), flow, steps=30)
    s = capi.Solver(H.small_mesh(), H.capi_params(flow))
    s.set_state(U)
    s.residual()
    x = np.random.default_rng(5).standard_normal((4, s.nc))
    solves = [
        lambda: s.bicgstab(0.1, tol=1e-30, chunk=100, max_total=300, check=False),
        lambda: s.bicgstab_block_jacobi(0.1, tol=1e-30, chunk=100, max_total=300, check=False),
        lambda: s.bicgstab_exact(0.1, tol=1e-30, chunk=100, max_total=200, check=False),
    ]
    for solve in solves:
        runs = []
        for _ in range(3):
            runs.append(solve())
            s.apply_jt(x)  # eager work on the same stream between the calls
        (l0, it0, rr0, st0) = runs[0]
        assert it0 > 0 and np.isfinite(rr0) and np.all(np.isfinite(l0))
        for lam, it, rr, st in runs[1:]:
            assert it == it0 and st == st0
            assert np.array_equal([rr], [rr0], equal_nan=True)
            assert np.array_equal(lam, l0, equal_nan=True)


def test_shape_gradient_matches():
    s, m, p = pair(H.shipped_mesh(), H.SHIPPED, H.shipped_flow())
    m.residual(p)
    st, lam, it, rr = m.bicgstab(p, 50.0, tol=1e-6, chunk=100, max_total=100)
    names = ["pos_x", "c", "m", "p", "t", "beta"]
    vals = [0.355, 1.0, 0.0, 0.4, 0.12, 50.0]
    g = s.shape_gradient(names, vals, ["m", "p", "t"], lam, 50.0)
    gm = m.shape_gradient(p, lam, "m", m0=0.0, p0=0.4, t=0.12, beta=50.0)
    gp = m.shape_gradient(p, lam, "p", m0=0.0, p0=0.4, t=0.12, beta=50.0)
    assert abs(g[0] - gm) <= 1e-8 * abs(gm)
    assert g[1] == gp == 0.0 and g[2] == 0.0
    # the device leaves the face states as computeResidual(mesh) leaves them
    fg, fo = s.faces(), m.faces()
    assert same(fg[0], fo[0])


def test_shape_gradient_cambered_triangles():
    flow = H.transonic()
    desc = H.small_mesh()
    from optimisation2forme_amd import mesher
    desc = mesher.naca_mesh(m=0.03, p=0.45, t=0.11, n_around=68, n_radial=40, wall_spacing=2e-3,
                            growth=1.15)
    U = H.developed_state(desc, flow, steps=10)
    s, m, p = pair(desc, flow, U)
    m.residual(p)
    lam = np.random.default_rng(5).standard_normal((4, m.nc)) * 1e-6
    names = ["m", "p", "t"]
    vals = [0.03, 0.45, 0.11]
    g = s.shape_gradient(names, vals, ["m", "p"], lam, 0.1)
    for k, nm in enumerate(["m", "p"]):
        go = m.shape_gradient(p, lam, nm, m0=0.03, p0=0.45, t=0.11, beta=0.1)
        assert abs(g[k] - go) <= 1e-8 * abs(go), nm


# ------------------------------------------------------------ host paths ---

def test_state_file_round_trip(tmp_path):
    s = capi.Solver(H.shipped_mesh(), H.capi_params(H.SHIPPED))
    s.load_state_file(H.shipped_flow_path())
    assert same(s.state(), H.shipped_flow())
    path = str(tmp_path / "state.txt")
    s.save_state_file(path)
    from oracle import msh2
    assert same(msh2.read_state(path), H.shipped_flow())
    bad = capi.Solver(H.small_mesh(), H.capi_params(H.SHIPPED))
    with pytest.raises(capi.O2FError) as ei:
        bad.load_state_file(H.shipped_flow_path())
    assert "Taille incoh" in str(ei.value)


def test_unknown_boundary_conditions_fail_like_the_reference():
    flow = dict(H.SHIPPED)
    par = H.capi_params(flow)
    par.bc_inflow_outflow_x = 0
    s = capi.Solver(H.tiny_mesh(), par)
    s.init_freestream()
    with pytest.raises(capi.O2FError) as ei:
        s.residual()
    assert ei.value.code == -2


def test_freestream_init_matches():
    s, m, p = pair(H.shipped_mesh(), H.SHIPPED)
    s.init_freestream()
    assert same(s.state(), m.state())


def test_time_kernel_leaves_the_state():
    """o2f_time_kernel (bench.py's per-kernel timing) recomputes the stage
    kernels' outputs at the current state (kernel 5, the fused tile stage,
    writes a spare state buffer): the state, the residual and the face
    states afterwards are those before; bad arguments are refused."""
    s, m, p = pair(H.shipped_mesh(), H.SHIPPED, H.shipped_flow())
    U0 = s.state()
    R0 = s.residual()
    f0 = s.faces()
    for k in range(6):
        assert s.time_kernel(k, 5) > 0.0
    assert same(s.state(), U0)
    assert same(s.residual(), R0)
    f1 = s.faces()
    assert same(f1[0], f0[0]) and same(f1[1], f0[1])
    for k, n in ((6, 5), (-1, 5), (0, 0)):
        with pytest.raises(RuntimeError):
            s.time_kernel(k, n)


def test_adjoint_rhs_and_forces_with_a_cell_of_two_wall_edges():
    """assembleAdjointRHS adds every wall edge into its left cell in wall
    order; the device adds each cell's wall edges in that order with one
    thread per cell.  On a box whose corner triangle owns two wall edges the
    right-hand side and C_L / C_D equal the oracle's bit for bit."""
    desc = H.corner_mesh()
    flow = H.subsonic_small()
    p = H.oracle_params(flow)
    m = O.Mesh(desc)
    m.init_freestream(p)
    m.run_steps(p, 5)
    s, m2, _ = pair(desc, flow, m.state())
    s.residual()
    m2.residual(p)
    assert same(s.adjoint_rhs(50.0), m2.adjoint_rhs(p, 50.0))
    assert tuple(s.lift_drag()) == tuple(m2.lift_drag(p))
