"""Batched shape candidates and the run-to-residual loop on the GPU.

Bar: every member of a batch (meshes of different sizes, flows of their
own) ends bit-identical to its own single-mesh run, whatever the launch
shapes (threads per workgroup, persistent grid), with or without hipGraph
replay; the autotuner may change only speed.  The run-to-residual loop
stops at the same step as the oracle's and leaves the same state.
"""
import sys

import numpy as np
import pytest

from optimisation2forme_amd import capi, mesher
from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _announce(request):
    print("\n[gpu case] %s" % request.node.nodeid, flush=True)
    sys.stdout.flush()
    yield


# Three candidates: different NACA shapes, sizes and flow conditions.
CANDIDATES = [
    dict(m=0.0, p=0.4, t=0.12, n_around=40, n_radial=14, aoa=2.0, mach=0.6),
    dict(m=0.03, p=0.45, t=0.11, n_around=72, n_radial=26, aoa=1.0, mach=0.85),
    dict(m=0.05, p=0.3, t=0.14, n_around=56, n_radial=40, aoa=3.5, mach=0.5),
]


def candidate(k, limiter=0):
    c = CANDIDATES[k]
    desc = mesher.naca_mesh(m=c["m"], p=c["p"], t=c["t"], n_around=c["n_around"], n_radial=c["n_radial"],
                            wall_spacing=4e-3, radius=20.0, aoa_deg=c["aoa"])
    # Fourier heat flux: the reference's (R x Fourier) diverges at Re 5000
    flow = dict(viscosity=H.re_viscosity(c["mach"]), Pr=0.72, cfl=0.5, mach=c["mach"], venkat_k=5.0,
                limiter=limiter, interp_mode=1, heat_flux=1)
    return desc, flow


def solvers(limiter=0):
    out = []
    for k in range(len(CANDIDATES)):
        desc, flow = candidate(k, limiter)
        s = capi.Solver(desc, H.capi_params(flow))
        s.init_freestream()
        out.append(s)
    return out


def single_states(nsteps, limiter=0, faces=False):
    """Each candidate alone, on the four-kernel stages: its state after
    nsteps steps (and its last stage's face states)."""
    out = []
    for s in solvers(limiter):
        s.set_tile_cells(0)
        s.run_steps(nsteps)
        out.append((s.state(), s.faces()) if faces else s.state())
    return out


SHAPES = {
    "default": [(256, 0)] * 4,
    "small_blocks_persistent": [(64, 1024), (64, 2048), (128, 1024), (64, 512)],
    "mixed": [(128, 0), (256, 1024), (64, 0), (128, 2048)],
}


@pytest.mark.parametrize("shape", sorted(SHAPES))
@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_members_bit_identical_to_single_runs(shape, use_graph):
    """The four-kernel batched stages at several launch shapes."""
    n = 23  # two graph replays of 10 steps and a tail of 3
    ref = single_states(n)
    ss = solvers()
    b = capi.Batch(ss)
    b.set_tile_cells(0)
    assert b.tile_info()["tiles"] == 0
    b.set_shapes(SHAPES[shape])
    assert b.shapes() == SHAPES[shape]
    steps, res = b.run(-1.0, n, check_every=10, use_graph=use_graph)
    assert list(steps) == [n] * len(ss)
    for k, s in enumerate(ss):
        assert np.array_equal(s.state(), ref[k]), (shape, use_graph, k)
    b.close()


@pytest.mark.parametrize("cells", [-1, 40, 300])
@pytest.mark.parametrize("use_graph", [True, False])
def test_tiled_batch_bit_identical_to_single_runs(cells, use_graph):
    """The fused tile stages over every member's tiles (the members' tables
    concatenated, one launch per stage): every member's state and the face
    states its last stage leaves equal its own four-kernel run, at the
    automatic tile size and at small and large caps."""
    n = 23
    ref = single_states(n, faces=True)
    ss = solvers()
    b = capi.Batch(ss)
    b.set_tile_cells(cells)
    info = b.tile_info()
    nc = [len(candidate(k)[0]["cell_nv"]) for k in range(len(CANDIDATES))]
    assert info["tiles"] >= sum(-(-c // info["cells"]) for c in nc), info
    assert info["tiles_per_cu"] >= 1 and 0 < info["lds_bytes"] <= 160 * 1024, info
    if cells > 0:
        assert info["cells"] == cells
    steps, _ = b.run(-1.0, n, check_every=10, use_graph=use_graph)
    assert list(steps) == [n] * len(ss)
    for k, s in enumerate(ss):
        U, (fl, fr) = ref[k]
        assert np.array_equal(s.state(), U), (cells, use_graph, k)
        got_l, got_r = s.faces()
        assert np.array_equal(got_l, fl) and np.array_equal(got_r, fr), (cells, use_graph, k)
    b.close()


def test_tiled_batch_with_quads_and_triangles():
    """The reference's shipped quad mesh (its own flow and state) in one
    batch with two triangle candidates: the four-slot batched tile stage
    over all three, each member bitwise its own four-kernel run."""
    n = 9
    members = [(H.shipped_mesh(), H.SHIPPED, H.shipped_flow())]
    for k in (0, 2):
        desc, flow = candidate(k)
        members.append((desc, flow, None))
    ref, ss = [], []
    for desc, flow, U in members:
        for out in (ref, ss):
            s = capi.Solver(desc, H.capi_params(flow))
            if U is None:
                s.init_freestream()
            else:
                s.set_state(U)
            out.append(s)
    for s in ref:
        s.set_tile_cells(0)
        s.run_steps(n)
    b = capi.Batch(ss)
    assert b.tile_info()["tiles"] > 0
    b.run(-1.0, n, check_every=4, use_graph=True)
    for k, (s, r) in enumerate(zip(ss, ref)):
        assert np.array_equal(s.state(), r.state()), k
        assert np.array_equal(s.faces()[0], r.faces()[0]) and np.array_equal(s.faces()[1], r.faces()[1]), k
    b.close()


def test_tiled_batch_falls_back_when_tiles_do_not_fit():
    """A cap whose tiles need more LDS than a workgroup has: the batch runs
    the four-kernel batched stages (tile_info reports none, with the reason)
    and the members are unchanged."""
    n = 5
    ref = single_states(n)
    ss = solvers()
    b = capi.Batch(ss)
    b.set_tile_cells(100000)
    assert b.tile_info()["tiles"] == 0
    assert "four-kernel" in capi.lib().o2f_batch_last_error(b.h).decode()
    b.run(-1.0, n, check_every=5, use_graph=False)
    for k, s in enumerate(ss):
        assert np.array_equal(s.state(), ref[k]), k
    b.close()


def test_tiled_batch_follows_a_change_of_stencils():
    """A member's interpolation mode changed between two batch runs: the
    batch rebuilds its tiles on the new stencils (bitwise the four-kernel
    stages on the new mode)."""
    ss = solvers()
    b = capi.Batch(ss)
    b.run(-1.0, 3, check_every=3, use_graph=False)
    desc, flow = candidate(1)
    flow = dict(flow, interp_mode=0)
    ss[1].set_params(H.capi_params(flow))
    ref = capi.Solver(desc, H.capi_params(flow))
    ref.set_tile_cells(0)
    ref.set_state(ss[1].state())
    b.run(-1.0, 4, check_every=4, use_graph=False)
    ref.run_steps(4)
    assert np.array_equal(ss[1].state(), ref.state())
    b.close()


@pytest.mark.parametrize("exact", [False, True])
def test_batch_adjoint_solves_bitwise_each_members_own(exact):
    """o2f_batch_bicgstab: the members' adjoint solves driven together from
    several host threads, each on its member's stream; every member's
    lambda, iteration count, residual and status equal its own solve's."""
    ss = solvers()
    b = capi.Batch(ss)
    b.run(-1.0, 30, check_every=10)
    kw = dict(tol=1e-8, chunk=40, max_total=120)
    lams, it, rr, code = b.bicgstab(0.1, exact=exact, **kw)
    assert np.all(it > 0)
    for k, s in enumerate(ss):
        one = s.bicgstab_exact(0.1, check=False, **kw) if exact else s.bicgstab(0.1, check=False, **kw)
        lam, i, r, st = one
        assert np.all(np.isfinite(lam))
        assert np.array_equal(lam, lams[k]) and i == it[k] and r == rr[k] and st == code[k], (exact, k)
    b.close()


def test_batch_barth_jespersen_and_consecutive_runs():
    ref = single_states(30, limiter=1)
    ss = solvers(limiter=1)
    b = capi.Batch(ss)
    b.set_shapes(SHAPES["mixed"])
    b.run(-1.0, 12, check_every=4, use_graph=True)
    b.run(-1.0, 18, check_every=6, use_graph=True)   # a new graph, same trajectory
    for k, s in enumerate(ss):
        assert np.array_equal(s.state(), ref[k]), k


def test_tile_size_autotuner_changes_only_speed():
    """o2f_batch_autotune on automatic tile stages times three tile sizes
    and keeps one: the state is left alone and the run afterwards is
    bitwise the members' own four-kernel runs."""
    n = 12
    ref = single_states(n)
    ss = solvers()
    b = capi.Batch(ss)
    before = b.tile_info()
    U0 = [s.state() for s in ss]
    b.autotune(reps=2)
    after = b.tile_info()
    assert after["tiles"] > 0 and after["cells"] in (before["cells"], before["cells"] * 9 // 10,
                                                     before["cells"] * 4 // 5), (before, after)
    for k, s in enumerate(ss):
        assert np.array_equal(s.state(), U0[k]), k
    b.run(-1.0, n, check_every=6, use_graph=True)
    for k, s in enumerate(ss):
        assert np.array_equal(s.state(), ref[k]), k
    b.close()


def test_autotuner_changes_only_speed():
    ref = single_states(15)
    ss = solvers()
    b = capi.Batch(ss)
    shapes = b.autotune(reps=2)
    assert len(shapes) == 4
    for blk, grid in shapes:
        assert blk in (64, 128, 256) and grid >= 0
    U0 = [s.state() for s in ss]
    for k, s in enumerate(ss):   # residual-only timing passes leave the state alone
        s2 = capi.Solver(*[candidate(k)[0], H.capi_params(candidate(k)[1])])
        s2.init_freestream()
        assert np.array_equal(U0[k], s2.state())
    b.run(-1.0, 15, check_every=5, use_graph=True)
    for k, s in enumerate(ss):
        assert np.array_equal(s.state(), ref[k]), k


def test_fourier_heat_flux_residual_and_steps_bit_identical():
    desc, flow = candidate(1)
    p = H.oracle_params(flow)
    m = O.Mesh(desc)
    m.init_freestream(p)
    m.run_steps(p, 5)
    s = capi.Solver(desc, H.capi_params(flow))
    s.set_state(m.state())
    assert np.array_equal(s.residual(), m.residual(p))
    s.run_steps(4)
    m.run_steps(p, 4)
    assert H.rel_err(s.state(), m.state()) <= 1e-10


def test_run_to_residual_matches_the_oracle():
    desc, flow = candidate(0)
    p = H.oracle_params(flow)
    m = O.Mesh(desc)
    m.init_freestream(p)
    _, r0 = m.run_to_residual(p, 1e30, 5, 1)      # stops after step 0: r(U^0)
    m.init_freestream(p)
    tol = 0.6 * r0
    n_cpu, r_cpu = m.run_to_residual(p, tol, 400, 5)
    assert 1 < n_cpu < 400, (n_cpu, r_cpu, r0)
    s = capi.Solver(desc, H.capi_params(flow))
    s.init_freestream()
    n_gpu, r_gpu = s.run_to_residual(tol, 400, check_every=5, use_graph=True)
    assert n_gpu == n_cpu
    assert abs(r_gpu - r_cpu) <= 1e-12 * r_cpu
    assert H.rel_err(s.state(), m.state()) <= 1e-10


def test_batch_members_stop_independently():
    ss = solvers()
    b = capi.Batch(ss)
    # r at the cold start of every member, then a tolerance each reaches at its own step
    steps1, r0 = b.run(1e30, 3, check_every=1, use_graph=False)
    assert list(steps1) == [1, 1, 1]
    for s in ss:
        s.init_freestream()
    tol = float(np.min(r0)) * 0.7
    steps, res = b.run(tol, 300, check_every=3, use_graph=True)
    for k, s in enumerate(ss):
        ref = capi.Solver(candidate(k)[0], H.capi_params(candidate(k)[1]))
        ref.init_freestream()
        ref.run_steps(int(steps[k]))
        assert np.array_equal(s.state(), ref.state()), k
        if steps[k] < 300:
            assert res[k] <= tol


def test_batch_run_until_matches_each_members_own_time_loop():
    """o2f_batch_run_until: every member stops at its own step once its t
    reaches the final time, bitwise where its own o2f_run_until stops."""
    descs = [mesher.naca_mesh(m=0.02, p=0.4, t=0.12, n_around=n, n_radial=r, wall_spacing=2e-3, radius=10.0)
             for n, r in ((48, 20), (64, 30), (56, 24))]
    flow = dict(H.transonic(limiter=0), heat_flux=1)
    solvers = [capi.Solver(d, H.capi_params(flow)) for d in descs]
    for s in solvers:
        s.init_freestream()
    final = 40.5 * solvers[1].dt()
    b = capi.Batch(solvers)
    steps, t = b.run_until(final, 10000)
    b.close()
    assert len(set(steps.tolist())) > 1          # the members stop at different steps
    for s, d, n, tt in zip(solvers, descs, steps, t):
        one = capi.Solver(d, H.capi_params(flow))
        one.init_freestream()
        n1, t1, _, _ = one.run_until(final)
        assert n1 == n and t1 == tt and tt >= final
        assert np.array_equal(one.state(), s.state())
