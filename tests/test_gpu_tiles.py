"""Fused tile stages (csrc/o2f_tiles.hip) against the four-kernel stages,
on the GPU, through the C ABI.

The time-stepping entry points evaluate each SSP-RK3 stage in one launch
over compact cell tiles (o2f_set_tile_cells: -1 automatic, the default; 0
the four-kernel stages; n a cap of n cells).  Every entity a tile forms uses
the four-kernel stage's arithmetic, so the bar is bit-identity: the state,
the dt sequence, the time, and the intermediates the last stage leaves (face
states, stored residual) after the same steps from the same state.  The
four-kernel stages are themselves bitwise the CPU oracle
(tests/test_gpu_parity.py); one test closes the loop against the oracle
directly.
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


def same(a, b):
    return np.array_equal(np.asarray(a), np.asarray(b))


def small_flow(limiter=0):
    """The 10k-triangle case (NACA 0012, Mach 0.6, Re 5000) with the Fourier
    heat flux, which keeps these states physical (with the reference's R
    times Fourier flux its wall cells lose positivity within ~40 steps)."""
    return dict(H.subsonic_small(limiter=limiter), heat_flux=1)


def run(desc, flow, U, nsteps, tile_cells):
    s = capi.Solver(desc, H.capi_params(flow))
    s.set_tile_cells(tile_cells)
    s.set_state(U)
    dts, t = s.run_steps(nsteps)
    out = dict(U=s.state(), dts=dts, t=t, faces=s.faces(), rnorm=s.residual_norm(), cl_cd=s.lift_drag(),
               info=s.tile_info())
    s.close()
    return out


def assert_same_run(a, b):
    assert np.all(np.isfinite(a["U"]))
    assert same(a["U"], b["U"])
    assert same(a["dts"], b["dts"]) and a["t"] == b["t"]
    assert same(a["faces"][0], b["faces"][0]) and same(a["faces"][1], b["faces"][1])
    assert a["rnorm"] == b["rnorm"] and a["cl_cd"] == b["cl_cd"]


@pytest.fixture(scope="module")
def small_developed():
    return H.developed_state(H.small_mesh(), small_flow(), steps=20)


@pytest.mark.parametrize("limiter", [0, 1])
def test_tiled_steps_bitwise_four_kernel_steps_10k(small_developed, limiter):
    flow = small_flow(limiter)
    a = run(H.small_mesh(), flow, small_developed, 12, -1)
    b = run(H.small_mesh(), flow, small_developed, 12, 0)
    assert a["info"]["tiles"] > 0 and b["info"]["tiles"] == 0
    assert_same_run(a, b)


def test_tiled_steps_bitwise_four_kernel_steps_shipped_quads():
    U = H.shipped_flow()
    a = run(H.shipped_mesh(), H.SHIPPED, U, 25, -1)
    b = run(H.shipped_mesh(), H.SHIPPED, U, 25, 0)
    assert a["info"]["tiles"] > 0
    assert_same_run(a, b)


@pytest.mark.parametrize("cells", [37, 64, 150, 400])
def test_tile_sizes_bitwise(small_developed, cells):
    """Tiles of a few dozen cells (large rims, many tiles per CU) up to a few
    hundred (two rounds of workgroups): the same steps bit for bit."""
    flow = small_flow()
    a = run(H.small_mesh(), flow, small_developed, 6, cells)
    b = run(H.small_mesh(), flow, small_developed, 6, 0)
    assert 0 < a["info"]["tiles"] and a["info"]["cells"] == cells
    assert_same_run(a, b)


def test_oversized_tiles_fall_back_to_the_four_kernels(small_developed):
    """A tile whose tables need more LDS than a workgroup has is refused:
    the stages stay the four-kernel ones and the run is unchanged."""
    flow = small_flow()
    a = run(H.small_mesh(), flow, small_developed, 3, 100000)
    b = run(H.small_mesh(), flow, small_developed, 3, 0)
    assert a["info"]["tiles"] == 0
    assert_same_run(a, b)


@pytest.mark.parametrize("tiles", [-1, 37])
@pytest.mark.parametrize("max_steps", [-1, 7])
def test_tiled_run_until_bitwise(small_developed, max_steps, tiles):
    """o2f_run_until on tiles: the same stopping step, times and state, and
    the last stage's intermediates (formed after the loop from that stage's
    input) equal the four-kernel loop's, which stores them every step.
    tiles -1: one tile per CU (the four-lane stages, which start each step
    themselves); 37 cells: more tiles than CUs (the 256-thread stages, each
    step after the loop's first started from the speed max the previous
    step's last stage folded, the gated steps after the stop included)."""
    flow = small_flow()
    res = []
    for cells in (tiles, 0):
        s = capi.Solver(H.small_mesh(), H.capi_params(flow))
        s.set_tile_cells(cells)
        s.set_state(small_developed)
        dt0 = s.dt()
        r = s.run_until(10.5 * dt0, max_steps)
        res.append((r, s.state(), s.faces(), s.residual_norm(), s.lift_drag(), s.tile_info()["tiles"]))
        s.close()
    (ra, Ua, fa, na, la, ta), (rb, Ub, fb, nb, lb, tb) = res
    assert ta > 0 and tb == 0
    assert tiles == -1 or ta > 256  # more tiles than the MI355X's 256 CUs
    assert ra == rb
    assert ra[0] == 7 if max_steps == 7 else ra[0] >= 10
    assert np.all(np.isfinite(Ua))
    assert same(Ua, Ub) and same(fa[0], fb[0]) and same(fa[1], fb[1])
    assert na == nb and la == lb


@pytest.mark.parametrize("tiles", [-1, 37])
def test_tiled_steps_with_a_nan_cell_bitwise_four_kernel_steps(small_developed, tiles):
    """A state with one NaN cell (a blown-up solve): every dt reduction skips
    the NaN speeds (the one-block step start, the four-lane stages' start,
    and the 256-thread stages' carried 64-bit atomic max alike), so the dts
    stay finite and equal, and the NaN spreads over the same cells with the
    finite values bit for bit the same."""
    flow = small_flow()
    U = np.array(small_developed, copy=True)
    nc = U.shape[-1]
    U[..., nc // 2] = np.nan
    a = run(H.small_mesh(), flow, U, 4, tiles)
    b = run(H.small_mesh(), flow, U, 4, 0)
    assert a["info"]["tiles"] > 0 and (tiles == -1 or a["info"]["tiles"] > 256)
    assert np.all(np.isfinite(a["dts"])) and same(a["dts"], b["dts"]) and a["t"] == b["t"]
    assert np.isnan(a["U"]).any() and np.isfinite(a["U"]).any()
    assert np.array_equal(a["U"], b["U"], equal_nan=True)


@pytest.mark.parametrize("limiter", [0, 1])
def test_tiled_steps_bitwise_four_kernel_steps_200k_developed(limiter):
    """The benchmark mesh (NACA 2412, 199,680 triangles, Mach 0.85) after
    300 steps from the cold start: 20 more steps on tiles and on the four
    kernels, bit for bit, with either limiter; and the tiling is the
    automatic one the benchmark uses."""
    flow = dict(H.transonic(limiter=limiter), heat_flux=1)
    s = capi.Solver(H.big_mesh(), H.capi_params(flow))
    s.set_tile_cells(0)
    s.init_freestream()
    s.run_steps(300)
    U = s.state()
    s.close()
    a = run(H.big_mesh(), flow, U, 20, -1)
    b = run(H.big_mesh(), flow, U, 20, 0)
    info = a["info"]
    assert info["tiles"] > 0 and info["tiles"] <= 256 * info["tiles_per_cu"]
    assert_same_run(a, b)


def test_tiled_steps_match_the_oracle_10k(small_developed):
    """Four tiled steps against the CPU oracle's: within 1e-10 of each
    variable's scale (the step's dt differs in its last bits, the device's
    hypot against glibc's; every stage is otherwise bitwise)."""
    flow = small_flow()
    s = capi.Solver(H.small_mesh(), H.capi_params(flow))
    s.set_state(small_developed)
    dts_g, _ = s.run_steps(4)
    assert s.tile_info()["tiles"] > 0
    m = O.Mesh(H.small_mesh())
    p = H.oracle_params(flow)
    m.set_state(small_developed)
    dts_o = m.run_steps(p, 4)
    assert np.allclose(dts_g, dts_o, rtol=1e-14, atol=0)
    assert max(H.var_rel_err(s.state(), m.state())) < 1e-10


@pytest.mark.parametrize("tiles", [-1, 37])
@pytest.mark.parametrize("check_every", [1, 7])
def test_tiled_run_to_residual_bitwise_batched_stages(small_developed, check_every, tiles):
    """o2f_run_to_residual on tiles against the batched four-kernel stages
    (o2f_set_tile_cells(0)): the same stopping step and residual, state and
    last-stage face states, checking every step or every seventh, on the
    four-lane stages (tiles -1) and the 256-thread ones (37 cells)."""
    flow = small_flow()
    # the residual of step 49's start state (checked whether every step or
    # every seventh is): a tolerance some step up to 49 reaches
    s0 = capi.Solver(H.small_mesh(), H.capi_params(flow))
    s0.set_tile_cells(0)
    s0.set_state(small_developed)
    n0, r49 = s0.run_to_residual(1e-300, 50, check_every=1)
    s0.close()
    assert n0 == 50 and r49 > 0
    tol = r49
    res = []
    for cells in (tiles, 0):
        s = capi.Solver(H.small_mesh(), H.capi_params(flow))
        s.set_tile_cells(cells)
        s.set_state(small_developed)
        out = s.run_to_residual(tol, 400, check_every=check_every)
        res.append((out, s.state(), s.faces(), s.tile_info()["tiles"]))
        s.close()
    (oa, Ua, fa, ta), (ob, Ub, fb, tb) = res
    assert ta > 0 and tb == 0
    assert tiles == -1 or ta > 256  # more tiles than the MI355X's 256 CUs
    assert oa == ob and 0 < oa[0] <= 50 and oa[1] <= tol
    assert np.all(np.isfinite(Ua)) and same(Ua, Ub)
    assert same(fa[0], fb[0]) and same(fa[1], fb[1])


def test_tiled_run_to_residual_without_tolerance(small_developed):
    """With no tolerance every step runs: the same end state either way."""
    flow = small_flow()
    states = []
    for cells in (-1, 0):
        s = capi.Solver(H.small_mesh(), H.capi_params(flow))
        s.set_tile_cells(cells)
        s.set_state(small_developed)
        n, r = s.run_to_residual(0.0, 9)
        assert n == 9 and r == -1.0
        states.append(s.state())
        s.close()
    assert same(states[0], states[1])
