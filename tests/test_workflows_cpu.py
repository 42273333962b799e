"""CPU checks of the workload builders: shape-candidate sampling, the
network inputs (wall distance, normalisation), the benchmark's batch
candidates and its one-GPU guard."""
import os
import subprocess
import sys

import numpy as np

import bench
from optimisation2forme_amd import mesher, warmstart as W
from tests import helpers as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_candidates_are_seeded_and_in_range():
    a, b = W.candidates(50, seed=3), W.candidates(50, seed=3)
    assert a == b and a != W.candidates(50, seed=4)
    for c in a:
        assert 0 <= c["m"] <= 0.06 and 0.2 <= c["p"] <= 0.6 and 0.08 <= c["t"] <= 0.15
        assert 0.5 <= c["mach"] <= 0.85 and 0 <= c["aoa"] <= 4
        assert c["n_around"] % 4 == 0


def test_wall_distance_matches_brute_force():
    desc = H.tiny_mesh()
    cx, cy = W.cell_centres(desc)
    d = W.wall_distance(desc, cx, cy)
    xy = np.asarray(desc["xy"])
    bv = np.asarray(desc["bnd_vtx"])[np.asarray(desc["bnd_tag"]) == ord("w")]
    for i in range(0, len(cx), 7):
        best = np.inf
        for a, b in bv:
            A, B = xy[a], xy[b]
            t = np.clip(np.dot([cx[i] - A[0], cy[i] - A[1]], B - A) / np.dot(B - A, B - A), 0, 1)
            best = min(best, np.hypot(*(A + t * (B - A) - [cx[i], cy[i]])))
        assert abs(d[i] - best) <= 1e-12 * max(1.0, best)
    assert np.all(d > 0)


def test_normalised_freestream_is_zero():
    for mach in (0.5, 0.85):
        u = mach * W.A_INF
        E = W.P_INF / (W.GAMMA - 1) + 0.5 * W.RHO * u * u
        n = 5
        U = np.array([np.full(n, W.RHO), np.full(n, W.RHO * u), np.zeros(n), np.full(n, E)])
        assert np.max(np.abs(W.normalise(U, mach))) < 1e-6


def test_raw_inputs_layout():
    c = W.candidates(1, seed=9)[0]
    desc, _ = W.build(c)
    raw = W.raw_inputs(desc, c)
    assert raw.shape == (len(desc["cell_nv"]), 8) and raw.dtype == np.float32
    assert np.allclose(raw[:, 3:], [c["m"], c["p"], c["t"], c["mach"], c["aoa"]], rtol=1e-6)


def test_batch_candidates_cover_the_specified_ranges():
    cands = bench.batch_candidates()
    assert len(cands) == 64
    tri = [2 * c["n_around"] * (c["n_radial"] - 1) for c in cands]
    assert 19000 < min(tri) and max(tri) < 122000
    assert 4.3e6 < sum(tri) < 4.8e6
    for c in cands[:8]:          # every candidate meshes without folding
        mesher.naca_mesh(m=c["m"], p=c["p"], t=c["t"], n_around=c["n_around"], n_radial=c["n_radial"],
                         wall_spacing=1e-3, radius=20.0, aoa_deg=c["aoa"])


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_wall_distance_tree_equals_brute_force():
    """warmstart.wall_distance (a k-d tree over the wall vertices, exact
    distances to their edges) equals the distance to every wall edge, on an
    O-mesh with a blunt and one with a sharp trailing edge and on the
    reference's shipped quads."""
    from optimisation2forme_amd import mesher, warmstart as W
    from tests import helpers as H
    for desc in (H.small_mesh(), mesher.naca_mesh(m=0.04, p=0.3, t=0.1, n_around=96, n_radial=30,
                                                  wall_spacing=1e-3, closed_te=True), H.shipped_mesh()):
        cx, cy = W.cell_centres(desc)
        assert np.array_equal(W.wall_distance(desc, cx, cy), W.wall_distance_brute(desc, cx, cy))
