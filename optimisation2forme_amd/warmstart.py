"""Learned warm start: training fields from the solver, training, and the
warm-versus-cold comparison.

    fields = solve_fields(cands, tol, max_steps)      # the solver's own fields (batched)
    net = train(fields, steps, batch)                 # BF16 MFMA training on the GPU
    report = compare(net, held_out, tol, max_steps)   # steps to tol: warm vs cold start

Every candidate is a NACA 4-digit shape with its own O-mesh and flow
(Mach, angle of attack); the solver runs with the Fourier heat flux (the
reference's R x Fourier flux diverges once viscosity matters).  The
network's prediction only replaces the freestream cold start: both starts
iterate to the same density-residual tolerance, so results are those of
the solver and only the step count changes.
"""
import time

import numpy as np

from . import capi, mesher

RHO, P_INF, GAMMA = 1.225, 101325.0, 1.4
A_INF = float(np.sqrt(GAMMA * P_INF / RHO))


def candidates(n, seed, n_around=(56, 96), n_radial=(24, 40)):
    """Seeded shapes and flows in the ranges m ~ U[0, 0.06], p ~ U[0.2, 0.6],
    t ~ U[0.08, 0.15], Mach ~ U[0.5, 0.85], AoA ~ U[0, 4] deg."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        out.append(dict(m=float(rng.uniform(0, 0.06)), p=float(rng.uniform(0.2, 0.6)),
                        t=float(rng.uniform(0.08, 0.15)), mach=float(rng.uniform(0.5, 0.85)),
                        aoa=float(rng.uniform(0, 4)),
                        n_around=int(rng.integers(n_around[0] // 4, n_around[1] // 4 + 1)) * 4,
                        n_radial=int(rng.integers(n_radial[0], n_radial[1] + 1))))
    return out


def candidates_sized(n, seed, triangles=(30e3, 100e3), mach=(0.5, 0.85)):
    """Seeded shapes and flows as candidates() (Mach range given), each with
    an O-mesh of a target size drawn uniformly from `triangles` (about 2.6
    cells around per radial layer, as the batched workload's meshes)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        m, p, t = float(rng.uniform(0, 0.06)), float(rng.uniform(0.2, 0.6)), float(rng.uniform(0.08, 0.15))
        tri = rng.uniform(*triangles) if triangles[1] > triangles[0] else triangles[0]
        n_around = max(16, int(round(np.sqrt(tri / 2 * 2.6) / 4)) * 4)
        n_radial = int(round(tri / (2 * n_around))) + 1
        out.append(dict(m=m, p=p, t=t, mach=float(rng.uniform(*mach)) if mach[1] > mach[0] else mach[0],
                        aoa=float(rng.uniform(0, 4)), n_around=n_around, n_radial=n_radial))
    return out


def develop(cands, final_time, max_steps, wall_spacing=1e-3, use_graph=True):
    """Each candidate's time loop from its freestream cold start to the
    physical time final_time (at most max_steps steps), all in one batch;
    returns (candidate, mesh, state, steps, residual)."""
    solvers, descs = [], []
    for c in cands:
        desc, prm = build(c, wall_spacing=wall_spacing)
        s = capi.Solver(desc, prm)
        s.init_freestream()
        solvers.append(s)
        descs.append(desc)
    b = capi.Batch(solvers)
    steps, _ = b.run_until(final_time, max_steps, use_graph=use_graph)
    out = []
    for c, d, s, n in zip(cands, descs, solvers, steps):
        s.residual()
        out.append((c, d, s.state(), int(n), s.residual_norm()))
    b.close()
    for s in solvers:
        s.close()
    return out


def build(c, wall_spacing=4e-3, radius=20.0):
    desc = mesher.naca_mesh(m=c["m"], p=c["p"], t=c["t"], n_around=c["n_around"], n_radial=c["n_radial"],
                            wall_spacing=wall_spacing, radius=radius, aoa_deg=c["aoa"])
    mu = RHO * c["mach"] * A_INF / 5000.0
    prm = capi.make_params(viscosity=mu, Pr=0.72, cfl_number=0.5, inflow_velocity=c["mach"], venkat_k=5.0,
                           limiter="venkat", interp_mode="mvc", heat_flux="fourier")
    return desc, prm


def cell_centres(desc):
    """Vertex means of the cells (Cell::centre, src/mesh.cpp:123-130),
    triangles and quads."""
    cv = np.asarray(desc["cell_vtx"])
    nv = np.asarray(desc["cell_nv"])
    xy = np.asarray(desc["xy"])
    ctr = xy[cv[:, :3]].sum(axis=1)
    quad = nv == 4
    ctr[quad] += xy[cv[quad, 3]]
    ctr /= nv[:, None]
    return ctr[:, 0], ctr[:, 1]


def wall_distance(desc, cx, cy, k_near=8):
    """Distance of each point to the nearest wall edge segment: the exact
    segment distances to the wall edges at the k_near wall vertices nearest
    to the point (a k-d tree over the wall vertices), which hold the nearest
    segment on any mesh whose wall edges are short against the distance to
    the opposite surface."""
    from scipy.spatial import cKDTree
    xy = np.asarray(desc["xy"], dtype=np.float64)
    bv = np.asarray(desc["bnd_vtx"])
    wall = bv[np.asarray(desc["bnd_tag"]) == ord("w")]
    verts, inv = np.unique(wall.ravel(), return_inverse=True)
    inv = inv.reshape(wall.shape)
    # the (at most two) wall edges of each wall vertex
    edges_of = [[] for _ in range(len(verts))]
    for e, (i, j) in enumerate(inv):
        edges_of[i].append(e)
        edges_of[j].append(e)
    width = max(len(x) for x in edges_of)
    adj = np.array([x + [x[0]] * (width - len(x)) for x in edges_of])
    P = np.stack([cx, cy], 1)
    k = min(k_near, len(verts))
    _, near = cKDTree(xy[verts]).query(P, k=k)
    near = near.reshape(len(P), k)
    a, b = xy[wall[:, 0]], xy[wall[:, 1]]
    d = np.full(len(P), np.inf)
    for j in range(k):
        for w in range(width):
            e = adj[near[:, j], w]
            A, AB = a[e], b[e] - a[e]
            l2 = np.maximum(np.sum(AB * AB, 1), 1e-300)
            t = np.clip(np.sum((P - A) * AB, 1) / l2, 0.0, 1.0)
            Q = A + t[:, None] * AB
            d = np.minimum(d, np.sqrt(np.sum((Q - P) ** 2, 1)))
    return d


def wall_distance_brute(desc, cx, cy):
    """The same by testing every wall edge (the reference for small meshes)."""
    xy = np.asarray(desc["xy"])
    bv = np.asarray(desc["bnd_vtx"])
    wall = bv[np.asarray(desc["bnd_tag"]) == ord("w")]
    a, b = xy[wall[:, 0]], xy[wall[:, 1]]
    d = np.full(cx.shape, np.inf)
    P = np.stack([cx, cy], 1)
    for k in range(0, len(a), 64):
        A, B = a[k:k + 64], b[k:k + 64]
        AB = B - A
        l2 = np.maximum(np.sum(AB * AB, 1), 1e-300)
        s = np.clip(((P[:, None, :] - A[None]) * AB[None]).sum(2) / l2[None], 0, 1)
        Q = A[None] + s[..., None] * AB[None]
        d = np.minimum(d, np.sqrt(((Q - P[:, None, :]) ** 2).sum(2)).min(1))
    return d


def raw_inputs(desc, c):
    cx, cy = cell_centres(desc)
    wd = wall_distance(desc, cx, cy)
    n = cx.shape[0]
    return np.stack([cx, cy, wd, np.full(n, c["m"]), np.full(n, c["p"]), np.full(n, c["t"]),
                     np.full(n, c["mach"]), np.full(n, c["aoa"])], 1).astype(np.float32)


def normalise(U, mach):
    """Network targets of a state (4 x n_cells)."""
    u = mach * A_INF
    Einf = P_INF / (GAMMA - 1.0) + 0.5 * RHO * u * u
    return np.stack([U[0] / RHO - 1.0, U[1] / (RHO * A_INF) - mach, U[2] / (RHO * A_INF),
                     (U[3] - Einf) / P_INF], 1).astype(np.float32)


def solve_fields(cands, tol, max_steps, check_every=50, use_graph=True):
    """Cold-start every candidate and run it (batched, one launch per stage
    kernel) until its density residual is <= tol; returns a list of
    (candidate, mesh, state, steps, residual)."""
    solvers, descs = [], []
    for c in cands:
        desc, prm = build(c)
        s = capi.Solver(desc, prm)
        s.init_freestream()
        solvers.append(s)
        descs.append(desc)
    b = capi.Batch(solvers)
    steps, res = b.run(tol, max_steps, check_every=check_every, use_graph=use_graph)
    out = [(c, d, s.state(), int(n), float(r)) for c, d, s, n, r in zip(cands, descs, solvers, steps, res)]
    b.close()
    for s in solvers:
        s.close()
    return out


def converge(cases, tol, max_steps, check_every=100, use_graph=True):
    """Each case (desc, params, shape / flow labels m, p, t, mach, aoa) from
    its freestream cold start, batched, until its density residual is <=
    tol; returns (case, desc, state, steps, residual) per case, the
    training fields of train()."""
    solvers = []
    for c in cases:
        s = capi.Solver(c["desc"], c["params"])
        s.init_freestream()
        solvers.append(s)
    b = capi.Batch(solvers)
    steps, res = b.run(tol, max_steps, check_every=check_every, use_graph=use_graph)
    out = [(c, c["desc"], s.state(), int(n), float(r)) for c, s, n, r in zip(cases, solvers, steps, res)]
    b.close()
    for s in solvers:
        s.close()
    return out


def dataset(fields):
    raw = np.concatenate([raw_inputs(d, c) for c, d, _, _, _ in fields])
    tgt = np.concatenate([normalise(U, c["mach"]) for c, _, U, _, _ in fields])
    return raw, tgt


def train(fields, steps, batch=65536, lr=1e-3, seed=0):
    raw, tgt = dataset(fields)
    net = capi.WarmStartNet(seed=seed)
    net.set_data(raw, tgt)
    t0 = time.perf_counter()
    losses = net.train(steps, batch=batch, lr=lr, seed=seed + 1)
    return net, losses, time.perf_counter() - t0


def compare(net, cands, tol, max_steps, check_every=50, use_graph=True):
    """Per candidate: steps and wall time to the density-residual tolerance
    from the freestream cold start and from the network's prediction
    (inference included), C_L / C_D of both end states, and the
    surrogate's error against the cold run's end state."""
    rows = []
    for c in cands:
        desc, prm = build(c)
        cold = capi.Solver(desc, prm)
        warm = capi.Solver(desc, prm)
        cold.init_freestream()
        t0 = time.perf_counter()
        n_cold, r_cold = cold.run_to_residual(tol, max_steps, check_every=check_every, use_graph=use_graph)
        t_cold = time.perf_counter() - t0
        t0 = time.perf_counter()
        net.warm_start(warm, c["m"], c["p"], c["t"], c["aoa"])
        t_inf = time.perf_counter() - t0
        U0 = warm.state()
        t0 = time.perf_counter()
        n_warm, r_warm = warm.run_to_residual(tol, max_steps, check_every=check_every, use_graph=use_graph)
        t_warm = time.perf_counter() - t0 + t_inf
        Uc, Uw = cold.state(), warm.state()
        cold.residual()
        warm.residual()
        target = normalise(Uc, c["mach"])
        pred = normalise(U0, c["mach"])
        rows.append(dict(candidate=c, cells=cold.nc, cold_steps=int(n_cold), warm_steps=int(n_warm),
                         cold_res=float(r_cold), warm_res=float(r_warm), cold_s=t_cold, warm_s=t_warm,
                         inference_s=t_inf, cold_cl_cd=list(cold.lift_drag()), warm_cl_cd=list(warm.lift_drag()),
                         surrogate_rmse=[float(v) for v in np.sqrt(np.mean((pred - target) ** 2, 0))],
                         state_rel_diff=float(np.max(np.abs(Uw - Uc)) / np.max(np.abs(Uc)))))
        cold.close()
        warm.close()
    return rows
