"""Structured-topology triangle meshes of NACA 4-digit airfoils.

Gmsh is not available here, so the triangle meshes of the transonic and
batched workloads come from this generator: an O-grid between the airfoil
and a circular far field, every quad split into two triangles (interior
vertex valence 6).  The topology is fixed for given (n_around, n_radial), so
node positions move smoothly with (m, p, t, AoA), which central finite
differences on re-meshed shapes need.

Conventions follow the reference: chord 1 with the leading edge at
x = pos_x = 0.355, y = 0 (src/shape_optimization.cpp:830-834); thickness is
added vertically to the mean line, y = yc +/- yt (the same construction as
applyShapePerturbationNACA12, :868-880); the trailing edge is open
(coefficient -0.1015) and closed by a flat base that belongs to the wall.
Boundary tags: 'w' on the airfoil (physical name "Airfoil"), 'f' on the far
field ("FreeStream").  The angle of attack rotates the airfoil about its
leading edge, nose up, so the freestream stays along +x as in the reference
(src/boundary.cpp:13-15).
"""
import math

import numpy as np


def naca4(m, p, t, xr, closed_te=False):
    """Mean line and half thickness at chord fraction xr (vectorised); with
    closed_te the last thickness coefficient is -0.1036, which closes the
    trailing edge (yt(1) = 0)."""
    xr = np.clip(xr, 0.0, 1.0)
    a4 = -0.1036 if closed_te else -0.1015
    yt = 5.0 * t * (0.2969 * np.sqrt(xr) - 0.1260 * xr - 0.3516 * xr ** 2 + 0.2843 * xr ** 3
                    + a4 * xr ** 4)
    if m == 0.0:
        return np.zeros_like(xr), yt
    yc = np.where(xr <= p, m / (p * p) * (2.0 * p * xr - xr * xr),
                  m / ((1.0 - p) ** 2) * ((1.0 - 2.0 * p) + 2.0 * p * xr - xr * xr))
    return yc, yt


def airfoil_contour(m, p, t, n_around, n_base=4, pos_x=0.355, chord=1.0, closed_te=False):
    """Closed clockwise contour: lower surface TE->LE, upper LE->TE, then the
    flat trailing-edge base back to the start (n_base segments); with
    closed_te the two surfaces meet at a sharp trailing edge (no base)."""
    if closed_te:
        n_base = 0
    n_side = (n_around - n_base) // 2
    if n_side < 4 or 2 * n_side + n_base != n_around:
        raise ValueError("n_around must be n_base + 2*k with k >= 4")
    beta = np.linspace(0.0, math.pi, n_side + 1)
    xr = 0.5 * (1.0 - np.cos(beta))            # 0 .. 1, clustered at LE and TE
    yc, yt = naca4(m, p, t, xr, closed_te)
    lower = np.stack([xr[::-1], (yc - yt)[::-1]], 1)      # TE -> LE
    upper = np.stack([xr, yc + yt], 1)                    # LE -> TE
    pts = np.concatenate([lower[:-1], upper[:-1]])        # 2*n_side points, LE once
    if n_base:
        te_up, te_lo = upper[-1], lower[0]
        base = [te_up + (te_lo - te_up) * (k / n_base) for k in range(n_base)]
        pts = np.concatenate([pts, np.array(base)])
    pts[:, 0] = pos_x + pts[:, 0] * chord
    pts[:, 1] = pts[:, 1] * chord
    return pts


def radial_distances(h0, n_radial, radius=None, growth=None):
    """Wall distances d_0 = 0 < d_1 = h0 < ... of a geometric layer growth;
    either the growth ratio is given, or it is solved so that d_last = radius."""
    n = n_radial - 1
    if growth is None:
        lo, hi = 1.0 + 1e-12, 2.0
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            if h0 * (mid ** n - 1.0) / (mid - 1.0) < radius:
                lo = mid
            else:
                hi = mid
        growth = 0.5 * (lo + hi)
    k = np.arange(n_radial, dtype=np.float64)
    d = h0 * (growth ** k - 1.0) / (growth - 1.0)
    return d, growth


def _unit(v):
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def march_layers(wall, d, centre, R, d_switch=0.25):
    """Layer-by-layer normal marching from the wall (distances d), with
    light normal smoothing and tangential point redistribution, up to the
    first layer at distance >= d_switch (a quarter chord: well inside the
    radius of curvature of any concave part of a NACA 4-digit surface, so
    the normals do not cross); beyond it every point moves out along its ray
    from `centre` (a point inside the airfoil), its radius interpolated
    between that layer's and R with the same layer distances.  The switch
    layer is star-shaped around the centre, so the outer layers are nested
    and cannot fold."""
    P = wall.copy()
    outward = np.sign(np.mean(np.sum(_unit(np.stack([(np.roll(P, -1, 0) - np.roll(P, 1, 0))[:, 1],
                                                        -(np.roll(P, -1, 0) - np.roll(P, 1, 0))[:, 0]], 1))
                                     * (P - centre), 1)))
    layers = [P.copy()]
    d_last = d[-1]
    j_sw = len(d) - 1
    for j in range(1, len(d)):
        dd = d[j] - d[j - 1]
        T = np.roll(P, -1, 0) - np.roll(P, 1, 0)
        N = outward * _unit(np.stack([T[:, 1], -T[:, 0]], 1))
        frac = d[j] / d_last
        for _ in range(2 + int(30 * frac)):
            N = _unit(0.5 * N + 0.25 * (np.roll(N, 1, 0) + np.roll(N, -1, 0)))
        Pn = P + dd * N
        nu = 0.1 + 0.4 * min(1.0, d[j] / 1.0)
        for _ in range(3):
            Lp = 0.5 * (np.roll(Pn, 1, 0) + np.roll(Pn, -1, 0)) - Pn
            Tn = _unit(np.roll(Pn, -1, 0) - np.roll(Pn, 1, 0))
            Pn = Pn + nu * np.sum(Lp * Tn, 1, keepdims=True) * Tn
        P = Pn
        layers.append(P.copy())
        if d[j] >= d_switch:
            j_sw = j
            break
    rel = layers[-1] - centre
    r_sw = np.linalg.norm(rel, axis=1)
    for j in range(j_sw + 1, len(d)):
        s = (d[j] - d[j_sw]) / (d_last - d[j_sw])
        rj = (1.0 - s) * r_sw + s * R
        layers.append(centre + rel * (rj / r_sw)[:, None])
    return np.array(layers)


def naca_mesh(m=0.0, p=0.4, t=0.12, n_around=128, n_radial=40, wall_spacing=2e-3, radius=20.0,
              growth=None, aoa_deg=0.0, pos_x=0.355, chord=1.0, closed_te=False):
    """Triangle O-mesh; returns the mesh description (the arrays of
    o2f_mesh_desc) plus metadata.  With `growth` given the far-field radius
    follows from it; otherwise the growth is solved to reach `radius`."""
    wall = airfoil_contour(m, p, t, n_around, pos_x=pos_x, chord=chord, closed_te=closed_te)
    if aoa_deg:
        a = math.radians(aoa_deg)
        ca, sa = math.cos(a), math.sin(a)
        dx, dy = wall[:, 0] - pos_x, wall[:, 1]
        wall = np.stack([pos_x + ca * dx + sa * dy, -sa * dx + ca * dy], 1)
    d, g = radial_distances(wall_spacing, n_radial, radius, growth)
    R = d[-1] if growth is not None else radius
    # centre: the mean-line point at mid chord (inside the airfoil), rotated with it
    yc_mid = float(naca4(m, p, t, np.array([0.5]))[0][0]) * chord
    a = math.radians(aoa_deg)
    centre = np.array([pos_x + math.cos(a) * 0.5 * chord + math.sin(a) * yc_mid,
                       -math.sin(a) * 0.5 * chord + math.cos(a) * yc_mid])
    xy = march_layers(wall, d, centre, R, d_switch=0.25 * chord)   # (n_radial, n_around, 2)
    ni, nj = n_around, n_radial
    idx = lambda j, i: j * ni + (i % ni)
    cells = []
    for j in range(nj - 1):
        for i in range(ni):
            a, b, c, dd = idx(j, i), idx(j, i + 1), idx(j + 1, i + 1), idx(j + 1, i)
            cells.append((a, b, c))
            cells.append((a, c, dd))
    cells = np.array(cells, dtype=np.int32)
    xyf = xy.reshape(-1, 2)
    # orient counter-clockwise (the reference takes |area| anyway)
    v0, v1, v2 = xyf[cells[:, 0]], xyf[cells[:, 1]], xyf[cells[:, 2]]
    cross = (v1[:, 0] - v0[:, 0]) * (v2[:, 1] - v0[:, 1]) - (v2[:, 0] - v0[:, 0]) * (v1[:, 1] - v0[:, 1])
    if np.any(np.abs(cross) < 1e-300):
        raise RuntimeError("degenerate triangle")
    if not (np.all(cross > 0) or np.all(cross < 0)):
        raise RuntimeError("folded mesh: %d inverted triangles" % int(min((cross > 0).sum(), (cross < 0).sum())))
    if cross[0] < 0:
        cells = cells[:, [0, 2, 1]]
    bnd, tags = [], []
    for i in range(ni):
        bnd.append((idx(0, i), idx(0, i + 1)))
        tags.append(ord("w"))
    for i in range(ni):
        bnd.append((idx(nj - 1, i), idx(nj - 1, i + 1)))
        tags.append(ord("f"))
    cell_vtx = np.full((len(cells), 4), -1, dtype=np.int32)
    cell_vtx[:, :3] = cells
    return {
        "xy": np.ascontiguousarray(xyf),
        "cell_nv": np.full(len(cells), 3, dtype=np.int32),
        "cell_vtx": cell_vtx,
        "bnd_vtx": np.array(bnd, dtype=np.int32),
        "bnd_tag": np.array(tags, dtype=np.uint8),
        "meta": {"m": m, "p": p, "t": t, "n_around": ni, "n_radial": nj, "growth": g,
                 "radius": float(R), "wall_spacing": wall_spacing, "aoa_deg": aoa_deg},
    }


# The named triangle workloads (NACA geometry, mesh sizes and flow).
def transonic_200k(aoa_deg=2.5):
    """NACA 2412, ~200k triangles / ~100k vertices, wall spacing 5e-4,
    circular far field of 20 chords."""
    return naca_mesh(m=0.02, p=0.4, t=0.12, n_around=512, n_radial=196, wall_spacing=5e-4,
                     radius=20.0, aoa_deg=aoa_deg)


def small_10k(aoa_deg=2.0):
    """NACA 0012, ~10k triangles / ~5k vertices, wall spacing 2e-3, growth 1.1."""
    return naca_mesh(m=0.0, p=0.4, t=0.12, n_around=68, n_radial=74, wall_spacing=2e-3,
                     growth=1.1, aoa_deg=aoa_deg)


def write_msh2(path, desc):
    """MSH 2.2 ASCII with physical groups Airfoil (1, tag 44), FreeStream
    (1, tag 42) and Fluid (2, tag 45), readable by o2f_msh2_read."""
    xy = desc["xy"]
    with open(path, "w") as f:
        f.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n")
        f.write('$PhysicalNames\n3\n1 42 "FreeStream"\n1 44 "Airfoil"\n2 45 "Fluid"\n$EndPhysicalNames\n')
        f.write("$Nodes\n%d\n" % len(xy))
        for k, (x, y) in enumerate(xy):
            f.write("%d %.17g %.17g 0\n" % (k + 1, x, y))
        f.write("$EndNodes\n")
        lines = []
        phys = {ord("w"): 44, ord("f"): 42, ord("i"): 41, ord("o"): 43}
        for (a, b), tg in zip(desc["bnd_vtx"], desc["bnd_tag"]):
            lines.append("1 2 %d 1 %d %d" % (phys[int(tg)], a + 1, b + 1))
        for nv, v in zip(desc["cell_nv"], desc["cell_vtx"]):
            typ = 2 if nv == 3 else 3
            lines.append("%d 2 45 1 %s" % (typ, " ".join(str(x + 1) for x in v[:nv])))
        f.write("$Elements\n%d\n" % len(lines))
        for k, s in enumerate(lines):
            f.write("%d %s\n" % (k + 1, s))
        f.write("$EndElements\n")
