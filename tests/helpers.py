"""Shared fixtures of the test suite: meshes, states and flow conditions."""
import functools
import gzip
import os
import tempfile

import numpy as np

from oracle import msh2
from oracle import oracle as O
from optimisation2forme_amd import mesher

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# Flow of the shipped optimisation run (config.txt:14-75): Mach 0.3,
# mu = 1.8e-5, Pr = 0.71, CFL 0.5, other SimulationParams defaults.
SHIPPED = dict(viscosity=1.8e-5, Pr=0.71, cfl=0.5, mach=0.3)


def re_viscosity(mach, re=5000.0, rho=1.225, p=101325.0, gamma=1.4, chord=1.0):
    """mu = rho_inf U_inf c / Re."""
    return rho * mach * np.sqrt(gamma * p / rho) * chord / re


def transonic(limiter=0, interp_mode=1, venkat_k=5.0):
    """NACA 2412 transonic flow: Mach 0.85, Re 5000, Pr 0.72, CFL 0.5, K = 5."""
    return dict(viscosity=re_viscosity(0.85), Pr=0.72, cfl=0.5, mach=0.85, venkat_k=venkat_k,
                limiter=limiter, interp_mode=interp_mode)


def subsonic_small(limiter=0, interp_mode=1):
    """NACA 0012: Mach 0.6, Re 5000, Pr 0.72, CFL 0.5, K = 5."""
    return dict(viscosity=re_viscosity(0.6), Pr=0.72, cfl=0.5, mach=0.6, venkat_k=5.0,
                limiter=limiter, interp_mode=interp_mode)


def oracle_params(d):
    return O.params(**d)


def capi_params(d):
    from optimisation2forme_amd import capi
    kw = dict(d)
    kw["cfl_number"] = kw.pop("cfl")
    kw["inflow_velocity"] = kw.pop("mach")
    return capi.make_params(**kw)


@functools.lru_cache(maxsize=None)
def _gunzip(name):
    src = os.path.join(GOLDEN, name)
    fd, path = tempfile.mkstemp(suffix="_" + name.replace(".gz", ""))
    with gzip.open(src, "rb") as f, os.fdopen(fd, "wb") as g:
        g.write(f.read())
    return path


@functools.lru_cache(maxsize=None)
def shipped_mesh():
    """The reference's shipped quad mesh (25,274 cells)."""
    return msh2.read_msh2(_gunzip("naca_2D_shape_0.msh.gz"))


def shipped_mesh_path():
    return _gunzip("naca_2D_shape_0.msh.gz")


@functools.lru_cache(maxsize=None)
def shipped_flow():
    """State naca_2D_flow_1, produced by the reference on shape 0."""
    return msh2.read_state(_gunzip("naca_2D_flow_1.gz"))


def shipped_flow_path():
    return _gunzip("naca_2D_flow_1.gz")


# (mesh, start, end) triples of the reference's shipped optimisation chain:
# run k loads naca_2D_flow_k on naca_2D_shape_k.msh, integrates it for
# final_time = 0.005 s with mesh_opti/config.txt's parameters and saves
# naca_2D_flow_{k+1} (mesh_opti/config.txt:62-70, run_opti.sh:211-254).
CHAIN = (0, 15, 29)
CHAIN_FINAL_TIME = 0.005


@functools.lru_cache(maxsize=None)
def chain_mesh(k):
    return msh2.read_msh2(_gunzip("naca_2D_shape_%d.msh.gz" % k))


def chain_mesh_path(k):
    return _gunzip("naca_2D_shape_%d.msh.gz" % k)


@functools.lru_cache(maxsize=None)
def chain_flow(k):
    return msh2.read_state(_gunzip("naca_2D_flow_%d.gz" % k))


def chain_flow_path(k):
    return _gunzip("naca_2D_flow_%d.gz" % k)


def var_rel_err(U, Uref):
    """Per conservative variable: max |U - Uref| / max |Uref|."""
    U, Uref = np.asarray(U), np.asarray(Uref)
    return [float(np.max(np.abs(U[q] - Uref[q])) / np.max(np.abs(Uref[q]))) for q in range(4)]


@functools.lru_cache(maxsize=None)
def tiny_mesh():
    """160-triangle O-mesh (brute-force sizes)."""
    return mesher.naca_mesh(m=0.02, p=0.4, t=0.12, n_around=20, n_radial=5, wall_spacing=0.05,
                            radius=6.0)


def corner_mesh(nx=8, ny=4, lx=2.0, ly=1.0):
    """A structured triangulated box whose bottom and left sides are walls
    (top far field, right outlet): the corner triangle owns two wall edges,
    which no O-mesh has."""
    xs = np.linspace(0.0, lx, nx + 1)
    ys = np.linspace(0.0, ly, ny + 1)
    xy = np.array([(x, y) for y in ys for x in xs], np.float64)

    def vid(i, j):
        return j * (nx + 1) + i

    cells = []
    for j in range(ny):
        for i in range(nx):
            a, b, c, d = vid(i, j), vid(i + 1, j), vid(i, j + 1), vid(i + 1, j + 1)
            cells += [[a, b, c, -1], [b, d, c, -1]]
    bnd, tag = [], []
    for i in range(nx):
        bnd.append((vid(i, 0), vid(i + 1, 0)))
        tag.append(ord("w"))
    for j in range(ny):
        bnd.append((vid(0, j), vid(0, j + 1)))
        tag.append(ord("w"))
    for i in range(nx):
        bnd.append((vid(i, ny), vid(i + 1, ny)))
        tag.append(ord("f"))
    for j in range(ny):
        bnd.append((vid(nx, j), vid(nx, j + 1)))
        tag.append(ord("o"))
    return dict(xy=xy, cell_nv=np.full(len(cells), 3, np.int32), cell_vtx=np.array(cells, np.int32),
                bnd_vtx=np.array(bnd, np.int32), bnd_tag=np.array(tag, np.uint8))


@functools.lru_cache(maxsize=None)
def small_mesh():
    return mesher.small_10k()


@functools.lru_cache(maxsize=None)
def big_mesh():
    return mesher.transonic_200k()


def developed_state(desc, flow, steps=20, seed=None):
    """Freestream start advanced `steps` SSP-RK3 steps by the oracle: a state
    with realistic limiter / HLLC branch mixes."""
    m = O.Mesh(desc)
    p = oracle_params(flow)
    m.init_freestream(p)
    m.run_steps(p, steps)
    U = m.state()
    if seed is not None:
        rng = np.random.default_rng(seed)
        U = U * (1.0 + 1e-3 * rng.standard_normal(U.shape))
    return U


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    den = np.maximum(np.abs(b), 1e-300)
    scale = np.max(np.abs(b)) if b.size else 1.0
    return float(np.max(np.abs(a - b) / np.maximum(den, 1e-12 * scale))) if a.size else 0.0


def quad_box_mesh(nx=16, ny=16, lx=1.0, ly=1.0, tag="o"):
    """A uniform nx x ny grid of rectangles over [0, lx] x [0, ly], every
    boundary line tagged `tag` (outlet: state and stresses pass through)."""
    xs = np.linspace(0.0, lx, nx + 1)
    ys = np.linspace(0.0, ly, ny + 1)
    xy = np.array([(x, y) for y in ys for x in xs], np.float64)

    def vid(i, j):
        return j * (nx + 1) + i

    cells = [[vid(i, j), vid(i + 1, j), vid(i + 1, j + 1), vid(i, j + 1)] for j in range(ny) for i in range(nx)]
    bnd = [(vid(i, 0), vid(i + 1, 0)) for i in range(nx)] + [(vid(i, ny), vid(i + 1, ny)) for i in range(nx)]
    bnd += [(vid(0, j), vid(0, j + 1)) for j in range(ny)] + [(vid(nx, j), vid(nx, j + 1)) for j in range(ny)]
    return dict(xy=xy, cell_nv=np.full(len(cells), 4, np.int32), cell_vtx=np.array(cells, np.int32),
                bnd_vtx=np.array(bnd, np.int32), bnd_tag=np.full(len(bnd), ord(tag), np.uint8))
