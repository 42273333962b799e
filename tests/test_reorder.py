"""o2f_mesh_reorder (host only): the renumbered mesh is the same mesh, its
numbering has locality, and the flow on it is the flow on the original.

The reference keeps the mesher's numbering (src/mesh.cpp:9-111); the
renumbering is this library's, so there is no reference output to compare
with.  What pins it: the permutations are bijections and map every array of
the copy onto the original; reverse Cuthill-McKee's vertex-id spread within
a cell stays within a small factor of SciPy's RCM on the same graph (a
published implementation), far below a random numbering's; the interleave
places block j of chunk x at 8 j + x; and the oracle's residual and forces
on the renumbered mesh equal the original's up to the summation order."""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import reverse_cuthill_mckee

from optimisation2forme_amd import capi, mesher
from oracle import oracle as O
from tests import helpers as H


def scramble(desc, seed, with_perm=False):
    """The same mesh with cells and vertices numbered at random (and the
    cell permutation, new -> old)."""
    rng = np.random.default_rng(seed)
    nc, nv = len(desc["cell_nv"]), len(desc["xy"])
    cperm = rng.permutation(nc)
    vnew = rng.permutation(nv)            # old -> new
    vold = np.empty(nv, np.int64)
    vold[vnew] = np.arange(nv)
    cv = np.asarray(desc["cell_vtx"])[cperm].copy()
    cv[cv >= 0] = vnew[cv[cv >= 0]]
    out = dict(xy=np.asarray(desc["xy"])[vold], cell_nv=np.asarray(desc["cell_nv"])[cperm],
               cell_vtx=cv.astype(np.int32), bnd_vtx=vnew[np.asarray(desc["bnd_vtx"])].astype(np.int32),
               bnd_tag=np.asarray(desc["bnd_tag"]))
    return (out, cperm) if with_perm else out


def spread(desc):
    """Per cell, the largest difference of its vertex ids."""
    cv = np.asarray(desc["cell_vtx"])
    v = np.where(cv >= 0, cv, cv[:, :1])
    return v.max(1) - v.min(1)


def scipy_rcm_spread(desc):
    cv = np.asarray(desc["cell_vtx"])
    nc, nv = len(cv), len(desc["xy"])
    rows = np.repeat(np.arange(nc), 4)
    cols = cv.ravel()
    m = cols >= 0
    A = sp.csr_matrix((np.ones(m.sum()), (rows[m], cols[m])), shape=(nc, nv))
    order = reverse_cuthill_mckee((A @ A.T).tocsr(), symmetric_mode=True)
    vnew = -np.ones(nv, np.int64)
    k = 0
    for c in order:
        for v in cv[c]:
            if v >= 0 and vnew[v] < 0:
                vnew[v] = k
                k += 1
    w = cv[order].copy()
    w[w >= 0] = vnew[w[w >= 0]]
    return spread(dict(cell_vtx=w))


@pytest.fixture(scope="module")
def scrambled():
    return scramble(H.small_mesh(), 4)


@pytest.mark.parametrize("method", ["rcm", "rcm_xcd"])
def test_the_copy_is_the_same_mesh(scrambled, method):
    out, cperm, vperm = capi.reorder_mesh(scrambled, method)
    nc, nv = len(scrambled["cell_nv"]), len(scrambled["xy"])
    assert sorted(cperm) == list(range(nc)) and sorted(vperm) == list(range(nv))
    vnew = np.empty(nv, np.int64)
    vnew[vperm] = np.arange(nv)
    assert np.array_equal(out["xy"], np.asarray(scrambled["xy"])[vperm])
    assert np.array_equal(out["cell_nv"], np.asarray(scrambled["cell_nv"])[cperm])
    cv = np.asarray(scrambled["cell_vtx"])[cperm].copy()
    cv[cv >= 0] = vnew[cv[cv >= 0]]
    assert np.array_equal(out["cell_vtx"], cv)       # orientation kept, -1 padding kept
    assert np.array_equal(out["bnd_vtx"], vnew[np.asarray(scrambled["bnd_vtx"])])
    assert np.array_equal(out["bnd_tag"], scrambled["bnd_tag"])
    # vertices in order of first use by the renumbered cells
    first = []
    seen = set()
    for row in out["cell_vtx"]:
        for v in row:
            if v >= 0 and v not in seen:
                seen.add(v)
                first.append(v)
    assert first == list(range(len(first)))
    # deterministic
    out2, cperm2, _ = capi.reorder_mesh(scrambled, method)
    assert np.array_equal(cperm, cperm2)


def test_rcm_restores_locality(scrambled):
    out, _, _ = capi.reorder_mesh(scrambled, "rcm")
    s_scr, s_rcm, s_ref = spread(scrambled), spread(out), scipy_rcm_spread(scrambled)
    s_gen = spread(H.small_mesh())
    print("vertex-id spread within a cell (max / mean): scrambled %d / %.0f, rcm %d / %.0f, scipy rcm %d / %.0f, "
          "generated %d / %.0f" % (s_scr.max(), s_scr.mean(), s_rcm.max(), s_rcm.mean(), s_ref.max(), s_ref.mean(),
                                    s_gen.max(), s_gen.mean()))
    assert s_rcm.max() < s_scr.max() / 20
    assert s_rcm.max() <= 1.5 * s_ref.max() and s_rcm.mean() <= 1.5 * s_ref.mean()


def test_xcd_interleave_places_block_j_of_chunk_x_at_8j_plus_x(scrambled):
    _, base, _ = capi.reorder_mesh(scrambled, "rcm")
    _, inter, _ = capi.reorder_mesh(scrambled, "rcm_xcd")
    nc = len(base)
    nb = -(-nc // 64)
    per = -(-nb // 8)
    expect = []
    for j in range(per):
        for x in range(8):
            b = x * per + j
            if b < nb:
                expect.extend(base[b * 64:(b + 1) * 64])
    assert np.array_equal(inter, expect)


def test_flow_on_the_renumbered_mesh_is_the_flow_on_the_original():
    """On a physical state.  (With the reference's R-times-Fourier heat flux
    the 10k mesh's trailing-edge cells reach negative pressure within 15
    steps; HLLC's sound speed is then NaN, and min / max with std::min /
    std::max semantics give results that depend on the operands' order, so
    an edge whose left and right cells swap gets another flux.)"""
    desc = H.small_mesh()
    flow = dict(H.subsonic_small(limiter=0, interp_mode=1), heat_flux=1)
    p = H.oracle_params(flow)
    U = H.developed_state(desc, flow, steps=15)
    rho, ru, rv, E = U
    assert np.all(0.4 * (E - 0.5 * (ru * ru + rv * rv) / rho) > 0)
    out, cperm, _ = capi.reorder_mesh(desc, "rcm_xcd")
    a, b = O.Mesh(desc), O.Mesh(out)
    a.set_state(U)
    b.set_state(U[:, cperm])
    Ra, Rb = a.residual(p), b.residual(p)
    back = np.empty_like(Rb)
    back[:, cperm] = Rb
    scale = np.max(np.abs(Ra), axis=1, keepdims=True)
    err = np.max(np.abs(back - Ra) / scale)
    print("residual: max |difference| / max |R| per variable %.3g" % err)
    assert err < 1e-13
    (cla, cda), (clb, cdb) = a.lift_drag(p), b.lift_drag(p)
    assert abs(cla - clb) <= 1e-12 * abs(cla) and abs(cda - cdb) <= 1e-12 * abs(cda)
    # and over time: 20 steps from the same state stay together (measured
    # 5e-16 .. 1e-15 of each variable's scale)
    a.run_steps(p, 20)
    b.run_steps(p, 20)
    back = np.empty_like(U)
    back[:, cperm] = b.state()
    assert max(H.var_rel_err(back, a.state())) < 1e-13


def test_bad_arguments_fail():
    desc = H.small_mesh()
    with pytest.raises(capi.O2FError):
        capi.reorder_mesh(desc, 7)
    bad = dict(desc)
    cv = np.asarray(desc["cell_vtx"]).copy()
    cv[3, 1] = len(desc["xy"])
    bad["cell_vtx"] = cv
    with pytest.raises(capi.O2FError):
        capi.reorder_mesh(bad, "rcm")
