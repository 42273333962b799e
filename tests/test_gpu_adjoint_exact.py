"""The exact discrete adjoint (o2f_apply_jt_exact, o2f_bicgstab_exact) on
the GPU against finite differences of the oracle's residual.

Pins: every entry of (dR/dU) on a 160-triangle mesh against central
differences of the oracle residual (a dropped term, a wrong sign or a
transposed block shows up in whole rows); sampled rows around wall,
far-field and interior cells of a 10k-triangle mesh the same way; and the
BiCGSTAB solution against the dense finite-difference operator.  Entries next to a min / max or limiter
switch have no derivative, so finite differences across them disagree;
the tests bound how many entries may do so.
"""
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


def flow(limiter=0, heat_flux=1):
    f = dict(H.subsonic_small(limiter=limiter))
    f["heat_flux"] = heat_flux
    return f


def state(desc, fl, steps=15, noise=1e-3):
    """A developed state with seeded relative noise: every cell then has a
    gradient well above round-off.  (Near a uniform state the
    Barth-Jespersen ratio (M - u) / (pred - u) has round-off-sized terms,
    its derivative ~ 1 / (pred - u) holds only within round-off of the point,
    and no finite difference can check it.)"""
    m = O.Mesh(desc)
    p = H.oracle_params(fl)
    m.init_freestream(p)
    m.run_steps(p, steps)
    U = m.state()
    if noise:
        U = U * (1.0 + noise * np.random.default_rng(5).standard_normal(U.shape))
    m.set_state(U)
    return m, p, U


def fd_jacobian(m, p, U, rel=1e-6, sided=False):
    """J[:, j] = (R(U + h e_j) - R(U - h e_j)) / 2h, h relative to each
    variable's scale; with sided, also the forward and backward quotients."""
    nc = U.shape[1]
    scale = np.max(np.abs(U), axis=1)
    m.set_state(U)
    R0 = m.residual(p).copy().reshape(-1)
    J = np.zeros((4 * nc, 4 * nc))
    Jf, Jb = (np.zeros_like(J), np.zeros_like(J)) if sided else (None, None)
    for q in range(4):
        h = rel * scale[q]
        for c in range(nc):
            Up = U.copy()
            Up[q, c] += h
            m.set_state(Up)
            Rp = m.residual(p).copy().reshape(-1)
            Up[q, c] -= 2 * h
            m.set_state(Up)
            Rm = m.residual(p).copy().reshape(-1)
            J[:, q * nc + c] = (Rp - Rm) / (2 * h)
            if sided:
                Jf[:, q * nc + c] = (Rp - R0) / h
                Jb[:, q * nc + c] = (R0 - Rm) / h
    m.set_state(U)
    return (J, Jf, Jb) if sided else J


@pytest.mark.parametrize("limiter,heat_flux", [(0, 1), (1, 1), (0, 0)])
def test_exact_adjoint_every_entry_against_finite_differences(limiter, heat_flux):
    desc = H.tiny_mesh()
    fl = flow(limiter, heat_flux)
    m, p, U = state(desc, fl)
    J1, Jf, Jb = fd_jacobian(m, p, U, 1e-6, sided=True)
    J2 = fd_jacobian(m, p, U, 1e-7)
    s = capi.Solver(desc, H.capi_params(fl))
    s.set_state(U)
    n = J1.shape[0]
    M = np.zeros_like(J1)
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        M[i] = s.apply_jt_exact(e.reshape(4, -1)).reshape(-1)   # (J^T e_i)_j = J_ij
    rows = np.maximum(np.max(np.abs(J1), axis=1, keepdims=True), 1e-300)
    err = np.abs(M - J1) / rows
    nz = (np.abs(J1) > 1e-12 * rows) | (np.abs(M) > 1e-12 * rows)
    # smooth: forward and backward quotients agree (no min / max / limiter
    # switch at the point: there the central quotient averages two slopes)
    # and the central quotients of two step sizes agree
    smooth = nz & (np.abs(Jf - Jb) / rows < 1e-5) & (np.abs(J1 - J2) / rows < 1e-7)
    frac = smooth.sum() / nz.sum()
    p99, mx = np.percentile(err[smooth], 99), err[smooth].max()
    print("nonzeros %d smooth %.4f p99 %.3g max %.3g" % (nz.sum(), frac, p99, mx), flush=True)
    assert frac > 0.9
    assert p99 < 1e-6
    if limiter == 1:
        # Barth-Jespersen: alpha = (M - u) / (pred - u) has a 1 / (pred - u)
        # derivative; a few cells with a small pred - u sit within a finite
        # difference step of their switch
        assert np.mean(err[smooth] > 1e-4) < 1e-3
    else:
        assert mx < 1e-4
    # (almost) no entry outside the finite-difference pattern
    zero = (np.abs(J1) + np.abs(J2)) == 0.0
    stray = np.mean(np.abs(M[zero]) > 1e-12 * np.broadcast_to(rows, M.shape)[zero])
    assert stray < (1e-3 if limiter == 1 else 1e-12)


def vertex_rings(desc, c0, depth):
    """Cells within `depth` vertex-sharing rings of cell c0."""
    cv = np.asarray(desc["cell_vtx"])
    v2c = {}
    for c, row in enumerate(cv):
        for v in row:
            if v >= 0:
                v2c.setdefault(int(v), []).append(c)
    ring = {c0}
    for _ in range(depth):
        ring |= {c2 for c in ring for v in cv[c] if v >= 0 for c2 in v2c[int(v)]}
    return sorted(ring)


def test_sampled_rows_10k_triangles():
    """Rows of dR/dU around a wall cell, a far-field cell and interior cells of
    the 10k-triangle mesh, against one-sided / central finite differences."""
    desc = H.small_mesh()
    fl = flow()
    m, p, U = state(desc, fl, steps=10)
    s = capi.Solver(desc, H.capi_params(fl))
    s.set_state(U)
    nc = U.shape[1]
    scale = np.max(np.abs(U), axis=1)
    m.set_state(U)
    R0 = m.residual(p).copy()
    edges = m.edges()
    wall_cell = int(edges["cl"][np.argmax(edges["bc"] == ord("w"))])
    far_cell = int(edges["cl"][np.argmax(edges["bc"] == ord("f"))])
    rng = np.random.default_rng(3)
    picks = [wall_cell, far_cell] + [int(c) for c in rng.integers(0, nc, 2)]
    total, bad = 0, 0
    for c0 in picks:
        rows = vertex_rings(desc, c0, 1)
        cols = vertex_rings(desc, c0, 2)
        M = {}
        for c in rows:
            for q in range(4):
                e = np.zeros((4, nc))
                e[q, c] = 1.0
                M[(q, c)] = s.apply_jt_exact(e)          # row (q, c) of dR/dU
        for c in cols:
            for q in range(4):
                def R_at(h):
                    Up = U.copy()
                    Up[q, c] += h
                    m.set_state(Up)
                    return m.residual(p).copy()
                h1, h2 = 1e-6 * scale[q], 1e-7 * scale[q]
                Rp1, Rm1, Rp2, Rm2 = R_at(h1), R_at(-h1), R_at(h2), R_at(-h2)
                for (qr, cr), row in M.items():
                    rs = max(np.max(np.abs(row)), 1e-300)
                    jc1 = (Rp1[qr, cr] - Rm1[qr, cr]) / (2 * h1)
                    jc2 = (Rp2[qr, cr] - Rm2[qr, cr]) / (2 * h2)
                    jf = (Rp1[qr, cr] - R0[qr, cr]) / h1
                    jb = (R0[qr, cr] - Rm1[qr, cr]) / h1
                    if abs(jf - jb) > 1e-5 * rs or abs(jc1 - jc2) > 1e-7 * rs:
                        continue                          # a switch within the step
                    total += 1
                    if abs(row[q, c] - jc1) > 1e-5 * rs:
                        bad += 1
    m.set_state(U)
    print("checked %d entries, %d off by more than 1e-5 of their row" % (total, bad), flush=True)
    assert total > 2000
    assert bad == 0


def test_exact_bicgstab_solves_the_dense_system():
    """A viscous state whose exact Jacobian BiCGSTAB can solve (the operator
    of a non-converged explicit-scheme state has eigenvalues of both signs
    spread over six decades; at Re 5000 BiCGSTAB stalls on it).  The
    solution is checked against the dense operator assembled row by row from
    o2f_apply_jt_exact (whose entries the test above pins) and, more
    loosely, against the finite-difference operator (kink entries times a
    large lambda)."""
    desc = H.tiny_mesh()
    fl = flow()
    fl["viscosity"] = 1.0
    m, p, U = state(desc, fl, noise=0.0)
    J = fd_jacobian(m, p, U)
    s = capi.Solver(desc, H.capi_params(fl))
    s.set_state(U)
    beta = 0.1
    lam, iters, rr, st = s.bicgstab_exact(beta, tol=1e-10, chunk=20000, max_total=20000, check=False)
    b = s.adjoint_rhs(beta).reshape(-1)    # face states of the state, set by the solve
    print("iters %d relres %.3g status %d" % (iters, rr, st), flush=True)
    assert st == 0 and rr <= 1e-10
    n = J.shape[0]
    M = np.zeros_like(J)
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        M[i] = s.apply_jt_exact(e.reshape(4, -1)).reshape(-1)
    lv = lam.reshape(-1)
    assert np.linalg.norm(M.T @ lv - b) <= 1e-8 * np.linalg.norm(b)
    assert np.linalg.norm(J.T @ lv - b) <= 5e-2 * np.linalg.norm(b)


@pytest.mark.parametrize("limiter", [0, 1])
def test_exact_rhs_every_entry_against_finite_differences(limiter):
    """(dJ/dU)^T of J = -C_L + beta C_D (pressure of the wall face states,
    centre x > 1.3 skipped) against central differences of the oracle's
    objective over every state entry of the 160-triangle mesh, whose
    trailing edge crosses x = 1.3.  The reference's right-hand side keeps
    only the face-state derivative of every wall edge, so it must differ."""
    desc = H.tiny_mesh()
    fl = flow(limiter)
    m, p, U = state(desc, fl)
    beta = 0.3
    nc = U.shape[1]
    scale = np.max(np.abs(U), axis=1)

    def J_at(Us):
        m.set_state(Us)
        m.residual(p)
        return m.objective(p, beta)

    J0 = J_at(U)
    g1, gf, gb, g2 = (np.zeros((4, nc)) for _ in range(4))
    for q in range(4):
        for c in range(nc):
            vals = {}
            for h in (1e-6, 1e-7):
                Up = U.copy()
                Up[q, c] += h * scale[q]
                Jp = J_at(Up)
                Up[q, c] -= 2 * h * scale[q]
                Jm = J_at(Up)
                vals[h] = (Jp, Jm, h * scale[q])
            Jp, Jm, hh = vals[1e-6]
            g1[q, c] = (Jp - Jm) / (2 * hh)
            gf[q, c] = (Jp - J0) / hh
            gb[q, c] = (J0 - Jm) / hh
            Jp, Jm, hh = vals[1e-7]
            g2[q, c] = (Jp - Jm) / (2 * hh)
    m.set_state(U)
    s = capi.Solver(desc, H.capi_params(fl))
    s.set_state(U)
    b = s.adjoint_rhs_exact(beta)
    b_ref = s.adjoint_rhs(beta)
    sc = np.max(np.abs(g1), axis=1, keepdims=True)
    err = np.abs(b - g1) / sc
    nz = (np.abs(g1) > 1e-12 * sc) | (np.abs(b) > 1e-12 * sc)
    smooth = nz & (np.abs(gf - gb) / sc < 1e-5) & (np.abs(g1 - g2) / sc < 1e-6)
    print("nonzeros %d smooth %d max err %.3g median %.3g; reference rhs differs by %.3g" %
          (nz.sum(), smooth.sum(), err[smooth].max(), np.median(err[smooth]),
           np.max(np.abs(b_ref - g1) / sc)), flush=True)
    # the limiter ratio's second derivative splits the one-sided quotients of
    # many entries by more than the filter allows; a kink whose slope jump
    # passes the filter (< 1e-5) can shift a central quotient by that much
    assert smooth.sum() > 0.6 * nz.sum()
    assert err[smooth].max() < 1e-5
    assert np.median(err[smooth]) < 1e-8
    # where J's differences are exactly zero, the derivative is below what
    # they resolve: a few ulps of J over the step
    zero = (g1 == 0) & (g2 == 0)
    resolution = 4 * np.spacing(abs(J0)) / (2e-7 * scale)[:, None]
    print("off-pattern max |b| / resolution %.3g" %
          np.max(np.abs(b[zero]) / np.broadcast_to(resolution, b.shape)[zero]), flush=True)
    assert np.all(np.abs(b[zero]) <= np.broadcast_to(resolution, b.shape)[zero])
    # the exact right-hand side reaches cells off the wall row (the
    # reconstruction stencil); the reference's does not
    assert np.count_nonzero(np.any(b != 0, axis=0)) > np.count_nonzero(np.any(b_ref != 0, axis=0))
    assert np.max(np.abs(b_ref - g1) / sc) > 1e-3


def test_exact_bicgstab_with_the_exact_rhs():
    """The solve with exact_rhs: M^T lambda = (dJ/dU)^T with M the dense exact
    operator (viscous state, as above)."""
    desc = H.tiny_mesh()
    fl = flow()
    fl["viscosity"] = 1.0
    m, p, U = state(desc, fl, noise=0.0)
    s = capi.Solver(desc, H.capi_params(fl))
    s.set_state(U)
    beta = 0.1
    lam, iters, rr, st = s.bicgstab_exact(beta, tol=1e-8, chunk=20000, max_total=80000, check=False,
                                          exact_rhs=True)
    print("iters %d relres %.3g status %d" % (iters, rr, st), flush=True)
    assert st == 0 and rr <= 1e-8
    b = s.adjoint_rhs_exact(beta).reshape(-1)
    n = b.shape[0]
    M = np.zeros((n, n))
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        M[i] = s.apply_jt_exact(e.reshape(4, -1)).reshape(-1)
    assert np.linalg.norm(M.T @ lam.reshape(-1) - b) <= 1e-7 * np.linalg.norm(b)


def test_block_diagonal_probing_is_the_operators_diagonal():
    """o2f_adjoint_block_diagonal (probing J^T with cells coloured two vertex
    rings apart) against the diagonal blocks of the dense operators
    assembled row by row: bit for bit, exact and reference form."""
    desc = H.tiny_mesh()
    fl = flow()
    m, p, U = state(desc, fl)
    s = capi.Solver(desc, H.capi_params(fl))
    s.set_state(U)
    nc = U.shape[1]
    n = 4 * nc
    for exact in (True, False):
        M = np.zeros((n, n))
        for i in range(n):
            e = np.zeros(n)
            e[i] = 1.0
            M[i] = (s.apply_jt_exact(e.reshape(4, -1)) if exact else s.apply_jt(e.reshape(4, -1))).reshape(-1)
        D = s.adjoint_block_diagonal(exact)
        Dd = np.array([[[M[q * nc + c, r * nc + c] for r in range(4)] for q in range(4)] for c in range(nc)])
        assert np.array_equal(D, Dd), exact


@pytest.mark.parametrize("viscosity", [None, 1.0])
def test_block_jacobi_reference_bicgstab(viscosity):
    """applyJacobianTranspose's system right-preconditioned by the inverses
    of its diagonal blocks (which map its spectrum into the right
    half-plane): the same solution as the reference's unpreconditioned
    BiCGSTAB where that converges, in an order of magnitude fewer
    iterations; at mu = 1 the unpreconditioned loop does not converge within
    5000 iterations at all."""
    desc = H.tiny_mesh()
    fl = flow()
    if viscosity:
        fl["viscosity"] = viscosity
    m = O.Mesh(desc)
    p = H.oracle_params(fl)
    m.init_freestream(p)
    m.run_steps(p, 2000)
    s = capi.Solver(desc, H.capi_params(fl))
    s.set_state(m.state())
    s.residual()
    lam_bj, it_bj, rr_bj, st_bj = s.bicgstab_block_jacobi(0.1, tol=1e-8, chunk=5000, max_total=5000, check=False)
    lam, it, rr, st = s.bicgstab(0.1, tol=1e-8, chunk=5000, max_total=5000, check=False)
    print("block-Jacobi: %d iterations relres %.3g status %d; identity: %d relres %.3g status %d" %
          (it_bj, rr_bj, st_bj, it, rr, st), flush=True)
    assert st_bj == 0 and rr_bj <= 1e-8 and it_bj < 500
    if viscosity:
        assert st != 0
    else:
        assert st == 0 and it > 5 * it_bj
        assert np.linalg.norm(lam_bj - lam) <= 1e-6 * np.linalg.norm(lam)


@pytest.mark.parametrize("mesh,limiter,heat_flux", [("tiny", 0, 1), ("tiny", 1, 1), ("tiny", 0, 0),
                                                    ("small", 0, 1), ("small", 1, 1), ("shipped", 0, 0)])
def test_transpose_identity_against_the_forward_jvp(mesh, limiter, heat_flux):
    """<J^T x, y> (the hand-written reverse-mode kernels) against <x, J y>
    (the dual-number forward-mode residual, o2f_residual_jvp) for seeded
    random x, y: two implementations that share no derivative code agree to
    rounding, entry pattern and branch choices included."""
    if mesh == "shipped":
        desc, fl = H.shipped_mesh(), dict(H.SHIPPED)
        U0 = H.shipped_flow()
        U = U0 * (1.0 + 1e-3 * np.random.default_rng(6).standard_normal(U0.shape))
    else:
        desc = H.tiny_mesh() if mesh == "tiny" else H.small_mesh()
        fl = flow(limiter, heat_flux)
        _, _, U = state(desc, fl, steps=15 if mesh == "tiny" else 10)
    s = capi.Solver(desc, H.capi_params(fl))
    s.set_state(U)
    R_fwd = s.residual()
    rng = np.random.default_rng(11)
    for trial in range(3):
        x = rng.standard_normal(U.shape)
        y = rng.standard_normal(U.shape) * np.max(np.abs(U), axis=1, keepdims=True)
        R, Jy = s.residual_jvp(y)
        assert np.array_equal(R, R_fwd)
        JTx = s.apply_jt_exact(x)
        a, b = np.sum(JTx * y), np.sum(x * Jy)
        scale = np.sum(np.abs(x * Jy))
        print("%s trial %d: <J^T x, y> %.15e  <x, J y> %.15e  diff / scale %.3g" %
              (mesh, trial, a, b, abs(a - b) / scale), flush=True)
        assert abs(a - b) <= 1e-11 * scale


def test_jvp_against_finite_differences():
    """o2f_residual_jvp along a random direction against central differences
    of the oracle's residual (160-triangle mesh; entries next to a limiter
    or min / max switch excluded as above)."""
    desc = H.tiny_mesh()
    fl = flow()
    m, p, U = state(desc, fl)
    s = capi.Solver(desc, H.capi_params(fl))
    s.set_state(U)
    y = np.random.default_rng(12).standard_normal(U.shape) * np.max(np.abs(U), axis=1, keepdims=True)
    _, Jy = s.residual_jvp(y)

    def R_at(h):
        m.set_state(U + h * y)
        return m.residual(p).copy()
    R0 = R_at(0.0)
    c1 = (R_at(1e-7) - R_at(-1e-7)) / 2e-7
    c2 = (R_at(1e-8) - R_at(-1e-8)) / 2e-8
    f1, b1 = (R_at(1e-7) - R0) / 1e-7, (R0 - R_at(-1e-7)) / 1e-7
    sc = np.max(np.abs(c1), axis=1, keepdims=True)
    smooth = (np.abs(f1 - b1) / sc < 1e-4) & (np.abs(c1 - c2) / sc < 1e-5)
    err = np.abs(Jy - c1) / sc
    print("smooth %.3f  p99 %.3g  max %.3g" % (smooth.mean(), np.percentile(err[smooth], 99), err[smooth].max()),
          flush=True)
    assert smooth.mean() > 0.9
    assert np.percentile(err[smooth], 99) < 1e-6
