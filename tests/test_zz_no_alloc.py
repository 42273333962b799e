"""o2f_hllc_flux_batch_noalloc (the faces staged through the library's fixed
per-GPU array, no device allocation) against o2f_hllc_flux_batch and the
oracle: bit-identical fluxes over all four HLLC branches, on two calls in a
row -- the maximum batch, then a smaller one -- and a clear error beyond the
maximum."""
import sys

import numpy as np
import pytest

from optimisation2forme_amd import capi
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def faces(n, seed):
    rng = np.random.default_rng(seed)
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
    return qL, qR, np.array([np.cos(a), np.sin(a)])


def test_noalloc_flux_equals_the_allocating_entry_point():
    for n, seed in ((capi.NOALLOC_MAX_FACES, 1), (1000, 2)):
        print("\n[gpu case] noalloc HLLC batch of %d faces" % n, flush=True)
        sys.stdout.flush()
        qL, qR, nrm = faces(n, seed)
        F = capi.hllc_flux_batch_noalloc(qL, qR, nrm)
        assert np.array_equal(F, capi.hllc_flux_batch(qL, qR, nrm))
        assert np.array_equal(F, O.hllc_batch(qL, qR, nrm))
        assert np.all(np.isfinite(F))


def test_noalloc_rejects_a_batch_beyond_the_maximum():
    n = capi.NOALLOC_MAX_FACES + 1
    qL, qR, nrm = faces(n, 3)
    with pytest.raises(capi.O2FError) as info:
        capi.hllc_flux_batch_noalloc(qL, qR, nrm)
    assert "exceed the maximum of %d" % capi.NOALLOC_MAX_FACES in str(info.value)
    # and a valid call still works after the rejected one
    qL, qR, nrm = faces(64, 4)
    assert np.array_equal(capi.hllc_flux_batch_noalloc(qL, qR, nrm), capi.hllc_flux_batch(qL, qR, nrm))
