"""The device on renumbered meshes (o2f_mesh_reorder): bitwise the oracle on
the same renumbered mesh, whatever the numbering, and the same flow as on
the original numbering up to the summation order."""
import sys

import numpy as np
import pytest

from optimisation2forme_amd import capi
from oracle import oracle as O
from tests import helpers as H
from tests.test_reorder import scramble

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _announce(request):
    print("\n[gpu case] %s" % request.node.nodeid, flush=True)
    sys.stdout.flush()
    yield


@pytest.mark.parametrize("method", ["scrambled", "rcm", "rcm_xcd"])
def test_renumbered_mesh_bit_identical_to_the_oracle(method):
    flow = dict(H.subsonic_small(limiter=0, interp_mode=1), heat_flux=1)
    p = H.oracle_params(flow)
    U = H.developed_state(H.small_mesh(), flow, steps=15)
    scr, cperm = scramble(H.small_mesh(), 4, with_perm=True)
    Ud = U[:, cperm]                     # the state on the scrambled numbering
    desc = scr
    if method != "scrambled":
        desc, rperm, _ = capi.reorder_mesh(scr, method)
        Ud = Ud[:, rperm]
    s = capi.Solver(desc, H.capi_params(flow))
    m = O.Mesh(desc)
    s.set_state(Ud)
    m.set_state(Ud)
    assert np.array_equal(s.residual(), m.residual(p))
    assert np.array_equal(s.gradients(), m.gradients())
    s.run_steps(10)
    m.run_steps(p, 10)
    # dt: the device's hypot and glibc's differ in the last bit (as on every mesh)
    assert max(H.var_rel_err(s.state(), m.state())) <= 1e-12
