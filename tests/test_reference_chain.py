"""The oracle against the reference's own shipped run: mesh_opti/config.txt
loads naca_2D_flow_0 on naca_2D_shape_0.msh, integrates it with the time
loop of src/simulation.cpp:68-126 for final_time = 0.005 s (Mach 0.3,
mu 1.8e-5, Pr 0.71, CFL 0.5: the last occurrence of each key wins) and saves
naca_2D_flow_1.  Integrating the same start on the same mesh must end on the
same state.

This pins every part of the explicit residual at once, on the reference's
own output: the vertex interpolation (mean-value weights at valence 4), the
Green-Gauss gradients, the Venkatakrishnan limiter with k_eps = 1e-2, the
interface and ghost states, HLLC, the viscous stresses, the reference's
heat flux (R times Fourier's law), the wall / far-field boundary states,
the dt reduction and SSP-RK3.  The shipped files were written by the
author's arm64 build (FMA allowed), so rounding differs; after 3348 steps
the oracle ends 5e-14 .. 5e-13 of each variable's scale from flow_1, while
the start is 7e-2 away.  Measured sensitivity of this pin, each run to the
same final time: Venkatakrishnan k_eps 0.05 instead of 0.01 misses by 1e-5,
mu x 1.01 by 1e-4, Barth-Jespersen by 1e-2, Fourier's heat flux by 2e-2
(the GPU test repeats these against the device).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.slow


def test_oracle_integrates_flow_0_to_the_shipped_flow_1():
    desc = H.chain_mesh(0)
    U0, U1 = H.chain_flow(0), H.chain_flow(1)
    m = O.Mesh(desc)
    p = H.oracle_params(H.SHIPPED)
    m.set_state(U0)
    steps, t = m.time_loop(p, H.CHAIN_FINAL_TIME)
    assert t >= H.CHAIN_FINAL_TIME and steps > 3000
    err = H.var_rel_err(m.state(), U1)
    start = H.var_rel_err(U0, U1)
    print("steps %d, t %.17g, end error per variable %s (start %s)" % (steps, t, err, start))
    assert max(err) <= 1e-12, err
    assert min(start) >= 1e-3       # the pin is not the identity
