"""CPU checks of the outer shape-optimisation loop's host logic against
run_opti.sh's rules: the signed, clamped step, the parameter-file rewrite
and the shape-log format it parses."""
import os
import re

from optimisation2forme_amd import optimize as OPT
from tests import helpers as H

SHIPPED_PARAMS = os.path.join(H.GOLDEN, "naca_2D_shape_params_0.txt")


def test_sign_step_follows_run_opti():
    # negative gradient -> step up, positive or zero -> step down (run_opti.sh:105-106)
    assert OPT.sign_step(0.02, 0.4, -1.0, 2.0) == (0.0205, 0.398)
    assert OPT.sign_step(0.02, 0.4, 0.0, 0.0) == (0.0195, 0.398)
    # clamps (run_opti.sh:111-112)
    assert OPT.sign_step(0.0, 0.0005, 1.0, 1.0) == (0.0, 0.001)
    assert OPT.sign_step(0.12, 0.999, -1.0, -1.0) == (0.12, 0.999)
    # the shipped trajectory's first step: m 0 -> 0.0005 needs dJ/dm < 0
    assert OPT.sign_step(0.0, 0.4, -0.5, 0.0)[0] == 0.0005


def test_param_file_round_trip(tmp_path):
    vals, marked = OPT.read_params(SHIPPED_PARAMS)
    assert "m" in vals and "p" in vals and "beta" in vals
    out = tmp_path / "p1.txt"
    OPT.write_params(SHIPPED_PARAMS, str(out), 0.0205, 0.398)
    v2, m2 = OPT.read_params(str(out))
    assert m2 == marked
    assert v2["m"] == 0.0205 and v2["p"] == 0.398
    assert {k: v for k, v in v2.items() if k not in ("m", "p")} == {k: v for k, v in vals.items() if k not in ("m", "p")}


def test_log_line_is_what_run_opti_parses():
    line = OPT.log_line(-0.12345678912, ["m", "p"], [-1.5e-3, 2.25])
    assert line.startswith(" J=-0.12345679 ")
    # parse_log_J_and_grads (run_opti.sh:32-78): tokens J=, m=, p=
    tok = dict(t.split("=", 1) for t in line.split())
    assert float(tok["J"]) == -0.12345679 and float(tok["m"]) == -1.5e-3 and float(tok["p"]) == 2.25
    assert re.fullmatch(r" J=-?\d+\.\d{8}( \w+=-?\d\.\d{6}e[+-]\d\d)+\n", line)
