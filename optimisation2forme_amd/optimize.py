"""The shape-optimisation outer loop of run_opti.sh on the device.

run_opti.sh (:211-254) repeats: run the solver (flow, adjoint, shape
gradient; the gradient goes to the shape log), read J, dJ/dm, dJ/dp from the
log, take a fixed signed step on m and p with clamps (update_geo_from_grads,
:91-121), re-mesh with Gmsh, rewrite the parameter file (:123-129) and bump
the file names in config.txt.  Here the same loop runs in one process on
the GPU:

- the mesh comes from the project mesher, whose topology does not depend
  on (m, p), so each iteration starts from the previous iteration's state
  instead of a state file (the reference chains state files the same way);
- the flow runs a fixed number of steps or to a density-residual tolerance;
- the adjoint is the reference's operator (BiCGSTAB on the FD HLLC blocks)
  or the exact one, with the reference's right-hand side or the exact
  (dJ/dU)^T;
- the gradient's geometric part is the reference's central difference or
  the forward-mode tangent (its eps -> 0 limit, no mesh copies);
- the parameter file and the shape log keep the reference's formats.
"""
import os

import numpy as np

from . import capi, mesher

# run_opti.sh:6-11
STEP_M, STEP_P = 0.0005, 0.002
P_MIN, P_MAX = 1e-3, 0.999
M_MIN, M_MAX = 0.0, 0.12


def _clamp(v, lo, hi):
    return lo if v < lo else hi if v > hi else v


def sign_step(m, p, gm, gp):
    """update_geo_from_grads (run_opti.sh:91-121): a step against the sign of
    each gradient (a zero gradient steps down), printed with 8 decimals,
    clamped, printed again."""
    s_m = 1 if gm < 0 else -1
    s_p = 1 if gp < 0 else -1
    m_new = float("%.8f" % (m + s_m * STEP_M))
    p_new = float("%.8f" % (p + s_p * STEP_P))
    return float("%.8f" % _clamp(m_new, M_MIN, M_MAX)), float("%.8f" % _clamp(p_new, P_MIN, P_MAX))


def read_params(path):
    """Shape-parameter file (loadParamsGeneric, src/shape_optimization.cpp:19-99):
    `key = value`, `key *= value` marks an optimised key, `#` comments.
    Returns (ordered dict of all values, list of marked keys in file order)."""
    values, marked = {}, []
    with open(path) as f:
        for line in f:
            line = line.split("#", 1)[0].strip()
            if not line:
                continue
            if "*=" in line:
                k, v = line.split("*=", 1)
                k = k.strip()
                if k not in marked:
                    marked.append(k)
            elif "=" in line:
                k, v = line.split("=", 1)
            else:
                parts = line.split()
                if len(parts) < 2:
                    continue
                k, v = parts[0], parts[1]
            k, v = k.strip(), v.strip()
            if not k or not v:
                continue
            try:
                values[k] = float(v)
            except ValueError:
                continue
    return values, marked


def write_params(path_in, path_out, m_new, p_new):
    """update_params_txt (run_opti.sh:123-129): the `m` and `p` lines of the
    input file get the new values, every other line is copied."""
    out = []
    with open(path_in) as f:
        for line in f:
            stripped = line.lstrip()
            for key, val in (("m", m_new), ("p", p_new)):
                rest = stripped[len(key):]
                if stripped.startswith(key) and rest.lstrip().startswith(("*=", "=")):
                    sep = "*=" if rest.lstrip().startswith("*=") else "="
                    lead = line[:len(line) - len(stripped)]
                    pre = stripped[:len(stripped) - len(rest.lstrip())] + sep
                    after = rest.lstrip()[len(sep):]
                    comment = ""
                    if "#" in after:
                        comment = "  #" + after.split("#", 1)[1].rstrip("\n")
                    line = "%s%s %.8f%s\n" % (lead, pre, val, comment)
                    break
            out.append(line)
    with open(path_out, "w") as f:
        f.writelines(out)


def log_line(J, names, grads):
    """The shape log of the driver (appendShapeOptLog): ' J=%.8f name=%e ...'."""
    return " J=%.8f" % J + "".join(" %s=%.6e" % (n, g) for n, g in zip(names, grads)) + "\n"


def run(param_file, iterations, flow, out_dir, t=None, mesh=None, steps=2000, tol=None, adjoint="reference",
        adjoint_tol=1e-6, adjoint_chunk=500, adjoint_max=100000, require_adjoint=True, exact_rhs=False,
        gradient="reference"):
    """`iterations` outer iterations from the parameter file `param_file`
    (its `m`, `p`, `t`, `beta`; the marked keys are differentiated).
    flow: make_params keyword arguments; mesh: naca_mesh keyword arguments
    (n_around, n_radial, wall_spacing, radius).  adjoint: "reference" or
    "exact" (exact_rhs: the exact right-hand side with it); gradient:
    "reference" (central differences) or "tangent".  Writes
    shape_params_<i>.txt and shape_opt_<i>.log under out_dir and returns the
    history [(i, m, p, J, {name: grad}, adjoint iterations, relres)]."""
    os.makedirs(out_dir, exist_ok=True)
    mesh = dict(n_around=96, n_radial=32, wall_spacing=2e-3, radius=20.0) if mesh is None else mesh
    values, marked = read_params(param_file)
    m, p = values.get("m", 0.02), values.get("p", 0.4)
    t = values.get("t", 0.12) if t is None else t
    beta = values.get("beta", 1.0)
    prm = capi.make_params(**flow)
    state, history, cur = None, [], param_file
    for i in range(iterations):
        desc = mesher.naca_mesh(m=m, p=p, t=t, **mesh)
        s = capi.Solver(desc, prm)
        if state is None:
            s.init_freestream()
        else:
            s.set_state(state)          # same topology: the previous shape's flow
        if tol is None:
            s.run_steps(steps)
        else:
            s.run_to_residual(tol, steps)
        s.residual()                    # computeResidual before the adjoint
        J = s.objective(beta)
        if adjoint == "exact":
            lam, its, rr, st = s.bicgstab_exact(beta, adjoint_tol, adjoint_chunk, adjoint_max, check=require_adjoint,
                                                exact_rhs=exact_rhs)
        else:
            lam, its, rr, st = s.bicgstab(beta, 1e-6, adjoint_tol, adjoint_chunk, adjoint_max, check=require_adjoint)
        vals = dict(values)
        vals.update(m=m, p=p, t=t)
        names_all = list(vals)
        if gradient == "tangent":
            g = s.shape_gradient_tangent(names_all, [vals[k] for k in names_all], marked, lam, beta)
        else:
            g = s.shape_gradient(names_all, [vals[k] for k in names_all], marked, lam, beta)
        grads = dict(zip(marked, g))
        with open(os.path.join(out_dir, "shape_opt_%d.log" % i), "w") as f:
            f.write(log_line(J, marked, g))
        history.append((i, m, p, J, grads, its, rr))
        state = s.state()
        s.close()
        m, p = sign_step(m, p, grads.get("m", 0.0), grads.get("p", 0.0))
        nxt = os.path.join(out_dir, "shape_params_%d.txt" % (i + 1))
        write_params(cur, nxt, m, p)
        cur = nxt
    return history
