"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  Arrays are
float64 SoA in the reference's cell / edge order.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

D = C.POINTER(C.c_double)
I = C.POINTER(C.c_int)


class Params(C.Structure):
    _fields_ = [
        ("gamma", C.c_double), ("viscosity", C.c_double), ("Pr", C.c_double), ("cp", C.c_double),
        ("cfl", C.c_double), ("rho_ref", C.c_double), ("p_ref", C.c_double), ("mach", C.c_double),
        ("venkat_k", C.c_double), ("limiter", C.c_int), ("interp_mode", C.c_int),
        ("bc_known", C.c_int), ("heat_flux", C.c_int),
    ]


def params(gamma=1.4, viscosity=1e-6, Pr=0.71, cp=1005.0, cfl=0.1, rho_ref=1.225,
           p_ref=101325.0, mach=1.0, venkat_k=1e-2, limiter=0, interp_mode=0, bc_known=1,
           heat_flux=0):
    """Defaults are SimulationParams' (include/params.hpp:15-73)."""
    return Params(gamma, viscosity, Pr, cp, cfl, rho_ref, p_ref, mach, venkat_k, limiter,
                  interp_mode, bc_known, heat_flux)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        vp = C.c_void_p
        L.orc_mesh_create.restype = vp
        L.orc_mesh_create.argtypes = [C.c_int, D, C.c_int, I, I, C.c_int, I, C.c_char_p]
        L.orc_mesh_clone.restype = vp
        L.orc_mesh_clone.argtypes = [vp]
        L.orc_mesh_destroy.argtypes = [vp]
        L.orc_counts.argtypes = [vp, I, I, I]
        for name in ("orc_set_state", "orc_get_state", "orc_get_residual", "orc_get_gradients"):
            getattr(L, name).argtypes = [vp, D]
        L.orc_get_faces.argtypes = [vp, D, D]
        L.orc_set_faces.argtypes = [vp, D, D]
        L.orc_get_edges.argtypes = [vp, I, I, I, I, C.c_char_p, D]
        L.orc_get_cells.argtypes = [vp, D, I]
        L.orc_get_vertex_bc.argtypes = [vp, C.c_char_p, D]
        P = C.POINTER(Params)
        L.orc_init_freestream.argtypes = [vp, P]
        for name in ("orc_advection", "orc_diffusion", "orc_residual"):
            getattr(L, name).argtypes = [vp, P]
            getattr(L, name).restype = C.c_int
        L.orc_dt.argtypes = [vp, P]
        L.orc_dt.restype = C.c_double
        L.orc_rk3_step.argtypes = [vp, P, C.c_double]
        L.orc_rk3_step.restype = C.c_int
        L.orc_save_old.argtypes = [vp]
        L.orc_rk3_substep.argtypes = [vp, P, C.c_double, C.c_double, C.c_double]
        L.orc_rk3_substep.restype = C.c_int
        L.orc_run_steps.argtypes = [vp, P, C.c_int, D]
        L.orc_run_steps.restype = C.c_int
        L.orc_time_loop.argtypes = [vp, P, C.c_double, C.c_int, I, D]
        L.orc_time_loop.restype = C.c_int
        L.orc_residual_norm.argtypes = [vp, P]
        L.orc_residual_norm.restype = C.c_double
        L.orc_run_to_residual.argtypes = [vp, P, C.c_double, C.c_int, C.c_int, I, D]
        L.orc_run_to_residual.restype = C.c_int
        L.orc_lift_drag.argtypes = [vp, P, D, D]
        L.orc_objective.argtypes = [vp, P, C.c_double]
        L.orc_objective.restype = C.c_double
        L.orc_adjoint_rhs.argtypes = [vp, P, C.c_double, D]
        L.orc_apply_jt.argtypes = [vp, P, D, D, C.c_double]
        L.orc_edge_blocks.argtypes = [vp, P, C.c_int, C.c_double, D, D]
        L.orc_bicgstab.argtypes = [vp, P, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, D, I, D]
        L.orc_bicgstab.restype = C.c_int
        L.orc_bicgstab_config.argtypes = [C.c_int, C.c_double]
        L.orc_bicgstab_restarts.restype = C.c_int
        L.orc_perturb_shape.argtypes = [vp, C.c_int] + [C.c_double] * 6
        L.orc_perturb_shape.restype = C.c_int
        L.orc_shape_gradient.argtypes = [vp, P, D, C.c_int] + [C.c_double] * 7 + [D]
        L.orc_shape_gradient.restype = C.c_int
        L.orc_hllc_batch.argtypes = [D, D, D, C.c_double, D, C.c_int]
        L.orc_num_threads.restype = C.c_int
        _LIB = L
    return _LIB


def _d(a):
    return a.ctypes.data_as(D)


def _i(a):
    return a.ctypes.data_as(I)


def hllc_batch(qL, qR, normal, gamma=1.4):
    qL = np.ascontiguousarray(qL, dtype=np.float64)
    qR = np.ascontiguousarray(qR, dtype=np.float64)
    normal = np.ascontiguousarray(normal, dtype=np.float64)
    n = qL.shape[1]
    F = np.zeros((4, n))
    lib().orc_hllc_batch(_d(qL), _d(qR), _d(normal), gamma, _d(F), n)
    return F


PARAM_CODES = {"m": 0, "cambrure": 0, "p": 1, "position_cambrure": 1}


class Mesh:
    """One reference Mesh plus its cell state (the reference's AoS Mesh)."""

    def __init__(self, desc=None, _handle=None):
        L = lib()
        if _handle is not None:
            self.h = _handle
        else:
            xy = np.ascontiguousarray(desc["xy"], dtype=np.float64)
            cnv = np.ascontiguousarray(desc["cell_nv"], dtype=np.int32)
            cv = np.ascontiguousarray(desc["cell_vtx"], dtype=np.int32)
            bv = np.ascontiguousarray(desc["bnd_vtx"], dtype=np.int32)
            bt = bytes(np.asarray(desc["bnd_tag"], dtype=np.uint8))
            self.h = L.orc_mesh_create(xy.shape[0], _d(xy), cnv.shape[0], _i(cnv), _i(cv),
                                       bv.shape[0], _i(bv), bt)
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        L.orc_counts(self.h, C.byref(a), C.byref(b), C.byref(c))
        self.nv, self.ne, self.nc = a.value, b.value, c.value

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.orc_mesh_destroy(self.h)
            self.h = None

    def clone(self):
        return Mesh(_handle=lib().orc_mesh_clone(self.h))

    # state
    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(4, self.nc)
        lib().orc_set_state(self.h, _d(U))

    def state(self):
        U = np.zeros((4, self.nc))
        lib().orc_get_state(self.h, _d(U))
        return U

    def residual_array(self):
        R = np.zeros((4, self.nc))
        lib().orc_get_residual(self.h, _d(R))
        return R

    def gradients(self):
        G = np.zeros((8, self.nc))
        lib().orc_get_gradients(self.h, _d(G))
        return G

    def faces(self):
        L_, R_ = np.zeros((4, self.ne)), np.zeros((4, self.ne))
        lib().orc_get_faces(self.h, _d(L_), _d(R_))
        return L_, R_

    def set_faces(self, fl, fr):
        fl = np.ascontiguousarray(fl, dtype=np.float64)
        fr = np.ascontiguousarray(fr, dtype=np.float64)
        lib().orc_set_faces(self.h, _d(fl), _d(fr))

    def edges(self):
        n = self.ne
        va, vb, cl, cr = (np.zeros(n, np.int32) for _ in range(4))
        bc = C.create_string_buffer(n)
        geo = np.zeros((5, n))
        lib().orc_get_edges(self.h, _i(va), _i(vb), _i(cl), _i(cr), bc, _d(geo))
        return {"va": va, "vb": vb, "cl": cl, "cr": cr,
                "bc": np.frombuffer(bc.raw, dtype=np.uint8).copy(), "len": geo[0],
                "nx": geo[1], "ny": geo[2], "mx": geo[3], "my": geo[4]}

    def cells(self):
        geo = np.zeros((3, self.nc))
        ce = np.zeros((self.nc, 4), np.int32)
        lib().orc_get_cells(self.h, _d(geo), _i(ce))
        return {"vol": geo[0], "cx": geo[1], "cy": geo[2], "edges": ce}

    def vertices(self):
        bc = C.create_string_buffer(self.nv)
        xy = np.zeros((self.nv, 2))
        lib().orc_get_vertex_bc(self.h, bc, _d(xy))
        return {"bc": np.frombuffer(bc.raw, dtype=np.uint8).copy(), "xy": xy}

    # operators
    def init_freestream(self, p):
        lib().orc_init_freestream(self.h, C.byref(p))

    def advection(self, p):
        if lib().orc_advection(self.h, C.byref(p)):
            raise ValueError("Unknown boundary condition")
        return self.residual_array()

    def diffusion(self, p):
        lib().orc_diffusion(self.h, C.byref(p))
        return self.residual_array()

    def residual(self, p):
        if lib().orc_residual(self.h, C.byref(p)):
            raise ValueError("Unknown boundary condition")
        return self.residual_array()

    def dt(self, p):
        return lib().orc_dt(self.h, C.byref(p))

    def rk3_step(self, p, dt):
        if lib().orc_rk3_step(self.h, C.byref(p), dt):
            raise ValueError("Unknown boundary condition")

    def save_old(self):
        lib().orc_save_old(self.h)

    def rk3_substep(self, p, c0, c1, dt):
        if lib().orc_rk3_substep(self.h, C.byref(p), c0, c1, dt):
            raise ValueError("Unknown boundary condition")

    def run_steps(self, p, n):
        dts = np.zeros(n)
        if lib().orc_run_steps(self.h, C.byref(p), n, _d(dts)):
            raise ValueError("Unknown boundary condition")
        return dts

    def time_loop(self, p, final_time, max_steps=10 ** 9):
        """Simulator::time_loop: steps until t >= final_time; (steps, t)."""
        n, t = C.c_int(), C.c_double()
        if lib().orc_time_loop(self.h, C.byref(p), final_time, max_steps, C.byref(n), C.byref(t)):
            raise ValueError("Unknown boundary condition")
        return n.value, t.value

    def residual_norm(self, p):
        return lib().orc_residual_norm(self.h, C.byref(p))

    def run_to_residual(self, p, tol, max_steps, check_every=10):
        n, r = C.c_int(), C.c_double()
        if lib().orc_run_to_residual(self.h, C.byref(p), tol, max_steps, check_every, C.byref(n),
                                     C.byref(r)):
            raise ValueError("Unknown boundary condition")
        return n.value, r.value

    def lift_drag(self, p):
        cl, cd = C.c_double(), C.c_double()
        lib().orc_lift_drag(self.h, C.byref(p), C.byref(cl), C.byref(cd))
        return cl.value, cd.value

    def objective(self, p, beta):
        return lib().orc_objective(self.h, C.byref(p), beta)

    def adjoint_rhs(self, p, beta):
        b = np.zeros((4, self.nc))
        lib().orc_adjoint_rhs(self.h, C.byref(p), beta, _d(b))
        return b

    def apply_jt(self, p, x, eps=1e-6):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(4, self.nc)
        y = np.zeros((4, self.nc))
        lib().orc_apply_jt(self.h, C.byref(p), _d(x), _d(y), eps)
        return y

    def edge_blocks(self, p, e, eps=1e-6):
        ful, fur = np.zeros((4, 4)), np.zeros((4, 4))
        lib().orc_edge_blocks(self.h, C.byref(p), e, eps, _d(ful), _d(fur))
        return ful, fur

    def bicgstab(self, p, beta, eps=1e-6, tol=1e-6, chunk=500, max_total=100000):
        lam = np.zeros((4, self.nc))
        it, rr = C.c_int(), C.c_double()
        st = lib().orc_bicgstab(self.h, C.byref(p), beta, eps, tol, chunk, max_total, _d(lam),
                                C.byref(it), C.byref(rr))
        return st, lam, it.value, rr.value

    def perturb_shape(self, name, pos_x, c, m0, p0, t, eps):
        return lib().orc_perturb_shape(self.h, PARAM_CODES.get(name, -1), pos_x, c, m0, p0, t, eps)

    def shape_gradient(self, p, lam, name, pos_x=0.355, c=1.0, m0=0.02, p0=0.4, t=0.12,
                       beta=1.0, eps=1e-6):
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        g = C.c_double()
        lib().orc_shape_gradient(self.h, C.byref(p), _d(lam), PARAM_CODES.get(name, -1), pos_x, c,
                                 m0, p0, t, beta, eps, C.byref(g))
        return g.value


def bicgstab_config(device_order=False, restart_eps=2.220446049250313e-16):
    """Dot-product order of the oracle's BiCGSTAB (sequential, or the
    device's reduction order for bitwise comparison) and the epsilon of its
    orthogonality restart test (default DBL_EPSILON, Eigen's)."""
    lib().orc_bicgstab_config(int(device_order), restart_eps)


def bicgstab_restarts():
    """Orthogonality restarts taken by the last Mesh.bicgstab call."""
    return lib().orc_bicgstab_restarts()


def num_threads():
    return lib().orc_num_threads()
