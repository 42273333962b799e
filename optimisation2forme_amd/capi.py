"""ctypes binding of the C ABI in include/o2f.h (libo2f.so, built in-tree).

The product path: every call goes to the gfx950 library.  There is no CPU
fallback -- loading fails loudly when libo2f.so is missing, and creating a
context fails loudly without a GPU.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libo2f.so")

D = C.POINTER(C.c_double)
I = C.POINTER(C.c_int)
F = C.POINTER(C.c_float)

O2F_OK = 0
ERRORS = {-1: "O2F_ERR_ARG", -2: "O2F_ERR_BC", -3: "O2F_ERR_HIP", -4: "O2F_ERR_IO",
          -5: "O2F_ERR_DIVERGED", -6: "O2F_ERR_NOCONV"}

LIMITERS = {"venkat": 0, "venkatakrishnan": 0, "barth_jespersen": 1, "bj": 1}
INTERP_MODES = {"ref": 0, "mvc": 1}
HEAT_FLUX = {"reference": 0, "ref": 0, "fourier": 1}
GEOMETRY = {"reference": 0, "full": 1}  # O2F_GEOM_*

# Every symbol include/o2f.h declares (checked by the CPU test suite).
EXPORTS = [
    "o2f_params_default", "o2f_msh2_read", "o2f_msh2_free", "o2f_ctx_create", "o2f_ctx_destroy",
    "o2f_last_error", "o2f_global_error", "o2f_counts", "o2f_set_params", "o2f_set_state",
    "o2f_get_state", "o2f_init_freestream", "o2f_load_state_file", "o2f_save_state_file",
    "o2f_advection", "o2f_diffusion", "o2f_residual", "o2f_get_gradients", "o2f_get_face_states",
    "o2f_set_face_states", "o2f_hllc_flux_batch", "o2f_compute_dt", "o2f_rk3_step",
    "o2f_run_steps", "o2f_time_loop", "o2f_residual_norm", "o2f_lift_drag", "o2f_objective",
    "o2f_adjoint_rhs", "o2f_apply_jt", "o2f_bicgstab", "o2f_shape_gradient", "o2f_time_steps",
    "o2f_mesh_topology", "o2f_mesh_reorder", "o2f_run_until", "o2f_run_to_residual", "o2f_batch_create",
    "o2f_batch_destroy", "o2f_batch_last_error", "o2f_batch_set_shapes", "o2f_batch_get_shapes",
    "o2f_batch_autotune", "o2f_batch_run", "o2f_batch_time_steps", "o2f_apply_jt_exact",
    "o2f_bicgstab_exact", "o2f_mlp_create", "o2f_mlp_destroy", "o2f_mlp_last_error", "o2f_mlp_num_params",
    "o2f_mlp_get", "o2f_mlp_set", "o2f_mlp_set_data", "o2f_mlp_train", "o2f_mlp_forward",
    "o2f_mlp_loss_grad", "o2f_mlp_warm_start", "o2f_time_adjoint", "o2f_time_kernel", "o2f_naca_displacement",
    "o2f_residual_tangent", "o2f_shape_gradient_tangent", "o2f_adjoint_rhs_exact",
    "o2f_adjoint_block_diagonal", "o2f_bicgstab_options", "o2f_mlp_set_gemm_variant", "o2f_residual_jvp",
    "o2f_rk3_save_old", "o2f_rk3_substep", "o2f_bicgstab_set_restart_eps",
    "o2f_hllc_flux_batch_noalloc", "o2f_batch_run_until", "o2f_set_tile_cells", "o2f_tile_info",
    "o2f_batch_set_tile_cells", "o2f_batch_tile_info", "o2f_batch_bicgstab",
]


class Params(C.Structure):
    """o2f_params (include/o2f.h)."""
    _fields_ = [
        ("gamma", C.c_double), ("viscosity", C.c_double), ("Pr", C.c_double), ("cp", C.c_double),
        ("cfl_number", C.c_double), ("rho_ref", C.c_double), ("p_ref", C.c_double),
        ("inflow_velocity", C.c_double), ("venkat_k", C.c_double), ("limiter", C.c_int),
        ("interp_mode", C.c_int), ("bc_inflow_outflow_x", C.c_int), ("heat_flux", C.c_int),
    ]


class MeshDesc(C.Structure):
    """o2f_mesh_desc (include/o2f.h)."""
    _fields_ = [
        ("n_vertices", C.c_int), ("xy", D), ("n_cells", C.c_int), ("cell_nv", I),
        ("cell_vtx", I), ("n_bnd", C.c_int), ("bnd_vtx", I), ("bnd_tag", C.c_char_p),
    ]


class O2FError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (ERRORS.get(code, code), msg))
        self.code = code


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libo2f.so is not built: run `python -m optimisation2forme_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        P = C.POINTER(Params)
        sig = {
            "o2f_params_default": (None, [P]),
            "o2f_msh2_read": (C.c_int, [C.c_char_p, C.POINTER(MeshDesc)]),
            "o2f_msh2_free": (None, [C.POINTER(MeshDesc)]),
            "o2f_ctx_create": (C.c_int, [C.POINTER(MeshDesc), P, C.POINTER(vp)]),
            "o2f_ctx_destroy": (None, [vp]),
            "o2f_last_error": (C.c_char_p, [vp]),
            "o2f_global_error": (C.c_char_p, []),
            "o2f_counts": (C.c_int, [vp, I, I, I]),
            "o2f_set_params": (C.c_int, [vp, P]),
            "o2f_set_state": (C.c_int, [vp, D]),
            "o2f_get_state": (C.c_int, [vp, D]),
            "o2f_init_freestream": (C.c_int, [vp]),
            "o2f_load_state_file": (C.c_int, [vp, C.c_char_p]),
            "o2f_save_state_file": (C.c_int, [vp, C.c_char_p]),
            "o2f_advection": (C.c_int, [vp, D]),
            "o2f_diffusion": (C.c_int, [vp, D]),
            "o2f_residual": (C.c_int, [vp, D]),
            "o2f_get_gradients": (C.c_int, [vp, D]),
            "o2f_get_face_states": (C.c_int, [vp, D, D]),
            "o2f_set_face_states": (C.c_int, [vp, D, D]),
            "o2f_hllc_flux_batch": (C.c_int, [D, D, D, C.c_double, D, C.c_int]),
            "o2f_hllc_flux_batch_noalloc": (C.c_int, [D, D, D, C.c_double, D, C.c_int]),
            "o2f_compute_dt": (C.c_int, [vp, D]),
            "o2f_rk3_step": (C.c_int, [vp, C.c_double]),
            "o2f_rk3_save_old": (C.c_int, [vp]),
            "o2f_bicgstab_set_restart_eps": (C.c_int, [vp, C.c_double]),
            "o2f_rk3_substep": (C.c_int, [vp, C.c_double, C.c_double, C.c_double]),
            "o2f_run_steps": (C.c_int, [vp, C.c_int, D, D]),
            "o2f_time_loop": (C.c_int, [vp, C.c_double, C.c_int, I, D]),
            "o2f_residual_norm": (C.c_int, [vp, D]),
            "o2f_lift_drag": (C.c_int, [vp, D, D]),
            "o2f_objective": (C.c_int, [vp, C.c_double, D]),
            "o2f_adjoint_rhs": (C.c_int, [vp, C.c_double, D]),
            "o2f_apply_jt": (C.c_int, [vp, D, D, C.c_double]),
            "o2f_bicgstab": (C.c_int, [vp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, D, I, D]),
            "o2f_shape_gradient": (C.c_int, [vp, C.POINTER(C.c_char_p), D, C.c_int,
                                             C.POINTER(C.c_char_p), C.c_int, D, C.c_double,
                                             C.c_double, D]),
            "o2f_naca_displacement": (C.c_int, [vp, C.c_char_p, C.POINTER(C.c_char_p), D, C.c_int, D]),
            "o2f_residual_tangent": (C.c_int, [vp, D, C.c_int, C.c_double, D, D, D]),
            "o2f_residual_jvp": (C.c_int, [vp, D, D, D]),
            "o2f_shape_gradient_tangent": (C.c_int, [vp, C.POINTER(C.c_char_p), D, C.c_int,
                                                     C.POINTER(C.c_char_p), C.c_int, D, C.c_double,
                                                     C.c_int, D]),
            "o2f_time_steps": (C.c_int, [vp, C.c_int, D, D]),
            "o2f_mesh_topology": (C.c_int, [C.POINTER(MeshDesc), I, I, I, C.c_char_p, D, D, I]),
            "o2f_mesh_reorder": (C.c_int, [C.POINTER(MeshDesc), C.c_int, C.POINTER(MeshDesc), I, I]),
            "o2f_run_until": (C.c_int, [vp, C.c_double, C.c_int, I, D, D, D]),
            "o2f_run_to_residual": (C.c_int, [vp, C.c_double, C.c_int, C.c_int, C.c_int, I, D]),
            "o2f_apply_jt_exact": (C.c_int, [vp, D, D]),
            "o2f_time_adjoint": (C.c_int, [vp, C.c_int, C.c_int, D]),
            "o2f_time_kernel": (C.c_int, [vp, C.c_int, C.c_int, D]),
            "o2f_set_tile_cells": (C.c_int, [vp, C.c_int]),
            "o2f_tile_info": (C.c_int, [vp, I, I, I, I, I]),
            "o2f_mlp_create": (C.c_int, [C.c_ulonglong, C.POINTER(vp)]),
            "o2f_mlp_destroy": (None, [vp]),
            "o2f_mlp_last_error": (C.c_char_p, [vp]),
            "o2f_mlp_num_params": (C.c_int, [vp, C.POINTER(C.c_long), I]),
            "o2f_mlp_get": (C.c_int, [vp, F, F]),
            "o2f_mlp_set": (C.c_int, [vp, F]),
            "o2f_mlp_set_data": (C.c_int, [vp, F, F, C.c_long]),
            "o2f_mlp_train": (C.c_int, [vp, C.c_int, C.c_int, C.c_float, C.c_ulonglong, F]),
            "o2f_mlp_forward": (C.c_int, [vp, F, C.c_long, F]),
            "o2f_mlp_loss_grad": (C.c_int, [vp, F, F, C.c_long, F, F]),
            "o2f_mlp_warm_start": (C.c_int, [vp, vp, C.c_double, C.c_double, C.c_double, C.c_double]),
            "o2f_mlp_set_gemm_variant": (C.c_int, [vp, C.c_int]),
            "o2f_bicgstab_exact": (C.c_int, [vp, C.c_double, C.c_double, C.c_int, C.c_int, D, I, D, C.c_int]),
            "o2f_adjoint_rhs_exact": (C.c_int, [vp, C.c_double, D]),
            "o2f_adjoint_block_diagonal": (C.c_int, [vp, C.c_int, D]),
            "o2f_bicgstab_options": (C.c_int, [vp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                               D, I, D]),
            "o2f_batch_create": (C.c_int, [C.POINTER(vp), C.c_int, C.POINTER(vp)]),
            "o2f_batch_destroy": (None, [vp]),
            "o2f_batch_last_error": (C.c_char_p, [vp]),
            "o2f_batch_set_shapes": (C.c_int, [vp, I]),
            "o2f_batch_get_shapes": (C.c_int, [vp, I]),
            "o2f_batch_autotune": (C.c_int, [vp, C.c_int, I]),
            "o2f_batch_run": (C.c_int, [vp, C.c_double, C.c_int, C.c_int, C.c_int, I, D]),
            "o2f_batch_time_steps": (C.c_int, [vp, C.c_int, C.c_int, D]),
            "o2f_batch_set_tile_cells": (C.c_int, [vp, C.c_int]),
            "o2f_batch_tile_info": (C.c_int, [vp, I, I, I, I, I]),
            "o2f_batch_bicgstab": (C.c_int, [vp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                             C.POINTER(D), I, D, I]),
            "o2f_batch_run_until": (C.c_int, [vp, C.c_double, C.c_int, C.c_int, I, D]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _d(a):
    return a.ctypes.data_as(D)


def _i(a):
    return a.ctypes.data_as(I)


def make_params(gamma=1.4, viscosity=1e-6, Pr=0.71, cp=1005.0, cfl_number=0.1, rho_ref=1.225,
                p_ref=101325.0, inflow_velocity=1.0, venkat_k=1e-2, limiter="venkat",
                interp_mode="ref", bc_inflow_outflow_x=True, heat_flux="reference"):
    """o2f_params; defaults are SimulationParams' (include/params.hpp:15-73)."""
    if isinstance(limiter, str):
        limiter = LIMITERS[limiter]
    if isinstance(interp_mode, str):
        interp_mode = INTERP_MODES[interp_mode]
    if isinstance(heat_flux, str):
        heat_flux = HEAT_FLUX[heat_flux]
    return Params(gamma, viscosity, Pr, cp, cfl_number, rho_ref, p_ref, inflow_velocity, venkat_k,
                  int(limiter), int(interp_mode), int(bool(bc_inflow_outflow_x)), int(heat_flux))


def _mesh_desc(desc):
    xy = np.ascontiguousarray(desc["xy"], dtype=np.float64)
    cnv = np.ascontiguousarray(desc["cell_nv"], dtype=np.int32)
    cv = np.ascontiguousarray(desc["cell_vtx"], dtype=np.int32)
    bv = np.ascontiguousarray(desc["bnd_vtx"], dtype=np.int32)
    bt = bytes(np.asarray(desc["bnd_tag"], dtype=np.uint8))
    keep = (xy, cnv, cv, bv, bt)
    return MeshDesc(xy.shape[0], _d(xy), cnv.shape[0], _i(cnv), _i(cv), bv.shape[0], _i(bv), bt), keep


def mesh_topology(desc):
    """o2f_mesh_topology: the library's host-side topology and geometry."""
    L = lib()
    md, keep = _mesh_desc(desc)
    ne = C.c_int()
    st = L.o2f_mesh_topology(C.byref(md), C.byref(ne), None, None, None, None, None, None)
    if st != O2F_OK:
        raise O2FError(st, L.o2f_global_error().decode())
    n, nc = ne.value, md.n_cells
    ev, ec = np.zeros(2 * n, np.int32), np.zeros(2 * n, np.int32)
    bc = C.create_string_buffer(max(n, 1))
    eg, cg = np.zeros(5 * n), np.zeros(3 * nc)
    ce = np.zeros(4 * nc, np.int32)
    st = L.o2f_mesh_topology(C.byref(md), C.byref(ne), _i(ev), _i(ec), bc, _d(eg), _d(cg), _i(ce))
    if st != O2F_OK:
        raise O2FError(st, L.o2f_global_error().decode())
    eg, cg = eg.reshape(5, n), cg.reshape(3, nc)
    return {"va": ev[0::2], "vb": ev[1::2], "cl": ec[0::2], "cr": ec[1::2],
            "bc": np.frombuffer(bc.raw[:n], dtype=np.uint8).copy(), "len": eg[0], "nx": eg[1],
            "ny": eg[2], "mx": eg[3], "my": eg[4], "vol": cg[0], "cx": cg[1], "cy": cg[2],
            "edges": ce.reshape(nc, 4)}


def _owned_desc(L, d):
    """A library-owned mesh description -> dict of numpy arrays (then freed)."""
    try:
        nv, nc, nb = d.n_vertices, d.n_cells, d.n_bnd
        out = {
            "xy": np.ctypeslib.as_array(d.xy, (nv * 2,)).reshape(nv, 2).copy(),
            "cell_nv": np.ctypeslib.as_array(d.cell_nv, (nc,)).copy(),
            "cell_vtx": np.ctypeslib.as_array(d.cell_vtx, (nc * 4,)).reshape(nc, 4).copy(),
            "bnd_vtx": (np.ctypeslib.as_array(d.bnd_vtx, (nb * 2,)).reshape(nb, 2).copy()
                        if nb else np.zeros((0, 2), np.int32)),
            "bnd_tag": np.frombuffer(C.string_at(d.bnd_tag, nb), dtype=np.uint8).copy(),
        }
    finally:
        L.o2f_msh2_free(C.byref(d))
    return out


def read_msh2(path):
    """o2f_msh2_read -> dict of numpy arrays (the Gmsh API element order)."""
    L = lib()
    d = MeshDesc()
    st = L.o2f_msh2_read(path.encode(), C.byref(d))
    if st != O2F_OK:
        raise O2FError(st, L.o2f_global_error().decode())
    return _owned_desc(L, d)


REORDER = {"rcm": 0, "rcm_xcd": 1}


def reorder_mesh(desc, method="rcm"):
    """o2f_mesh_reorder -> (renumbered mesh, cell_perm, vertex_perm), both
    permutations new -> old: a state maps as U_new = U_old[:, cell_perm]."""
    L = lib()
    md, keep = _mesh_desc(desc)
    out = MeshDesc()
    nc, nv = md.n_cells, md.n_vertices
    cperm = np.zeros(nc, np.int32)
    vperm = np.zeros(nv, np.int32)
    st = L.o2f_mesh_reorder(C.byref(md), REORDER.get(method, method), C.byref(out), _i(cperm), _i(vperm))
    if st != O2F_OK:
        raise O2FError(st, L.o2f_global_error().decode())
    return _owned_desc(L, out), cperm, vperm


class Solver:
    """One o2f context: a mesh resident on the GPU plus its state.

    Arrays in and out are float64 SoA in the reference's cell / edge order."""

    def __init__(self, desc, params):
        L = lib()
        self._keep = []
        xy = np.ascontiguousarray(desc["xy"], dtype=np.float64)
        cnv = np.ascontiguousarray(desc["cell_nv"], dtype=np.int32)
        cv = np.ascontiguousarray(desc["cell_vtx"], dtype=np.int32)
        bv = np.ascontiguousarray(desc["bnd_vtx"], dtype=np.int32)
        bt = bytes(np.asarray(desc["bnd_tag"], dtype=np.uint8))
        md = MeshDesc(xy.shape[0], _d(xy), cnv.shape[0], _i(cnv), _i(cv), bv.shape[0], _i(bv), bt)
        self.params = params
        h = C.c_void_p()
        st = L.o2f_ctx_create(C.byref(md), C.byref(params), C.byref(h))
        if st != O2F_OK:
            raise O2FError(st, L.o2f_global_error().decode())
        self.h = h
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        L.o2f_counts(self.h, C.byref(a), C.byref(b), C.byref(c))
        self.nv, self.ne, self.nc = a.value, b.value, c.value

    def close(self):
        if getattr(self, "h", None):
            lib().o2f_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != O2F_OK:
            raise O2FError(st, lib().o2f_last_error(self.h).decode())

    def set_params(self, params):
        self.params = params
        self._check(lib().o2f_set_params(self.h, C.byref(params)))

    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(4, self.nc)
        self._check(lib().o2f_set_state(self.h, _d(U)))

    def state(self):
        U = np.zeros((4, self.nc))
        self._check(lib().o2f_get_state(self.h, _d(U)))
        return U

    def init_freestream(self):
        self._check(lib().o2f_init_freestream(self.h))

    def load_state_file(self, path):
        self._check(lib().o2f_load_state_file(self.h, path.encode()))

    def save_state_file(self, path):
        self._check(lib().o2f_save_state_file(self.h, path.encode()))

    def advection(self):
        R = np.zeros((4, self.nc))
        self._check(lib().o2f_advection(self.h, _d(R)))
        return R

    def diffusion(self):
        R = np.zeros((4, self.nc))
        self._check(lib().o2f_diffusion(self.h, _d(R)))
        return R

    def residual(self):
        R = np.zeros((4, self.nc))
        self._check(lib().o2f_residual(self.h, _d(R)))
        return R

    def gradients(self):
        G = np.zeros((8, self.nc))
        self._check(lib().o2f_get_gradients(self.h, _d(G)))
        return G

    def faces(self):
        fl, fr = np.zeros((4, self.ne)), np.zeros((4, self.ne))
        self._check(lib().o2f_get_face_states(self.h, _d(fl), _d(fr)))
        return fl, fr

    def set_faces(self, fl, fr):
        fl = np.ascontiguousarray(fl, dtype=np.float64)
        fr = np.ascontiguousarray(fr, dtype=np.float64)
        self._check(lib().o2f_set_face_states(self.h, _d(fl), _d(fr)))

    def dt(self):
        v = C.c_double()
        self._check(lib().o2f_compute_dt(self.h, C.byref(v)))
        return v.value

    def rk3_step(self, dt):
        self._check(lib().o2f_rk3_step(self.h, dt))

    def rk3_save_old(self):
        self._check(lib().o2f_rk3_save_old(self.h))

    def rk3_substep(self, c0, c1, dt):
        self._check(lib().o2f_rk3_substep(self.h, c0, c1, dt))

    def run_steps(self, n, t=0.0):
        dts = np.zeros(max(n, 1))
        tt = C.c_double(t)
        self._check(lib().o2f_run_steps(self.h, n, _d(dts), C.byref(tt)))
        return dts[:n], tt.value

    def run_until(self, final_time, max_steps=-1, t=0.0):
        n, tt, ts, dl = C.c_int(), C.c_double(t), C.c_double(), C.c_double()
        self._check(lib().o2f_run_until(self.h, final_time, max_steps, C.byref(n), C.byref(tt),
                                        C.byref(ts), C.byref(dl)))
        return n.value, tt.value, ts.value, dl.value

    def run_to_residual(self, tol, max_steps, check_every=10, use_graph=True):
        """Steps until the density residual of a checked step is <= tol;
        returns (steps, last residual)."""
        n, r = C.c_int(), C.c_double()
        self._check(lib().o2f_run_to_residual(self.h, tol, max_steps, check_every, int(use_graph),
                                              C.byref(n), C.byref(r)))
        return n.value, r.value

    def time_loop(self, final_time, max_steps=-1):
        n, t = C.c_int(), C.c_double()
        self._check(lib().o2f_time_loop(self.h, final_time, max_steps, C.byref(n), C.byref(t)))
        return n.value, t.value

    def residual_norm(self):
        v = C.c_double()
        self._check(lib().o2f_residual_norm(self.h, C.byref(v)))
        return v.value

    def lift_drag(self):
        a, b = C.c_double(), C.c_double()
        self._check(lib().o2f_lift_drag(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def objective(self, beta):
        v = C.c_double()
        self._check(lib().o2f_objective(self.h, beta, C.byref(v)))
        return v.value

    def adjoint_rhs(self, beta):
        b = np.zeros((4, self.nc))
        self._check(lib().o2f_adjoint_rhs(self.h, beta, _d(b)))
        return b

    def apply_jt(self, x, eps=1e-6):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(4, self.nc)
        y = np.zeros((4, self.nc))
        self._check(lib().o2f_apply_jt(self.h, _d(x), _d(y), eps))
        return y

    def apply_jt_exact(self, x):
        """y = (dR/dU)^T x, the exact discrete adjoint of the full residual."""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(4, self.nc)
        y = np.zeros((4, self.nc))
        self._check(lib().o2f_apply_jt_exact(self.h, _d(x), _d(y)))
        return y

    def bicgstab_exact(self, beta, tol=1e-6, chunk=500, max_total=100000, check=True, exact_rhs=False,
                       block_jacobi=False):
        lam = np.zeros((4, self.nc))
        it, rr = C.c_int(), C.c_double()
        opts = (1 if exact_rhs else 0) | (2 if block_jacobi else 0)
        st = lib().o2f_bicgstab_exact(self.h, beta, tol, chunk, max_total, _d(lam), C.byref(it), C.byref(rr), opts)
        if check:
            self._check(st)
        return lam, it.value, rr.value, st

    def bicgstab_block_jacobi(self, beta, eps=1e-6, tol=1e-6, chunk=500, max_total=100000, check=True):
        """o2f_bicgstab right-preconditioned by the operator's diagonal blocks."""
        lam = np.zeros((4, self.nc))
        it, rr = C.c_int(), C.c_double()
        st = lib().o2f_bicgstab_options(self.h, beta, eps, tol, chunk, max_total, 2, _d(lam), C.byref(it),
                                        C.byref(rr))
        if check:
            self._check(st)
        return lam, it.value, rr.value, st

    def set_restart_eps(self, restart_eps):
        self._check(lib().o2f_bicgstab_set_restart_eps(self.h, restart_eps))

    def bicgstab(self, beta, eps=1e-6, tol=1e-6, chunk=500, max_total=100000, check=True):
        lam = np.zeros((4, self.nc))
        it, rr = C.c_int(), C.c_double()
        st = lib().o2f_bicgstab(self.h, beta, eps, tol, chunk, max_total, _d(lam), C.byref(it),
                                C.byref(rr))
        if check:
            self._check(st)
        return lam, it.value, rr.value, st

    def shape_gradient(self, names_all, values_all, opt_names, lam, beta, eps=1e-6):
        na = (C.c_char_p * len(names_all))(*[n.encode() for n in names_all])
        va = np.ascontiguousarray(values_all, dtype=np.float64)
        no = (C.c_char_p * max(len(opt_names), 1))(*[n.encode() for n in opt_names])
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        g = np.zeros(max(len(opt_names), 1))
        self._check(lib().o2f_shape_gradient(self.h, na, _d(va), len(names_all), no, len(opt_names),
                                             _d(lam), beta, eps, _d(g)))
        return g[:len(opt_names)]

    def adjoint_block_diagonal(self, exact=True):
        """dR_c/dU_c per cell, shape (n_cells, 4, 4)."""
        D = np.zeros((self.nc, 4, 4))
        self._check(lib().o2f_adjoint_block_diagonal(self.h, int(exact), _d(D)))
        return D

    def adjoint_rhs_exact(self, beta):
        """(dJ/dU)^T through the reconstruction (4 x n_cells)."""
        b = np.zeros((4, self.nc))
        self._check(lib().o2f_adjoint_rhs_exact(self.h, beta, _d(b)))
        return b

    def naca_displacement(self, name, names_all, values_all):
        """d(x, y)/d(name) of the wall vertices, shape (n_vertices, 2)."""
        na = (C.c_char_p * max(len(names_all), 1))(*[n.encode() for n in names_all])
        va = np.ascontiguousarray(values_all if len(values_all) else [0.0], dtype=np.float64)
        d = np.zeros((self.nv, 2))
        self._check(lib().o2f_naca_displacement(self.h, name.encode(), na, _d(va), len(names_all), _d(d)))
        return d

    def residual_tangent(self, dxy, geometry="full", beta=1.0):
        """(R, dR, dJ): the residual, its derivative along the vertex
        displacement dxy (n_vertices x 2) and that of J = -C_L + beta C_D,
        at the current state."""
        dxy = np.ascontiguousarray(dxy, dtype=np.float64).reshape(self.nv, 2)
        R, dR, dJ = np.zeros((4, self.nc)), np.zeros((4, self.nc)), C.c_double()
        self._check(lib().o2f_residual_tangent(self.h, _d(dxy), GEOMETRY[geometry], beta, _d(R), _d(dR),
                                               C.byref(dJ)))
        return R, dR, dJ.value

    def residual_jvp(self, dU):
        """(R, (dR/dU) dU) at the current state, forward mode."""
        dU = np.ascontiguousarray(dU, dtype=np.float64).reshape(4, self.nc)
        R, dR = np.zeros((4, self.nc)), np.zeros((4, self.nc))
        self._check(lib().o2f_residual_jvp(self.h, _d(dU), _d(R), _d(dR)))
        return R, dR

    def shape_gradient_tangent(self, names_all, values_all, opt_names, lam, beta, geometry="reference"):
        na = (C.c_char_p * len(names_all))(*[n.encode() for n in names_all])
        va = np.ascontiguousarray(values_all, dtype=np.float64)
        no = (C.c_char_p * max(len(opt_names), 1))(*[n.encode() for n in opt_names])
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        g = np.zeros(max(len(opt_names), 1))
        self._check(lib().o2f_shape_gradient_tangent(self.h, na, _d(va), len(names_all), no, len(opt_names),
                                                     _d(lam), beta, GEOMETRY[geometry], _d(g)))
        return g[:len(opt_names)]

    def time_adjoint(self, nprod, exact):
        ms = C.c_double()
        self._check(lib().o2f_time_adjoint(self.h, nprod, int(exact), C.byref(ms)))
        return ms.value

    def set_tile_cells(self, cells):
        """Stages of the time-stepping entry points: -1 fused tile stages
        sized automatically (the default), 0 the four-kernel stages, n > 0
        fused tile stages of at most n cells per tile."""
        self._check(lib().o2f_set_tile_cells(self.h, int(cells)))

    def tile_info(self):
        """The tiling in use: dict of tiles, cells (per-tile cap), lds_bytes,
        tiles_per_cu, vgprs (tiles 0: the four-kernel stages)."""
        v = [C.c_int() for _ in range(5)]
        self._check(lib().o2f_tile_info(self.h, *[C.byref(x) for x in v]))
        keys = ("tiles", "cells", "lds_bytes", "tiles_per_cu", "vgprs")
        return {k: x.value for k, x in zip(keys, v)}

    def time_kernel(self, kernel, launches):
        """Microseconds per launch of one stage kernel (0 vertex gather,
        1 reconstruction, 2 viscous gather + edge flux, 3 flux sum in
        residual form, 4 those four in order, 5 one fused tile stage),
        launched back to back at the current state."""
        us = C.c_double()
        self._check(lib().o2f_time_kernel(self.h, int(kernel), int(launches), C.byref(us)))
        return us.value

    def time_steps(self, n, per_kernel=False):
        tot = C.c_double()
        km = np.zeros(4)
        self._check(lib().o2f_time_steps(self.h, n, C.byref(tot), _d(km) if per_kernel else None))
        return tot.value, (km if per_kernel else None)


def hllc_flux_batch(qL, qR, normal, gamma=1.4):
    qL = np.ascontiguousarray(qL, dtype=np.float64)
    qR = np.ascontiguousarray(qR, dtype=np.float64)
    normal = np.ascontiguousarray(normal, dtype=np.float64)
    n = qL.shape[1]
    F = np.zeros((4, n))
    st = lib().o2f_hllc_flux_batch(_d(qL), _d(qR), _d(normal), gamma, _d(F), n)
    if st != O2F_OK:
        raise O2FError(st, lib().o2f_global_error().decode())
    return F


class Batch:
    """o2f_batch: several Solvers advanced together, one launch per stage
    kernel over all members.  The Solvers must stay alive while the batch
    exists."""

    KINDS = ("vertex_gather", "cell_reconstruct", "viscous_gather_edge_flux", "flux_update")

    def __init__(self, solvers):
        L = lib()
        self.solvers = list(solvers)
        arr = (C.c_void_p * len(self.solvers))(*[s.h.value for s in self.solvers])
        h = C.c_void_p()
        st = L.o2f_batch_create(arr, len(self.solvers), C.byref(h))
        if st != O2F_OK:
            raise O2FError(st, "o2f_batch_create failed")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().o2f_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != O2F_OK:
            raise O2FError(st, lib().o2f_batch_last_error(self.h).decode())

    def shapes(self):
        a = np.zeros(8, np.int32)
        self._check(lib().o2f_batch_get_shapes(self.h, _i(a)))
        return [(int(a[2 * k]), int(a[2 * k + 1])) for k in range(4)]

    def set_shapes(self, shapes):
        a = np.ascontiguousarray(np.asarray(shapes, np.int32).reshape(8))
        self._check(lib().o2f_batch_set_shapes(self.h, _i(a)))

    def autotune(self, reps=3):
        a = np.zeros(8, np.int32)
        self._check(lib().o2f_batch_autotune(self.h, reps, _i(a)))
        return [(int(a[2 * k]), int(a[2 * k + 1])) for k in range(4)]

    def run(self, tol, max_steps, check_every=10, use_graph=True):
        n = len(self.solvers)
        steps, res = np.zeros(n, np.int32), np.zeros(n)
        self._check(lib().o2f_batch_run(self.h, tol, max_steps, check_every, int(use_graph), _i(steps),
                                        _d(res)))
        return steps, res

    def run_until(self, final_time, max_steps, use_graph=True):
        """Every member's time loop to final_time: (steps, t) per member."""
        n = len(self.solvers)
        steps, t = np.zeros(n, np.int32), np.zeros(n)
        self._check(lib().o2f_batch_run_until(self.h, final_time, max_steps, int(use_graph), _i(steps), _d(t)))
        return steps, t

    def time_steps(self, nsteps, use_graph=True):
        ms = C.c_double()
        self._check(lib().o2f_batch_time_steps(self.h, nsteps, int(use_graph), C.byref(ms)))
        return ms.value

    def set_tile_cells(self, cells):
        """Stages of the batch's steps: -1 fused tile stages over every
        member's tiles, sized automatically (the default), 0 the four-kernel
        batched stages, n > 0 tiles of at most n cells."""
        self._check(lib().o2f_batch_set_tile_cells(self.h, int(cells)))

    def bicgstab(self, beta, eps=1e-6, tol=1e-6, chunk=500, max_total=100000, exact=False):
        """Every member's adjoint solve (Solver.bicgstab, or bicgstab_exact),
        several at once: (lambdas, iters, relres, status) per member."""
        lams = [np.zeros((4, s.nc)) for s in self.solvers]
        ptrs = (D * len(lams))(*[_d(l) for l in lams])
        n = len(self.solvers)
        it, rr, code = np.zeros(n, np.int32), np.zeros(n), np.zeros(n, np.int32)
        self._check(lib().o2f_batch_bicgstab(self.h, beta, eps, tol, chunk, max_total, int(bool(exact)), ptrs,
                                             _i(it), _d(rr), _i(code)))
        return lams, it, rr, code

    def tile_info(self):
        """The tiling in use: dict of tiles (over all members), cells (per-tile
        cap), lds_bytes, tiles_per_cu, vgprs (tiles 0: the four-kernel stages)."""
        v = [C.c_int() for _ in range(5)]
        self._check(lib().o2f_batch_tile_info(self.h, *[C.byref(x) for x in v]))
        keys = ("tiles", "cells", "lds_bytes", "tiles_per_cu", "vgprs")
        return {k: x.value for k, x in zip(keys, v)}


def _f(a):
    return a.ctypes.data_as(F)


class WarmStartNet:
    """o2f_mlp: the learned warm start (BF16 MFMA forward / backward, Adam)."""

    def __init__(self, seed=0):
        h = C.c_void_p()
        st = lib().o2f_mlp_create(seed, C.byref(h))
        if st != O2F_OK:
            raise O2FError(st, "o2f_mlp_create failed")
        self.h = h
        n = C.c_long()
        lay = np.zeros(20, np.int32)
        self._check(lib().o2f_mlp_num_params(self.h, C.byref(n), _i(lay)))
        self.n_params = n.value
        self.layout = [tuple(int(v) for v in lay[4 * l:4 * l + 4]) for l in range(5)]  # K, N, w_off, b_off

    def close(self):
        if getattr(self, "h", None):
            lib().o2f_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != O2F_OK:
            raise O2FError(st, lib().o2f_mlp_last_error(self.h).decode())

    def set_gemm_variant(self, variant):
        """-1: per-shape tuned GEMM kernel (default); 0-6: force one."""
        self._check(lib().o2f_mlp_set_gemm_variant(self.h, int(variant)))

    def weights(self):
        w = np.zeros(self.n_params, np.float32)
        fb = np.zeros(32, np.float32)
        self._check(lib().o2f_mlp_get(self.h, _f(w), _f(fb)))
        return w, fb.reshape(2, 16)

    def set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        self._check(lib().o2f_mlp_set(self.h, _f(w)))

    def set_data(self, raw, targets):
        raw = np.ascontiguousarray(raw, dtype=np.float32)
        targets = np.ascontiguousarray(targets, dtype=np.float32)
        self._check(lib().o2f_mlp_set_data(self.h, _f(raw), _f(targets), raw.shape[0]))

    def train(self, steps, batch=65536, lr=1e-3, seed=1):
        losses = np.zeros(max(steps, 1), np.float32)
        self._check(lib().o2f_mlp_train(self.h, steps, batch, lr, seed, _f(losses)))
        return losses[:steps]

    def forward(self, raw):
        raw = np.ascontiguousarray(raw, dtype=np.float32)
        out = np.zeros((raw.shape[0], 4), np.float32)
        self._check(lib().o2f_mlp_forward(self.h, _f(raw), raw.shape[0], _f(out)))
        return out

    def loss_grad(self, raw, targets):
        raw = np.ascontiguousarray(raw, dtype=np.float32)
        targets = np.ascontiguousarray(targets, dtype=np.float32)
        loss = np.zeros(1, np.float32)
        g = np.zeros(self.n_params, np.float32)
        self._check(lib().o2f_mlp_loss_grad(self.h, _f(raw), _f(targets), raw.shape[0], _f(loss), _f(g)))
        return float(loss[0]), g

    def warm_start(self, solver, m, p, t, aoa_deg):
        self._check(lib().o2f_mlp_warm_start(self.h, solver.h, m, p, t, aoa_deg))


NOALLOC_MAX_FACES = 16384  # O2F_NOALLOC_MAX_FACES


def hllc_flux_batch_noalloc(qL, qR, normal, gamma=1.4):
    """o2f_hllc_flux_batch_noalloc: the same fluxes, staged through the
    library's fixed per-GPU array (at most NOALLOC_MAX_FACES faces)."""
    qL = np.ascontiguousarray(qL, dtype=np.float64)
    qR = np.ascontiguousarray(qR, dtype=np.float64)
    normal = np.ascontiguousarray(normal, dtype=np.float64)
    n = qL.shape[1]
    F = np.zeros((4, n))
    st = lib().o2f_hllc_flux_batch_noalloc(_d(qL), _d(qR), _d(normal), gamma, _d(F), n)
    if st != O2F_OK:
        raise O2FError(st, lib().o2f_global_error().decode())
    return F
