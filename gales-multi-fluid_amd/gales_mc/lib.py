"""ctypes binding of libgales_mc.so (the HIP library) and a thin Context class.

The binding loads only the in-tree library built by `make -C
gales-multi-fluid_amd`; there is no fallback: a missing library raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi
from .cases import mesh_struct

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libgales_mc.so")

_lib = None


class GmcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"gales_mc error {code}: {msg}")
        self.code = code


def load():
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C {PKG_DIR}`")
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER
    d = C.c_double
    vp = C.c_void_p
    sig = {
        "gmc_create": (C.c_int, [P(abi.Mesh), P(abi.Params), P(vp)]),
        "gmc_destroy": (C.c_int, [vp]),
        "gmc_last_error": (C.c_char_p, [vp]),
        "gmc_get_info": (C.c_int, [vp, P(abi.Info)]),
        "gmc_set_stream": (C.c_int, [vp, vp]),
        "gmc_set_dt": (C.c_int, [vp, d]),
        "gmc_set_material": (C.c_int, [vp, P(d)]),
        "gmc_set_state": (C.c_int, [vp, P(d), P(d)]),
        "gmc_get_state": (C.c_int, [vp, P(d), P(d)]),
        "gmc_set_state_device": (C.c_int, [vp, vp, vp]),
        "gmc_assemble": (C.c_int, [vp, C.c_int, P(d)]),
        "gmc_residual": (C.c_int, [vp, P(d)]),
        "gmc_get_timing": (C.c_int, [vp, P(abi.Timing)]),
        "gmc_element": (C.c_int, [vp, C.c_int, C.c_int, P(d), P(d)]),
        "gmc_get_pattern": (C.c_int, [vp, P(C.c_int), P(C.c_int)]),
        "gmc_get_system": (C.c_int, [vp, P(d), P(d)]),
        "gmc_spmv": (C.c_int, [vp, P(d), P(d)]),
        "gmc_jvp": (C.c_int, [vp, P(d), P(d)]),
        "gmc_solve": (C.c_int, [vp, P(abi.Solver), P(abi.SolveInfo)]),
        "gmc_get_solution": (C.c_int, [vp, P(d)]),
        "gmc_newton_update": (C.c_int, [vp]),
        "gmc_predictor": (C.c_int, [vp]),
        "gmc_refine": (C.c_int, [vp, P(abi.Solver), C.c_int, P(d)]),
        "gmc_vjp": (C.c_int, [vp, P(d), P(d), P(d), P(d), C.c_int]),
        "gmc_vjp_device": (C.c_int, [vp, vp, vp, vp, vp, C.c_int]),
        "gmc_residual_device": (C.c_int, [vp, vp, vp, vp]),
        "gmc_corrector": (C.c_int, [vp]),
        "gmc_dof_trimmer": (C.c_int, [vp, C.c_int, d, d]),
        "gmc_rhs_norm": (C.c_int, [vp, P(d)]),
        "gmc_rhs_norm_noalloc": (C.c_int, [vp, P(d)]),
        "gmc_adaptive_dt": (C.c_int, [vp, d, d, d, P(d)]),
        "gmc_hydrostatic_profile": (C.c_int, [P(abi.Params), d, d, d, C.c_int, P(d), P(d), P(d), P(d),
                                              P(d), P(d)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


class Context:
    """One gmc_ctx: mesh, parameters, device state and assembled system."""

    def __init__(self, mesh, params):
        self.lib = load()
        self._mesh = mesh_struct(mesh)
        self.params = params
        h = C.c_void_p()
        st = self.lib.gmc_create(C.byref(self._mesh), C.byref(params), C.byref(h))
        if st != abi.GMC_OK:
            raise GmcError(st, self.lib.gmc_last_error(None).decode())
        self.h = h
        info = abi.Info()
        self._check(self.lib.gmc_get_info(h, C.byref(info)))
        self.info = info
        self.ndof = info.ndof
        self.n_nodes = info.n_nodes
        self.nen = info.nen

    def _check(self, st):
        if st != abi.GMC_OK:
            raise GmcError(st, self.lib.gmc_last_error(self.h).decode())

    def _vec(self, x, name):
        """x as a contiguous float64 vector of n_nodes * ndof entries: the C
        side reads and writes exactly that many, so any other length is
        rejected here."""
        x = np.ascontiguousarray(x, np.float64).reshape(-1)
        if x.size != self.n_nodes * self.ndof:
            raise ValueError(f"{name} has {x.size} entries, expected n_nodes * ndof = "
                             f"{self.n_nodes * self.ndof}")
        return x

    def close(self):
        if getattr(self, "h", None):
            self.lib.gmc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # state -----------------------------------------------------------------
    def set_state(self, I=None, P=None):
        I = None if I is None else self._vec(I, "I")
        P = None if P is None else self._vec(P, "P")
        self._check(self.lib.gmc_set_state(self.h, _dp(I), _dp(P)))

    def get_state(self):
        n = self.n_nodes * self.ndof
        I = np.empty(n)
        P = np.empty(n)
        self._check(self.lib.gmc_get_state(self.h, _dp(I), _dp(P)))
        return I.reshape(self.n_nodes, self.ndof), P.reshape(self.n_nodes, self.ndof)

    def set_stream(self, stream_handle):
        """Orders the context's work on a HIP stream (an int handle, e.g.
        torch.cuda.current_stream().cuda_stream; 0 restores its own stream)."""
        self._check(self.lib.gmc_set_stream(self.h, C.c_void_p(stream_handle or None)))

    def set_state_device(self, I_ptr, P_ptr):
        self._check(self.lib.gmc_set_state_device(self.h, C.c_void_p(I_ptr), C.c_void_p(P_ptr)))

    def residual_device(self, I_ptr, P_ptr, R_ptr):
        """R(I, P) = -rhs with device pointers (I or P 0: keep the current)."""
        self._check(self.lib.gmc_residual_device(self.h, C.c_void_p(I_ptr or None),
                                                 C.c_void_p(P_ptr or None), C.c_void_p(R_ptr)))

    def vjp_device(self, v_ptr, gs_ptr, gp_ptr, gpar_ptr, mode=abi.GMC_VJP_EXACT):
        self._check(self.lib.gmc_vjp_device(self.h, C.c_void_p(v_ptr), C.c_void_p(gs_ptr or None),
                                            C.c_void_p(gp_ptr or None), C.c_void_p(gpar_ptr or None),
                                            mode))

    def residual_vector(self):
        """R(I, P) = -rhs of the context's state, on the host."""
        self.residual()
        return -self.system()[1]

    def set_dt(self, dt):
        self._check(self.lib.gmc_set_dt(self.h, dt))

    def set_material(self, theta):
        th = np.ascontiguousarray(theta, np.float64).reshape(-1)
        if th.size != 6:
            raise ValueError(f"theta has {th.size} entries, expected 6")
        self._check(self.lib.gmc_set_material(self.h, _dp(th)))

    # assembly --------------------------------------------------------------
    def assemble(self, flags=abi.GMC_ASM_MATRIX | abi.GMC_ASM_RHS):
        s = C.c_double()
        self._check(self.lib.gmc_assemble(self.h, flags, C.byref(s)))
        return s.value

    def residual(self):
        n = C.c_double()
        self._check(self.lib.gmc_residual(self.h, C.byref(n)))
        return n.value

    def timing(self):
        t = abi.Timing()
        self._check(self.lib.gmc_get_timing(self.h, C.byref(t)))
        return t

    def element(self, e, boundary=False):
        ne = self.nen * self.ndof
        m = np.empty(ne * ne)
        r = np.empty(ne)
        self._check(self.lib.gmc_element(self.h, int(e), 1 if boundary else 0, _dp(m), _dp(r)))
        return m.reshape(ne, ne), r

    def pattern(self):
        rowptr = np.empty(self.n_nodes + 1, np.int32)
        colidx = np.empty(self.info.nnzb, np.int32)
        self._check(self.lib.gmc_get_pattern(self.h, rowptr.ctypes.data_as(C.POINTER(C.c_int)),
                                             colidx.ctypes.data_as(C.POINTER(C.c_int))))
        return rowptr, colidx

    def system(self):
        vals = np.empty(self.info.nnzb * self.ndof * self.ndof)
        rhs = np.empty(self.n_nodes * self.ndof)
        self._check(self.lib.gmc_get_system(self.h, _dp(vals), _dp(rhs)))
        return vals.reshape(-1, self.ndof, self.ndof), rhs

    def spmv(self, x):
        x = self._vec(x, "x")
        y = np.empty_like(x)
        self._check(self.lib.gmc_spmv(self.h, _dp(x), _dp(y)))
        return y

    def jvp(self, x):
        """Matrix-free FP64 product J(I, P) x (equals spmv after an FP64 assembly)."""
        x = self._vec(x, "x")
        y = np.empty_like(x)
        self._check(self.lib.gmc_jvp(self.h, _dp(x), _dp(y)))
        return y

    # solver ----------------------------------------------------------------
    def solve(self, restart=30, max_iters=30, rel_tol=1e-8, precond=1):
        s = abi.Solver(restart, max_iters, rel_tol, precond)
        info = abi.SolveInfo()
        self._check(self.lib.gmc_solve(self.h, C.byref(s), C.byref(info)))
        return info

    def solution(self):
        x = np.empty(self.n_nodes * self.ndof)
        self._check(self.lib.gmc_get_solution(self.h, _dp(x)))
        return x

    def newton_update(self):
        self._check(self.lib.gmc_newton_update(self.h))

    def predictor(self):
        self._check(self.lib.gmc_predictor(self.h))

    def corrector(self):
        """I += Delta without the trim (fluid_updater::corrector)."""
        self._check(self.lib.gmc_corrector(self.h))

    def dof_trimmer(self, dof_index, lo, hi):
        self._check(self.lib.gmc_dof_trimmer(self.h, int(dof_index), float(lo), float(hi)))

    def rhs_norm(self):
        """||rhs||_2 of the last assembled rhs (no re-assembly)."""
        n = C.c_double()
        self._check(self.lib.gmc_rhs_norm(self.h, C.byref(n)))
        return n.value

    def adaptive_dt(self, cfl, dt_min, dt_max):
        """adaptive_time_criteria::f_mc + set_dt on the I state; the result
        becomes the context's dt and is returned."""
        dt = C.c_double()
        self._check(self.lib.gmc_adaptive_dt(self.h, float(cfl), float(dt_min), float(dt_max),
                                             C.byref(dt)))
        return dt.value

    def rhs_norm_noalloc(self):
        """rhs_norm through gmc_rhs_norm_noalloc (no device allocation)."""
        n = C.c_double()
        self._check(self.lib.gmc_rhs_norm_noalloc(self.h, C.byref(n)))
        return n.value

    def refine(self, n_refine=2, restart=30, max_iters=30, rel_tol=1e-8, precond=1):
        s = abi.Solver(restart, max_iters, rel_tol, precond)
        res = np.empty(n_refine + 1)
        self._check(self.lib.gmc_refine(self.h, C.byref(s), n_refine, _dp(res)))
        return res

    def vjp(self, v, mode=abi.GMC_VJP_EXACT, want_prev=False, want_params=None):
        """(g_state, g_prev, g_params) of v^T dR; the frozen mode returns
        g_state = J^T v only."""
        v = self._vec(v, "v")
        if want_params is None:
            want_params = mode == abi.GMC_VJP_EXACT
        gs = np.empty_like(v)
        gp = np.empty_like(v) if want_prev else None
        gpar = np.empty(6) if want_params else None
        self._check(self.lib.gmc_vjp(self.h, _dp(v), _dp(gs), _dp(gp), _dp(gpar), mode))
        return gs, gp, gpar


def hydrostatic_profile(params, p_top, y_top, h, n_steps, T, T_half, Y, Y_half):
    """(y, p) tables of gmc_hydrostatic_profile (pressure_profile.hpp:487-501).
    T: n_steps + 1 values at y_top - h i, T_half: n_steps values half a step
    lower; Y, Y_half: the first nc - 1 mass fractions at the same points."""
    lib = load()
    nc1 = params.nc - 1
    T = np.ascontiguousarray(T, np.float64).reshape(-1)
    T_half = np.ascontiguousarray(T_half, np.float64).reshape(-1)
    Y = np.ascontiguousarray(Y, np.float64).reshape(-1)
    Y_half = np.ascontiguousarray(Y_half, np.float64).reshape(-1)
    if T.size != n_steps + 1 or T_half.size != n_steps or Y.size != (n_steps + 1) * nc1 \
            or Y_half.size != n_steps * nc1:
        raise ValueError("T, T_half, Y, Y_half sizes do not match n_steps and nc")
    y = np.empty(n_steps + 1)
    p = np.empty(n_steps + 1)
    st = lib.gmc_hydrostatic_profile(C.byref(params), float(p_top), float(y_top), float(h), int(n_steps),
                                     _dp(T), _dp(T_half), _dp(Y), _dp(Y_half), _dp(y), _dp(p))
    if st != abi.GMC_OK:
        raise GmcError(st, "gmc_hydrostatic_profile: bad arguments")
    return y, p
