"""ctypes wrapper of the FP64 CPU oracle (oracle/liboracle_gmc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.  The product library never calls it.
"""
import ctypes as C
import os
import subprocess
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# GMC_ORACLE_LIB names another build of gmc_oracle.c (the mutation tests load
# copies with a planted mistake, to show that the pins catch it)
LIB = os.environ.get("GMC_ORACLE_LIB") or os.path.join(HERE, "liboracle_gmc.so")
_lib = None
_lock = threading.Lock()


class OrcProps(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("rho", "mu", "cp", "cv", "kappa", "c", "alpha", "beta", "D")] + \
               [("rho_comp", C.c_double * 4), ("e_comp", C.c_double * 4), ("D_comp", C.c_double * 4),
                ("st", C.c_int)]


class OrcMagmaOut(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("rho", "mu", "cp", "cv", "kappa", "c", "alpha", "beta", "D")] + \
               [("wf", C.c_double * 5), ("vf", C.c_double * 5), ("rho_ch", C.c_double * 5),
                ("vf_gas", C.c_double), ("wl_h2o", C.c_double), ("wl_co2", C.c_double),
                ("mu_melt", C.c_double)] + \
               [(k, C.c_double * 5) for k in ("alpha_ch", "beta_ch", "c_ch", "cp_ch", "cv_ch")] + \
               [("st", C.c_int)]


_MD = 8


class OrcProbe(C.Structure):
    _fields_ = [("J", C.c_double), ("JxW", C.c_double),
                ("sh", C.c_double * 4), ("dsh", (C.c_double * 4) * 3), ("G", (C.c_double * 3) * 3),
                ("I", C.c_double * _MD), ("dI", (C.c_double * _MD) * 3),
                ("pr", OrcProps),
                ("U", C.c_double * _MD), ("S", C.c_double * _MD), ("Res", C.c_double * _MD),
                ("DUDY", (C.c_double * _MD) * _MD),
                ("A", ((C.c_double * _MD) * _MD) * 3),
                ("K", (((C.c_double * _MD) * _MD) * 3) * 3),
                ("Fadv", (C.c_double * _MD) * 3), ("Fdif", (C.c_double * _MD) * 3),
                ("h", C.c_double), ("td", C.c_double * _MD), ("tau", (C.c_double * _MD) * _MD),
                ("nu", C.c_double), ("h_dc", C.c_double)]


class OrcMesh(C.Structure):
    _fields_ = [("coords", C.POINTER(C.c_double)), ("conn", C.POINTER(C.c_int)),
                ("dmask", C.POINTER(C.c_uint)), ("side_start", C.POINTER(C.c_int)),
                ("side_local", C.POINTER(C.c_int)), ("side_flag", C.POINTER(C.c_int))]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB):
                build()
            _lib = C.CDLL(LIB)
            d, i, vp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_void_p
            _lib.orc_properties.argtypes = [vp, C.c_double, C.c_double, d, C.c_double, C.POINTER(OrcProps)]
            _lib.orc_magma.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.POINTER(OrcMagmaOut)]
            _lib.orc_pal_2003.argtypes = [C.c_double, C.c_double, C.c_double, C.POINTER(C.c_int)]
            _lib.orc_bilu0.argtypes = [C.c_int, C.c_int, i, i, d, d, d]
            _lib.orc_bilu0_apply.argtypes = [C.c_int, C.c_int, i, i, d, d, d, d]
            _lib.orc_pal_2003.restype = C.c_double
            _lib.orc_inverse.argtypes = [C.c_int, d, d]
            _lib.orc_uref_inv.argtypes = [vp, d]
            _lib.orc_element_interior.argtypes = [vp, d, d, d, d, d, d]
            _lib.orc_element_boundary.argtypes = [vp, d, d, C.c_int, i, i, d, d]
            _lib.orc_point_probe.argtypes = [vp, d, d, d, d, C.c_int, C.POINTER(OrcProbe)]
            _lib.orc_compute_h.restype = C.c_double
            _lib.orc_compute_h.argtypes = [C.c_int, d, d, d]
            _lib.orc_pattern.restype = C.c_longlong
            _lib.orc_pattern.argtypes = [C.c_int, C.c_int, C.c_int, i, i, i]
            _lib.orc_assemble_range.argtypes = [vp, C.POINTER(OrcMesh), d, d, i, i, d, d,
                                                C.c_longlong, C.c_longlong]
            _lib.orc_spmv.argtypes = [C.c_int, C.c_int, i, i, d, d, d]
            _lib.orc_gmres.argtypes = [C.c_int, C.c_int, i, i, d, d, d, C.c_int, C.c_int,
                                       C.c_double, C.c_int, i, d]
            _lib.orc_gmres_ordered.argtypes = _lib.orc_gmres.argtypes
            _lib.orc_adaptive_dt.restype = C.c_double
            _lib.orc_adaptive_dt.argtypes = [vp, C.POINTER(OrcMesh), C.c_longlong, d, C.c_double,
                                             C.c_double, C.c_double, C.c_double, d]
            _lib.orc_hydrostatic_profile.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_int,
                                                     d, d, d, d, d, d]
            _lib.orc_pressure_at.argtypes = [d, d, C.c_int, C.c_double, d]
            _lib.orc_residual_check.argtypes = [d, C.c_longlong, C.c_int, d, d]
            _lib.orc_newton_update.argtypes = [C.c_longlong, C.c_int, C.c_int, d, d]
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def props(params, p, T, Y, sr=0.0):
    o = OrcProps()
    Y = np.ascontiguousarray(Y, np.float64)
    lib().orc_properties(C.byref(params), p, T, _d(Y), sr, C.byref(o))
    return o


def bilu0(rowptr, colidx, vals, n):
    """(lu, dinv) of the oracle's block ILU(0) of a BSR matrix (n x n blocks)."""
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    colidx = np.ascontiguousarray(colidx, np.int32)
    vals = np.ascontiguousarray(vals, np.float64).reshape(-1)
    nn = rowptr.size - 1
    lu = np.empty_like(vals)
    dinv = np.empty(nn * n * n)
    st = lib().orc_bilu0(nn, n, _i(rowptr), _i(colidx), _d(vals), _d(lu), _d(dinv))
    if st:
        raise ArithmeticError(f"block ILU(0) status {st}")
    return lu, dinv


def bilu0_apply(rowptr, colidx, lu, dinv, n, r):
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    colidx = np.ascontiguousarray(colidx, np.int32)
    r = np.ascontiguousarray(r, np.float64)
    x = np.empty_like(r)
    lib().orc_bilu0_apply(rowptr.size - 1, n, _i(rowptr), _i(colidx), _d(lu), _d(dinv), _d(r), _d(x))
    return x


def pal_2003(vf_g, mu, sr):
    """(mu after the Pal 2003 correction, status)."""
    st = C.c_int(0)
    v = lib().orc_pal_2003(vf_g, mu, sr, C.byref(st))
    return v, st.value


def magma(m, p, T, sr=0.0):
    """orc_magma on an abi.Magma struct (e.g. params.comp[c].magma)."""
    o = OrcMagmaOut()
    lib().orc_magma(C.byref(m), p, T, sr, C.byref(o))
    return o


def inverse(A):
    A = np.ascontiguousarray(A, np.float64)
    n = A.shape[0]
    out = np.empty_like(A)
    st = lib().orc_inverse(n, _d(A), _d(out))
    if st:
        raise ArithmeticError("LU factorization fails")
    return out


def uref_inv(params):
    u = np.zeros(8)
    lib().orc_uref_inv(C.byref(params), _d(u))
    return u


def ndof_of(params):
    return params.dim + 2 + params.nc - 1


def element_interior(params, xe, Ie, Pe, uref=None):
    n = ndof_of(params)
    ne = (params.dim + 1) * n
    xe = np.ascontiguousarray(xe, np.float64).reshape(-1)
    Ie = np.ascontiguousarray(Ie, np.float64).reshape(-1)
    Pe = np.ascontiguousarray(Pe, np.float64).reshape(-1)
    u = uref_inv(params) if uref is None else uref
    m = np.empty(ne * ne)
    r = np.empty(ne)
    st = lib().orc_element_interior(C.byref(params), _d(xe), _d(Ie), _d(Pe), _d(u), _d(m), _d(r))
    if st:
        raise ArithmeticError(f"oracle element status {st}")
    return m.reshape(ne, ne), r


def point_probe(params, xe, Ie, Pe, q, uref=None):
    """Per-point quantities of interior point q (numpy arrays trimmed to
    ndof / dim)."""
    n = ndof_of(params)
    d = params.dim
    xe = np.ascontiguousarray(xe, np.float64).reshape(-1)
    Ie = np.ascontiguousarray(Ie, np.float64).reshape(-1)
    Pe = np.ascontiguousarray(Pe, np.float64).reshape(-1)
    u = uref_inv(params) if uref is None else np.ascontiguousarray(uref, np.float64)
    o = OrcProbe()
    st = lib().orc_point_probe(C.byref(params), _d(xe), _d(Ie), _d(Pe), _d(u), q, C.byref(o))
    if st:
        raise ArithmeticError(f"oracle probe status {st}")
    a = lambda x: np.array(x)
    return dict(J=o.J, JxW=o.JxW, sh=a(o.sh)[:d + 1], dsh=a(o.dsh)[:d, :d + 1], G=a(o.G)[:d, :d],
                I=a(o.I)[:n], dI=a(o.dI)[:d, :n], pr=o.pr, U=a(o.U)[:n], S=a(o.S)[:n],
                Res=a(o.Res)[:n], DUDY=a(o.DUDY)[:n, :n], A=a(o.A)[:d, :n, :n],
                K=a(o.K)[:d, :d, :n, :n], Fadv=a(o.Fadv)[:d, :n], Fdif=a(o.Fdif)[:d, :n],
                h=o.h, td=a(o.td)[:n], tau=a(o.tau)[:n, :n], nu=o.nu, h_dc=o.h_dc)


def compute_h(dim, G, v, dv):
    """quad::compute_h on explicit G (dim x dim), v and dv[i][j] = dv_i/dx_j."""
    G3 = np.zeros((3, 3))
    G3[:dim, :dim] = G
    dv3 = np.zeros((3, 3))
    dv3[:dim, :dim] = dv
    v3 = np.zeros(3)
    v3[:dim] = v
    return lib().orc_compute_h(dim, _d(np.ascontiguousarray(G3)), _d(v3), _d(np.ascontiguousarray(dv3)))


def element_boundary(params, xe, Ie, side_local, side_flag):
    n = ndof_of(params)
    ne = (params.dim + 1) * n
    xe = np.ascontiguousarray(xe, np.float64).reshape(-1)
    Ie = np.ascontiguousarray(Ie, np.float64).reshape(-1)
    sl = np.ascontiguousarray(side_local, np.int32)
    sf = np.ascontiguousarray(side_flag, np.int32)
    m = np.empty(ne * ne)
    r = np.empty(ne)
    st = lib().orc_element_boundary(C.byref(params), _d(xe), _d(Ie), len(sl), _i(sl), _i(sf),
                                    _d(m), _d(r))
    if st:
        raise ArithmeticError(f"oracle boundary status {st}")
    return m.reshape(ne, ne), r


def element_data(mesh, I, P, e):
    """Element coordinates and I/P element vectors (node-major)."""
    en = mesh["conn"][e]
    return mesh["coords"][en].reshape(-1), I[en].reshape(-1), P[en].reshape(-1)


def element_sides(mesh, e):
    s = np.nonzero(mesh["side_elem"] == e)[0]
    return mesh["side_local"][s], mesh["side_flag"][s]


def pattern(n_nodes, conn):
    conn = np.ascontiguousarray(conn, np.int32)
    nen = conn.shape[1]
    rowptr = np.empty(n_nodes + 1, np.int32)
    nnzb = lib().orc_pattern(n_nodes, conn.shape[0], nen, _i(conn), _i(rowptr), None)
    colidx = np.empty(nnzb, np.int32)
    lib().orc_pattern(n_nodes, conn.shape[0], nen, _i(conn), _i(rowptr), _i(colidx))
    return rowptr, colidx


def dmask_of(params, mesh):
    tab = np.array([params.dirichlet_mask[f] for f in range(16)], np.uint32)
    return tab[mesh["node_flag"]]


def side_csr(mesh):
    E = mesh["conn"].shape[0]
    order = np.argsort(mesh["side_elem"], kind="stable")
    se = mesh["side_elem"][order]
    start = np.zeros(E + 1, np.int32)
    np.add.at(start, se + 1, 1)
    start = np.cumsum(start).astype(np.int32)
    return start, np.ascontiguousarray(mesh["side_local"][order], np.int32), \
        np.ascontiguousarray(mesh["side_flag"][order], np.int32)


class Assembler:
    """fluid_assembly::execute over a mesh, into the oracle's own BSR."""

    def __init__(self, params, mesh):
        self.params = params
        self.mesh = mesh
        self.n = ndof_of(params)
        self.nen = params.dim + 1
        self.coords = np.ascontiguousarray(mesh["coords"], np.float64)
        self.conn = np.ascontiguousarray(mesh["conn"], np.int32)
        self.n_nodes = self.coords.shape[0]
        self.dmask = np.ascontiguousarray(dmask_of(params, mesh), np.uint32)
        self.side_start, self.side_local, self.side_flag = side_csr(mesh)
        self.rowptr, self.colidx = pattern(self.n_nodes, self.conn)
        self.view = OrcMesh(_d(self.coords), _i(self.conn),
                            self.dmask.ctypes.data_as(C.POINTER(C.c_uint)),
                            _i(self.side_start), _i(self.side_local), _i(self.side_flag))

    def new_system(self):
        vals = np.zeros((len(self.colidx), self.n, self.n))
        rhs = np.zeros(self.n_nodes * self.n)
        return vals, rhs

    def assemble(self, I, P, e0=0, e1=None, vals=None, rhs=None):
        if vals is None:
            vals, rhs = self.new_system()
        e1 = self.conn.shape[0] if e1 is None else e1
        I = np.ascontiguousarray(I, np.float64).reshape(-1)
        P = np.ascontiguousarray(P, np.float64).reshape(-1)
        st = lib().orc_assemble_range(C.byref(self.params), C.byref(self.view), _d(I), _d(P),
                                      _i(self.rowptr), _i(self.colidx), _d(vals), _d(rhs), e0, e1)
        if st:
            raise ArithmeticError(f"oracle assembly status {st}")
        return vals, rhs

    def spmv(self, vals, x):
        x = np.ascontiguousarray(x, np.float64).reshape(-1)
        y = np.empty_like(x)
        lib().orc_spmv(self.n_nodes, self.n, _i(self.rowptr), _i(self.colidx), _d(vals), _d(x), _d(y))
        return y

    def gmres(self, vals, b, restart=30, max_iters=30, rel_tol=1e-8, precond=1, ordered=False):
        """ordered: the device solver's summation order (orc_gmres_ordered)."""
        b = np.ascontiguousarray(b, np.float64)
        vals = np.ascontiguousarray(vals, np.float64)
        x = np.empty_like(b)
        it = C.c_int()
        rr = C.c_double()
        f = lib().orc_gmres_ordered if ordered else lib().orc_gmres
        st = f(self.n_nodes, self.n, _i(self.rowptr), _i(self.colidx), _d(vals), _d(b),
               _d(x), restart, max_iters, rel_tol, precond, C.byref(it), C.byref(rr))
        if st:
            raise ValueError(f"oracle gmres status {st}")
        return x, it.value, rr.value

    def dense(self, vals):
        n = self.n
        N = self.n_nodes * n
        A = np.zeros((N, N))
        for i in range(self.n_nodes):
            for t in range(self.rowptr[i], self.rowptr[i + 1]):
                j = self.colidx[t]
                A[i * n:(i + 1) * n, j * n:(j + 1) * n] = vals[t]
        return A


def newton_update(I, delta, ndof, trim):
    I = np.ascontiguousarray(I, np.float64).reshape(-1).copy()
    d = np.ascontiguousarray(delta, np.float64).reshape(-1)
    lib().orc_newton_update(I.size // ndof, ndof, trim, _d(I), _d(d))
    return I


def residual_check(b, it, state):
    b = np.ascontiguousarray(b, np.float64)
    res = C.c_double()
    conv = lib().orc_residual_check(_d(b), b.size, it, _d(state), C.byref(res))
    return bool(conv), res.value


def adaptive_dt(params, mesh, I, cfl, dt_min, dt_max, dt_prev):
    """(new dt, element minima) of adaptive_time_criteria::f_mc + set_dt."""
    A = Assembler(params, mesh)
    I = np.ascontiguousarray(I, np.float64).reshape(-1)
    E = A.conn.shape[0]
    el = np.empty(E)
    dt = lib().orc_adaptive_dt(C.byref(params), C.byref(A.view), E, _d(I), cfl, dt_min, dt_max,
                               dt_prev, _d(el))
    return dt, el


def hydrostatic_profile(params, p_top, y_top, h, n_steps, T, T_half, Y, Y_half):
    arrs = [np.ascontiguousarray(a, np.float64).reshape(-1) for a in (T, T_half, Y, Y_half)]
    y = np.empty(n_steps + 1)
    p = np.empty(n_steps + 1)
    lib().orc_hydrostatic_profile(C.byref(params), p_top, y_top, h, n_steps, *[_d(a) for a in arrs],
                                  _d(y), _d(p))
    return y, p


def pressure_at(y_tab, p_tab, y):
    y_tab = np.ascontiguousarray(y_tab, np.float64)
    p_tab = np.ascontiguousarray(p_tab, np.float64)
    out = C.c_double()
    if lib().orc_pressure_at(_d(y_tab), _d(p_tab), len(y_tab), y, C.byref(out)):
        raise ValueError("y outside the pressure table")
    return out.value
