"""The adjoint's per-direction dual-number evaluation (the code the VJP kernels
run, gmc_dual.hpp + gmc_element.hpp compiled for the CPU) against the FP64
oracle: central finite differences of the oracle's element residual, and the
reference's own element matrix where it is the exact Jacobian.  CPU only."""
import ctypes as C

import numpy as np
import pytest

import orc
from helpers import rel, residual_nonsmooth

from gales_mc import abi, cases, material

EPS = 2.220446049250313e-16


def _d(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _dir(shim, P, xe, Ie, Pe, dI=None, dP=None, k=-1):
    ne = Ie.size
    r = np.empty(ne)
    dr = np.empty(ne)
    assert shim.shim_interior_dir(C.byref(P), _d(xe), _d(Ie), _d(Pe), _d(dI), _d(dP), k, _d(r), _d(dr)) == 0
    return r, dr


def _fd_ok(dr, fd, r, h):
    """Central differences carry truncation O(h^2) and rounding ~eps |r| / h,
    amplified by the cancellation inside the residual (terms of 1e7 Pa summing
    to far less); the dual derivative must agree with them to 1e-6 relative on
    top of that rounding floor."""
    floor = 1000.0 * EPS * np.linalg.norm(r) / h
    return np.linalg.norm(dr - fd) <= 1e-6 * np.linalg.norm(fd) + floor


@pytest.mark.parametrize("cid,scale", [(0, 1), (2, 16), (4, 16)])
def test_dual_matches_finite_differences(shim, cid, scale):
    """State, previous-step and material directions on elements where the
    residual is differentiable (residual_nonsmooth: the compute_h switch and
    the |v| = 0 tau switch make the reference's residual jump there)."""
    mesh, P, I, Pst = cases.config(cid, scale=scale)
    n = P.dim + 2 + P.nc - 1
    ne = (P.dim + 1) * n
    sc = np.tile([1e4] + [1.0] * P.dim + [5.0] + [0.05] * (P.nc - 1), P.dim + 1)
    u = orc.uref_inv(P)
    th = np.array(material.theta_of(P))
    rng = np.random.default_rng(17 + cid)
    E = mesh["conn"].shape[0]
    checked = 0
    for e in rng.integers(0, E, 12):
        xe, Ie, Pe = orc.element_data(mesh, I, Pst, e)
        if residual_nonsmooth(P, xe, Ie):
            continue
        checked += 1
        h = 1e-6
        dI = sc * rng.standard_normal(ne)
        r, dr = _dir(shim, P, xe, Ie, Pe, dI=dI)
        fd = (orc.element_interior(P, xe, Ie + h * dI, Pe, u)[1]
              - orc.element_interior(P, xe, Ie - h * dI, Pe, u)[1]) / (2 * h)
        assert _fd_ok(dr, fd, r, h), ("I", e, rel(dr, fd))
        dP = sc * rng.standard_normal(ne)
        r, dr = _dir(shim, P, xe, Ie, Pe, dP=dP)
        fd = (orc.element_interior(P, xe, Ie, Pe + h * dP, u)[1]
              - orc.element_interior(P, xe, Ie, Pe - h * dP, u)[1]) / (2 * h)
        assert _fd_ok(dr, fd, r, h), ("P", e, rel(dr, fd))
        for k in range(6):
            r, dr = _dir(shim, P, xe, Ie, Pe, k=k)
            for hr in (1e-6, 1e-4):
                # h = 1e-6 relative, and 1e-4 relative where the rounding
                # floor is 100x lower and truncation still ~1e-8
                hk = hr * th[k]
                tp, tm = th.copy(), th.copy()
                tp[k] += hk
                tm[k] -= hk
                Pp, Pm = material.with_theta(P, tp), material.with_theta(P, tm)
                fd = (orc.element_interior(Pp, xe, Ie, Pe, orc.uref_inv(Pp))[1]
                      - orc.element_interior(Pm, xe, Ie, Pe, orc.uref_inv(Pm))[1]) / (2 * hk)
                assert _fd_ok(dr, fd, r, hk), ("theta", hr, k, e, rel(dr, fd))
    assert checked >= 6


@pytest.mark.parametrize("dim,comps", [(2, [cases.MELT, cases.GAS_R461]),
                                       (3, [cases.MELT, cases.CRYSTAL, cases.H2O])])
def test_dual_equals_reference_jacobian_at_uniform_state(shim, dim, comps):
    """At a uniform state with I = P and no gravity the reference's element
    matrix (frozen coefficients) is the exact Jacobian: every dropped term
    multiplies a vanishing residual or gradient.  The dual derivative along
    each unit direction must reproduce its columns (m = dr_/dI, r = -r_), and
    the boundary integral's likewise up to one D-scaled block."""
    P = cases.make_params(dim, comps, gravity=(0.0, 0.0, 0.0), dt=1e-5)
    n = dim + 2 + P.nc - 1
    nen = dim + 1
    y = np.concatenate([[1.2e7], [0.3, -0.2, 0.1][:dim], [1320.0], np.full(P.nc - 1, 0.5 / (P.nc - 1))])
    xe = (np.array([0.1, 0.2, 0.16, 0.21, 0.12, 0.27]) if dim == 2 else
          np.array([0.1, 0.2, 0.3, 0.16, 0.21, 0.31, 0.12, 0.27, 0.29, 0.11, 0.22, 0.36]))
    Ie = np.tile(y, nen)
    m, _ = orc.element_interior(P, xe, Ie, Ie)
    N = nen * n
    Jd = np.zeros((N, N))
    for j in range(N):
        e = np.zeros(N)
        e[j] = 1.0
        _, dr = _dir(shim, P, xe, Ie, Ie, dI=e)
        Jd[:, j] = -dr
    assert rel(Jd, m) < 1e-12
    sides = np.arange(dim + 1, dtype=np.int32)
    for flag in (3, 5):
        flags = np.full(dim + 1, flag, dtype=np.int32)
        mb, _ = orc.element_boundary(P, xe, Ie, sides, flags)
        Jb = np.zeros((N, N))
        for j in range(N):
            e = np.zeros(N)
            e[j] = 1.0
            r = np.empty(N)
            dr = np.empty(N)
            assert shim.shim_boundary_dir(C.byref(P), _d(xe), _d(Ie), _d(e), -1, dim + 1,
                                          sides.ctypes.data_as(C.POINTER(C.c_int)),
                                          flags.ctypes.data_as(C.POINTER(C.c_int)), _d(r), _d(dr)) == 0
            Jb[:, j] = -dr
        # The boundary matrix keeps the interior K entry +rho D_i H_i in the
        # energy row / species columns (fluid_2d.hpp:549,570), while the
        # rebuilt boundary flux carries -H.J (fluid_2d.hpp:803,808), whose
        # derivative is -rho D_i H_i: those entries differ by a D-scaled term
        # (D = 1e-10), 1e-8 relative; every other entry agrees exactly.
        ie = dim + 1
        ey = np.zeros((N, N), bool)
        for b in range(nen):
            for a in range(nen):
                ey[b * n + ie, a * n + dim + 2:a * n + n] = True
        scale = np.abs(mb).max()
        assert np.abs(Jb - mb)[~ey].max() < 1e-12 * scale, flag
        assert rel(Jb, mb) < 1e-8, flag


def test_dual_is_nonzero_only_where_the_input_enters(shim):
    """A previous-step direction changes only the time term N_b (U - U_P)/dt
    and the time part of the least-squares residual; with P = I that part is
    linear in U_P, so doubling the direction doubles the derivative exactly
    and flipping it flips the sign."""
    mesh, P, I, Pst = cases.config(0)
    xe, Ie, Pe = orc.element_data(mesh, I, Pst, 1234)
    dP = np.random.default_rng(4).standard_normal(Ie.size) * np.tile([1e4, 1, 1, 5, 0.05], 3)
    _, d1 = _dir(shim, P, xe, Ie, Pe, dP=dP)
    _, d2 = _dir(shim, P, xe, Ie, Pe, dP=2 * dP)
    _, d3 = _dir(shim, P, xe, Ie, Pe, dP=-dP)
    assert np.abs(d1).max() > 0
    assert rel(d2, 2 * d1) < 1e-15 and rel(d3, -d1) < 1e-15


def test_theta_mapping_round_trip():
    for cid in (0, 2, 4):
        _, P, _, _ = cases.config(cid, scale=16)
        th = material.theta_of(P)
        P2 = material.with_theta(P, th)
        assert np.allclose(material.theta_of(P2), th, rtol=1e-14, atol=0)
        for k in range(2):
            c, c2 = P.comp[k], P2.comp[k]
            if c.kind == abi.GMC_COMP_IDEAL_GAS:
                assert abs(c2.gamma - c.gamma) <= 1e-14 * c.gamma
                assert abs(c2.molar_mass - c.molar_mass) <= 1e-14 * c.molar_mass


@pytest.mark.parametrize("cid,scale,sharp", [(0, 1, 2.0), (0, 1, 1.0), (0, 1, 1.5), (1, 16, 2.0),
                                             (2, 16, 2.0), (2, 16, 1.5), (4, 16, 2.0)])
def test_gradient_pieces_closed_form_equal_dual(shim, cid, scale, sharp):
    """The gradient-direction adjoint (k_vjp_grad): d nu/d dI in closed form
    (dc_nu_grad) equals dc_nu differentiated by dual numbers, and d tau/d dI =
    (d tau/dh) (dh/d dv) equals tau_matrix differentiated directly, at every
    quadrature point of a sample of interior elements."""
    mesh, P, I, Pst = cases.config(cid, scale=scale)
    P.dc_sharp = sharp
    dim, n = P.dim, P.dim + 2 + P.nc - 1
    nq = dim + 1
    E = mesh["conn"].shape[0]
    nu2 = np.empty(2)
    dc, dd = np.empty(dim * n), np.empty(dim * n)
    ta, tb = np.empty(dim * dim * n * n), np.empty(dim * dim * n * n)
    seen_nu = seen_tau = 0
    for e in sorted(set(np.linspace(0, E - 1, 24).astype(int).tolist())):
        xe, Ie, Pe = orc.element_data(mesh, I, Pst, e)
        for q in range(nq):
            assert shim.shim_grad_pieces(C.byref(P), _d(xe), _d(Ie), _d(Pe), q, _d(nu2), _d(dc), _d(dd),
                                         _d(ta), _d(tb)) == 0
            assert nu2[0] == nu2[1], (e, q)
            if np.abs(dd).max() > 0:
                seen_nu += 1
                assert np.linalg.norm(dc - dd) <= 1e-11 * np.linalg.norm(dd), (e, q, rel(dc, dd))
            else:
                assert not np.any(dc), (e, q)
            if np.abs(ta).max() > 0:
                seen_tau += 1
                assert np.linalg.norm(ta - tb) <= 1e-11 * np.linalg.norm(ta), (e, q, rel(ta, tb))
            else:
                assert not np.any(tb), (e, q)
    # the sample exercises both (shock capturing and tau's length are active)
    assert seen_nu >= 8 and seen_tau >= 8, (seen_nu, seen_tau)
