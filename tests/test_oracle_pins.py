"""Pins of the oracle's per-point coefficients and boundary fluxes.  CPU only.

Every expected value here comes from something the mathematics fixes, not from
the oracle's own formulas:
  - quad::compute_h (quad.hpp:356-383, 701-739) on right triangles and right
    tetrahedra, where the metric G is diagonal and h has a closed form;
  - tau_non_diag_comp_2019 (fluid_2d.hpp:932-960): each min() branch forced
    in turn by the material, the diagonal evaluated by hand, and
    tau = DUDY^-1 diag(t) checked against an exact rational solve; the
    low-Mach correction of row 0 by hand;
  - dc_2006 (fluid_2d.hpp:1027-1065) on states that vary along x only, with
    Z = d/dx F_adv(I(x)) - S from central differences of the closed-form flux,
    for dc_sharp 2, 1 and 1.5;
  - the boundary integral (fluid_2d.hpp:610-868, RM_b :341-364) against the
    closed-form fluxes (advective, viscous stress, Fourier, species with the
    -H.J energy term), against global conservation on a mesh, and the Neumann
    wall overrides;
  - the interior diffusive flux K.grad(I) (fluid_2d.hpp:531-571, 601-602),
    whose energy row carries +H.J;
  - the 3D element on a z-invariant state with w = 0 reduces to the 2D one.
The mixture properties these use (orc.props) are pinned by hand values in
test_oracle.py.
"""
from fractions import Fraction

import numpy as np
import pytest

import orc
from helpers import rel

from gales_mc import cases

QP2 = np.array([[1 / 6, 1 / 6], [4 / 6, 1 / 6], [1 / 6, 4 / 6]])
_A, _B = 0.1381966011250105, 0.5854101966249685
QP3 = np.array([[_A, _A, _A], [_B, _A, _A], [_A, _B, _A], [_A, _A, _B]])


# ---------------------------------------------------------------------------
# closed-form physics from the primitive state (fluid_2d.hpp:372-385,
# 531-571, 581-599, 776-818)
# ---------------------------------------------------------------------------
def split(P, y):
    d = P.dim
    p, v, T = y[0], np.asarray(y[1:1 + d], float), y[1 + d]
    Yd = list(y[2 + d:])
    return p, v, T, Yd + [1.0 - sum(Yd)]


def U_of(P, y):
    p, v, T, Y = split(P, y)
    o = orc.props(P, p, T, Y)
    E = o.cv * T + 0.5 * v @ v
    return np.concatenate([[o.rho], o.rho * v, [o.rho * E], o.rho * np.array(Y[:-1])])


def F_adv(P, y):
    """F_k = [rho v_k, rho v_k v + p e_k, v_k (rho E + p), rho v_k Y_i]."""
    d = P.dim
    p, v, T, Y = split(P, y)
    o = orc.props(P, p, T, Y)
    E = o.cv * T + 0.5 * v @ v
    F = []
    for k in range(d):
        mom = o.rho * v[k] * v
        mom[k] += p
        F.append(np.concatenate([[o.rho * v[k]], mom, [v[k] * (o.rho * E + p)],
                                 o.rho * v[k] * np.array(Y[:-1])]))
    return np.array(F)


def F_dif(P, y, grad, hj_sign, q_over=None, J_over=None):
    """Viscous stress, heat and species fluxes of the state y with gradients
    grad[k] = d y / d x_k.  hj_sign = +1: interior K.grad(I); -1: boundary.
    q_over / J_over: {k: value} Neumann overrides."""
    d = P.dim
    p, v, T, Y = split(P, y)
    o = orc.props(P, p, T, Y)
    nc = P.nc
    dv = np.array([[grad[j][1 + i] for j in range(d)] for i in range(d)])   # dv_i/dx_j
    mu, lam = o.mu, -2.0 * o.mu / 3.0
    tau = mu * (dv + dv.T) + lam * np.trace(dv) * np.eye(d)
    q = np.array([-o.kappa * grad[k][1 + d] for k in range(d)])
    J = np.array([[o.rho * P.comp[i].diffusivity * grad[k][2 + d + i] for i in range(nc - 1)]
                  for k in range(d)])
    if q_over:
        for k, val in q_over.items():
            q[k] = val
    if J_over:
        for k, val in J_over.items():
            J[k, 0] = val
    H = np.array([o.e_comp[i] + p / o.rho_comp[i] for i in range(nc - 1)])
    F = []
    for k in range(d):
        F.append(np.concatenate([[0.0], tau[:, k], [tau[k] @ v - q[k] + hj_sign * H @ J[k]], J[k]]))
    return np.array(F)


def S_of(P, y):
    d = P.dim
    p, v, T, Y = split(P, y)
    o = orc.props(P, p, T, Y)
    g = np.array([P.gravity[k] for k in range(d)])
    return np.concatenate([[0.0], o.rho * g, [o.rho * g @ v], np.zeros(P.nc - 1)])


# ---------------------------------------------------------------------------
# elements and linear fields
# ---------------------------------------------------------------------------
def right_tri(hx, hy, theta=0.0, origin=(0.3, 0.1)):
    c, s = np.cos(theta), np.sin(theta)
    ex, ey = np.array([c, s]), np.array([-s, c])
    o = np.array(origin)
    return np.concatenate([o, o + hx * ex, o + hy * ey])


def right_tet(hx, hy, hz, origin=(0.3, 0.1, 0.2)):
    o = np.array(origin)
    return np.concatenate([o, o + [hx, 0, 0], o + [0, hy, 0], o + [0, 0, hz]])


def qp_x(xe, dim, q):
    X = xe.reshape(dim + 1, dim)
    xi = (QP2 if dim == 2 else QP3)[q]
    return X[0] + sum(xi[k] * (X[k + 1] - X[0]) for k in range(dim))


def linear_nodes(xe, dim, x0, y0, grad):
    """Nodal values of the linear field y(x) = y0 + sum_k grad[k] (x - x0)_k."""
    X = xe.reshape(dim + 1, dim)
    return np.concatenate([y0 + sum(grad[k] * (X[a] - x0)[k] for k in range(dim))
                           for a in range(dim + 1)])


def probe_linear(P, xe, y0, grad, q=1, Pe=None):
    d = P.dim
    x0 = qp_x(xe, d, q)
    Ie = linear_nodes(xe, d, x0, y0, grad)
    return orc.point_probe(P, xe, Ie, Ie if Pe is None else Pe, q), Ie


def exact_solve_diag(D, t):
    """X with D X = diag(t), in exact rational arithmetic."""
    n = len(t)
    X = np.zeros((n, n))
    for j in range(n):
        A = [[Fraction(float(D[i, k])) for k in range(n)] + [Fraction(float(t[j])) if i == j else Fraction(0)]
             for i in range(n)]
        for c in range(n):
            piv = next(r for r in range(c, n) if A[r][c] != 0)
            A[c], A[piv] = A[piv], A[c]
            for r in range(n):
                if r != c and A[r][c] != 0:
                    f = A[r][c] / A[c][c]
                    A[r] = [x - f * y for x, y in zip(A[r], A[c])]
        X[:, j] = [float(A[i][n] / A[i][i]) for i in range(n)]
    return X


GAS_RICH = np.array([1.0e7, 300.0, 0.0, 1300.0, 0.05])    # Mach ~ 0.34
SLOW = np.array([1.0e7, 1.0, 0.0, 1300.0, 0.5])           # Mach ~ 1.5e-3


# ---------------------------------------------------------------------------
# compute_h
# ---------------------------------------------------------------------------
def test_compute_h_closed_forms_2d():
    hx, hy = 0.03, 0.05
    G = np.diag([1 / hx ** 2, 1 / hy ** 2])
    v = np.array([2.0, 0.0])
    z = np.zeros((2, 2))
    along = np.array([[0.0, 0.0], [0.0, 0.0]])
    across = np.array([[0.0, 0.0], [3.0, 0.0]])        # acc = (0, 6) perpendicular to v
    parallel = np.array([[3.0, 0.0], [0.0, 0.0]])      # acc = (6, 0) along v
    assert orc.compute_h(2, G, v, z) == pytest.approx(2 * hx, rel=1e-14)
    assert orc.compute_h(2, G, v, along) == pytest.approx(2 * hx, rel=1e-14)
    assert orc.compute_h(2, G, v, across) == pytest.approx(2 * np.hypot(hx, hy), rel=1e-14)
    assert orc.compute_h(2, G, v, parallel) == pytest.approx(2 * hx, rel=1e-14)
    assert orc.compute_h(2, G, [0.0, -1.5], z) == pytest.approx(2 * hy, rel=1e-14)
    # v at 45 degrees, no acceleration: h = 2 / sqrt(s0' G s0)
    s0 = np.array([1.0, 1.0]) / np.sqrt(2.0)
    assert orc.compute_h(2, G, [1.0, 1.0], z) == pytest.approx(2 / np.sqrt(s0 @ G @ s0), rel=1e-14)


def test_compute_h_closed_forms_3d():
    hx, hy, hz = 0.03, 0.05, 0.02
    G = np.diag([1 / hx ** 2, 1 / hy ** 2, 1 / hz ** 2])
    v = np.array([2.0, 0.0, 0.0])
    z = np.zeros((3, 3))
    across = np.zeros((3, 3))
    across[1, 0] = 3.0                                 # acc along y: s1 = y, s2 = z
    assert orc.compute_h(3, G, v, z) == pytest.approx(2 * hx, rel=1e-14)
    assert orc.compute_h(3, G, v, across) == pytest.approx(2 * np.sqrt(hx ** 2 + hy ** 2 + hz ** 2),
                                                           rel=1e-14)
    across_z = np.zeros((3, 3))
    across_z[2, 0] = -1.0                              # acc along z: s1 = -z, s2 = y
    assert orc.compute_h(3, G, v, across_z) == pytest.approx(2 * np.sqrt(hx ** 2 + hy ** 2 + hz ** 2),
                                                             rel=1e-14)


@pytest.mark.parametrize("theta", [0.0, 0.7, -2.1])
def test_compute_h_through_the_element(theta):
    """The element path (geometry G of a rotated right triangle + compute_h):
    uniform velocity along the rotated x axis gives h = 2 hx; a velocity field
    whose acceleration is perpendicular to v gives 2 sqrt(hx^2 + hy^2)."""
    hx, hy = 0.03, 0.05
    P = cases.make_params(2, [cases.MELT, cases.GAS_R461], dt=1e-3)
    xe = right_tri(hx, hy, theta)
    ex, ey = np.array([np.cos(theta), np.sin(theta)]), np.array([-np.sin(theta), np.cos(theta)])
    y0 = SLOW.copy()
    y0[1:3] = 1.0 * ex
    zero = [np.zeros(5), np.zeros(5)]
    pr, _ = probe_linear(P, xe, y0, zero)
    if theta == 0.0:
        # axis-aligned: sum_a dN_a v_a is exactly 0, so acc = 0.  On a rotated
        # element the rounding of that sum leaves acc ~ 1e-16, which
        # quad.hpp:368 compares with an absolute epsilon: the second
        # direction then switches on, as in the reference
        assert pr["h"] == pytest.approx(2 * hx, rel=1e-13)
    # v(x) = ex + a ((x - x_q).ex) ey: at the point v = ex, acc = a ey
    a = 2.0
    grad = [np.zeros(5), np.zeros(5)]
    for k in range(2):
        grad[k][1:3] = a * ex[k] * ey
    pr, _ = probe_linear(P, xe, y0, grad)
    assert pr["h"] == pytest.approx(2 * np.hypot(hx, hy), rel=1e-13)
    assert np.trace(pr["G"]) == pytest.approx(1 / hx ** 2 + 1 / hy ** 2, rel=1e-13)


# ---------------------------------------------------------------------------
# tau_non_diag_comp_2019
# ---------------------------------------------------------------------------
def test_tau_vanishes_at_rest():
    P = cases.make_params(2, [cases.MELT, cases.GAS_R461], dt=1e-3)
    xe = right_tri(0.03, 0.05)
    y0 = SLOW.copy()
    y0[1:3] = 0.0
    grad = [np.array([1e3, 0, 0, 2.0, 0.1]), np.array([-5e2, 0, 0, 1.0, 0.0])]
    pr, _ = probe_linear(P, xe, y0, grad)
    assert pr["h"] == 0.0 and not np.any(pr["tau"])


def _hand_td(P, y, h, steady=False):
    p, v, T, Y = split(P, y)
    o = orc.props(P, p, T, Y)
    vn = np.linalg.norm(v)
    adv = h / (2 * (vn + o.c)) + h / (2 * vn)
    use = adv if steady else min(P.dt * 0.5, adv)
    cand = dict(v=h * h * o.rho / (12 * o.mu), E=h * h * o.rho * o.cv / (12 * o.kappa),
                Y=h * h / (12 * P.comp[0].diffusivity))
    td = np.array([use, min(use, cand["v"]), min(use, cand["v"]), min(use, cand["E"]), min(use, cand["Y"])])
    return td, use, adv, cand, vn / o.c


TAU_CASES = {
    # name: (melt overrides, dt, steady, branch that must be the one taken)
    "advective": (dict(), 1e-3, 0, "adv"),
    "time_step": (dict(), 1e-5, 0, "dt"),
    "viscous": (dict(mu=1e5), 1e-3, 0, "v"),
    "thermal": (dict(kappa=1e7), 1e-3, 0, "E"),
    "species": (dict(diffusivity=10.0), 1e-3, 0, "Y"),
    "steady": (dict(), 1e-5, 1, "adv"),
}


@pytest.mark.parametrize("case", list(TAU_CASES))
def test_tau_2019_branches(case):
    """Mach >= 0.1 (no correction): t = diag of tau by hand with the branch of
    every min() fixed by the material, and tau = DUDY^-1 diag(t) exactly."""
    over, dt, steady, branch = TAU_CASES[case]
    melt = dict(cases.MELT)
    melt.update(over)
    P = cases.make_params(2, [melt, cases.GAS_R461], dt=dt)
    P.steady_state = steady
    hx, hy = 0.03, 0.05
    xe = right_tri(hx, hy)
    pr, _ = probe_linear(P, xe, GAS_RICH, [np.zeros(5), np.zeros(5)])
    h = 2 * hx
    td, use, adv, cand, mach = _hand_td(P, GAS_RICH, h, steady=bool(steady))
    assert mach >= 0.1
    # the case really exercises its branch
    if branch == "adv":
        assert use == adv and all(c > use for c in cand.values())
        if steady:
            assert adv > 0.5 * dt
    elif branch == "dt":
        assert use == 0.5 * dt < adv
    else:
        assert cand[branch] < use and all(c > use for k, c in cand.items() if k != branch)
    assert pr["h"] == pytest.approx(h, rel=1e-13)
    np.testing.assert_allclose(pr["td"], td, rtol=1e-13, atol=0)
    X = exact_solve_diag(pr["DUDY"], td)
    assert rel(pr["tau"], X) < 1e-12


def test_tau_2019_low_mach_correction():
    """Mach < 0.1: rows 1.. of tau are DUDY^-1 diag(t); tau_00 is replaced by
    (1/tau_00 + 1/(rho tau_11 tr G))^-1 and the rest of row 0 is zero."""
    P = cases.make_params(2, [cases.MELT, cases.GAS_R461], dt=1e-3)
    hx, hy = 0.03, 0.05
    xe = right_tri(hx, hy)
    pr, _ = probe_linear(P, xe, SLOW, [np.zeros(5), np.zeros(5)])
    td, use, adv, cand, mach = _hand_td(P, SLOW, 2 * hx)
    assert mach < 0.1
    np.testing.assert_allclose(pr["td"], td, rtol=1e-13, atol=0)
    X = exact_solve_diag(pr["DUDY"], td)
    tau = pr["tau"]
    assert rel(tau[1:], X[1:]) < 1e-12
    rho = pr["pr"].rho
    trG = 1 / hx ** 2 + 1 / hy ** 2
    t00 = 1.0 / (1.0 / X[0, 0] + 1.0 / (rho * X[1, 1] * trG))
    assert tau[0, 0] == pytest.approx(t00, rel=1e-12)
    assert not np.any(tau[0, 1:])
    assert np.abs(X[0, 1:]).max() > 0          # the zeroing removed something


# ---------------------------------------------------------------------------
# dc_2006
# ---------------------------------------------------------------------------
def _deriv(f, y0, g, eps):
    """d/ds f(y0 + s g) at 0: central differences, Richardson-extrapolated."""
    d1 = (f(y0 + eps * g) - f(y0 - eps * g)) / (2 * eps)
    d2 = (f(y0 + 0.5 * eps * g) - f(y0 - 0.5 * eps * g)) / eps
    return (4 * d2 - d1) / 3


X_GRAD = np.array([4.0e4, 0.5, -0.3, 20.0, 0.8])          # d/dx of (p, u, v, T, Y1)


@pytest.mark.parametrize("sharp", [2.0, 1.0, 1.5])
@pytest.mark.parametrize("scale", [1.0, 2.5])
def test_dc_nu_on_a_state_varying_along_x(sharp, scale):
    P = cases.make_params(2, [cases.MELT, cases.GAS_R461], dt=1e-3)
    P.dc_sharp = sharp
    P.dc_scale_fact = scale
    hx, hy = 0.03, 0.05
    xe = right_tri(hx, hy)
    y0 = np.array([1.0e7, 0.7, 0.2, 1300.0, 0.4])
    pr, _ = probe_linear(P, xe, y0, [X_GRAD, np.zeros(5)])
    # j = grad(rho)/|grad(rho)| = +-e_x, so h = 1 / sum_a |dN_a/dx| = hx / 2
    h = hx / 2
    assert pr["h_dc"] == pytest.approx(h, rel=1e-13)
    Z = _deriv(lambda y: F_adv(P, y)[0], y0, X_GRAD, 1e-4) - S_of(P, y0)
    dU = _deriv(lambda y: U_of(P, y), y0, X_GRAD, 1e-4)
    # U_ref at (p_ref, v_ref = 0, T_ref, Y_ref = 0.5 for every component), 0 where U_ref < eps
    o = orc.props(P, P.p_ref, P.T_ref, [0.5, 0.5])
    Uref = np.array([o.rho, 0.0, 0.0, o.rho * o.cv * P.T_ref, 0.5 * o.rho])
    uinv = np.where(Uref < np.finfo(float).eps, 0.0, 1.0 / np.where(Uref == 0, 1, Uref))
    zn = np.linalg.norm(uinv * Z)
    un = np.linalg.norm(uinv * U_of(P, y0))
    use1 = np.linalg.norm(uinv * dU) ** 2
    nu = {2.0: zn * h * h / un, 1.0: zn * h / np.sqrt(use1),
          1.5: zn * h * (1 / np.sqrt(use1) + h / un) / 2}[sharp]
    assert pr["nu"] == pytest.approx(scale * nu, rel=1e-8)


def test_dc_off_when_density_is_uniform():
    """Only the velocity varies: grad(rho) = 0, no shock capturing at all."""
    P = cases.make_params(2, [cases.MELT, cases.GAS_R461], dt=1e-3)
    Poff = cases.make_params(2, [cases.MELT, cases.GAS_R461], dt=1e-3, dc=False)
    xe = right_tri(0.03, 0.05)
    y0 = np.array([1.0e7, 0.7, 0.2, 1300.0, 0.4])
    grad = [np.array([0, 0.5, -0.3, 0, 0.0]), np.array([0, 0.1, 0.2, 0, 0.0])]
    pr, Ie = probe_linear(P, xe, y0, grad)
    assert pr["nu"] == 0.0 and pr["h_dc"] == 0.0
    m1, r1 = orc.element_interior(P, xe, Ie, Ie)
    m0, r0 = orc.element_interior(Poff, xe, Ie, Ie)
    assert np.array_equal(m1, m0) and np.array_equal(r1, r0)


# ---------------------------------------------------------------------------
# diffusive fluxes
# ---------------------------------------------------------------------------
def _diffusive_params(dim, nc=2):
    comps = [cases.custom(2400.0, 1e2, 1200.0, 1.5, D=1e-3), cases.ideal_gas(8.314462 / 461.0, D=2e-3)]
    if nc == 3:
        comps.insert(1, cases.custom(2800.0, 1e4, 1000.0, 2.0, D=5e-4))
    return cases.make_params(dim, comps, dt=1e-3)


def test_interior_diffusive_flux_closed_form():
    """F_dif = K.grad(I): stress, Fourier, species flux, energy with +H.J."""
    P = _diffusive_params(2)
    xe = right_tri(0.03, 0.05, 0.4)
    y0 = np.array([1.0e7, 0.7, 0.2, 1300.0, 0.4])
    grad = [np.array([4e4, 0.5, -0.3, 20.0, 0.8]), np.array([-1e4, 0.2, 0.6, -7.0, -0.5])]
    pr, _ = probe_linear(P, xe, y0, grad)
    expect = F_dif(P, y0, grad, hj_sign=+1)
    assert rel(pr["Fdif"], expect) < 1e-12
    # the H.J term is a visible part of the energy flux
    assert np.abs(expect[:, 3] - F_dif(P, y0, grad, hj_sign=-1)[:, 3]).max() > 1e-3 * np.abs(expect[:, 3]).max()


def _side_points(xe, dim, side):
    """Physical side points, shape values and area vector JN x w_bd of the
    reference's side rule (2-point Gauss per edge, element.hpp:145-164)."""
    X = xe.reshape(dim + 1, dim)
    assert dim == 2
    a, b = [(1, 2), (2, 0), (0, 1)][side]
    t = 0.5 * (1 + np.array([-1, 1]) / np.sqrt(3.0))
    edge = X[b] - X[a]
    JNw = 0.5 * np.array([edge[1], -edge[0]])      # outward (CCW element) normal x length, weight 1/2
    pts = []
    for ti in t:
        sh = np.zeros(3)
        sh[a], sh[b] = 1 - ti, ti
        pts.append((X[a] + ti * edge, sh))
    return pts, JNw


@pytest.mark.parametrize("side", [0, 1, 2])
def test_boundary_residual_one_edge_kat(side):
    """execute_b on one edge of a natural (outlet) side: r = -sum over the two
    side points of N_b (F_adv - F_dif).n w, with the boundary energy flux
    tau.v - q - H.J and J = rho D grad(Y)."""
    P = _diffusive_params(2)
    xe = right_tri(0.03, 0.05, 0.4)
    y0 = np.array([1.0e7, 0.7, 0.2, 1300.0, 0.4])
    grad = [np.array([4e4, 0.5, -0.3, 20.0, 0.8]), np.array([-1e4, 0.2, 0.6, -7.0, -0.5])]
    x0 = qp_x(xe, 2, 0)
    Ie = linear_nodes(xe, 2, x0, y0, grad)
    m, r = orc.element_boundary(P, xe, Ie, [side], [cases.OUTLET])
    pts, JNw = _side_points(xe, 2, side)
    expect = np.zeros(15)
    for x, sh in pts:
        y = y0 + grad[0] * (x - x0)[0] + grad[1] * (x - x0)[1]
        flux = (F_adv(P, y) - F_dif(P, y, grad, hj_sign=-1)).T @ JNw
        for b in range(3):
            expect[5 * b:5 * b + 5] -= sh[b] * flux
    assert rel(r, expect) < 1e-12


def test_neumann_wall_overrides():
    """Side flag 5 prescribes q = 0 and J_1 = 0: relative to the natural side,
    the energy rows lose -(q + H.J).n and the species rows lose -J.n; a
    prescribed nonzero q replaces the Fourier flux."""
    P = _diffusive_params(2)
    xe = right_tri(0.03, 0.05, 0.4)
    y0 = np.array([1.0e7, 0.7, 0.2, 1300.0, 0.4])
    grad = [np.array([4e4, 0.5, -0.3, 20.0, 0.8]), np.array([-1e4, 0.2, 0.6, -7.0, -0.5])]
    x0 = qp_x(xe, 2, 0)
    Ie = linear_nodes(xe, 2, x0, y0, grad)
    side = 1
    _, r3 = orc.element_boundary(P, xe, Ie, [side], [cases.OUTLET])
    _, r5 = orc.element_boundary(P, xe, Ie, [side], [cases.WALL])
    pts, JNw = _side_points(xe, 2, side)
    diff = np.zeros(15)
    for x, sh in pts:
        y = y0 + grad[0] * (x - x0)[0] + grad[1] * (x - x0)[1]
        f_nat = F_dif(P, y, grad, hj_sign=-1) @ np.eye(5)
        f_wall = F_dif(P, y, grad, hj_sign=-1, q_over={0: 0.0, 1: 0.0}, J_over={0: 0.0, 1: 0.0})
        d = (f_wall - f_nat).T @ JNw            # change of F_dif . n
        for b in range(3):
            diff[5 * b:5 * b + 5] += sh[b] * d   # r = -(F_adv - F_dif).n: +change of F_dif
    assert not np.any(diff[[0, 1, 2, 5, 6, 7, 10, 11, 12]])
    assert rel(r5 - r3, diff) < 1e-10
    # a prescribed heat flux replaces -kappa dT/dx_k
    qv = (123.0, -45.0)
    P.neumann[7].q_mask = 3
    P.neumann[7].q[0], P.neumann[7].q[1] = qv
    _, r7 = orc.element_boundary(P, xe, Ie, [side], [7])
    expect = np.zeros(15)
    for x, sh in pts:
        y = y0 + grad[0] * (x - x0)[0] + grad[1] * (x - x0)[1]
        flux = (F_adv(P, y) - F_dif(P, y, grad, hj_sign=-1, q_over={0: qv[0], 1: qv[1]})).T @ JNw
        for b in range(3):
            expect[5 * b:5 * b + 5] -= sh[b] * flux
    assert rel(r7, expect) < 1e-12


def test_global_conservation_on_a_mesh():
    """Without Dirichlet rows and gravity, the assembled rhs summed over all
    nodes is minus the discrete balance law of every conserved variable:
    sum_e sum_q JxW (U(I) - U(P))/dt + sum over boundary sides of the closed-
    form normal flux (F_adv - F_dif).n at the side points.  The interior flux,
    least-squares and shock-capturing terms cancel (sum_b N_b = 1)."""
    mesh = cases.channel2d(6, 3, jitter=0.2, seed=9)
    P = _diffusive_params(2)
    for f in range(16):
        P.dirichlet_mask[f] = 0
        P.gravity[1] = 0.0
    I, Pst = cases.fields(mesh, 2, 31)
    A = orc.Assembler(P, mesh)
    _, rhs = A.assemble(I, Pst)
    total = rhs.reshape(-1, 5).sum(0)
    vol = np.zeros(5)
    mag = np.zeros(5)
    for e, en in enumerate(mesh["conn"]):
        xe = mesh["coords"][en].reshape(-1)
        X = xe.reshape(3, 2)
        e1, e2 = X[1] - X[0], X[2] - X[0]
        J = e1[0] * e2[1] - e1[1] * e2[0]
        for xi in QP2:
            sh = np.array([1 - xi.sum(), xi[0], xi[1]])
            t = (U_of(P, sh @ I[en]) - U_of(P, sh @ Pst[en])) * (J / 6) / P.dt
            vol += t
            mag += np.abs(t)
    bnd = np.zeros(5)
    for e, s, f in zip(mesh["side_elem"], mesh["side_local"], mesh["side_flag"]):
        en = mesh["conn"][e]
        xe = mesh["coords"][en].reshape(-1)
        X = xe.reshape(3, 2)
        Jm = np.array([X[1] - X[0], X[2] - X[0]]).T            # dx/dxi
        dsh = np.linalg.solve(Jm.T, np.array([[-1.0, 1.0, 0.0], [-1.0, 0.0, 1.0]]))
        grad = [dsh[k] @ I[en] for k in range(2)]
        pts, JNw = _side_points(xe, 2, s)
        wall = f == cases.WALL
        for x, sh in pts:
            y = sh @ I[en]
            fl = (F_adv(P, y) - F_dif(P, y, grad, hj_sign=-1,
                                      q_over={0: 0.0, 1: 0.0} if wall else None,
                                      J_over={0: 0.0, 1: 0.0} if wall else None)).T @ JNw
            bnd += fl
            mag += np.abs(fl)
    np.testing.assert_array_less(np.abs(total + vol + bnd), 1e-11 * mag)
    assert np.all(np.abs(vol + bnd) > 1e-6 * mag)        # a nonzero balance is checked


# ---------------------------------------------------------------------------
# the 3D element reduces to the 2D one
# ---------------------------------------------------------------------------
def test_3d_point_reduces_to_2d():
    """A z-invariant state with w = 0 at a point of a right tetrahedron gives
    the 2D point's DUDY, A_1, A_2, K_kl (k, l < 2), U, fluxes and DC viscosity;
    compute_h gains exactly the binormal term 4 hz^2 when the acceleration is
    off the streamline, and nothing when it is zero."""
    comps = [cases.custom(2400.0, 1e2, 1200.0, 1.5, D=1e-3), cases.ideal_gas(8.314462 / 461.0, D=2e-3)]
    P2 = cases.make_params(2, comps, dt=1e-3, gravity=(0.0, -9.81, 0.0))
    P3 = cases.make_params(3, comps, dt=1e-3, gravity=(0.0, -9.81, 0.0))
    hx, hy, hz = 0.03, 0.05, 0.02
    xe2, xe3 = right_tri(hx, hy), right_tet(hx, hy, hz)
    r3 = [0, 1, 2, 4, 5]                                   # (p, u, v, T, Y1) in the 3D dofs
    y2 = np.array([1.0e7, 0.7, 0.2, 1300.0, 0.4])
    y3 = np.insert(y2, 3, 0.0)
    for grad2, same_h in (([np.array([4e4, 0.0, 0.0, 20.0, 0.8]), np.array([-1e4, 0.0, 0.0, -7.0, -0.5])], True),
                          ([np.array([4e4, 0.5, -0.3, 20.0, 0.8]), np.array([-1e4, 0.2, 0.6, -7.0, -0.5])], False)):
        grad3 = [np.insert(g, 3, 0.0) for g in grad2] + [np.zeros(6)]
        p2, _ = probe_linear(P2, xe2, y2, grad2, q=1)
        p3, _ = probe_linear(P3, xe3, y3, grad3, q=1)
        ix = np.ix_(r3, r3)
        assert rel(p3["DUDY"][ix], p2["DUDY"]) < 1e-14
        for k in range(2):
            assert rel(p3["A"][k][ix], p2["A"][k]) < 1e-14
            for l in range(2):
                assert rel(p3["K"][k][l][ix], p2["K"][k][l]) < 1e-14
        assert not np.any(p3["DUDY"][3, r3]) and not np.any(p3["DUDY"][r3, 3] * (np.arange(5) != 1))
        for key in ("U", "S"):
            assert rel(p3[key][r3], p2[key]) < 1e-14 and p3[key][3] == 0.0
        assert rel(p3["Fadv"][:2][:, r3], p2["Fadv"]) < 1e-14
        assert rel(p3["Fdif"][:2][:, r3], p2["Fdif"]) < 1e-13
        assert p3["h_dc"] == pytest.approx(p2["h_dc"], rel=1e-13)
        assert p3["nu"] == pytest.approx(p2["nu"], rel=1e-12)
        if same_h:
            assert p3["h"] == pytest.approx(p2["h"], rel=1e-13)
        else:
            assert p3["h"] ** 2 - p2["h"] ** 2 == pytest.approx(4 * hz ** 2, rel=1e-10)
    # tau: same h, Mach >= 0.1, so no trace(G) enters: the 3D tau restricted
    # to (p, u, v, T, Y1) is the 2D tau
    y2 = GAS_RICH.copy()
    y3 = np.insert(y2, 3, 0.0)
    zero2, zero3 = [np.zeros(5)] * 2, [np.zeros(6)] * 3
    p2, _ = probe_linear(P2, xe2, y2, zero2)
    p3, _ = probe_linear(P3, xe3, y3, zero3)
    assert rel(p3["tau"][np.ix_(r3, r3)], p2["tau"]) < 1e-12
    # low Mach: the 3D correction uses trace(G) = 1/hx^2 + 1/hy^2 + 1/hz^2
    p3, _ = probe_linear(P3, xe3, np.insert(SLOW, 3, 0.0), zero3)
    X = exact_solve_diag(p3["DUDY"], p3["td"])
    trG = 1 / hx ** 2 + 1 / hy ** 2 + 1 / hz ** 2
    t00 = 1.0 / (1.0 / X[0, 0] + 1.0 / (p3["pr"].rho * X[1, 1] * trG))
    assert p3["tau"][0, 0] == pytest.approx(t00, rel=1e-12)


# ---------------------------------------------------------------------------
# the other tau forms (fluid_2d.hpp:895-929, 963-1021) and mass-flux sides
# (fluid_2d.hpp:623-758)
# ---------------------------------------------------------------------------
def _jac(f, y, scale):
    n = len(y)
    J = []
    for j in range(n):
        # 1e-5: at 1e-6 the rounding of rho E ~ 5e7 leaves 4e-8 in d(rho E)/dT
        h = 1e-5 * scale[j]
        yp, ym = y.copy(), y.copy()
        yp[j] += h
        ym[j] -= h
        J.append((f(yp) - f(ym)) / (2 * h))
    return np.array(J).T


@pytest.mark.parametrize("steady", [0, 1])
def test_tau_2007_closed_form(steady):
    """Diagonal incompressible form on a right triangle: G = diag(1/hx^2,
    1/hy^2), g = (1/hx, 1/hy), |G|_F^2 = 1/hx^4 + 1/hy^4."""
    P = cases.make_params(2, [cases.MELT, cases.GAS_R461], dt=1e-3)
    P.tau_variant = 2007
    P.steady_state = steady
    hx, hy = 0.03, 0.05
    xe = right_tri(hx, hy)
    y0 = np.array([1.0e7, 0.7, -0.4, 1300.0, 0.4])
    pr, _ = probe_linear(P, xe, y0, [np.zeros(5), np.zeros(5)])
    p, v, T, Y = split(P, y0)
    o = orc.props(P, p, T, Y)
    vGv = v[0] ** 2 / hx ** 2 + v[1] ** 2 / hy ** 2
    frob = 1 / hx ** 4 + 1 / hy ** 4
    gg = 1 / hx ** 2 + 1 / hy ** 2
    extra = 0.0 if steady else 4.0 / P.dt ** 2
    tm = 1 / np.sqrt(vGv + 9 * frob * (o.mu / o.rho) ** 2 + extra)
    te = 1 / np.sqrt(vGv + 9 * frob * (o.kappa / (o.rho * o.cv)) ** 2 + extra)
    expect = np.diag([1 / (o.rho * tm * gg), tm / o.rho, tm / o.rho, te / (o.rho * o.cv), tm])
    assert rel(pr["tau"], expect) < 1e-13
    np.testing.assert_allclose(np.diag(pr["tau"]), np.diag(expect), rtol=1e-13)


def test_tau_2014_closed_form():
    """Diagonal form from the diagonals of DUDY, A_k (central differences of
    the closed-form U and F_adv) and K_kk (lambda + 2 mu or mu on velocity,
    kappa, rho D); tau_00 replaced by (tau_00 + rho tau_11 tr G)^-1."""
    P = _diffusive_params(2)
    P.tau_variant = 2014
    hx, hy = 0.03, 0.05
    xe = right_tri(hx, hy)
    y0 = np.array([1.0e7, 0.7, -0.4, 1300.0, 0.4])
    pr, _ = probe_linear(P, xe, y0, [np.zeros(5), np.zeros(5)])
    p, v, T, Y = split(P, y0)
    o = orc.props(P, p, T, Y)
    sc = np.abs(y0) + 1e-3
    D = np.diag(_jac(lambda y: U_of(P, y), y0, sc))
    A = [np.diag(_jac(lambda y: F_adv(P, y)[k], y0, sc)) for k in range(2)]
    mu, lam = o.mu, -2 * o.mu / 3
    K = [np.array([0.0, lam + 2 * mu, mu, o.kappa, o.rho * P.comp[0].diffusivity]),
         np.array([0.0, mu, lam + 2 * mu, o.kappa, o.rho * P.comp[0].diffusivity])]
    G = np.diag([1 / hx ** 2, 1 / hy ** 2])
    t = np.sqrt(4 / P.dt ** 2 * D ** 2)
    a = np.sqrt(sum(G[k, l] * A[k] * A[l] for k in range(2) for l in range(2)))
    d = np.sqrt(sum(G[k, k] ** 2 * K[k] ** 2 for k in range(2)))
    td = 1 / (t + a + d + 1e-8)
    td[0] = 1 / (td[0] + o.rho * td[1] * np.trace(G))
    np.testing.assert_allclose(np.diag(pr["tau"]), td, rtol=1e-7)
    assert not np.any(pr["tau"] - np.diag(np.diag(pr["tau"])))


def test_tau_2001_unsteady():
    """The 2001 form never sets its alpha (fluid_2d.hpp:885), so lambda =
    1/sqrt(4/dt^2) + h/(2|v|) > dt/2 and every diagonal entry is
    min(dt/2, ...); tau = DUDY^-1 diag(t) and the low-Mach row-0 correction."""
    P = cases.make_params(2, [cases.MELT, cases.GAS_R461], dt=1e-3)
    P.tau_variant = 2001
    hx, hy = 0.03, 0.05
    xe = right_tri(hx, hy)
    pr, _ = probe_linear(P, xe, SLOW, [np.zeros(5), np.zeros(5)])
    p, v, T, Y = split(P, SLOW)
    o = orc.props(P, p, T, Y)
    h = 2 * hx
    use = 0.5 * P.dt
    td = np.array([use, min(use, o.rho * h * h / (o.mu * 12)), min(use, o.rho * h * h / (o.mu * 12)),
                   min(use, o.rho * h * h * o.cv / (o.kappa * 12)), min(use, h * h / (1e-10 * 12))])
    assert td[1] < use                     # the viscous branch is active
    np.testing.assert_allclose(pr["td"], td, rtol=1e-13)
    X = exact_solve_diag(pr["DUDY"], td)
    assert rel(pr["tau"][1:], X[1:]) < 1e-12
    t00 = 1 / (1 / X[0, 0] + 1 / (o.rho * X[1, 1] * (1 / hx ** 2 + 1 / hy ** 2)))
    assert pr["tau"][0, 0] == pytest.approx(t00, rel=1e-12) and not np.any(pr["tau"][0, 1:])


def _mflux_params(m, diffusive=True):
    P = _diffusive_params(2) if diffusive else cases.make_params(2, [cases.MELT, cases.GAS_R461], dt=1e-3)
    nb = P.neumann[9]
    nb.mflux_mask = 3
    nb.mflux[0], nb.mflux[1] = m
    return P


@pytest.mark.parametrize("side", [0, 1])
def test_mass_flux_side_kat(side):
    """A side with prescribed mass fluxes (m1, m2): F_k^adv = [m_k, m_k v +
    p e_k, m_k E + p v_k, m_k Y] (fluid_2d.hpp:672-677, 739-744) replaces the
    advective flux; the diffusive flux is untouched."""
    m = (3.5, -1.25)
    P = _mflux_params(m)
    xe = right_tri(0.03, 0.05, 0.4)
    y0 = np.array([1.0e7, 0.7, 0.2, 1300.0, 0.4])
    grad = [np.array([4e4, 0.5, -0.3, 20.0, 0.8]), np.array([-1e4, 0.2, 0.6, -7.0, -0.5])]
    x0 = qp_x(xe, 2, 0)
    Ie = linear_nodes(xe, 2, x0, y0, grad)
    _, r = orc.element_boundary(P, xe, Ie, [side], [9])
    pts, JNw = _side_points(xe, 2, side)
    expect = np.zeros(15)
    for x, sh in pts:
        y = y0 + grad[0] * (x - x0)[0] + grad[1] * (x - x0)[1]
        pp, v, T, Y = split(P, y)
        o = orc.props(P, pp, T, Y)
        E = o.cv * T + 0.5 * v @ v
        Fm = np.array([np.concatenate([[m[k]], m[k] * v + pp * np.eye(2)[k], [m[k] * E + pp * v[k]],
                                       [m[k] * Y[0]]]) for k in range(2)])
        flux = (Fm - F_dif(P, y, grad, hj_sign=-1)).T @ JNw
        for b in range(3):
            expect[5 * b:5 * b + 5] -= sh[b] * flux
    assert rel(r, expect) < 1e-12


def test_mass_flux_side_matrix_is_jacobian():
    """At a uniform state the mass-flux side's matrix is the exact Jacobian of
    its residual (A_k = dF_k/dY with m fixed).  With the cases' D = 1e-10: the
    boundary matrix carries K(E, Y) = +rho D H while the boundary flux carries
    -H.J (fluid_2d.hpp:549,570 vs :803,808), so with a large D the energy-
    species entries differ from the Jacobian by that quirk."""
    P = _mflux_params((3.5, -1.25), diffusive=False)
    xe = right_tri(0.03, 0.05, 0.4)
    Ie = np.tile(np.array([1.0e7, 0.7, 0.2, 1300.0, 0.4]), 3)
    mb, _ = orc.element_boundary(P, xe, Ie, [0, 1, 2], [9, 9, 9])
    Jb = np.zeros((15, 15))
    for j in range(15):
        h = 1e-6 * max(abs(Ie[j]), 1e-3)
        Ip, Im = Ie.copy(), Ie.copy()
        Ip[j] += h
        Im[j] -= h
        Jb[:, j] = -(orc.element_boundary(P, xe, Ip, [0, 1, 2], [9] * 3)[1]
                     - orc.element_boundary(P, xe, Im, [0, 1, 2], [9] * 3)[1]) / (2 * h)
    assert rel(mb, Jb) < 2e-6
