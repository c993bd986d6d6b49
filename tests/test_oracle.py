"""Pins of the CPU oracle (oracle/gmc_oracle.c).  CPU only.

The reference ships no tests or golden vectors (its only "tests" are the
production cases, which need a 192-rank MPI run and a missing mesh blob), and
it cannot be built here (Boost uBLAS / Trilinos absent).  The oracle is pinned
instead by facts the mathematics fixes:
  - hand-derived property values and pure-component limits;
  - DUDY = dU/dY and A_k = dF_k/dY by central differences of the closed-form
    U and F (on the independent NumPy restatement, which agrees with the C
    oracle element to rounding);
  - at a uniform state with I = P and g = 0 the element matrix is the exact
    Jacobian of the element residual (the least-squares residual and the
    gradients vanish, so the frozen coefficients carry no error), interior
    and boundary, checked by central differences;
  - patch tests (uniform flow has zero interior-node residual);
  - LU inverse against LAPACK; GMRES against a dense direct solve;
  - BSR pattern and scatter (Dirichlet avg_val rule) against brute force.
"""
import numpy as np
import pytest

import element_np
import orc
from helpers import h_is_fragile, rel

from gales_mc import abi, cases


@pytest.fixture(scope="module")
def c0():
    return cases.config(0)


# ---------------------------------------------------------------------------
# properties
# ---------------------------------------------------------------------------
def test_properties_kat():
    P = cases.make_params(2, [cases.MELT, cases.GAS_R461])
    o = orc.props(P, 1e7, 1300.0, [0.5, 0.5])
    # hand values (p = 1e7 Pa, T = 1300 K, Y = 0.5/0.5, melt rho 2400 beta 0, gas R 461)
    assert o.rho_comp[1] == pytest.approx(16.6861, rel=1e-5)
    assert o.rho == pytest.approx(33.1418, rel=1e-5)
    assert o.c == pytest.approx(652.199, rel=1e-5)
    assert o.beta == pytest.approx(9.93095e-8, rel=1e-5)
    assert o.alpha == pytest.approx(7.63920e-4, rel=1e-5)
    assert o.cv == pytest.approx(0.5 * 1200 + 0.5 * 1152.5, rel=1e-12)
    assert o.cp == pytest.approx(0.5 * 1200 + 0.5 * 1613.5, rel=1e-12)
    assert o.mu == pytest.approx(0.5 * 1e2 + 0.5 * 1e-5, rel=1e-15)
    assert o.kappa == pytest.approx(0.5 * 1.5 + 0.5 * 1e-5 * 1613.5 / 0.72, rel=1e-12)
    assert o.e_comp[0] == pytest.approx(1200 * 1300.0) and o.e_comp[1] == pytest.approx(1152.5 * 1300.0)


def test_properties_limits():
    P = cases.make_params(2, [cases.MELT, cases.GAS_R461])
    p, T = 2e7, 1250.0
    liq = orc.props(P, p, T, [1.0, 0.0])
    assert liq.rho == pytest.approx(2400.0, rel=1e-15)
    assert liq.alpha == 0.0 and liq.beta == 0.0
    assert liq.c == pytest.approx(1e20, rel=1e-12)
    gas = orc.props(P, p, T, [0.0, 1.0])
    R = 461.0
    assert gas.rho == pytest.approx(p / (R * T), rel=1e-12)
    assert gas.c == pytest.approx(np.sqrt(1.4 * R * T), rel=1e-12)
    assert gas.beta == pytest.approx(1.0 / p, rel=1e-14)
    assert gas.alpha == pytest.approx(1.0 / T, rel=1e-14)


def test_properties_mixture_rules():
    """Volume-additive density, Wood sound speed, mass-weighted closures."""
    P = cases.make_params(3, [cases.MELT, cases.CRYSTAL, cases.H2O])
    Y = np.array([0.3, 0.45, 0.25])
    p, T = 1.5e7, 1400.0
    o = orc.props(P, p, T, Y)
    Rg = 8.314462 / 0.01802
    rc = np.array([2400.0, 2800.0, p / (Rg * T)])
    cc = np.array([1e20, 1e20, np.sqrt(1.4 * Rg * T)])
    rho = 1.0 / np.sum(Y / rc)
    assert o.rho == pytest.approx(rho, rel=1e-14)
    # Wood: 1/(rho c^2) = sum phi_c / (rho_c c_c^2), phi_c = Y_c rho / rho_c
    phi = Y * rho / rc
    c = np.sqrt(1.0 / (rho * np.sum(phi / (rc * cc ** 2))))
    assert o.c == pytest.approx(c, rel=1e-12)
    assert o.mu == pytest.approx(0.3 * 1e2 + 0.45 * 1e4 + 0.25 * 1e-5, rel=1e-14)
    # for these EOS beta p - alpha T vanishes (de/dp = 0 of the reference formula)
    assert abs(o.beta * p - o.alpha * T) <= 1e-14 * o.beta * p


# ---------------------------------------------------------------------------
# flux Jacobians by finite differences (NumPy restatement == C oracle)
# ---------------------------------------------------------------------------
def _flux_point(P, y):
    d = P.dim
    n = d + 2 + P.nc - 1
    grad = [np.zeros(n) for _ in range(d)]
    return element_np.flux_mats(P, y, grad)


@pytest.mark.parametrize("dim,comps", [(2, [cases.MELT, cases.GAS_R461]),
                                       (3, [cases.MELT, cases.CRYSTAL, cases.H2O])])
def test_dudy_and_a_are_jacobians(dim, comps):
    P = cases.make_params(dim, comps)
    n = dim + 2 + P.nc - 1
    rng = np.random.default_rng(5)
    y = np.concatenate([[1e7 + 3e4], rng.uniform(-1, 1, dim), [1310.0],
                        rng.dirichlet(np.ones(P.nc))[:-1]])
    f = _flux_point(P, y)
    scale = np.abs(y) + 1e-3
    J_U = np.zeros((n, n))
    J_F = [np.zeros((n, n)) for _ in range(dim)]
    for j in range(n):
        h = 1e-6 * scale[j]
        yp, ym = y.copy(), y.copy()
        yp[j] += h
        ym[j] -= h
        fp, fm = _flux_point(P, yp), _flux_point(P, ym)
        J_U[:, j] = (fp["U"] - fm["U"]) / (2 * h)
        for k in range(dim):
            J_F[k][:, j] = (fp["Fadv"][k] - fm["Fadv"][k]) / (2 * h)
    assert rel(f["DUDY"], J_U) < 1e-7
    for k in range(dim):
        assert rel(f["A"][k], J_F[k]) < 1e-7


@pytest.mark.parametrize("cid,scale", [(0, 1), (1, 16), (2, 16), (4, 16)])
def test_numpy_restatement_matches_oracle(cid, scale):
    mesh, P, I, Pst = cases.config(cid, scale=scale)
    u = orc.uref_inv(P)
    n = P.dim + 2 + P.nc - 1
    worst = 0.0
    E = mesh["conn"].shape[0]
    checked = 0
    for e in [0, 3, E // 3, E // 2, E - 1, E - 2, 7, 11]:
        xe, Ie, Pe = orc.element_data(mesh, I, Pst, e)
        if h_is_fragile(P, xe, Ie):
            continue
        checked += 1
        m1, r1 = orc.element_interior(P, xe, Ie, Pe, u)
        m2, r2 = element_np.element_interior(P, xe, Ie, Pe, u[:n])
        worst = max(worst, rel(m1, m2), rel(r1, r2))
    assert checked >= 4
    assert worst < 1e-12


def test_dc_and_tau_are_active(c0):
    """The KAT elements exercise every term: turning DC off or zeroing the
    velocity (tau = 0) changes the element matrix at O(1e-6) or more."""
    mesh, P, I, Pst = c0
    xe, Ie, Pe = orc.element_data(mesh, I, Pst, 1234)
    m, r = orc.element_interior(P, xe, Ie, Pe)
    P2 = cases.make_params(2, [cases.MELT, cases.GAS_R461], dc=False)
    m2, r2 = orc.element_interior(P2, xe, Ie, Pe)
    assert rel(m2, m) > 1e-8 and rel(r2, r) > 1e-8
    Ie3 = Ie.reshape(3, 5).copy()
    Ie3[:, 1:3] = 0.0
    m3, _ = orc.element_interior(P2, xe, Ie3.ravel(), Pe)
    assert rel(m3, m2) > 1e-6


# ---------------------------------------------------------------------------
# element matrix = exact Jacobian at a uniform state (I = P, g = 0)
# ---------------------------------------------------------------------------
def _uniform_setup(dim, comps, v):
    # dt small enough that every diagonal tau entry is min(dt/2, ...) = dt/2:
    # quad::compute_h (quad.hpp:366-375) switches its second direction on when
    # a velocity perturbation creates a convective acceleration, so an
    # h-dependent tau would put a kink inside the finite-difference stencil.
    P = cases.make_params(dim, comps, gravity=(0.0, 0.0, 0.0), dt=1e-5)
    n = dim + 2 + P.nc - 1
    nen = dim + 1
    y = np.concatenate([[1.2e7], v, [1320.0], np.full(P.nc - 1, 0.5 / (P.nc - 1))])
    if dim == 2:
        xe = np.array([0.1, 0.2, 0.16, 0.21, 0.12, 0.27])
    else:
        xe = np.array([0.1, 0.2, 0.3, 0.16, 0.21, 0.31, 0.12, 0.27, 0.29, 0.11, 0.22, 0.36])
    Ie = np.tile(y, nen)
    return P, n, nen, xe, Ie


@pytest.mark.parametrize("dim,comps", [(2, [cases.MELT, cases.GAS_R461]),
                                       (3, [cases.MELT, cases.CRYSTAL, cases.H2O])])
def test_element_matrix_is_jacobian_at_uniform_state(dim, comps):
    v = [0.3, -0.2, 0.1][:dim]
    P, n, nen, xe, Ie = _uniform_setup(dim, comps, v)
    m, r = orc.element_interior(P, xe, Ie, Ie)
    N = nen * n
    Jfd = np.zeros((N, N))
    for j in range(N):
        h = 1e-6 * max(abs(Ie[j]), 1e-3)
        Ip, Im = Ie.copy(), Ie.copy()
        Ip[j] += h
        Im[j] -= h
        _, rp = orc.element_interior(P, xe, Ip, Ie)
        _, rm = orc.element_interior(P, xe, Im, Ie)
        Jfd[:, j] = -(rp - rm) / (2 * h)
    assert rel(m, Jfd) < 2e-6
    # boundary integral on every side, with and without Neumann overrides
    for flag in (3, 5):
        sides = list(range(dim + 1))
        mb, rb = orc.element_boundary(P, xe, Ie, sides, [flag] * len(sides))
        Jb = np.zeros((N, N))
        for j in range(N):
            h = 1e-6 * max(abs(Ie[j]), 1e-3)
            Ip, Im = Ie.copy(), Ie.copy()
            Ip[j] += h
            Im[j] -= h
            _, rp = orc.element_boundary(P, xe, Ip, sides, [flag] * len(sides))
            _, rm = orc.element_boundary(P, xe, Im, sides, [flag] * len(sides))
            Jb[:, j] = -(rp - rm) / (2 * h)
        assert rel(mb, Jb) < 2e-6, flag


def test_gravity_gap_is_the_missing_source_derivative():
    """With g != 0 the reference matrix omits dS/dY (S0 only carries the
    energy-velocity entries, fluid_2d.hpp:573-575): at a state at rest
    (tau = 0) without shock capturing, the p/T/Y columns of the FD Jacobian
    differ from m by exactly -N_b (g_k drho/dy) N_a JxW in the momentum rows."""
    P, n, nen, xe, Ie = _uniform_setup(2, [cases.MELT, cases.GAS_R461], [0.0, 0.0])
    # dt = 1 keeps the time term small so FD round-off stays far below the gap
    Pg = cases.make_params(2, [cases.MELT, cases.GAS_R461], gravity=(0.0, -9.81, 0.0), dt=1.0,
                           dc=False)
    m, r = orc.element_interior(Pg, xe, Ie, Ie)
    cols = [a * n + c for a in range(nen) for c in (0, 3, 4)]      # p, T, Y1 of every node
    gap = np.zeros((nen * n, nen * n))
    for j in cols:
        # the residual carries a ~1e5 pressure term, so a larger step keeps
        # round-off below the gap (the residual is smooth in p, T, Y here)
        h = 1e-5 * max(abs(Ie[j]), 1e-3)
        Ip, Im = Ie.copy(), Ie.copy()
        Ip[j] += h
        Im[j] -= h
        fd = -(orc.element_interior(Pg, xe, Ip, Ie)[1] - orc.element_interior(Pg, xe, Im, Ie)[1]) / (2 * h)
        gap[:, j] = fd - m[:, j]
    f = element_np.flux_mats(Pg, Ie[:n], [np.zeros(n), np.zeros(n)])
    drho = f["DUDY"][0]
    J, _, _ = element_np.geometry(xe, 2)
    mass = J / 2.0 / 12.0 * (np.ones((3, 3)) + np.eye(3))       # sum_q JxW N_b N_a (exact rule)
    expect = np.zeros_like(gap)
    g = np.array([0.0, -9.81])
    for b in range(nen):
        for a in range(nen):
            for k in range(2):
                for c in (0, 3, 4):
                    expect[b * n + 1 + k, a * n + c] = -mass[b, a] * g[k] * drho[c]
    assert np.abs(expect).max() > 0
    assert np.abs(gap - expect).max() <= 1e-4 * np.abs(expect).max()


# ---------------------------------------------------------------------------
# patch test and conservation
# ---------------------------------------------------------------------------
def test_patch_uniform_flow_has_zero_interior_residual():
    mesh = cases.channel2d(8, 4, jitter=0.2, seed=3)
    P = cases.make_params(2, [cases.MELT, cases.GAS_R461], gravity=(0.0, 0.0, 0.0))
    y = np.array([1.1e7, 0.4, -0.15, 1300.0, 0.4])
    I = np.tile(y, (mesh["coords"].shape[0], 1))
    A = orc.Assembler(P, mesh)
    vals, rhs = A.new_system()
    A.assemble(I, I, vals=vals, rhs=rhs)
    R = rhs.reshape(-1, 5)
    interior = mesh["node_flag"] == 0
    # compare with the size of a single element's flux term
    scale = np.abs(orc.element_interior(P, *orc.element_data(mesh, I, I, 0))[1]).max()
    assert np.abs(R[interior]).max() < 1e-12 * scale


# ---------------------------------------------------------------------------
# linear algebra
# ---------------------------------------------------------------------------
def test_inverse_matches_lapack():
    rng = np.random.default_rng(0)
    for n in (5, 7, 8):
        A = rng.normal(size=(n, n)) + n * np.eye(n)
        assert rel(orc.inverse(A), np.linalg.inv(A)) < 1e-14
    f = _flux_point(cases.make_params(2, [cases.MELT, cases.GAS_R461]),
                    np.array([1e7, 0.3, 0.1, 1300.0, 0.5]))
    D = f["DUDY"]
    # DUDY mixes Pa with mass fractions (cond ~ 1e13): backward-stable bound
    bound = 100 * np.finfo(float).eps * np.linalg.cond(D)
    assert rel(orc.inverse(D) @ D, np.eye(5)) < bound
    assert rel(orc.inverse(D), np.linalg.inv(D)) < bound
    with pytest.raises(ArithmeticError):
        orc.inverse(np.zeros((5, 5)))


def test_pattern_and_scatter_brute_force():
    mesh = cases.channel2d(3, 2, seed=1)
    P = cases.make_params(2, [cases.MELT, cases.GAS_R461])
    I, Pst = cases.fields(mesh, 2, 11)
    A = orc.Assembler(P, mesh)
    nn = mesh["coords"].shape[0]
    pairs = set()
    for en in mesh["conn"]:
        for a in en:
            for b in en:
                pairs.add((int(a), int(b)))
    got = {(i, int(j)) for i in range(nn) for j in A.colidx[A.rowptr[i]:A.rowptr[i + 1]]}
    assert got == pairs
    vals, rhs = A.assemble(I, Pst)
    Dn = A.dense(vals)
    n = 5
    Dref = np.zeros_like(Dn)
    rref = np.zeros(nn * n)
    dmask = orc.dmask_of(P, mesh)
    for e, en in enumerate(mesh["conn"]):
        xe, Ie, Pe = orc.element_data(mesh, I, Pst, e)
        parts = [orc.element_interior(P, xe, Ie, Pe)]
        sl, sf = orc.element_sides(mesh, e)
        if len(sl):
            parts.append(orc.element_boundary(P, xe, Ie, sl, sf))
        for m, r in parts:
            m = m.copy()
            r = r.copy()
            avg = (1.0 + np.trace(m)) / 15.0
            idx = np.concatenate([np.arange(a * n, a * n + n) for a in en])
            for i in range(15):
                if dmask[en[i // n]] >> (i % n) & 1:
                    m[i] = 0.0
                    m[i, i] = avg
                    r[i] = 0.0
            Dref[np.ix_(idx, idx)] += m
            rref[idx] += r
    assert rel(Dn, Dref) < 1e-15 and rel(rhs, rref) < 1e-15


@pytest.mark.parametrize("precond,ordered", [(1, False), (2, False), (1, True)])
def test_gmres_against_direct_solve(precond, ordered):
    """GMRES(k) with point (1) or block (2) Jacobi right preconditioning
    converges to the dense LAPACK solution, in the plain and in the device's
    summation order."""
    mesh = cases.channel2d(6, 3, seed=2)
    P = cases.make_params(2, [cases.MELT, cases.GAS_R461])
    I, Pst = cases.fields(mesh, 2, 12)
    A = orc.Assembler(P, mesh)
    vals, rhs = A.assemble(I, Pst)
    D = A.dense(vals)
    x_ref = np.linalg.solve(D, rhs)
    x, it, rr = A.gmres(vals, rhs, restart=40, max_iters=4000, rel_tol=1e-13, precond=precond,
                        ordered=ordered)
    assert rr < 1e-12
    assert rel(D @ x, rhs) < 1e-12
    assert rel(x, x_ref) < 1e-6


def test_newton_update_and_trim():
    I = np.array([[1.0, 0.1, 0.2, 1300.0, 0.95], [2.0, 0.0, 0.0, 1301.0, 0.02]])
    d = np.array([[0.5, 0.1, 0.0, 1.0, 0.1], [0.0, 0.0, 0.0, 0.0, -0.05]])
    out = orc.newton_update(I, d, 5, 4).reshape(2, 5)
    assert np.array_equal(out[:, :4], (I + d)[:, :4])
    assert out[0, 4] == 1.0 and out[1, 4] == 0.0


def test_residual_check_semantics():
    st = np.zeros(1)
    conv, res = orc.residual_check(np.full(4, 3.0), 0, st)
    assert not conv and res == 6.0 and st[0] == 6.0
    conv, res = orc.residual_check(np.full(4, 3e-9), 1, st)
    assert conv                                      # 6e-9 <= max(6e-8, 1e-5)
    conv, res = orc.residual_check(np.full(4, 1e-4), 1, st)
    assert not conv and res == pytest.approx(2e-4 / 6.0)


# ---------------------------------------------------------------------------
# meshes
# ---------------------------------------------------------------------------
def test_channel_mesh_counts_and_normals():
    nx, ny = 16, 8
    mesh = cases.channel2d(nx, ny, jitter=0.2, seed=4)
    N = (nx + 1) * (ny + 1)
    edges = nx * (ny + 1) + (nx + 1) * ny + nx * ny
    A = orc.Assembler(cases.make_params(2, [cases.MELT, cases.GAS_R461]), mesh)
    assert len(A.colidx) == N + 2 * edges
    x = mesh["coords"]
    area = 0.0
    for en in mesh["conn"]:
        a, b, c = x[en]
        J = (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])
        assert J > 0
        area += J / 2
    assert area == pytest.approx(2.0, rel=1e-13)
    assert len(mesh["side_elem"]) == 2 * nx + 2 * ny


def test_kuhn_mesh_volume_and_faces():
    n = 4
    mesh = cases.kuhn3d(n, n, n, 1.0, 2.0, 1.0)
    x = mesh["coords"]
    vol = 0.0
    for en in mesh["conn"]:
        J = np.linalg.det(np.array([x[en[1]] - x[en[0]], x[en[2]] - x[en[0]], x[en[3]] - x[en[0]]]).T)
        assert J > 0
        vol += J / 6
    assert vol == pytest.approx(2.0, rel=1e-13)
    assert len(mesh["side_elem"]) == 2 * 6 * n * n
