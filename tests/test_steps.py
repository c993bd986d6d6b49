"""Steps either side of the assembly: the CFL-adaptive time step
(adaptive_time_step.hpp:183-253, 380-394) and the hydrostatic pressure
initialisation (pressure_profile.hpp:26-41, 326-501).  The oracle restatements
are pinned by closed forms; the library's are compared with them (the
hydrostatic table on the CPU, the adaptive step on the GPU)."""
import sys

import numpy as np
import pytest

import orc
from helpers import rel

from gales_mc import cases, material
from gales_mc.lib import Context, hydrostatic_profile


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _one_triangle(hx, hy):
    """One right triangle with axis-aligned legs (G = diag(1/hx^2, 1/hy^2))."""
    return dict(dim=2, coords=np.array([[0.0, 0.0], [hx, 0.0], [0.0, hy]]),
                conn=np.array([[0, 1, 2]], np.int32), node_flag=np.zeros(3, np.int32),
                side_elem=np.zeros(0, np.int32), side_local=np.zeros(0, np.int32),
                side_flag=np.zeros(0, np.int32), extent=(hx, hy))


# ---------------------------------------------------------------------------
# adaptive time step
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("speed,Y1", [(2.0, 0.5), (400.0, 0.05)])
def test_adaptive_dt_closed_form(speed, Y1):
    """Uniform flow along x on a right triangle: h = 2 hx at every point, so
    dt = CFL / (2 max(2 mu/rho, kappa/(rho cv)) / h^2 + v_tau / h), with
    v_tau = |v| below Mach 0.3 and sqrt(v^2 + 1.5 c^2 + c sqrt(16 v^2 + c^2))
    above (2D)."""
    hx, hy = 0.03, 0.05
    mesh = _one_triangle(hx, hy)
    P = cases.make_params(2, [cases.MELT, cases.GAS_R461])
    y = np.array([1e7, speed, 0.0, 1300.0, Y1])
    I = np.tile(y, 3)
    o = orc.props(P, 1e7, 1300.0, [Y1, 1 - Y1])
    h = 2 * hx
    v, c = speed, o.c
    mach = v / c
    vt = v if mach <= 0.3 else np.sqrt(v * v + 1.5 * c * c + c * np.sqrt(16 * v * v + c * c))
    use = 2 * max(2 * o.mu / o.rho, o.kappa / (o.rho * o.cv)) / h ** 2 + vt / h
    cfl = 0.5
    expect = cfl / use
    dt, el = orc.adaptive_dt(P, mesh, I, cfl, 1e-12, 1.0, dt_prev=10 * expect)
    assert (mach > 0.3) == (speed > 100)
    assert el[0] == pytest.approx(expect, rel=1e-13)
    assert dt == el[0]                                  # below 1.01 dt_prev: taken as is


def test_adaptive_dt_growth_and_clamps():
    """At rest every point proposes 1.2 dt_prev; growth is capped at 1 %;
    the result is clamped to [dt_min, dt_max]."""
    mesh = cases.channel2d(4, 2)
    P = cases.make_params(2, [cases.MELT, cases.GAS_R461])
    I = np.tile([1e7, 0.0, 0.0, 1300.0, 0.5], (mesh["coords"].shape[0], 1))
    dt, el = orc.adaptive_dt(P, mesh, I, 0.5, 1e-9, 1.0, dt_prev=1e-3)
    assert np.all(el == 1.2 * 1e-3) and dt == 1.01 * 1e-3
    assert orc.adaptive_dt(P, mesh, I, 0.5, 1e-9, 5e-4, dt_prev=1e-3)[0] == 5e-4
    assert orc.adaptive_dt(P, mesh, I, 0.5, 2e-3, 1.0, dt_prev=1e-3)[0] == 2e-3


# ---------------------------------------------------------------------------
# hydrostatic pressure
# ---------------------------------------------------------------------------
def _columns(n, Tfun, Yfun, y_top, h):
    y = y_top - h * np.arange(n + 1)
    yh = y[:-1] - 0.5 * h
    return Tfun(y), Tfun(yh), Yfun(y), Yfun(yh)


def test_hydrostatic_incompressible_is_linear():
    """Pure melt (constant rho): p = p_top + 9.81 rho s, which RK4 gives
    exactly."""
    P = cases.make_params(2, [cases.MELT, cases.GAS_R461])
    n, h = 60, 10.0
    T, Th, Y, Yh = _columns(n, lambda y: np.full(y.size, 1300.0), lambda y: np.ones(y.size), -3000.0, h)
    y, p = orc.hydrostatic_profile(P, 7e7, -3000.0, h, n, T, Th, Y, Yh)
    np.testing.assert_allclose(p, 7e7 + 9.81 * 2400.0 * (-3000.0 - y), rtol=1e-14)


def test_hydrostatic_rk4_step_on_the_isothermal_gas():
    """Pure ideal gas at constant T: dp/ds = k p with k = 9.81/(R T); one RK4
    step multiplies p by 1 + kh + (kh)^2/2 + (kh)^3/6 + (kh)^4/24 exactly, and
    the table approaches p_top exp(k s)."""
    P = cases.make_params(2, [cases.MELT, cases.GAS_R461])
    n, h, Tc = 40, 2000.0, 1300.0
    T, Th, Y, Yh = _columns(n, lambda y: np.full(y.size, Tc), lambda y: np.zeros(y.size), 0.0, h)
    y, p = orc.hydrostatic_profile(P, 1e5, 0.0, h, n, T, Th, Y, Yh)
    kh = 9.81 / (461.0 * Tc) * h
    growth = 1 + kh + kh ** 2 / 2 + kh ** 3 / 6 + kh ** 4 / 24
    np.testing.assert_allclose(p[1:] / p[:-1], growth, rtol=1e-13)
    assert abs(p[-1] / (1e5 * np.exp(kh * n)) - 1) < 1e-6


def test_pressure_interpolation():
    ys = np.array([-3000.0, -3001.0, -3002.0])
    ps = np.array([7e7, 7.1e7, 7.3e7])
    assert orc.pressure_at(ys, ps, -3000.5) == pytest.approx(7.05e7, rel=1e-15)
    assert orc.pressure_at(ys, ps, -3001.75) == pytest.approx(7.25e7, rel=1e-15)
    assert orc.pressure_at(ys[::-1].copy(), ps[::-1].copy(), -3000.5) == pytest.approx(7.05e7, rel=1e-15)
    with pytest.raises(ValueError):
        orc.pressure_at(ys, ps, -3005.0)


@pytest.mark.parametrize("nc", [2, 3])
def test_library_hydrostatic_profile_matches_oracle(nc):
    """The library's table (its own property code, on the host) equals the
    oracle's on the reference case's set-up: p0 = 70 MPa at y = -3000, 1 m
    steps down to -9000, melt below -3420 (fluid_ic_bc.hpp:26-38, 71-75)."""
    comps = [cases.MELT, cases.GAS_R461] if nc == 2 else [cases.MELT, cases.CRYSTAL, cases.H2O]
    P = cases.make_params(2, comps)
    n, h = 6000, 1.0
    def Yfun(y):
        Y = np.zeros((y.size, nc - 1))
        Y[:, 0] = np.where(y <= -3420.0, 0.9, 0.3)
        if nc == 3:
            Y[:, 1] = 0.05
        return Y
    T, Th, Y, Yh = _columns(n, lambda y: 1300.0 - 0.01 * y, Yfun, -3000.0, h)
    y_or, p_or = orc.hydrostatic_profile(P, 70e6, -3000.0, h, n, T, Th, Y, Yh)
    y_lib, p_lib = hydrostatic_profile(P, 70e6, -3000.0, h, n, T, Th, Y, Yh)
    np.testing.assert_array_equal(y_lib, y_or)
    assert rel(p_lib, p_or) < 1e-14
    assert p_lib[-1] > 70e6


# ---------------------------------------------------------------------------
# the library's adaptive step (GPU)
# ---------------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("cid,scale", [(0, 1), (1, 8), (2, 16), (4, 16)])
def test_adaptive_dt_gpu(cid, scale):
    """gmc_adaptive_dt equals the oracle's criterion to 1e-12, becomes the
    context's dt (the next assembly matches the oracle's with that dt), and
    the three set_dt branches (measured, +1 %, clamped) all occur."""
    mesh, P, I, Pst = cases.config(cid, scale=scale)
    _say(f"adaptive dt c{cid}/{scale}")
    with Context(mesh, P) as ctx:
        ctx.set_state(I, Pst)
        for cfl, dt_min, dt_max in ((0.5, 1e-12, 1.0), (1e6, 1e-12, 1.0), (0.5, 1e-12, 1e-9)):
            ref, _ = orc.adaptive_dt(P, mesh, I, cfl, dt_min, dt_max, dt_prev=P.dt)
            ctx.set_dt(P.dt)
            got = ctx.adaptive_dt(cfl, dt_min, dt_max)
            _say(f"  cfl {cfl:g}: dt {got:.6e} (oracle {ref:.6e}, previous {P.dt:g})")
            assert abs(got - ref) <= 1e-12 * ref
        P2 = material.clone(P)
        ctx.set_dt(P.dt)
        dt = ctx.adaptive_dt(0.5, 1e-12, 1.0)
        P2.dt = dt
        A = orc.Assembler(P2, mesh)
        vals_ref, rhs_ref = A.assemble(I, Pst)
        ctx.assemble()
        vals, rhs = ctx.system()
        assert rel(vals, vals_ref) < 1e-12 and rel(rhs, rhs_ref) < 1e-12
