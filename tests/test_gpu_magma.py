"""Magma mixtures on the GPU against the oracle, through the C-ABI.

The components are the reference's shipped magmas (shoshonite and phonolite
of sim/fluid_mc/P0/props.txt with its solubility tables), evaluated per
quadrature point on the device (gmc_magma.hpp).  Bar as for the other FP64
parity tests: element blocks, assembled J and rhs within 1e-12 of the
oracle; an out-of-table pressure comes back as GMC_ERR_RANGE.
"""
import sys

import numpy as np
import pytest

import orc
from helpers import magma_case, magma_tables, phonolite, rel, shoshonite

from gales_mc import abi, cases
from gales_mc.lib import Context, GmcError

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _system_parity(mesh, P, I, Pst, ctx, tol=TOL, flags=None):
    A = orc.Assembler(P, mesh)
    vals_ref, rhs_ref = A.assemble(I, Pst)
    if flags is None:
        ctx.assemble()
    else:
        ctx.assemble(flags)
    vals, rhs = ctx.system()
    ev, er = rel(vals, vals_ref), rel(rhs, rhs_ref)
    _say(f"  J rel {ev:.3e}, rhs rel {er:.3e}")
    assert ev < tol and er < TOL
    return vals, rhs


def _comps(kind, tab):
    if kind == "p0":
        return [shoshonite(tab), phonolite(tab)]
    if kind == "crystals_sr":   # strain-rate dependent crystal viscosity
        return [shoshonite(tab, crystal_on_mu=abi.GMC_MU_CARICCHI, crystal_vf=0.3, crystal_rho=2800.0),
                phonolite(tab, crystal_on_mu=abi.GMC_MU_EINSTEIN_ROSCOE, crystal_vf=0.2,
                          crystal_rho=2700.0)]
    if kind == "bilinear":
        T = np.array([1250.0, 1300.0, 1350.0])
        p = tab["p"][::10]
        wh = np.stack([tab["shoshonite_h2o"][::10] * f for f in (1.03, 1.0, 0.96)])
        wc = np.stack([tab["shoshonite_co2"][::10] * f for f in (0.9, 1.0, 1.12)])
        return [cases.magma(cases.SHOSHONITE_OXIDES, 0.02, 0.01, p, wh, wc, T_table=T), phonolite(tab)]
    raise ValueError(kind)


@pytest.mark.parametrize("kind,scatter", [("p0", ""), ("p0", "ebuf"), ("p0", "color"),
                                          ("crystals_sr", ""), ("bilinear", "")])
def test_magma_assembly_2d(kind, scatter, monkeypatch):
    if scatter:
        monkeypatch.setenv("GMC_SCATTER", scatter)
    _say(f"magma 2D {kind} scatter={scatter or 'default'}")
    mesh = cases.channel2d(64, 32, 2.0, 1.0)
    P, I, Pst = magma_case(mesh, _comps(kind, magma_tables()))
    with Context(mesh, P) as ctx:
        ctx.set_state(I, Pst)
        u = orc.uref_inv(P)
        bd = sorted(set(mesh["side_elem"].tolist()))
        for e in (0, 17, 999, mesh["conn"].shape[0] - 1):
            xe, Ie, Pe = orc.element_data(mesh, I, Pst, e)
            m_ref, r_ref = orc.element_interior(P, xe, Ie, Pe, u)
            m, r = ctx.element(e)
            assert rel(m, m_ref) < TOL and rel(r, r_ref) < TOL, e
        for e in bd[:3] + bd[-3:]:
            xe, Ie, Pe = orc.element_data(mesh, I, Pst, e)
            sl, sf = orc.element_sides(mesh, e)
            m_ref, r_ref = orc.element_boundary(P, xe, Ie, sl, sf)
            m, r = ctx.element(e, boundary=True)
            assert rel(m, m_ref) < TOL and rel(r, r_ref) < TOL, e
        _system_parity(mesh, P, I, Pst, ctx)


def test_magma_assembly_3d():
    _say("magma 3D: shoshonite, phonolite, H2O gas on the 8^3 Kuhn box")
    mesh = cases.kuhn3d(8, 8, 8)
    tab = magma_tables()
    P, I, Pst = magma_case(mesh, [shoshonite(tab), phonolite(tab), cases.H2O])
    with Context(mesh, P) as ctx:
        ctx.set_state(I, Pst)
        _system_parity(mesh, P, I, Pst, ctx)


def test_magma_fp32_tier():
    _say("magma FP32 tier")
    mesh = cases.channel2d(64, 32, 2.0, 1.0)
    P, I, Pst = magma_case(mesh, _comps("p0", magma_tables()))
    P.precision = abi.GMC_PREC_FP32
    with Context(mesh, P) as ctx:
        ctx.set_state(I, Pst)
        _system_parity(mesh, P, I, Pst, ctx, tol=1e-6,
                       flags=abi.GMC_ASM_MATRIX | abi.GMC_ASM_RHS | abi.GMC_ASM_FP32)


def test_magma_adaptive_dt_and_adjoint_modes():
    _say("magma adaptive dt / vjp modes")
    mesh = cases.channel2d(32, 16, 2.0, 1.0)
    P, I, Pst = magma_case(mesh, _comps("p0", magma_tables()))
    with Context(mesh, P) as ctx:
        ctx.set_state(I, Pst)
        dt = ctx.adaptive_dt(0.5, 1e-9, 1.0)
        dt_ref = orc.adaptive_dt(P, mesh, I, 0.5, 1e-9, 1.0, P.dt)[0]
        assert rel(dt, dt_ref) < TOL
        v = np.random.default_rng(3).standard_normal(I.size)
        with pytest.raises(GmcError) as e:
            ctx.vjp(v, mode=abi.GMC_VJP_EXACT)
        assert e.value.code == abi.GMC_ERR_UNSUPPORTED
        g = ctx.vjp(v, mode=abi.GMC_VJP_FROZEN)[0]
        ctx.assemble()
        vals, _ = ctx.system()
        A = orc.Assembler(P, mesh)
        assert rel(g, A.dense(vals).T @ v) < 1e-12


def test_magma_pressure_outside_the_table():
    _say("magma out-of-table pressure")
    mesh = cases.channel2d(16, 8, 2.0, 1.0)
    P, I, Pst = magma_case(mesh, _comps("p0", magma_tables()))
    I[40, 0] = 5e8          # a quadrature point with weight 2/3 on it sees p > 3e8
    with Context(mesh, P) as ctx:
        ctx.set_state(I, Pst)
        with pytest.raises(GmcError) as e:
            ctx.assemble()
        assert e.value.code == abi.GMC_ERR_RANGE
        # the status word is cleared: a state back inside the table assembles
        I[40, 0] = Pst[40, 0]
        ctx.set_state(I, Pst)
        _system_parity(mesh, P, I, Pst, ctx)
