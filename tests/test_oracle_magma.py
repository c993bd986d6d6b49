"""Pins of the oracle's magma submixture (mixtures.hpp:115-509).  CPU only.

Expected values come from what the mathematics and the published models fix,
not from the oracle's formulas:
  - table lookup (mixtures.hpp:342-412): exact at the nodes and on functions
    the interpolation reproduces exactly (linear in p; bilinear in T, p,
    including the mirrored |p - p_0| cell below the table); the reference's
    range errors, including the last node;
  - the phase split (mixtures.hpp:208-224): mass balances of H2O and CO2 and
    the dissolved fractions of the liquid;
  - every chemical's alpha and beta as the logarithmic T and p derivatives of
    its molar volume (central differences of 1/rho), the isothermal sound
    speed 1/sqrt(rho beta), cp - cv = T alpha^2 / (rho beta) and the ideal gas
    law;
  - the melt's Lange-Carmichael volume (chemicals.hpp:63-129) for one- and
    two-oxide melts from molar sums written here;
  - the mixture rules: volume additivity, the crystal density blend, Wood's
    sound speed in volume fractions, mass-weighted cp and kappa, and the
    mixture alpha / beta as FD derivatives of the mixture volume;
  - Giordano et al. (2008): log10 eta = A + B / (T - C) for single oxides and
    SiO2-Na2O / SiO2-MgO / CaO-Na2O / SiO2-H2O melts, with the published
    coefficients (b1..b7, b11, b12, c1..c6, c11 of the paper's table 2), the
    reference's Fe total FeO(T) = 0.9309 (Fe2O3 + FeO) and the normalisation;
  - the viscosity corrections: Ishii-Zuber mu / (1 - vf_g), Einstein-Roscoe
    (1 - phi)^-2.5, Caricchi's phi -> 0 limit and strain-rate floor, and the
    reference's Pal (2003) Newton loop in its Ca -> 0 limit, where its two
    steps have a closed form;
  - the plumbing: an interior point's viscosity is evaluated at the shear rate
    of its velocity gradient (fluid_2d.hpp:193-194), and a magma component's
    D is its submixture's (fluid_properties.hpp:187).
The fitted constants of Caricchi's strain-rate laws and Pal's bubble model
are restated as read; only the limits above pin them.
"""
import math
import os

import numpy as np
import pytest

import orc
from helpers import rel

from gales_mc import abi, cases

HERE = os.path.dirname(os.path.abspath(__file__))
TAB = np.load(os.path.join(HERE, "golden", "magma_tables.npz"))
OX = ("SiO2", "TiO2", "Al2O3", "Fe2O3", "FeO", "MnO", "MgO", "CaO", "Na2O", "K2O")
MOLAR = dict(SiO2=60.084, TiO2=79.876, Al2O3=101.956, Fe2O3=159.692, FeO=71.846, MnO=70.937,
             MgO=40.304, CaO=56.078, Na2O=61.979, K2O=94.195)   # g/mol
# Lange & Carmichael partial molar volumes (1e-6 m^3/mol), dV/dT (1e-9), dV/dp (1e-15)
LC_V = dict(SiO2=26.86, TiO2=23.16, Al2O3=37.42, Fe2O3=42.13, FeO=13.65, MnO=0.0, MgO=11.69,
            CaO=16.53, Na2O=28.88, K2O=45.07)
LC_DVDT = dict(SiO2=0.0, TiO2=7.24, Al2O3=0.0, Fe2O3=9.09, FeO=2.92, MnO=0.0, MgO=3.27, CaO=3.74,
               Na2O=7.68, K2O=12.08)
LC_DVDP = dict(SiO2=-1.89, TiO2=-2.31, Al2O3=-2.26, Fe2O3=-2.53, FeO=-0.45, MnO=0.0, MgO=0.27,
               CaO=0.34, Na2O=-2.4, K2O=-6.75)
# Giordano et al. (2008), table 2
GB = dict(b1=159.6, b2=-173.3, b3=72.1, b4=75.7, b5=-39.0, b6=-84.1, b7=141.5, b11=-2.43,
          b12=-0.91, b13=17.6)
GC = dict(c1=2.75, c2=15.7, c3=8.3, c4=10.2, c5=-12.3, c6=-99.5, c11=0.3)


def flat(wh, wc, p0=5e7, p1=4e8, n=8):
    p = np.linspace(p0, p1, n)
    return p, np.full(n, wh), np.full(n, wc)


def comp(oxides, h2o=0.0, co2=0.0, tables=None, **kw):
    if isinstance(oxides, dict):
        oxides = [oxides.get(k, 0.0) for k in OX]
    p, wh, wc = tables if tables is not None else flat(0.0, 0.0)
    return cases.magma(oxides, h2o, co2, p, wh, wc, **kw)


def params(*comps):
    cs = list(comps)
    if len(cs) == 1:
        cs.append(cases.custom(2400.0, 1e2, 1200.0, 1.5))
    return cases.make_params(2, cs)


def mg(c, p, T, sr=0.0):
    P = params(c)
    o = orc.magma(P.comp[0].magma, p, T, sr)
    return o


def shoshonite(**kw):
    return comp(cases.SHOSHONITE_OXIDES, 0.02, 0.01,
                (TAB["p"], TAB["shoshonite_h2o"], TAB["shoshonite_co2"]), **kw)


def log10_mu_melt(o):
    return math.log10(o.mu_melt)


# ---------------------------------------------------------------------------
# tables
# ---------------------------------------------------------------------------
def test_lookup_linear_in_p_and_range_errors():
    p = 6e7 + 1e5 * np.arange(11)
    wh = 0.01 + 2e-11 * (p - 6e7)
    wc = 3e-4 - 1e-12 * (p - 6e7)
    c = comp({"SiO2": 1.0}, 0.05, 0.01, (p, wh, wc))
    for q in (6e7, 6.03e7, 6.0371e7, 6.0999e7, p[-2]):
        o = mg(c, q, 1300.0)
        assert o.st == 0
        assert rel(o.wl_h2o, 0.01 + 2e-11 * (q - 6e7)) < 1e-13
        assert rel(o.wl_co2, 3e-4 - 1e-12 * (q - 6e7)) < 1e-12
    # the last node has no right neighbour (mixtures.hpp:349-353); below p_0 the
    # index is negative
    for q in (p[-1], p[-1] + 1.0, 6e7 - 1.0, 5e7):
        assert mg(c, q, 1300.0).st == abi.GMC_ERR_RANGE


def test_lookup_bilinear_exact_and_mirrored_below_the_table():
    T = np.array([1100.0, 1150.0, 1200.0, 1250.0, 1300.0])
    p = 6e7 + 2e6 * np.arange(7)
    f = lambda TT, pp: 0.012 + 1e-6 * (TT - 1100.0) + 3e-11 * (pp - 6e7) + 2e-15 * (TT - 1100.0) * (pp - 6e7)
    g = lambda TT, pp: 4e-4 - 2e-7 * (TT - 1100.0) + 1e-12 * (pp - 6e7)
    TT, PP = np.meshgrid(T, p, indexing="ij")
    c = cases.magma([1.0] + [0.0] * 9, 0.05, 0.01, p, f(TT, PP), g(TT, PP), T_table=T)
    for Tq, pq in ((1123.4, 6.31e7), (1299.0, 7.19e7), (1100.0, 6e7), (1180.0, 6.4e7)):
        o = mg(c, pq, Tq)
        assert o.st == 0
        assert rel(o.wl_h2o, f(Tq, pq)) < 1e-12
        assert rel(o.wl_co2, g(Tq, pq)) < 1e-12
    # |p - p_0| mirrors a pressure below the table into its first cell: the
    # bilinear form is then evaluated outside that cell (no range error)
    o = mg(c, 5.9e7, 1123.4)
    assert o.st == 0 and rel(o.wl_h2o, f(1123.4, 5.9e7)) < 1e-12
    assert mg(c, p[-1], 1123.4).st == abi.GMC_ERR_RANGE
    assert mg(c, 6.3e7, 1300.0).st == abi.GMC_ERR_RANGE


# ---------------------------------------------------------------------------
# phase split and chemicals
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("p", [6e7, 1.234e8, 2.5e8])
def test_phase_split_balances(p):
    o = mg(shoshonite(), p, 1300.0)
    wf = np.array(o.wf)
    assert abs(wf.sum() - 1.0) < 1e-15
    assert rel(wf[1] + wf[3], 0.02) < 1e-14
    assert rel(wf[2] + wf[4], 0.01) < 1e-14
    liq = wf[0] + wf[1] + wf[2]
    assert rel(wf[1] / liq, o.wl_h2o) < 1e-14
    assert rel(wf[2] / liq, o.wl_co2) < 1e-14
    assert rel(wf[0], 0.97) < 1e-15


@pytest.mark.parametrize("p,T", [(7e7, 1250.0), (2.2e8, 1380.0)])
def test_chemicals_thermodynamics(p, T):
    c = shoshonite()
    o = mg(c, p, T)
    hT, hp = 1e-3 * T, 1e-4 * p
    V = lambda pp, TT: 1.0 / np.array(mg(c, pp, TT).rho_ch)
    dlnV_dT = (np.log(V(p, T + hT)) - np.log(V(p, T - hT))) / (2 * hT)
    dlnV_dp = (np.log(V(p + hp, T)) - np.log(V(p - hp, T))) / (2 * hp)
    for k in range(5):
        assert rel(o.alpha_ch[k], dlnV_dT[k]) < 1e-6, k
        assert rel(o.beta_ch[k], -dlnV_dp[k]) < 1e-6, k
    for k in range(3):  # isothermal sound speed of the condensed chemicals
        assert rel(o.c_ch[k], 1.0 / math.sqrt(o.rho_ch[k] * o.beta_ch[k])) < 1e-14
        assert rel(o.cp_ch[k] - o.cv_ch[k], T * o.alpha_ch[k] ** 2 / (o.rho_ch[k] * o.beta_ch[k])) < 1e-10
    for k, M in ((3, 0.01802), (4, 0.04401)):  # ideal gases, gamma 1.4
        R = 8.314462 / M
        assert rel(o.rho_ch[k], p / (R * T)) < 1e-14
        assert rel(o.cp_ch[k] - o.cv_ch[k], R) < 1e-13
        assert rel(o.cp_ch[k] / o.cv_ch[k], 1.4) < 1e-14
        assert rel(o.c_ch[k], math.sqrt(1.4 * R * T)) < 1e-14
    assert rel(o.cp_ch[1], 2278.0) == 0.0 and rel(o.cp_ch[2], 2278.0) == 0.0


def melt_closed_form(moles, p, T):
    """Lange-Carmichael melt from mole fractions: M = sum x M_i,
    V = sum x (V_i + dV/dT_i (T - 1673) + dV/dp_i (p - 1e5))."""
    M = sum(x * MOLAR[k] * 1e-3 for k, x in moles.items())
    V = sum(x * (LC_V[k] * 1e-6 + LC_DVDT[k] * 1e-9 * (T - 1673.0) + LC_DVDP[k] * 1e-15 * (p - 1e5))
            for k, x in moles.items())
    dVdT = sum(x * LC_DVDT[k] * 1e-9 for k, x in moles.items())
    dVdp = sum(x * LC_DVDP[k] * 1e-15 for k, x in moles.items())
    return M / V, dVdT / V, -dVdp / V


def wf_of_moles(moles):
    tot = sum(x * MOLAR[k] for k, x in moles.items())
    return {k: x * MOLAR[k] / tot for k, x in moles.items()}


@pytest.mark.parametrize("moles", [{"SiO2": 1.0}, {"SiO2": 0.7, "Na2O": 0.3},
                                   {"SiO2": 0.5, "MgO": 0.2, "CaO": 0.2, "K2O": 0.1}])
def test_melt_lange_carmichael(moles):
    p, T = 1.5e8, 1450.0
    o = mg(comp(wf_of_moles(moles)), p, T)
    rho, a, b = melt_closed_form(moles, p, T)
    assert rel(o.rho_ch[0], rho) < 1e-13
    assert abs(o.alpha_ch[0] - a) <= 1e-13 * abs(a) + 1e-300
    assert rel(o.beta_ch[0], b) < 1e-13
    cp = {"SiO2": 1331.0, "Na2O": 1651.0, "MgO": 2424.0, "CaO": 1781.0, "K2O": 1030.0}
    assert rel(o.cp_ch[0], sum(x * cp[k] for k, x in moles.items())) < 1e-14
    # without volatiles the magma is its melt
    assert rel(o.rho, rho) < 1e-13 and rel(o.cp, o.cp_ch[0]) < 1e-15
    assert rel(o.c, o.c_ch[0]) < 1e-13 and rel(o.kappa, 1.5) < 1e-15


# ---------------------------------------------------------------------------
# mixture rules
# ---------------------------------------------------------------------------
def test_mixture_rules_and_wood():
    o = mg(shoshonite(), 1.1e8, 1320.0)
    wf = np.array(o.wf)
    rc = np.array(o.rho_ch)
    assert rel(1.0 / o.rho, np.sum(wf / rc)) < 1e-14
    phi = wf * o.rho / rc
    assert rel(phi.sum(), 1.0) < 1e-14
    assert rel(o.vf_gas, phi[3] + phi[4]) < 1e-15
    cc = np.array(o.c_ch)
    assert rel(1.0 / (o.rho * o.c ** 2), np.sum(phi / (rc * cc ** 2))) < 1e-13
    assert rel(o.cp, np.sum(wf * np.array(o.cp_ch))) < 1e-14
    assert rel(o.cv, np.sum(wf * np.array(o.cv_ch))) < 1e-14
    assert rel(o.kappa, np.sum(wf * np.array([1.5, 0.64, 0.64, 0, 0]) +
                               wf * np.array([0, 0, 0, 1e-5 * o.cp_ch[3] / 0.72, 1e-5 * o.cp_ch[4] / 0.72]))) < 1e-14
    assert rel(o.D, 1e-10) < 1e-15


def test_mixture_alpha_beta_are_volume_derivatives():
    # flat tables: the mass fractions do not move with p or T
    o0 = mg(shoshonite(), 1e8, 1300.0)
    c = comp(cases.SHOSHONITE_OXIDES, 0.02, 0.01, flat(o0.wl_h2o, o0.wl_co2))
    p, T = 1e8, 1300.0
    o = mg(c, p, T)
    hT, hp = 0.5, 1e4
    lnV = lambda pp, TT: -math.log(mg(c, pp, TT).rho)
    a = (lnV(p, T + hT) - lnV(p, T - hT)) / (2 * hT)
    b = -(lnV(p + hp, T) - lnV(p - hp, T)) / (2 * hp)
    assert rel(o.alpha, a) < 1e-6
    assert rel(o.beta, b) < 1e-6


def test_crystal_density_blend():
    o0 = mg(shoshonite(), 1e8, 1300.0)
    o1 = mg(shoshonite(crystal_vf=0.25, crystal_rho=2900.0), 1e8, 1300.0)
    assert rel(o1.rho, 0.75 * o0.rho + 0.25 * 2900.0) < 1e-14


# ---------------------------------------------------------------------------
# Giordano et al. (2008)
# ---------------------------------------------------------------------------
def giordano(mol_pct, h2o_pct, T):
    """log10 eta of the paper's model from hydrous mole percents."""
    m = mol_pct
    Si, Ti, Al = m.get("SiO2", 0), m.get("TiO2", 0), m.get("Al2O3", 0)
    Fe, Mn, Mg, Ca = m.get("FeO", 0), m.get("MnO", 0), m.get("MgO", 0), m.get("CaO", 0)
    Na, K, W = m.get("Na2O", 0), m.get("K2O", 0), h2o_pct
    B = (GB["b1"] * (Si + Ti) + GB["b2"] * Al + GB["b3"] * (Fe + Mn) + GB["b4"] * Mg + GB["b5"] * Ca +
         GB["b6"] * (Na + W) + GB["b7"] * (W + math.log(1 + W)) + GB["b11"] * (Si + Ti) * (Fe + Mn + Mg) +
         GB["b12"] * (Si + Ti + Al) * (Na + K + W) + GB["b13"] * Al * (Na + K))
    C = (GC["c1"] * Si + GC["c2"] * (Ti + Al) + GC["c3"] * (Fe + Mn + Mg) + GC["c4"] * Ca +
         GC["c5"] * (Na + K) + GC["c6"] * math.log(1 + W) + GC["c11"] * (Al + Fe + Mn + Mg + Ca) * (Na + K + W))
    return -4.55 + B / (T - C)


@pytest.mark.parametrize("moles", [{"SiO2": 1.0}, {"Al2O3": 1.0}, {"CaO": 1.0}, {"MgO": 1.0},
                                   {"FeO": 1.0}, {"SiO2": 0.7, "Na2O": 0.3},
                                   {"SiO2": 0.6, "MgO": 0.4}, {"CaO": 0.55, "Na2O": 0.45},
                                   {"SiO2": 0.5, "Al2O3": 0.2, "K2O": 0.3}])
@pytest.mark.parametrize("T", [1350.0, 1900.0])
def test_giordano_anhydrous(moles, T):
    o = mg(comp(wf_of_moles(moles)), 1e8, T)
    expect = giordano({k: 100.0 * x for k, x in moles.items()}, 0.0, T)
    assert abs(log10_mu_melt(o) - expect) < 1e-11 * max(1.0, abs(expect))


def test_giordano_fe_total_and_normalisation():
    # GiordanoEtAlViscosityModel.hpp:45-46 takes FeO(T) = 0.9309 (Fe2O3 + FeO)
    # by mass, so Fe2O3 and FeO count alike, then renormalises
    base = {"SiO2": 0.6, "Al2O3": 0.15, "CaO": 0.1}
    a = dict(base, Fe2O3=0.15)
    b = dict(base, FeO=0.15)
    T = 1400.0
    mu_a = mg(comp(a), 1e8, T).mu_melt
    mu_b = mg(comp(b), 1e8, T).mu_melt
    mu_2a = mg(comp({k: 2.0 * v for k, v in a.items()}), 1e8, T).mu_melt
    assert rel(mu_a, mu_b) < 1e-12
    assert rel(mu_a, mu_2a) < 1e-13
    w = dict(base, FeO=0.9309 * 0.15)
    tot = sum(v / MOLAR[k] for k, v in w.items())
    pct = {k: 100.0 * (v / MOLAR[k]) / tot for k, v in w.items()}
    assert abs(log10_mu_melt(mg(comp(a), 1e8, T)) - giordano(pct, 0.0, T)) < 1e-11
    # Fe2O3 counts as FeO(T): the same melt with Fe2O3 replaced by MgO differs
    c = dict(base, MgO=0.15)
    assert rel(mu_a, mg(comp(c), 1e8, T).mu_melt) > 1e-3


@pytest.mark.parametrize("wh,wc", [(0.02, 0.0), (0.035, 0.0), (0.02, 0.01)])
def test_giordano_water(wh, wc):
    moles = {"SiO2": 0.75, "Na2O": 0.1, "CaO": 0.15}
    wf = wf_of_moles(moles)
    c = comp(wf, wh, wc, flat(wh, wc))
    T = 1250.0
    o = mg(c, 1e8, T)
    M = sum(x * MOLAR[k] for k, x in moles.items())        # anhydrous g/mol
    w = wh / (1.0 - wc)                                    # CO2-free water fraction
    x = (w / 18.015) / (w / 18.015 + (1.0 - w) / M)       # mole fraction of H2O
    hyd = {k: 100.0 * v * (1.0 - x) for k, v in moles.items()}
    expect = giordano(hyd, 100.0 * x, T)
    assert abs(log10_mu_melt(o) - expect) < 1e-11 * abs(expect)


# ---------------------------------------------------------------------------
# viscosity corrections
# ---------------------------------------------------------------------------
def test_gas_and_crystal_corrections():
    p, T = 7e7, 1300.0
    o_none = mg(shoshonite(gas_on_mu=abi.GMC_MU_NONE), p, T)
    assert o_none.mu == o_none.mu_melt
    o_iz = mg(shoshonite(), p, T)
    assert o_iz.vf_gas > 0.05
    assert rel(o_iz.mu, o_iz.mu_melt / (1.0 - o_iz.vf_gas)) < 1e-15
    o_er = mg(shoshonite(crystal_on_mu=abi.GMC_MU_EINSTEIN_ROSCOE, crystal_vf=0.2, crystal_rho=2800.0),
              p, T)
    assert rel(o_er.mu, o_er.mu_melt / (1.0 - o_er.vf_gas) * 0.8 ** -2.5) < 1e-14


def test_caricchi_limits():
    p, T = 7e7, 1300.0
    kw = dict(gas_on_mu=abi.GMC_MU_NONE, crystal_on_mu=abi.GMC_MU_CARICCHI, crystal_rho=2800.0)
    tiny = mg(shoshonite(crystal_vf=1e-12, **kw), p, T, 3.0)
    assert rel(tiny.mu, tiny.mu_melt) < 1e-10
    # strain rates below machine epsilon use 1e-6 s^-1
    a = mg(shoshonite(crystal_vf=0.3, **kw), p, T, 0.0)
    b = mg(shoshonite(crystal_vf=0.3, **kw), p, T, 1e-6)
    assert a.mu == b.mu
    rs = [mg(shoshonite(crystal_vf=f, **kw), p, T, 1.0).mu for f in (0.1, 0.2, 0.3, 0.4)]
    assert all(x < y for x, y in zip(rs, rs[1:]))


def test_pal_2003_newton_loop_small_capillary_limit():
    p, T = 7e7, 1300.0
    c = shoshonite(gas_on_mu=abi.GMC_MU_PAL_2003)
    o0 = mg(c, p, T, 0.0)
    assert o0.mu == o0.mu_melt          # zero strain rate: no correction
    # a magma's Ca is far above 1/sqrt(2.4): pow(1 - a, -0.8) of a negative
    # base is NaN, F is NaN and the loop ends with a NaN viscosity
    o = mg(c, p, T, 0.5)
    assert o.st == 0 and math.isnan(o.mu)
    # Ca = r mu sr / 0.36 with mu sr = 1e-120: a = 2.4 Ca^2 ~ 3e-247, so
    # 1 - a x^2 rounds to 1 and the loop is Newton on x^5 - T^5 from x = 1e-5;
    # it stops after the first iterate whose F >= 1e-9, the second
    for vf in (0.0, 0.12, 0.4):
        mu, st = orc.pal_2003(vf, 1e-120, 1.0)
        Tm = (1.0 - vf / 0.9999) ** -0.9999
        x0 = 1e-5
        x1 = x0 - (x0 ** 5 - Tm ** 5) / (5 * x0 ** 4)
        x2 = x1 - (x1 ** 5 - Tm ** 5) / (5 * x1 ** 4)
        assert st == 0 and rel(mu, 1e-120 * x2) < 1e-12


# ---------------------------------------------------------------------------
# plumbing into the mixture and the element
# ---------------------------------------------------------------------------
def test_magma_component_in_the_mixture():
    ph = comp(cases.PHONOLITE_OXIDES, 0.01, 0.001, (TAB["p"], TAB["phonolite_h2o"], TAB["phonolite_co2"]))
    P = params(shoshonite(), ph)
    p, T, Y = 1.3e8, 1310.0, [0.35, 0.65]
    pr = orc.props(P, p, T, Y)
    a = orc.magma(P.comp[0].magma, p, T)
    b = orc.magma(P.comp[1].magma, p, T)
    assert pr.st == 0
    assert rel(1.0 / pr.rho, Y[0] / a.rho + Y[1] / b.rho) < 1e-15
    assert rel(pr.mu, Y[0] * a.mu + Y[1] * b.mu) < 1e-15
    assert pr.D_comp[0] == a.D and pr.D_comp[1] == b.D
    assert pr.rho_comp[0] == a.rho and rel(pr.e_comp[1], b.cv * T) < 1e-15
    assert orc.props(P, 5e7, T, Y).st == abi.GMC_ERR_RANGE


def test_interior_point_viscosity_at_its_shear_rate():
    kw = dict(gas_on_mu=abi.GMC_MU_NONE, crystal_on_mu=abi.GMC_MU_CARICCHI, crystal_vf=0.35,
              crystal_rho=2800.0)
    P = params(shoshonite(**kw), shoshonite(**kw))
    P.p_ref = 1e8
    xe = np.array([[0.0, 0.0], [0.1, 0.0], [0.0, 0.1]])
    shear = 3.0  # v_x = shear * y
    Ie = np.array([[1e8, shear * y, 0.0, 1300.0, 0.4] for (x, y) in xe])
    Pe = Ie.copy()
    pt = orc.point_probe(P, xe, Ie, Pe, 0)
    p, T = pt["I"][0], pt["I"][3]
    Y = [pt["I"][4], 1.0 - pt["I"][4]]
    sr = math.sqrt(2.0) * shear          # sqrt(dv1_1^2 + dv2_2^2 + 2 (dv1_2 + dv2_1)^2)
    assert rel(pt["pr"].mu, orc.props(P, p, T, Y, sr).mu) < 1e-15
    assert rel(pt["pr"].mu, orc.props(P, p, T, Y, 0.0).mu) > 1e-3


def test_element_reports_out_of_table_pressure():
    P = params(shoshonite(), shoshonite())
    P.p_ref = 1e8
    xe = np.array([[0.0, 0.0], [0.1, 0.0], [0.0, 0.1]])
    Ie = np.array([[p, 0.0, 0.0, 1300.0, 0.4] for p in (1e8, 1e8, 4.5e8)])
    with pytest.raises(ArithmeticError, match=str(abi.GMC_ERR_RANGE)):
        orc.element_interior(P, xe, Ie, Ie)
