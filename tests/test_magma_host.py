"""The library's magma properties on the host, against the oracle.  CPU only.

gmc_hydrostatic_profile evaluates the mixture properties in host code built
from the same source as the device kernels (gmc_magma.hpp), so the shipped
P0 magma pair, a bilinear (T, p) table and the strain-rate models are checked
here without a GPU; gmc_create's and the profile's errors for bad magma input
come back as status codes.
"""
import numpy as np
import pytest

import orc
from helpers import magma_tables, phonolite, rel, shoshonite

from gales_mc import abi, cases
from gales_mc.lib import GmcError, hydrostatic_profile


def _profile_inputs(nc, n, T0=1300.0):
    T = T0 + 0.01 * np.arange(n + 1)
    Th = T[:-1] + 0.005
    Y = np.tile(np.linspace(0.3, 0.6, nc - 1), n + 1)
    Yh = np.tile(np.linspace(0.3, 0.6, nc - 1), n)
    return T, Th, Y, Yh


@pytest.mark.parametrize("case", ["p0", "p0_3comp", "bilinear", "crystals"])
def test_hydrostatic_profile_with_magma_matches_oracle(case):
    tab = magma_tables()
    if case == "p0":
        comps = [shoshonite(tab), phonolite(tab)]
    elif case == "p0_3comp":
        comps = [shoshonite(tab), phonolite(tab), cases.H2O]
    elif case == "bilinear":
        T = np.array([1200.0, 1300.0, 1400.0])
        p = tab["p"][::20]
        wh = np.stack([tab["shoshonite_h2o"][::20] * f for f in (1.02, 1.0, 0.97)])
        wc = np.stack([tab["shoshonite_co2"][::20] * f for f in (0.9, 1.0, 1.1)])
        comps = [cases.magma(cases.SHOSHONITE_OXIDES, 0.02, 0.01, p, wh, wc, T_table=T), phonolite(tab)]
    else:
        comps = [shoshonite(tab, crystal_on_mu=abi.GMC_MU_EINSTEIN_ROSCOE, crystal_vf=0.3,
                            crystal_rho=2900.0), phonolite(tab, crystal_on_mu=abi.GMC_MU_CARICCHI,
                                                          crystal_vf=0.2, crystal_rho=2800.0)]
    P = cases.make_params(2, comps)
    P.p_ref = 1e8
    n, h = 300, 10.0
    T, Th, Y, Yh = _profile_inputs(P.nc, n)
    y_or, p_or = orc.hydrostatic_profile(P, 7e7, -3000.0, h, n, T, Th, Y, Yh)
    y_lib, p_lib = hydrostatic_profile(P, 7e7, -3000.0, h, n, T, Th, Y, Yh)
    assert np.array_equal(y_or, y_lib)
    assert rel(p_lib, p_or) < 1e-14
    if case == "p0":   # a column of the two magmas, rho ~ 2.1-2.5 t/m^3
        assert 7e7 + 2000 * 9.81 * n * h < p_lib[-1] < 7e7 + 2600 * 9.81 * n * h


def test_hydrostatic_profile_leaving_the_table():
    P = cases.make_params(2, [shoshonite(), phonolite()])
    P.p_ref = 1e8
    n = 100
    T, Th, Y, Yh = _profile_inputs(2, n)
    # 1 km of magma below 2.9e8 Pa passes the table's last pressure, 3e8 Pa
    with pytest.raises(GmcError) as e:
        hydrostatic_profile(P, 2.9e8, 0.0, 10.0, n, T, Th, Y, Yh)
    assert e.value.code == abi.GMC_ERR_RANGE


def test_bad_magma_input_is_rejected():
    tab = magma_tables()
    n = 4
    T, Th, Y, Yh = _profile_inputs(2, n)
    P = cases.make_params(2, [shoshonite(tab), phonolite(tab)])
    P.p_ref = 2e7                      # u_ref's properties below the table
    with pytest.raises(GmcError) as e:
        hydrostatic_profile(P, 7e7, 0.0, 10.0, n, T, Th, Y, Yh)
    assert e.value.code == abi.GMC_ERR_RANGE
    P = cases.make_params(2, [shoshonite(tab, gas_on_mu=7), phonolite(tab)])
    P.p_ref = 1e8
    with pytest.raises(GmcError) as e:
        hydrostatic_profile(P, 7e7, 0.0, 10.0, n, T, Th, Y, Yh)
    assert e.value.code == abi.GMC_ERR_ARG
    P = cases.make_params(2, [shoshonite(tab), phonolite(tab)])
    P.p_ref = 1e8
    P.comp[0].magma.n_p = 1
    with pytest.raises(GmcError) as e:
        hydrostatic_profile(P, 7e7, 0.0, 10.0, n, T, Th, Y, Yh)
    assert e.value.code == abi.GMC_ERR_ARG
