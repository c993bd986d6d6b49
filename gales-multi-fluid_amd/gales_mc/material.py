"""Material parameters theta = (rho0, mu, cp) of components 0 and 1.

A custom component (chemicals.hpp:23-47) is parametrised by its constant
density, viscosity and heat capacity; an ideal gas (chemicals.hpp:201-264) by
its specific gas constant R (the "rho0" slot: rho = p/(RT)), viscosity and cp,
with molar_mass = 8.314462/R and gamma = cp/(cp - R).  The same mapping as
gmc_set_material, for host-side use (finite differences, copies of params).
"""
import ctypes as C

from . import abi

R_UNIVERSAL = 8.314462


def clone(params):
    """Independent copy of an abi.Params."""
    out = abi.Params()
    C.memmove(C.byref(out), C.byref(params), C.sizeof(abi.Params))
    return out


def theta_of(params):
    th = []
    for k in range(2):
        c = params.comp[k]
        if c.kind == abi.GMC_COMP_CUSTOM:
            th += [c.rho, c.mu, c.cp]
        else:
            R = R_UNIVERSAL / c.molar_mass
            th += [R, c.mu, c.gamma * R / (c.gamma - 1.0)]
    return th


def with_theta(params, theta):
    """Copy of params with theta applied."""
    out = clone(params)
    for k in range(2):
        c = out.comp[k]
        r, mu, cp = (float(x) for x in theta[3 * k:3 * k + 3])
        if c.kind == abi.GMC_COMP_CUSTOM:
            c.rho, c.mu, c.cp = r, mu, cp
        else:
            if not (cp > r > 0.0):
                raise ValueError("ideal gas needs cp > R > 0")
            c.molar_mass = R_UNIVERSAL / r
            c.mu = mu
            c.gamma = cp / (cp - r)
    return out
