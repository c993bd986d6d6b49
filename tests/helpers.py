"""Shared helpers of the test suite."""
import numpy as np


def rel(a, b):
    """Norm-wise relative difference ||a - b|| / ||b|| (0 when both vanish)."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    nb = np.linalg.norm(b)
    na = np.linalg.norm(a - b)
    if nb == 0.0:
        return na
    return na / nb


def entry_rel(a, b):
    """max |a - b| / max |b| (entrywise error scaled by the largest entry)."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    s = np.abs(b).max()
    return np.abs(a - b).max() / (s if s > 0 else 1.0)


def h_is_fragile(P, xe, Ie):
    """True when quad::compute_h's second direction is rounding noise at some
    point: s1 = acc - (acc.s0) s0 is compared with an absolute epsilon
    (quad.hpp:366-375), so when the convective acceleration is parallel to v
    (every wall element with two no-slip nodes) the reference's h -- and tau --
    depends on the last bits of s1.  Independent restatements with a different
    operation order may take the other branch there."""
    d = P.dim
    n = d + 2 + P.nc - 1
    import element_np
    J, dsh, G = element_np.geometry(xe, d)
    Ie = np.asarray(Ie).reshape(d + 1, n)
    grad = [dsh[k] @ Ie for k in range(d)]
    dv = np.array([[grad[j][1 + i] for j in range(d)] for i in range(d)])
    for q in range(d + 1):
        sh = np.full(d + 1, 1.0 / 6.0 if d == 2 else 0.1381966011250105)
        v = (sh @ Ie)[1:1 + d]
        acc = dv @ v
        if np.linalg.norm(v) == 0.0:
            continue
        s0 = v / np.linalg.norm(v)
        s1 = acc - (acc @ s0) * s0
        if np.linalg.norm(s1) < 1e-9 * (np.linalg.norm(acc) + 1e-300) and np.linalg.norm(acc) > 0:
            return True
    return False


def residual_nonsmooth(P, xe, Ie):
    """True when the reference's element residual has no derivative at I_e:
    the compute_h switch above, or a quadrature point with |v| = 0, where tau
    jumps from 0 (fluid_2d.hpp:882 sets tau = 0 for |v| < eps) to its value
    for any velocity perturbation (elements with every node on a no-slip
    wall)."""
    if h_is_fragile(P, xe, Ie):
        return True
    d = P.dim
    n = d + 2 + P.nc - 1
    Ie = np.asarray(Ie).reshape(d + 1, n)
    return bool(np.all(Ie[:, 1:1 + d] == 0.0))


def kink_margin(P, xe, Ie, Pe):
    """Smallest relative distance, over the element's interior points, of the
    state from a switch of the reference residual: the min() branches of tau
    (fluid_2d.hpp:934-945), the absolute values of the DC length
    (fluid_2d.hpp:1040-1043) and the second direction of compute_h
    (quad.hpp:366-375).  The residual is smooth only between such switches,
    so finite differences that cross one measure a jump, not a derivative."""
    import orc
    d = P.dim
    n = d + 2 + P.nc - 1
    out = np.inf
    for q in range(d + 1):
        pr = orc.point_probe(P, xe, Ie, Pe, q)
        o = pr["pr"]
        v = pr["I"][1:1 + d]
        vn = np.linalg.norm(v)
        if pr["h"] > 0.0:
            h = pr["h"]
            adv = h / (2 * (vn + o.c)) + h / (2 * vn)
            use = adv if P.steady_state else min(0.5 * P.dt, adv)
            if not P.steady_state:
                out = min(out, abs(adv - 0.5 * P.dt) / use)
            cands = [h * h * o.rho / (12 * o.mu), h * h * o.rho * o.cv / (12 * o.kappa)]
            cands += [h * h / (12 * P.comp[i].diffusivity) for i in range(P.nc - 1)]
            out = min(out, min(abs(c - use) / use for c in cands))
            dv = np.array([[pr["dI"][j][1 + i] for j in range(d)] for i in range(d)])
            acc = dv @ v
            na = np.linalg.norm(acc)
            if na > 0.0:
                s0 = v / vn
                out = min(out, np.linalg.norm(acc - (acc @ s0) * s0) / na)
        if pr["h_dc"] > 0.0:
            grad_rho = np.array([pr["DUDY"][0] @ pr["dI"][k] for k in range(d)])
            j = grad_rho / np.linalg.norm(grad_rho)
            t = np.abs(j @ pr["dsh"])
            out = min(out, t.min() / t.sum())
    return out


# ---------------------------------------------------------------------------
# magma cases: the reference's shipped magmas (sim/fluid_mc/P0/props.txt) with
# its solubility tables (tests/golden/magma_tables.npz)
# ---------------------------------------------------------------------------
def magma_tables():
    import os
    return np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "magma_tables.npz"))


def shoshonite(tab=None, **kw):
    from gales_mc import cases
    tab = magma_tables() if tab is None else tab
    return cases.magma(cases.SHOSHONITE_OXIDES, 0.02, 0.01, tab["p"], tab["shoshonite_h2o"],
                       tab["shoshonite_co2"], **kw)


def phonolite(tab=None, **kw):
    from gales_mc import cases
    tab = magma_tables() if tab is None else tab
    return cases.magma(cases.PHONOLITE_OXIDES, 0.01, 0.001, tab["p"], tab["phonolite_h2o"],
                       tab["phonolite_co2"], **kw)


def magma_case(mesh, comps, seed=7, dt=1e-3):
    """(params, I, P) on mesh with magma components: the seeded fields of
    cases.fields shifted to p ~ 1e8 Pa (inside the tables' 6e7..3e8 Pa)."""
    from gales_mc import cases
    P = cases.make_params(mesh["dim"], comps, dt=dt)
    P.p_ref = 1e8
    I, Pst = cases.fields(mesh, len(comps), seed)
    I[:, 0] += 9e7
    Pst[:, 0] += 9e7
    return P, I, Pst
