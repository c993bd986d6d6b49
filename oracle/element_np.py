"""Second, independent NumPy restatement of the fluid_mc element integral (2D, and
the build-defined 3D extension).

TEST INFRASTRUCTURE ONLY.  It exists to catch transcription errors in the C
oracle (oracle/gmc_oracle.c): it is written in matrix form rather than entry
by entry, so a dropped term, wrong sign or transposed operand in either one
shows up as an O(1) disagreement.

  DUDY = phi drho^T + rho Dphi         (U = rho phi(Y))
  A_k  = v_k DUDY + U e_{v_k}^T + e_{v_k} e_p^T + e_E (v_k e_p + p e_{v_k})^T
  K    from the stress / Fourier / species-diffusion fluxes
  m    = sum_q JxW Phi_q^T C_q Phi_q with the least-squares block
         C_LSQ[k][l] = L_k tau R_l
  tau  uses numpy.linalg.inv (LAPACK), not the uBLAS-order LU of the oracle.

Follows sim/fluid_mc/P0/fluid_2d.hpp:171-333,372-603,877-960,1027-1065 and
the property rules of fluid_properties.hpp:199-226 / chemicals.hpp.
"""
import numpy as np

EPS = np.finfo(float).eps


def comp_props(c, p, T):
    if c.kind == 0:
        cv = c.cp if c.beta == 0.0 else c.cp - T * c.alpha ** 2 / (c.rho * c.beta)
        snd = 1e20 if c.beta == 0.0 else 1.0 / np.sqrt(c.rho * c.beta)
        return dict(rho=c.rho, mu=c.mu, cp=c.cp, cv=cv, kappa=c.kappa, alpha=c.alpha,
                    beta=c.beta, c=snd, D=c.diffusivity)
    R = 8.314462 / c.molar_mass
    cv = R / (c.gamma - 1.0)
    cp = c.gamma * cv
    return dict(rho=p / (R * T), mu=c.mu, cp=cp, cv=cv, kappa=c.mu * cp / 0.72, alpha=1.0 / T,
                beta=1.0 / p, c=np.sqrt(c.gamma * R * T), D=c.diffusivity)


def mixture(P, p, T, Y):
    cs = [comp_props(P.comp[i], p, T) for i in range(P.nc)]
    Y = np.asarray(Y, float)
    rc = np.array([c["rho"] for c in cs])
    rho = 1.0 / np.sum(Y / rc)
    out = dict(rho=rho)
    for k in ("mu", "cp", "cv", "kappa"):
        out[k] = float(np.sum(Y * np.array([c[k] for c in cs])))
    snd = np.array([c["c"] for c in cs])
    out["c"] = np.sqrt(1.0 / np.sum(Y / (rc ** 2 * snd ** 2))) / rho
    out["alpha"] = float(np.sum(Y * np.array([c["alpha"] for c in cs]) * rho / rc))
    out["beta"] = float(np.sum(Y * np.array([c["beta"] for c in cs]) * rho / rc))
    out["rho_c"] = rc
    out["e_c"] = np.array([c["cv"] for c in cs]) * T
    out["D_c"] = np.array([c["D"] for c in cs])
    return out


def split(P, y):
    d = P.dim
    p = y[0]
    v = np.array(y[1:1 + d])
    T = y[1 + d]
    Yd = np.array(y[2 + d:])
    Y = np.concatenate([Yd, [1.0 - Yd.sum()]])
    return p, v, T, Y


def conservative(P, pr, v, T, Y):
    E = pr["cv"] * T + 0.5 * v @ v
    return pr["rho"] * np.concatenate([[1.0], v, [E], Y[:-1]]), E


def flux_mats(P, y, grad):
    """DUDY, A[k], K[k][l], S0, U, Fadv[k], Fdif[k], S at one point."""
    d, nc = P.dim, P.nc
    n = d + 2 + nc - 1
    p, v, T, Y = split(P, y)
    pr = mixture(P, p, T, Y)
    rho = pr["rho"]
    U, E = conservative(P, pr, v, T, Y)
    phi = U / rho
    drho = np.zeros(n)
    drho[0] = rho * pr["beta"]
    drho[d + 1] = -rho * pr["alpha"]
    drho[d + 2:] = -rho ** 2 * (1.0 / pr["rho_c"][:-1] - 1.0 / pr["rho_c"][-1])
    Dphi = np.zeros((n, n))
    for k in range(d):
        Dphi[1 + k, 1 + k] = 1.0
    Dphi[d + 1, 0] = (pr["beta"] * p - pr["alpha"] * T) / rho
    Dphi[d + 1, 1:1 + d] = v
    Dphi[d + 1, d + 1] = pr["cp"] - p * pr["alpha"] / rho
    Dphi[d + 1, d + 2:] = pr["e_c"][:-1] - pr["e_c"][-1]
    for i in range(nc - 1):
        Dphi[d + 2 + i, d + 2 + i] = 1.0
    DUDY = np.outer(phi, drho) + rho * Dphi
    ep = np.eye(n)[0]
    eE = np.eye(n)[d + 1]
    A = []
    for k in range(d):
        ev = np.eye(n)[1 + k]
        A.append(v[k] * DUDY + np.outer(U, ev) + np.outer(ev, ep) + np.outer(eE, v[k] * ep + p * ev))
    mu = pr["mu"]
    lam = -2.0 * mu / 3.0
    H = pr["e_c"] + p / pr["rho_c"]
    K = [[np.zeros((n, n)) for _ in range(d)] for _ in range(d)]
    for k in range(d):
        for l in range(d):
            # stress tau_ik = lam delta_ik div v + mu (d_k v_i + d_i v_k); F_k[v_i] = tau_ik
            for i in range(d):
                for j in range(d):
                    K[k][l][1 + i, 1 + j] = (lam * (i == k) * (j == l) + mu * (i == l) * (j == k)
                                             + mu * (i == j) * (k == l))
            K[k][l][d + 1, 1:1 + d] = v @ K[k][l][1:1 + d, 1:1 + d]
            if k == l:
                K[k][l][d + 1, d + 1] = pr["kappa"]
                for i in range(nc - 1):
                    K[k][l][d + 2 + i, d + 2 + i] = rho * pr["D_c"][i]
                    K[k][l][d + 1, d + 2 + i] = rho * pr["D_c"][i] * H[i]
    g = np.array(P.gravity[:d])
    S0 = np.zeros((n, n))
    S0[d + 1, 1:1 + d] = rho * g
    S = np.zeros(n)
    S[1:1 + d] = rho * g
    S[d + 1] = rho * g @ v
    Fadv = []
    for k in range(d):
        F = v[k] * U
        F[1 + k] += p
        F[d + 1] += p * v[k]
        Fadv.append(F)
    Fdif = [sum(K[k][l] @ grad[l] for l in range(d)) for k in range(d)]
    return dict(DUDY=DUDY, A=A, K=K, S0=S0, U=U, Fadv=Fadv, Fdif=Fdif, S=S, pr=pr, p=p, v=v,
                T=T, Y=Y)


def geometry(xe, dim=2):
    x = np.asarray(xe, float).reshape(dim + 1, dim)
    Jm = np.array([x[a] - x[0] for a in range(1, dim + 1)]).T   # d(x)/d(xi)
    J = np.linalg.det(Jm)
    inv = np.linalg.inv(Jm)                                     # d(xi)/d(x)
    dref = np.hstack([-np.ones((dim, 1)), np.eye(dim)])
    dsh = inv.T @ dref                                           # [k][a]
    G = inv.T @ inv
    return J, dsh, G


def compute_h(G, v, dv):
    s0 = v / np.linalg.norm(v)
    a0 = s0 @ G @ s0
    acc = dv @ v
    s1 = acc - (acc @ s0) * s0
    n1 = np.linalg.norm(s1)
    if n1 > EPS:
        s1 = s1 / n1
    a1 = s1 @ G @ s1
    al = [a0, a1]
    if len(v) == 3:
        s2 = np.cross(s0, s1)
        n2 = np.linalg.norm(s2)
        if n2 > EPS:
            s2 = s2 / n2
        al.append(s2 @ G @ s2)
    h = 0.0
    for a in al:
        if a != 0.0:
            h += 1.0 / a
    return 2.0 * np.sqrt(abs(h))


def tau_2019(P, f, G, grad):
    d, nc = P.dim, P.nc
    n = d + 2 + nc - 1
    v = f["v"]
    vn = np.linalg.norm(v)
    if vn < EPS:
        return np.zeros((n, n))
    pr = f["pr"]
    dv = np.array([[grad[j][1 + i] for j in range(d)] for i in range(d)])
    h = compute_h(G, v, dv)
    use = min(P.dt * 0.5, h / (2.0 * (vn + pr["c"])) + h / (2.0 * vn))
    td = np.empty(n)
    td[0] = use
    td[1:1 + d] = min(use, h * h * pr["rho"] / (12.0 * pr["mu"]))
    td[d + 1] = min(use, h * h * pr["rho"] * pr["cv"] / (12.0 * pr["kappa"]))
    td[d + 2:] = np.minimum(use, h * h / (12.0 * pr["D_c"][:-1]))
    tau = np.linalg.inv(f["DUDY"]) * td[None, :]
    if vn / pr["c"] < 0.1:
        tau[0, 0] = 1.0 / (1.0 / tau[0, 0] + 1.0 / (pr["rho"] * tau[1, 1] * np.trace(G)))
        tau[0, 1:] = 0.0
    return tau


def dc_shock(P, f, dsh, grad, uref):
    d = P.dim
    if not P.dc_2006:
        return 0.0
    dU = [f["DUDY"] @ grad[k] for k in range(d)]
    gr = np.array([dU[k][0] for k in range(d)])
    nr = np.linalg.norm(gr)
    if nr < EPS:
        return 0.0
    j = gr / nr
    h = np.sum(np.abs(j @ dsh))
    if not h > EPS:
        return 0.0
    h = 1.0 / h
    Z = sum(f["A"][k] @ grad[k] for k in range(d)) - f["S"]
    zn = np.linalg.norm(uref * Z)
    un = np.linalg.norm(uref * f["U"])
    assert P.dc_sharp == 2.0
    return P.dc_scale_fact * zn * h * h / un


def element_interior(P, xe, Ie, Pe, uref):
    d, nc = P.dim, P.nc
    n = d + 2 + nc - 1
    nen = d + 1
    J, dsh, G = geometry(xe, d)
    Ie = np.asarray(Ie, float).reshape(nen, n)
    Pe = np.asarray(Pe, float).reshape(nen, n)
    grad = [dsh[k] @ Ie for k in range(d)]
    m = np.zeros((nen * n, nen * n))
    r = np.zeros(nen * n)
    if d == 2:
        pts = [(1 / 6, 1 / 6), (4 / 6, 1 / 6), (1 / 6, 4 / 6)]
        w = 1.0 / 6.0
    else:
        a_, b_ = 0.1381966011250105, 0.5854101966249685
        pts = [(a_, a_, a_), (b_, a_, a_), (a_, b_, a_), (a_, a_, b_)]
        w = 1.0 / 24.0
    for xi in pts:
        sh = np.concatenate([[1.0 - sum(xi)], xi])
        JxW = J * w
        yP = sh @ Pe
        yI = sh @ Ie
        pP, vP, TP, YP = split(P, yP)
        UP, _ = conservative(P, mixture(P, pP, TP, YP), vP, TP, YP)
        f = flux_mats(P, yI, grad)
        tau = tau_2019(P, f, G, grad)
        nu = dc_shock(P, f, dsh, grad, uref)
        # test/trial function blocks: phi_a = (N_a, dN_a/dx, dN_a/dy)
        Phi = np.zeros(((d + 1) * n, nen * n))
        for a in range(nen):
            for kk, val in enumerate([sh[a]] + [dsh[k, a] for k in range(d)]):
                Phi[kk * n:(kk + 1) * n, a * n:(a + 1) * n] = val * np.eye(n)
        Cg = np.zeros(((d + 1) * n, (d + 1) * n))
        Cg[0:n, 0:n] = f["DUDY"] / P.dt - f["S0"]
        for k in range(d):
            Cg[(1 + k) * n:(2 + k) * n, 0:n] = -f["A"][k]
            for l in range(d):
                Cg[(1 + k) * n:(2 + k) * n, (1 + l) * n:(2 + l) * n] = f["K"][k][l]
            Cg[(1 + k) * n:(2 + k) * n, (1 + k) * n:(2 + k) * n] += nu * f["DUDY"]
        Lm = np.vstack([-f["S0"]] + [f["A"][k] for k in range(d)])          # rows (k, i)
        Rm = np.hstack([f["DUDY"] / P.dt - f["S0"]] + [f["A"][k] for k in range(d)])
        C = Cg + Lm @ tau @ Rm
        m += JxW * Phi.T @ C @ Phi
        Res = f["DUDY"] @ (yI - yP) / P.dt + sum(f["A"][k] @ grad[k] for k in range(d)) - f["S"]
        flux_vec = np.concatenate([(f["U"] - UP) / P.dt - f["S"]] +
                                  [-f["Fadv"][k] + f["Fdif"][k] + nu * f["DUDY"] @ grad[k]
                                   for k in range(d)])
        r += JxW * (Phi.T @ (flux_vec + Lm @ (tau @ Res)))
    return m, -r
