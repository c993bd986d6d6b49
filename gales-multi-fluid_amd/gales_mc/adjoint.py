"""Differentiable residual (torch.autograd.Function over the HIP library) and
multi-step sensitivities with checkpoint-and-recompute.

Residual.apply(I, P, theta, ctx) returns R(I, P; theta) = -rhs of
fluid_assembly::execute (fluid_assembly.hpp:39-68; Dirichlet-blocked rows are
0), for I, P float64 CUDA tensors of n_nodes*ndof and theta the 6 material
parameters (rho0, mu, cp of components 0 and 1; see gales_mc.material).  The
backward pass is gmc_vjp_device in exact mode: v^T dR/dI, v^T dR/dP and
v^T dR/dtheta, all formed on the device.  There is no CPU fallback: without
the HIP library or a GPU the calls raise.

The multi-step map used for sensitivities across time steps is a damped
block-Jacobi pseudo-time integrator built on the same residual,
    I_{n+1} = I_n - omega M B^-1 R(I_n, I_{n-1}; theta),
with B the block diagonal of the FP64 Jacobian at the initial state and M an
optional 0/1 mask of advanced dofs (both fixed operators of the scheme, not
differentiated).  rollout() runs n steps and
evaluates an objective of the final state; with checkpoint_every = k only
every k-th state is kept and each segment is recomputed in the backward pass
(torch.utils.checkpoint), so the peak memory holds n/k states plus one
segment instead of all n.
"""
import numpy as np
import torch
from torch.utils.checkpoint import checkpoint

from . import abi


def _require_cuda(*ts):
    for t in ts:
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
            raise ValueError("gales_mc.adjoint needs contiguous float64 CUDA tensors")


class Residual(torch.autograd.Function):
    @staticmethod
    def forward(fctx, I, P, theta, ctx):
        _require_cuda(I, P)
        th = theta.detach().to("cpu", torch.float64)
        ctx.set_material(th.numpy())
        R = torch.empty_like(I)
        torch.cuda.current_stream().synchronize()
        ctx.residual_device(I.data_ptr(), P.data_ptr(), R.data_ptr())
        fctx.save_for_backward(I, P, theta)
        fctx.gmc = ctx
        return R

    @staticmethod
    def backward(fctx, gR):
        I, P, theta = fctx.saved_tensors
        ctx = fctx.gmc
        ctx.set_material(theta.detach().to("cpu", torch.float64).numpy())
        gR = gR.contiguous()
        torch.cuda.current_stream().synchronize()
        ctx.set_state_device(I.data_ptr(), P.data_ptr())
        gI = torch.empty_like(I)
        gP = torch.empty_like(P)
        gT = torch.empty(6, dtype=torch.float64, device=I.device)
        ctx.vjp_device(gR.data_ptr(), gI.data_ptr(), gP.data_ptr(), gT.data_ptr(), abi.GMC_VJP_EXACT)
        return gI, gP, gT.to(theta.device), None


def residual(I, P, theta, ctx):
    return Residual.apply(I, P, theta, ctx)


def block_diag_inverse(ctx, I, P):
    """Inverses of the ndof x ndof diagonal blocks of the FP64 Jacobian at
    (I, P) (host arrays), as a [n_nodes, ndof, ndof] tensor on the GPU."""
    ctx.set_state(I, P)
    ctx.assemble()
    vals, _ = ctx.system()
    rowptr, colidx = ctx.pattern()
    n = ctx.n_nodes
    diag = np.empty(n, np.int64)
    for i in range(n):
        cols = colidx[rowptr[i]:rowptr[i + 1]]
        diag[i] = rowptr[i] + int(np.nonzero(cols == i)[0][0])
    Binv = np.linalg.inv(vals[diag])
    return torch.from_numpy(Binv).cuda()


class Stepper:
    """I_{n+1} = I_n - omega M B^-1 R(I_n, I_{n-1}; theta), M an optional
    fixed 0/1 mask of the dofs the scheme advances (held dofs keep their
    initial values)."""

    def __init__(self, ctx, Binv, omega=0.25, mask=None):
        self.ctx = ctx
        self.Binv = Binv
        self.omega = omega
        self.ndof = ctx.ndof
        self.mask = mask

    def step(self, I, Iprev, theta):
        R = residual(I, Iprev, theta, self.ctx)
        d = torch.bmm(self.Binv, R.view(-1, self.ndof, 1)).view(-1)
        if self.mask is not None:
            d = d * self.mask
        return I - self.omega * d

    def run(self, I, Iprev, theta, n):
        for _ in range(n):
            I, Iprev = self.step(I, Iprev, theta), I
        return I, Iprev


def rollout(stepper, I0, P0, theta, n_steps, objective, checkpoint_every=None):
    """objective(I_n, theta) after n_steps steps from (I0, P0)."""
    I, Ip = I0, P0
    if not checkpoint_every:
        I, Ip = stepper.run(I, Ip, theta, n_steps)
        return objective(I, theta)
    done = 0
    while done < n_steps:
        k = min(checkpoint_every, n_steps - done)
        I, Ip = checkpoint(lambda a, b, t, k=k: stepper.run(a, b, t, k), I, Ip, theta,
                           use_reentrant=False)
        done += k
    return objective(I, theta)


def outlet_mass_flux(mesh, params, ndof, outlet_flag=3):
    """Objective: mass flux of the last component through the outlet sides,
    sum over sides of |side| * mean over its two nodes of rho Y_nc (v . n)
    (2D, two components: component 0 custom, component 1 ideal gas, the
    mixture density by the volume-additive rule of fluid_properties.hpp:199-226).
    Returns f(I, theta) -> scalar tensor."""
    if mesh["dim"] != 2 or params.nc != 2:
        raise ValueError("outlet_mass_flux is written for the 2D two-component case")
    if params.comp[0].kind != abi.GMC_COMP_CUSTOM or params.comp[1].kind != abi.GMC_COMP_IDEAL_GAS:
        raise ValueError("outlet_mass_flux expects a custom component 0 and an ideal-gas component 1")
    conn = np.asarray(mesh["conn"])
    coords = np.asarray(mesh["coords"])
    sel = np.nonzero(np.asarray(mesh["side_flag"]) == outlet_flag)[0]
    loc = {0: (1, 2), 1: (2, 0), 2: (0, 1)}  # triangle side s joins the nodes other than s
    e = np.asarray(mesh["side_elem"])[sel]
    s = np.asarray(mesh["side_local"])[sel]
    a = np.array([conn[ei, loc[si][0]] for ei, si in zip(e, s)])
    b = np.array([conn[ei, loc[si][1]] for ei, si in zip(e, s)])
    t = coords[b] - coords[a]
    length = np.hypot(t[:, 0], t[:, 1])
    nrm = np.stack([t[:, 1], -t[:, 0]], 1) / length[:, None]
    # outward: away from the element's third node
    third = np.array([conn[ei, si] for ei, si in zip(e, s)])
    out = ((coords[a] - coords[third]) * nrm).sum(1) < 0
    nrm[out] *= -1.0
    dev = "cuda"
    A = torch.from_numpy(a).to(dev)
    Bn = torch.from_numpy(b).to(dev)
    L = torch.from_numpy(length).to(dev)
    Nv = torch.from_numpy(nrm).to(dev)

    def f(I, theta):
        S = I.view(-1, ndof)
        p, v, T, Y1 = S[:, 0], S[:, 1:3], S[:, 3], S[:, 4]
        rho_melt, R_gas = theta[0].to(dev), theta[3].to(dev)
        Y2 = 1.0 - Y1
        rho = 1.0 / (Y1 / rho_melt + Y2 * R_gas * T / p)
        g = rho * Y2
        fa = g[A] * (v[A] * Nv).sum(1)
        fb = g[Bn] * (v[Bn] * Nv).sum(1)
        return (L * 0.5 * (fa + fb)).sum()

    return f


# ---------------------------------------------------------------------------
# sensitivities through one implicit time step (the reference's Newton loop,
# sim/fluid_mc/P0/main.cpp:206-240: predictor P -> I, then per iteration
# assemble, solve J delta = -R, corrector I += delta, trim)
# ---------------------------------------------------------------------------
def dirichlet_dofs(mesh, params, ndof):
    """Boolean [n_nodes * ndof]: the dofs the Dirichlet rule of their node's
    flag fixes (fluid_dofs_ic_bc.hpp:86-117): their rows of the assembled
    system are blocked (scatter.hpp:65-78), so a Newton update never moves
    them from the step's initial values."""
    flags = np.asarray(mesh["node_flag"])
    by_flag = np.array([params.dirichlet_mask[f] for f in range(abi.GMC_MAX_FLAG)], np.uint64)
    m = by_flag[flags]
    return ((m[:, None] >> np.arange(ndof, dtype=np.uint64)) & 1).astype(bool).reshape(-1)


def newton_step(ctx, P, theta=None, I0=None, rtol=1e-12, max_newton=60, restart=100, max_iters=1000,
                lin_tol=1e-13, precond=2):
    """One implicit time step from the previous state P (host array): Newton
    iterations from I0 (default P, the predictor) on R(I, P; theta) = 0 with
    the assembled FP64 system and the device GMRES, until ||rhs|| <= rtol
    ||rhs_0||.  Returns (I*, history of ||rhs||).  Dirichlet dofs keep I0's
    values.  Raises if the iteration does not reach rtol."""
    if theta is not None:
        ctx.set_material(theta)
    P = np.asarray(P, np.float64).reshape(-1)
    I = P.copy() if I0 is None else np.asarray(I0, np.float64).reshape(-1).copy()
    ctx.set_state(I, P)
    hist = []
    for _ in range(max_newton):
        ctx.assemble()
        hist.append(ctx.rhs_norm())
        if hist[-1] <= rtol * hist[0]:
            return ctx.get_state()[0].reshape(-1), hist
        ctx.solve(restart=restart, max_iters=max_iters, rel_tol=lin_tol, precond=precond)
        ctx.newton_update()
    raise RuntimeError(f"newton_step: ||rhs|| {hist[-1]:.3e} after {max_newton} iterations "
                       f"(from {hist[0]:.3e}); history {hist}")


def implicit_step_vjp(ctx, mesh, params, I_star, P, theta, g, rtol=1e-9, restart=60, maxiter=400,
                      precond="lu"):
    """Gradients of f(I*) through one converged implicit step, given g = df/dI*.

    At convergence the free rows satisfy R_F(I*, P; theta) = 0 and the
    Dirichlet dofs hold the step's initial values, I*_D = P_D.  With J the
    exact dR/dI (gmc_vjp exact mode: tau, nu, properties and the source all
    differentiated), lambda_F solves J_FF^T lambda_F = g_F, and
        df/dtheta = -lambda^T dR/dtheta,
        df/dP     = -lambda^T dR/dP + E_D (g_D - (J^T lambda)_D).
    The transposed solve is restarted GMRES on the matrix-free exact J^T
    products.  Its preconditioner is built from the assembled
    (frozen-coefficient) Jacobian, which the Newton iteration shows to be close
    to the exact one: precond "lu" applies the transpose of a sparse LU of
    its free-free part (host, scipy's SuperLU; for systems of moderate size),
    "block" the transposed inverses of its diagonal blocks.  Returns (df/dP,
    df/dtheta, info with the final relative residual)."""
    from scipy.sparse import bsr_matrix
    from scipy.sparse.linalg import LinearOperator, gmres, splu
    ndof = ctx.ndof
    n = ctx.n_nodes * ndof
    I_star = np.asarray(I_star, np.float64).reshape(-1)
    P = np.asarray(P, np.float64).reshape(-1)
    g = np.asarray(g, np.float64).reshape(-1)
    if I_star.size != n or P.size != n or g.size != n:
        raise ValueError(f"I_star, P and g need n_nodes * ndof = {n} entries")
    if precond not in ("lu", "block"):
        raise ValueError("precond is 'lu' or 'block'")
    ctx.set_material(theta)
    D = dirichlet_dofs(mesh, params, ndof)
    F = ~D
    nF = int(F.sum())
    ctx.set_state(I_star, P)
    ctx.assemble()
    vals, _ = ctx.system()
    rowptr, colidx = ctx.pattern()

    def full(xF):
        x = np.zeros(n)
        x[F] = xF
        return x

    if precond == "lu":
        J = bsr_matrix((vals, colidx, rowptr), shape=(n, n)).tocsr()
        Fi = np.nonzero(F)[0]
        lu = splu(J[Fi][:, Fi].T.tocsc())

        def prec(xF):
            return lu.solve(xF)
    else:
        diag = np.array([rowptr[i] + int(np.nonzero(colidx[rowptr[i]:rowptr[i + 1]] == i)[0][0])
                         for i in range(ctx.n_nodes)])
        BinvT = np.transpose(np.linalg.inv(vals[diag]), (0, 2, 1))

        def prec(xF):
            return np.einsum("nij,nj->ni", BinvT, full(xF).reshape(-1, ndof)).reshape(-1)[F]

    ctx.set_state(I_star, P)

    def jt(xF):
        gs, _, _ = ctx.vjp(full(xF), mode=abi.GMC_VJP_EXACT, want_prev=False, want_params=False)
        return gs[F]

    A = LinearOperator((nF, nF), matvec=jt, dtype=np.float64)
    M = LinearOperator((nF, nF), matvec=prec, dtype=np.float64)
    its = [0]

    def count(_):
        its[0] += 1

    lamF, info = gmres(A, g[F], rtol=rtol, atol=0.0, restart=restart, maxiter=maxiter, M=M,
                       callback=count, callback_type="pr_norm")
    if info != 0:
        raise RuntimeError(f"implicit_step_vjp: GMRES on J^T did not converge (info {info})")
    lam = full(lamF)
    gs, gp, gth = ctx.vjp(lam, mode=abi.GMC_VJP_EXACT, want_prev=True, want_params=True)
    dP = -gp
    dP[D] += g[D] - gs[D]
    res = np.linalg.norm(gs[F] - g[F]) / np.linalg.norm(g[F])
    return dP, -gth, {"rel_residual": float(res), "n_free": nF, "gmres_iterations": its[0]}
