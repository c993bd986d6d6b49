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
