"""Adjoint on the GPU, through the C-ABI and the torch binding.

  frozen VJP  = J^T v of the assembled FP64 matrix (oracle's, scipy transpose)
                within 1e-12
  exact VJP   = central finite differences of the GPU FP64 residual, h = 1e-6
                relative, on 32 seeded directions: |g.d - fd| <= 1e-6 |fd| plus
                the differences' rounding floor; previous-step and all six
                material directions likewise
  exact = frozen where the reference's matrix is the exact Jacobian (uniform
                state, no gravity)
  torch       autograd through Residual and a 16-step checkpointed rollout
                against finite differences

Finite-difference directions leave out the dofs of elements where the
reference's residual is not differentiable (Dirichlet-blocked dofs, and the
nodes of elements with the compute_h or |v| = 0 switch, residual_nonsmooth).
"""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

import orc
from helpers import rel, residual_nonsmooth

from gales_mc import abi, cases, material
from gales_mc.lib import Context

pytestmark = pytest.mark.gpu
EPS = 2.220446049250313e-16


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def c0():
    mesh, P, I, Pst = cases.config(0)
    ctx = Context(mesh, P)
    ctx.set_state(I, Pst)
    yield mesh, P, I.reshape(-1).copy(), Pst.reshape(-1).copy(), ctx
    ctx.close()


def _bsr_to_csr(vals, rowptr, colidx, n):
    nb = np.diff(rowptr)
    rows = np.repeat(np.arange(len(nb)), nb)
    r = (rows[:, None, None] * n + np.arange(n)[None, :, None]) + np.zeros((1, 1, n), int)
    c = (np.asarray(colidx)[:, None, None] * n + np.arange(n)[None, None, :]) + np.zeros((1, n, 1), int)
    N = len(nb) * n
    return sp.csr_matrix((vals.reshape(-1), (r.reshape(-1), c.reshape(-1))), shape=(N, N))


def _smooth_mask(mesh, P, I):
    """dofs that finite-difference directions may move."""
    n = P.dim + 2 + P.nc - 1
    dmask = orc.dmask_of(P, mesh)
    free = np.ones((mesh["coords"].shape[0], n), bool)
    for k in range(n):
        free[:, k] = ((dmask >> k) & 1) == 0
    conn = mesh["conn"]
    Ir = np.asarray(I).reshape(-1, n)
    for e in range(conn.shape[0]):
        xe, Ie, _ = orc.element_data(mesh, Ir, Ir, e)
        if residual_nonsmooth(P, xe, Ie):
            free[conn[e]] = False
    return free.reshape(-1)


def _scale(P, n_nodes):
    return np.tile([1e4] + [1.0] * P.dim + [5.0] + [0.05] * (P.nc - 1), n_nodes)


def test_vjp_frozen_is_assembled_transpose(c0):
    mesh, P, I, Pst, ctx = c0
    _say("vjp frozen c0")
    ctx.set_state(I, Pst)
    v = np.random.default_rng(21).standard_normal(I.size)
    g, gp, gpar = ctx.vjp(v, mode=abi.GMC_VJP_FROZEN)
    assert gp is None and gpar is None
    A = orc.Assembler(P, mesh)
    vals_ref, _ = A.assemble(I, Pst)
    JT = _bsr_to_csr(vals_ref, A.rowptr, A.colidx, ctx.ndof).T.tocsr()
    e_or = rel(g, JT @ v)
    ctx.assemble()
    vals, _ = ctx.system()
    e_gpu = rel(g, _bsr_to_csr(vals, A.rowptr, A.colidx, ctx.ndof).T @ v)
    _say(f"  vs oracle J^T v {e_or:.2e}, vs GPU matrix {e_gpu:.2e}")
    assert e_or < 1e-12 and e_gpu < 1e-12
    g2, _, _ = ctx.vjp(v, mode=abi.GMC_VJP_FROZEN)
    assert np.array_equal(g, g2)
    with pytest.raises(Exception):
        ctx.vjp(v, mode=abi.GMC_VJP_FROZEN, want_prev=True)


def _fd_ok(gd, fd, vR, h):
    floor = 1000.0 * EPS * vR / h
    return abs(gd - fd) <= 1e-6 * abs(fd) + floor


def test_vjp_exact_finite_differences(c0):
    mesh, P, I, Pst, ctx = c0
    _say("vjp exact vs FD c0")
    rng = np.random.default_rng(22)
    v = rng.standard_normal(I.size)
    ctx.set_state(I, Pst)
    gs, gp, gpar = ctx.vjp(v, want_prev=True)
    gs2, gp2, gpar2 = ctx.vjp(v, want_prev=True)
    assert np.array_equal(gs, gs2) and np.array_equal(gp, gp2) and np.array_equal(gpar, gpar2)
    free = _smooth_mask(mesh, P, I)
    sc = _scale(P, mesh["coords"].shape[0])
    h = 1e-6

    def Rvec(Ii, Pi):
        ctx.set_state(Ii, Pi)
        return ctx.residual_vector()

    def fd_of(Rp, Rm, hs):
        # difference the residual vectors before the dot product: the sum
        # v.R of 1e4 terms of 1e7 magnitude would round away the difference
        return (v @ (Rp - Rm)) / (2 * hs), np.abs(v * Rp).sum()

    worst = [0.0, 0.0]
    for k in range(32):
        d = sc * rng.standard_normal(I.size) * free
        for j, hs in enumerate((h, 1e-7)):
            fd, mag = fd_of(Rvec(I + hs * d, Pst), Rvec(I - hs * d, Pst), hs)
            worst[j] = max(worst[j], abs(gs @ d - fd) / abs(fd))
            if j == 0:
                assert _fd_ok(gs @ d, fd, mag, hs), (k, gs @ d, fd)
    _say(f"  state: worst rel err over 32 directions {worst[0]:.2e} (h = 1e-6), {worst[1]:.2e} (h = 1e-7)")
    worst = 0.0
    for k in range(8):
        d = sc * rng.standard_normal(I.size)
        fd, mag = fd_of(Rvec(I, Pst + h * d), Rvec(I, Pst - h * d), h)
        assert _fd_ok(gp @ d, fd, mag, h), (k, gp @ d, fd)
        worst = max(worst, abs(gp @ d - fd) / abs(fd))
    _say(f"  previous step: worst rel {worst:.2e} over 8 directions")
    th = np.array(material.theta_of(P))
    ctx.set_state(I, Pst)
    def dtheta(k, hr):
        hk = hr * th[k]
        tp, tm = th.copy(), th.copy()
        tp[k] += hk
        tm[k] -= hk
        ctx.set_material(tp)
        Rp = Rvec(I, Pst)
        ctx.set_material(tm)
        Rm = Rvec(I, Pst)
        return fd_of(Rp, Rm, hk)[0]

    for k in range(6):
        # h = 1e-6 relative; where a parameter's effect is below the rounding
        # of v.R at that step (the gas viscosity, 1e-5 Pa s), also 1e-4 and
        # 1e-3 relative and the Richardson extrapolation of the 0.5 % and 1 %
        # steps (truncation O(h^4)): the best of them must agree to 1e-6
        fds = [dtheta(k, hr) for hr in (1e-6, 1e-4, 1e-3)]
        fds.append((4.0 * dtheta(k, 5e-3) - dtheta(k, 1e-2)) / 3.0)
        errs = [abs(gpar[k] - fd) / abs(fd) for fd in fds]
        _say(f"  theta[{k}]: vjp {gpar[k]:.8e}; rel err vs FD at h = 1e-6/1e-4/1e-3/Richardson(1e-2): "
             + " ".join(f"{x:.1e}" for x in errs))
        assert min(errs) <= 1e-6, (k, errs)
    ctx.set_material(th)
    ctx.set_state(I, Pst)


def test_vjp_exact_equals_frozen_at_uniform_state():
    mesh = cases.channel2d(16, 8, seed=5)
    P = cases.make_params(2, [cases.MELT, cases.GAS_R461], gravity=(0.0, 0.0, 0.0), dt=1e-5)
    y = np.array([1.2e7, 0.3, -0.2, 1320.0, 0.5])
    I = np.tile(y, mesh["coords"].shape[0])
    v = np.random.default_rng(23).standard_normal(I.size)
    dmask = orc.dmask_of(P, mesh)
    blocked = np.array([[(m >> k) & 1 for k in range(5)] for m in dmask], bool).reshape(-1)
    v[blocked] = 0.0
    with Context(mesh, P) as ctx:
        ctx.set_state(I, I)
        ge, _, _ = ctx.vjp(v)
        gf, _, _ = ctx.vjp(v, mode=abi.GMC_VJP_FROZEN)
    _say(f"uniform state: exact vs frozen {rel(ge, gf):.2e}")
    # only the boundary energy/species K block differs (D-scaled, ~1e-9)
    assert rel(ge, gf) < 1e-7


def test_vjp_3d_four_components():
    """3D, 4 components (the full-form shape at 1/16 resolution, 20,736
    tetrahedra): the frozen J^T v equals the oracle's assembled matrix
    transposed times v, and at a uniform state with g = 0 the exact VJP agrees
    with it as in 2D."""
    mesh, P, I, Pst = cases.config(4, scale=16)
    n = P.dim + 2 + P.nc - 1
    v = np.random.default_rng(29).standard_normal(I.size)
    with Context(mesh, P) as ctx:
        ctx.set_state(I, Pst)
        g, _, _ = ctx.vjp(v, mode=abi.GMC_VJP_FROZEN)
    A = orc.Assembler(P, mesh)
    vals_ref, _ = A.assemble(I, Pst)
    JT = _bsr_to_csr(vals_ref, A.rowptr, A.colidx, n).T.tocsr()
    e_or = rel(g, JT @ v)
    _say(f"3d nc=4 frozen J^T v vs oracle {e_or:.2e}")
    assert e_or < 1e-12

    P0 = cases.make_params(3, [cases.MELT, cases.CRYSTAL, cases.CRYSTAL2, cases.H2O],
                           gravity=(0.0, 0.0, 0.0), dt=1e-5)
    m3 = cases.kuhn3d(6, 6, 3, 2.0, 2.0, 1.0)
    y = np.array([1.2e7, 0.3, -0.2, 0.1, 1320.0, 0.3, 0.2, 0.2])
    U = np.tile(y, m3["coords"].shape[0])
    w = np.random.default_rng(31).standard_normal(U.size)
    dmask = orc.dmask_of(P0, m3)
    blocked = np.array([[(mk >> k) & 1 for k in range(n)] for mk in dmask], bool).reshape(-1)
    w[blocked] = 0.0
    with Context(m3, P0) as ctx:
        ctx.set_state(U, U)
        ge, _, _ = ctx.vjp(w)
        gf, _, _ = ctx.vjp(w, mode=abi.GMC_VJP_FROZEN)
    _say(f"3d nc=4 uniform state: exact vs frozen {rel(ge, gf):.2e}")
    assert rel(ge, gf) < 1e-7


def test_torch_residual_autograd(c0):
    torch = pytest.importorskip("torch")
    from gales_mc import adjoint
    mesh, P, I, Pst, ctx = c0
    _say("torch residual c0")
    dev = torch.device("cuda")
    It = torch.from_numpy(I).to(dev).requires_grad_(True)
    Pt = torch.from_numpy(Pst).to(dev).requires_grad_(True)
    th = torch.tensor(material.theta_of(P), dtype=torch.float64, requires_grad=True)
    R = adjoint.residual(It, Pt, th, ctx)
    ctx.set_state(I, Pst)
    assert rel(R.detach().cpu().numpy(), ctx.residual_vector()) < 1e-15
    v = torch.from_numpy(np.random.default_rng(24).standard_normal(I.size)).to(dev)
    (R * v).sum().backward()
    gs, gp, gpar = ctx.vjp(v.cpu().numpy(), want_prev=True)
    assert np.array_equal(It.grad.cpu().numpy(), gs)
    assert np.array_equal(Pt.grad.cpu().numpy(), gp)
    assert np.allclose(th.grad.numpy(), gpar, rtol=1e-15, atol=0)


def test_rollout_checkpointed_sensitivities(c0):
    """16 steps of the damped block-Jacobi integrator, objective = outlet mass
    flux of component 2 after the last step; gradients w.r.t. theta and I0
    with checkpoints every 4 steps equal the un-checkpointed ones bitwise and
    central differences to 1e-6.

    The scheme advances only the dofs of smooth elements (M = the same mask as
    the finite-difference directions): a wall element whose h is rounding
    noise would otherwise inject a jump into every later step.  Differences
    are taken at h = 1e-6, 1e-5 and 1e-4 relative (and, for parameters,
    Richardson-extrapolated from 0.5% and 1% steps); the best one counts
    (rounding dominates the small steps of weakly acting parameters, the
    min() switches of tau the large ones)."""
    torch = pytest.importorskip("torch")
    from gales_mc import adjoint
    mesh, P, I, Pst, ctx = c0
    _say("rollout c0")
    dev = torch.device("cuda")
    Binv = adjoint.block_diag_inverse(ctx, I, Pst)
    free = _smooth_mask(mesh, P, I)
    st = adjoint.Stepper(ctx, Binv, omega=0.25, mask=torch.from_numpy(free.astype(np.float64)).to(dev))
    obj = adjoint.outlet_mass_flux(mesh, P, ctx.ndof)
    th0 = np.array(material.theta_of(P))

    def run(I0, th, ck):
        It = torch.from_numpy(I0).to(dev).requires_grad_(True)
        Pt = torch.from_numpy(Pst).to(dev)
        tt = torch.tensor(th, dtype=torch.float64, requires_grad=True)
        L = adjoint.rollout(st, It, Pt, tt, 16, obj, checkpoint_every=ck)
        L.backward()
        return L.item(), It.grad.cpu().numpy(), tt.grad.numpy()

    def peak_of(ck):
        torch.cuda.synchronize()
        base = torch.cuda.memory_allocated()
        torch.cuda.reset_peak_memory_stats()
        out = run(I, th0, ck)
        torch.cuda.synchronize()
        return out, torch.cuda.max_memory_allocated() - base

    run(I, th0, 4)  # first use allocates library workspaces (bmm); keep them out of the peaks
    (L1, gI1, gT1), peak_full = peak_of(None)
    (L2, gI2, gT2), peak_ck = peak_of(4)
    _say(f"  L {L1:.10e}; peak memory above baseline: all 16 steps kept {peak_full / 2**20:.2f} MiB, "
         f"checkpoints every 4 {peak_ck / 2**20:.2f} MiB")
    assert L1 == L2 and np.array_equal(gI1, gI2) and np.array_equal(gT1, gT2)
    assert peak_ck < peak_full

    def L_of(I0, th):
        with torch.no_grad():
            It = torch.from_numpy(I0).to(dev)
            Pt = torch.from_numpy(Pst).to(dev)
            return adjoint.rollout(st, It, Pt, torch.tensor(th, dtype=torch.float64), 16, obj).item()

    def dtheta(k, hr):
        hk = hr * th0[k]
        tp, tm = th0.copy(), th0.copy()
        tp[k] += hk
        tm[k] -= hk
        return (L_of(I, tp) - L_of(I, tm)) / (2 * hk)

    for k in range(6):
        fds = [dtheta(k, hr) for hr in (1e-6, 1e-5, 1e-4)]
        # Richardson extrapolation of the 0.5% and 1% steps (error O(h^4)):
        # the only usable estimate for the gas viscosity, whose effect on L
        # is below the rounding of the rollout at small steps
        fds.append((4.0 * dtheta(k, 5e-3) - dtheta(k, 1e-2)) / 3.0)
        errs = [abs(gT1[k] - fd) / abs(fd) for fd in fds]
        _say(f"  dL/dtheta[{k}] = {gT1[k]:.8e}; rel err vs FD at h = 1e-6/1e-5/1e-4/Richardson(1e-2): "
             + " ".join(f"{x:.1e}" for x in errs))
        # the gas viscosity (1e-5 Pa s) moves L by ~1e-9 relative per 1%:
        # every difference quotient is rounding-limited to ~1e-4 there
        assert min(errs) <= (1e-3 if k == 4 else 1e-6), (k, errs)
    sc = _scale(P, mesh["coords"].shape[0])
    rng = np.random.default_rng(25)
    for k in range(4):
        d = sc * rng.standard_normal(I.size) * free
        errs = []
        for hs in (1e-6, 1e-5, 1e-4):
            fd = (L_of(I + hs * d, th0) - L_of(I - hs * d, th0)) / (2 * hs)
            errs.append(abs(gI1 @ d - fd) / abs(fd))
        _say(f"  dL/dI0 . d[{k}] = {gI1 @ d:.8e}; rel err vs FD: " + " ".join(f"{x:.1e}" for x in errs))
        assert min(errs) <= 1e-6, (k, errs)
    ctx.set_material(th0)
    ctx.set_state(I, Pst)
