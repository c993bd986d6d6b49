"""Adjoint on the GPU, through the C-ABI and the torch binding.

  frozen VJP  = J^T v of the assembled FP64 matrix (oracle's, scipy transpose)
                within 1e-12, on config 0 and the jittered config-3 mesh
  exact VJP   = finite differences of the GPU FP64 residual on 32 seeded
                directions, 5-point stencil: |g.d - fd| <= 1e-6 |fd|, nothing
                added; on config 0, the jittered config-3 family at 1/32 and
                1/8 resolution and the 3D 4-component box at 1/16; previous-
                step and all six material directions likewise
  exact = frozen where the reference's matrix is the exact Jacobian (uniform
                state, no gravity)
  torch       autograd through Residual and a 16-step checkpointed rollout
                against finite differences, on the jittered config-3 mesh

The reference's residual is only piecewise smooth: the min() branches of tau,
the absolute values of the DC length and the direction switch of compute_h
put kinks in it.  Finite-difference directions therefore leave out the dofs
of elements with an interior point within a relative margin of 1e-2 of such
a switch (helpers.kink_margin), of elements where the residual is not
differentiable at all (helpers.residual_nonsmooth), and Dirichlet-blocked
dofs.
"""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

import orc
from helpers import kink_margin, rel, residual_nonsmooth

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


KINK = 1e-2
CASES = {"c0": (0, 1), "c3/32": (3, 32), "c3/8": (3, 8), "c4/16": (4, 16)}


def _case(name):
    cid, scale = CASES[name]
    mesh, P, I, Pst = cases.config(cid, scale=scale)
    return mesh, P, I.reshape(-1).copy(), Pst.reshape(-1).copy()


def _smooth_mask(mesh, P, I, Pst=None):
    """dofs that finite-difference directions may move."""
    n = P.dim + 2 + P.nc - 1
    dmask = orc.dmask_of(P, mesh)
    free = np.ones((mesh["coords"].shape[0], n), bool)
    for k in range(n):
        free[:, k] = ((dmask >> k) & 1) == 0
    conn = mesh["conn"]
    Ir = np.asarray(I).reshape(-1, n)
    Pr = Ir if Pst is None else np.asarray(Pst).reshape(-1, n)
    dropped = 0
    for e in range(conn.shape[0]):
        xe, Ie, Pe = orc.element_data(mesh, Ir, Pr, e)
        if residual_nonsmooth(P, xe, Ie) or kink_margin(P, xe, Ie, Pe) < KINK:
            free[conn[e]] = False
            dropped += 1
    _say(f"  directions avoid {dropped} of {conn.shape[0]} elements (kinks within {KINK:g})")
    return free.reshape(-1)


def _fd5(Rfun, x, d, h, v):
    """v . dR/dx . d by the 5-point stencil; the residual vectors are
    differenced before the dot product (v.R alone sums 1e4 terms of 1e7
    magnitude and would round the difference away)."""
    d1 = Rfun(x + h * d) - Rfun(x - h * d)
    d2 = Rfun(x + 2 * h * d) - Rfun(x - 2 * h * d)
    return (v @ (8.0 * d1 - d2)) / (12.0 * h), (v @ d1) / (2.0 * h)


def _scale(P, n_nodes):
    return np.tile([1e4] + [1.0] * P.dim + [5.0] + [0.05] * (P.nc - 1), n_nodes)


@pytest.mark.parametrize("name", ["c0", "c3/32", "c3/8"])
def test_vjp_frozen_is_assembled_transpose(name):
    mesh, P, I, Pst = _case(name)
    _say(f"vjp frozen {name}")
    v = np.random.default_rng(21).standard_normal(I.size)
    A = orc.Assembler(P, mesh)
    vals_ref, _ = A.assemble(I, Pst)
    with Context(mesh, P) as ctx:
        ctx.set_state(I, Pst)
        g, gp, gpar = ctx.vjp(v, mode=abi.GMC_VJP_FROZEN)
        assert gp is None and gpar is None
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


@pytest.mark.parametrize("name", ["c0", "c3/32", "c3/8", "c4/16"])
def test_vjp_exact_state_finite_differences(name):
    """32 seeded directions in the current state (scaled by 1e4 Pa, 1 m/s,
    5 K, 0.05 per dof), 5-point stencil at h = 1e-6 and 1e-5; the better of
    the two must agree to 1e-6.  Each estimate has one dominant error:
    rounding at 1e-6 (the pressure terms of one element's residual are ~5e4
    and cancel between elements, leaving ~1e-10 of rounding per residual
    entry against a change of ~1e-4 per step, ~1e-6 relative), curvature at
    1e-5 (near a no-slip wall |v| is small, and a 1e-5 m/s step turns the
    streamline direction of compute_h by ~1%).  A wrong VJP fails both.  The
    central difference at 1e-6 is printed too."""
    mesh, P, I, Pst = _case(name)
    _say(f"vjp exact state FD {name}")
    rng = np.random.default_rng(22)
    v = rng.standard_normal(I.size)
    free = _smooth_mask(mesh, P, I, Pst)
    sc = _scale(P, mesh["coords"].shape[0])
    with Context(mesh, P) as ctx:
        ctx.set_state(I, Pst)
        gs, gp, gpar = ctx.vjp(v, want_prev=True)
        gs2, gp2, gpar2 = ctx.vjp(v, want_prev=True)
        assert np.array_equal(gs, gs2) and np.array_equal(gp, gp2) and np.array_equal(gpar, gpar2)

        def Rvec(Ii):
            ctx.set_state(Ii, Pst)
            return ctx.residual_vector()

        errs, errs6, errs_c = [], [], []
        for k in range(32):
            d = sc * rng.standard_normal(I.size) * free
            fd, _ = _fd5(Rvec, I, d, 1e-5, v)
            fd6, fdc = _fd5(Rvec, I, d, 1e-6, v)
            errs.append(abs(gs @ d - fd) / abs(fd))
            errs6.append(abs(gs @ d - fd6) / abs(fd6))
            errs_c.append(abs(gs @ d - fdc) / abs(fdc))
        best = np.minimum(errs, errs6)
        _say(f"  worst rel err over 32 directions: 5-point best of the two steps {best.max():.2e} "
             f"(median {np.median(best):.2e}); h = 1e-5 {max(errs):.2e}, h = 1e-6 {max(errs6):.2e}, "
             f"central h = 1e-6 {max(errs_c):.2e}")
        assert best.max() <= 1e-6, best


# 5-point steps of the material directions, relative to the parameter: the
# gas viscosity (1e-5 Pa s, slot 4) moves R by ~1e-7 of what the melt
# viscosity does, below the rounding of R at small steps, so it takes 1e-2
# (truncation of the 5-point stencil O(h^4) ~ 1e-8)
THETA_STEP = [1e-6, 1e-6, 1e-6, 1e-6, 1e-2, 1e-6]


@pytest.mark.parametrize("name", ["c0", "c3/32"])
def test_vjp_exact_prev_and_material_finite_differences(name):
    mesh, P, I, Pst = _case(name)
    _say(f"vjp exact previous-step / material FD {name}")
    rng = np.random.default_rng(23)
    v = rng.standard_normal(I.size)
    sc = _scale(P, mesh["coords"].shape[0])
    with Context(mesh, P) as ctx:
        ctx.set_state(I, Pst)
        gs, gp, gpar = ctx.vjp(v, want_prev=True)

        def Rprev(Pi):
            ctx.set_state(I, Pi)
            return ctx.residual_vector()

        # R is smooth in P (time terms only), so the step is 1e-4: at 1e-6 the
        # estimate is rounding-limited as for the state directions
        worst = worst6 = 0.0
        for k in range(8):
            d = sc * rng.standard_normal(I.size)
            fd, _ = _fd5(Rprev, Pst, d, 1e-4, v)
            fd6, _ = _fd5(Rprev, Pst, d, 1e-6, v)
            worst = max(worst, abs(gp @ d - fd) / abs(fd))
            worst6 = max(worst6, abs(gp @ d - fd6) / abs(fd6))
        _say(f"  previous step: worst rel err over 8 directions {worst:.2e} (5-point h = 1e-4), "
             f"{worst6:.2e} (h = 1e-6)")
        assert worst <= 1e-6
        th = np.array(material.theta_of(P))
        ctx.set_state(I, Pst)

        def Rtheta(t):
            ctx.set_material(t)
            ctx.set_state(I, Pst)
            return ctx.residual_vector()

        for k in range(6):
            e = np.zeros(6)
            e[k] = th[k]
            fd, _ = _fd5(Rtheta, th, e, THETA_STEP[k], v)
            err = abs(gpar[k] * th[k] - fd) / abs(fd)
            _say(f"  theta[{k}]: vjp {gpar[k]:.8e}; rel err vs 5-point FD (h = {THETA_STEP[k]:g}) {err:.1e}")
            assert err <= 1e-6, (k, err)
        ctx.set_material(th)


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


def test_rollout_checkpointed_sensitivities():
    """16 steps of the damped block-Jacobi integrator on the jittered config-3
    mesh (1/32 resolution), objective = outlet mass flux of component 2 after
    the last step; gradients w.r.t. theta and I0 with checkpoints every 4
    steps equal the un-checkpointed ones bitwise and finite differences to
    1e-6.

    The gas viscosity is raised from 1e-5 to 10 Pa s here: at 1e-5 it moves
    the objective by ~1e-9 relative per 1% change, below the rollout's
    rounding, so no difference quotient could check its gradient.  The time
    step is 1e-4, so that tau's time branch min(dt/2, ...) is the one taken at
    every point along the 16 steps: with dt = 1e-3 the viscous branch
    h^2 rho / (12 mu) switches inside the range the trajectory sweeps, and
    perturbed trajectories cross the switches.  The scheme
    advances only the dofs of elements away from the residual's kinks (M =
    the same mask as the finite-difference directions).  Each difference is
    the 5-point stencil at h = 1e-6, 1e-5 and 1e-4 relative; the best of the
    three counts (rounding dominates the smallest step, the min() switches of
    tau met along 16 steps the largest)."""
    torch = pytest.importorskip("torch")
    from gales_mc import adjoint
    mesh, P, I, Pst = _case("c3/32")
    P.dt = 1e-4
    _say("rollout c3/32")
    dev = torch.device("cuda")
    th0 = np.array(material.theta_of(P))
    th0[4] = 10.0
    with Context(mesh, P) as ctx:
        ctx.set_material(th0)
        Binv = adjoint.block_diag_inverse(ctx, I, Pst)
        free = _smooth_mask(mesh, P, I, Pst)
        st = adjoint.Stepper(ctx, Binv, omega=0.25, mask=torch.from_numpy(free.astype(np.float64)).to(dev))
        obj = adjoint.outlet_mass_flux(mesh, P, ctx.ndof)

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

        def fd5(f, x, d, h):
            return (8.0 * (f(x + h * d) - f(x - h * d)) - (f(x + 2 * h * d) - f(x - 2 * h * d))) / (12.0 * h)

        for k in range(6):
            e = np.zeros(6)
            e[k] = th0[k]
            errs = [abs(gT1[k] * th0[k] - fd) / abs(fd)
                    for fd in (fd5(lambda t: L_of(I, t), th0, e, h) for h in (1e-6, 1e-5, 1e-4))]
            _say(f"  dL/dtheta[{k}] = {gT1[k]:.8e}; rel err vs 5-point FD at h = 1e-6/1e-5/1e-4: "
                 + " ".join(f"{x:.1e}" for x in errs))
            assert min(errs) <= 1e-6, (k, errs)
        sc = _scale(P, mesh["coords"].shape[0])
        rng = np.random.default_rng(25)
        for k in range(4):
            d = sc * rng.standard_normal(I.size) * free
            errs = [abs(gI1 @ d - fd) / abs(fd)
                    for fd in (fd5(lambda x: L_of(x, th0), I, d, h) for h in (1e-6, 1e-5, 1e-4))]
            _say(f"  dL/dI0 . d[{k}] = {gI1 @ d:.8e}; rel err vs 5-point FD: " + " ".join(f"{x:.1e}" for x in errs))
            assert min(errs) <= 1e-6, (k, errs)


def test_rollout_checkpointed_sensitivities_3d():
    """The checkpointed rollout in 3D (config 4 at 1/16: tetrahedra, four
    components): 8 steps of the damped block-Jacobi integrator, objective a
    fixed linear functional w . I_8 of the final state.  Gradients with
    checkpoints every 4 steps equal the un-checkpointed ones bitwise; the
    theta gradient of the density and viscosity parameters agrees with
    5-point differences of the whole rollout (best of h = 1e-6, 1e-5, 1e-4
    relative), the directions that move the state without crossing the
    residual's kinks over 8 steps."""
    torch = pytest.importorskip("torch")
    from gales_mc import adjoint
    mesh, P, I, Pst = _case("c4/16")
    P.dt = 1e-4
    _say("rollout 3D c4/16")
    dev = torch.device("cuda")
    th0 = np.array(material.theta_of(P))
    rng = np.random.default_rng(31)
    with Context(mesh, P) as ctx:
        Binv = adjoint.block_diag_inverse(ctx, I, Pst)
        free = _smooth_mask(mesh, P, I, Pst)
        st = adjoint.Stepper(ctx, Binv, omega=0.25, mask=torch.from_numpy(free.astype(np.float64)).to(dev))
        w = torch.from_numpy(rng.standard_normal(I.size) / np.maximum(np.abs(I), 1.0)).to(dev)

        def obj(It, th):
            return (w * It).sum()

        def run(ck):
            It = torch.from_numpy(I).to(dev).requires_grad_(True)
            Pt = torch.from_numpy(Pst).to(dev)
            tt = torch.tensor(th0, dtype=torch.float64, requires_grad=True)
            L = adjoint.rollout(st, It, Pt, tt, 8, obj, checkpoint_every=ck)
            L.backward()
            return L.item(), It.grad.cpu().numpy(), tt.grad.numpy()

        L1, gI1, gT1 = run(None)
        L2, gI2, gT2 = run(4)
        _say(f"  L {L1:.10e}, |dL/dI0| {np.linalg.norm(gI1):.3e}, dL/dtheta {gT1}")
        assert L1 == L2 and np.array_equal(gI1, gI2) and np.array_equal(gT1, gT2)
        assert np.all(np.isfinite(gI1)) and np.all(np.isfinite(gT1)) and np.any(gI1 != 0.0)

        def L_of(th):
            with torch.no_grad():
                It = torch.from_numpy(I).to(dev)
                Pt = torch.from_numpy(Pst).to(dev)
                return adjoint.rollout(st, It, Pt, torch.tensor(th, dtype=torch.float64), 8, obj).item()

        def fd5(f, x, d, h):
            return (8.0 * (f(x + h * d) - f(x - h * d)) - (f(x + 2 * h * d) - f(x - 2 * h * d))) / (12.0 * h)

        for k in (0, 1, 3, 4):
            e = np.zeros(6)
            e[k] = th0[k]
            errs = [abs(gT1[k] * th0[k] - fd) / max(abs(fd), 1e-300)
                    for fd in (fd5(L_of, th0, e, h) for h in (1e-6, 1e-5, 1e-4))]
            _say(f"  dL/dtheta[{k}] = {gT1[k]:.8e}; rel err vs 5-point FD at h = 1e-6/1e-5/1e-4: "
                 + " ".join(f"{x:.1e}" for x in errs))
            assert min(errs) <= 1e-6, (k, errs)
