"""Sensitivities through one implicit time step (the reference's Newton loop,
sim/fluid_mc/P0/main.cpp:206-240): gales_mc.adjoint.newton_step solves
R(I*, P; theta) = 0 by Newton iterations on the assembled FP64 system, and
implicit_step_vjp gives df/dP and df/dtheta of f(I*) from one transposed
solve with the exact J^T (matrix-free gmc_vjp) and one more exact VJP.
Checked against central differences of the whole Newton solve."""
import numpy as np
import pytest


from gales_mc import abi, cases, material
from gales_mc.adjoint import dirichlet_dofs, implicit_step_vjp, newton_step
from gales_mc.lib import Context


def say(*a):
    print(*a, flush=True)


def _case(dt):
    mesh = cases.channel2d(16, 8, seed=5)
    P = cases.make_params(2, [cases.MELT, cases.GAS_R461], dt=dt)
    _, Pst = cases.fields(mesh, 2, 7)
    return mesh, P, Pst.reshape(-1)


def _richardson(fun, h):
    """Central difference extrapolated in h: (4 D(h/2) - D(h)) / 3."""
    d1 = (fun(h) - fun(-h)) / (2 * h)
    d2 = (fun(h / 2) - fun(-h / 2)) / h
    return (4 * d2 - d1) / 3


@pytest.mark.gpu
def test_newton_step_converges_and_keeps_dirichlet_values():
    mesh, P, Pst = _case(1e-4)
    D = dirichlet_dofs(mesh, P, 5)
    with Context(mesh, P) as ctx:
        Is, hist = newton_step(ctx, Pst, np.array(material.theta_of(P)))
        say(f"newton: {len(hist) - 1} iterations, ||rhs|| {hist[0]:.3e} -> {hist[-1]:.3e}")
        assert hist[-1] <= 1e-12 * hist[0]
        assert np.array_equal(Is[D], Pst[D])
        ctx.set_state(Is, Pst)
        ctx.assemble()
        assert ctx.rhs_norm() <= 1e-12 * hist[0]


@pytest.mark.gpu
@pytest.mark.parametrize("dt", [1e-4])
def test_implicit_step_vjp_finite_differences(dt):
    mesh, P, Pst = _case(dt)
    theta = np.array(material.theta_of(P))
    rng = np.random.default_rng(3)
    with Context(mesh, P) as ctx:
        n = ctx.n_nodes * ctx.ndof
        g = rng.standard_normal(n)
        Is, _ = newton_step(ctx, Pst, theta)
        dP, dth, info = implicit_step_vjp(ctx, mesh, P, Is, Pst, theta, g)
        say(f"implicit step adjoint: GMRES {info['gmres_iterations']} iterations, "
            f"J^T lambda residual {info['rel_residual']:.2e}")
        assert info["rel_residual"] <= 1e-9

        def f_theta(k, h):
            t = theta.copy()
            t[k] += h
            return float(g @ newton_step(ctx, Pst, t)[0])

        worst = 0.0
        for k in range(6):
            h = 1e-4 * max(abs(theta[k]), 1.0)
            fd = _richardson(lambda s: f_theta(k, s), h)
            e = abs(dth[k] - fd) / abs(fd)
            say(f"  theta[{k}]: adjoint {dth[k]:.9e}  difference {fd:.9e}  rel {e:.2e}")
            worst = max(worst, e)
        assert worst <= 1e-5
        scale = np.tile(np.array([1e4, 0.1, 0.1, 1.0, 1e-2]), ctx.n_nodes)
        for r in range(3):
            d = rng.standard_normal(n) * scale
            fd = _richardson(lambda s: float(g @ newton_step(ctx, Pst + s * d, theta)[0]), 1e-3)
            e = abs(dP @ d - fd) / abs(fd)
            say(f"  P direction {r}: adjoint {dP @ d:.9e}  difference {fd:.9e}  rel {e:.2e}")
            assert e <= 1e-6
        # the Dirichlet part: f depends on P_D through I*_D = P_D and through R_F
        assert np.any(dP[dirichlet_dofs(mesh, P, ctx.ndof)] != 0.0)


def test_dirichlet_dofs_follow_the_node_flags():
    mesh = cases.channel2d(8, 4)
    P = cases.make_params(2, [cases.MELT, cases.GAS_R461])
    D = dirichlet_dofs(mesh, P, 5).reshape(-1, 5)
    flags = np.asarray(mesh["node_flag"])
    for n in range(flags.size):
        m = P.dirichlet_mask[flags[n]]
        assert [bool((m >> k) & 1) for k in range(5)] == list(D[n])
    assert D[flags == cases.WALL][:, 1:3].all() and not D[flags == cases.WALL][:, 0].any()
    assert abi.GMC_MAX_FLAG == 16
