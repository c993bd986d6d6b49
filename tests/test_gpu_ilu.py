"""Device GMRES with the block ILU(0) preconditioner (Ifpack "ILU" of the
reference's solver set-up, belos_linear_solver_factory.hpp:141-160, at fill
level 0) against the oracle's GMRES with its own block ILU(0).

Both solves run to a relative residual of 1e-12; the solutions then agree to
the problem's conditioning (the bar is written in each test), and the device
solution's true residual meets the tolerance by itself.
"""
import sys

import numpy as np
import pytest

import orc
from helpers import rel

from gales_mc import abi, cases
from gales_mc.lib import Context

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("cid,scale", [(0, 1), (1, 16), (2, 16)])
def test_ilu_gmres_matches_oracle(cid, scale):
    _say(f"ILU(0) GMRES c{cid}/{scale}")
    mesh, P, I, Pst = cases.config(cid, scale=scale)
    A = orc.Assembler(P, mesh)
    vals, rhs = A.assemble(I, Pst)
    x_ref, it_ref, rr_ref = A.gmres(vals, rhs, restart=50, max_iters=2000, rel_tol=1e-12, precond=3)
    with Context(mesh, P) as ctx:
        ctx.set_state(I, Pst)
        ctx.assemble()
        si = ctx.solve(restart=50, max_iters=2000, rel_tol=1e-12, precond=abi.GMC_PRECOND_ILU0)
        x = ctx.solution()
        _, rhs_d = ctx.system()
    r_true = rel(A.spmv(vals, x), rhs_d)
    d = rel(x, x_ref)
    _say(f"  device {si.iterations} it, rel res {si.rel_residual:.2e} (true vs oracle matrix {r_true:.2e}); "
         f"oracle {it_ref} it, {rr_ref:.2e}; |dx|/|x| {d:.2e}")
    assert si.rel_residual <= 1e-12 and rr_ref <= 1e-12
    assert r_true < 1e-11
    # two solutions with residuals <= 1e-12 differ by at most cond(J) 2e-12
    assert d < 1e-7


def test_ilu_beats_block_jacobi_on_the_jittered_channel():
    mesh, P, I, Pst = cases.config(1, scale=16)
    _say(f"ILU(0) vs block Jacobi, c1/16 ({mesh['conn'].shape[0]} elements)")
    with Context(mesh, P) as ctx:
        ctx.set_state(I, Pst)
        ctx.assemble()
        sj = ctx.solve(restart=50, max_iters=300, rel_tol=1e-10, precond=2)
        si = ctx.solve(restart=50, max_iters=300, rel_tol=1e-10, precond=abi.GMC_PRECOND_ILU0)
    _say(f"  block Jacobi {sj.iterations} it -> {sj.rel_residual:.2e}; ILU(0) {si.iterations} it -> "
         f"{si.rel_residual:.2e} in {si.seconds:.2f} s")
    assert si.rel_residual <= 1e-10 and sj.rel_residual > 1e-6


def test_ilu_refinement_reaches_1e10_on_the_fp32_tier():
    mesh, P, I, Pst = cases.config(1, scale=16)
    P.precision = abi.GMC_PREC_FP32
    with Context(mesh, P) as ctx:
        ctx.set_state(I, Pst)
        res = ctx.refine(n_refine=4, restart=50, max_iters=300, rel_tol=1e-3, precond=abi.GMC_PRECOND_ILU0)
    hist = res / res[0]
    _say("  FP32-tier refinement with ILU(0) inner solves: " + " ".join(f"{h:.2e}" for h in hist))
    assert hist[-1] <= 1e-10
