"""Precision tier (FP32 element matrices, FP64 residual and accumulation), the
matrix-free FP64 Jacobian product and mixed-precision iterative refinement,
through the C-ABI, against the FP64 oracle.

Stated errors against the FP64 oracle, per output:
  R (rhs)               FP64 in every tier: <= 1e-12 relative
  J (assembled matrix)  FP32 elements, FP64 sums: <= 1e-6 relative (~2e-7 seen)
  J x (matrix-free)     FP64: <= 1e-12 relative
  Newton update         after refinement, ||rhs - J64 x|| / ||rhs|| <= 1e-10,
                        and within 1e-7 of the FP64 path's update
"""
import sys

import numpy as np
import pytest

import orc
from helpers import rel

from gales_mc import abi, cases
from gales_mc.lib import Context

pytestmark = pytest.mark.gpu
J32_TOL = 1e-6


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _case(cid, scale=1, precision=abi.GMC_PREC_FP64):
    mesh, P, I, Pst = cases.config(cid, scale=scale)
    P.precision = precision
    return mesh, P, I, Pst


@pytest.fixture(scope="module")
def c0():
    mesh, P, I, Pst = _case(0)
    ctx = Context(mesh, P)
    ctx.set_state(I, Pst)
    A = orc.Assembler(P, mesh)
    vals_ref, rhs_ref = A.assemble(I, Pst)
    yield mesh, P, I, Pst, ctx, A, vals_ref, rhs_ref
    ctx.close()


@pytest.mark.parametrize("cid,scale", [(0, 1), (2, 16), (4, 16)])
def test_fp32_element(cid, scale):
    mesh, P, I, Pst = _case(cid, scale, abi.GMC_PREC_FP32)
    _say(f"fp32 element c{cid}/{scale}")
    u = orc.uref_inv(P)
    E = mesh["conn"].shape[0]
    with Context(mesh, P) as ctx:
        ctx.set_state(I, Pst)
        worst = 0.0
        for e in sorted(set(np.random.default_rng(7).integers(0, E, 16).tolist() + [0, E - 1])):
            xe, Ie, Pe = orc.element_data(mesh, I, Pst, e)
            m_ref, r_ref = orc.element_interior(P, xe, Ie, Pe, u)
            m, r = ctx.element(e)
            em = rel(m, m_ref)
            worst = max(worst, em)
            assert em < J32_TOL, (e, em)
            assert rel(r, r_ref) < 1e-12, (e, rel(r, r_ref))
        _say(f"  worst element-matrix error {worst:.2e}")
        assert worst > 1e-9  # the matrix really is formed in float


@pytest.mark.parametrize("cid,scale,scatter", [(0, 1, ""), (2, 16, ""), (2, 16, "ebuf32"),
                                               (4, 16, "ebuf"), (4, 16, "ebuf32")])
def test_fp32_assembly(cid, scale, scatter, monkeypatch):
    """FP32-tier J and FP64 rhs against the FP64 oracle.  The 3D cases run the
    element buffer with FP64 entries (ebuf) and with FP32 entries (ebuf32,
    what config 4 uses at full size); 8 dofs (c4) take the unrolled FP32
    kernel, 7 dofs (c2) the rolled one."""
    mesh, P, I, Pst = _case(cid, scale)
    if scatter:
        monkeypatch.setenv("GMC_SCATTER", scatter)
    _say(f"fp32 assembly c{cid}/{scale} scatter={scatter or 'default'}")
    A = orc.Assembler(P, mesh)
    vals_ref, rhs_ref = A.assemble(I, Pst)
    with Context(mesh, P) as ctx:
        if scatter:
            # FP64 assemblies: element buffer (1); with the FP32-entry buffer
            # they take the colour scatter (0) and only the tier uses the buffer
            assert ctx.info.scatter == (1 if scatter == "ebuf" else 0)
        ctx.set_state(I, Pst)
        ctx.assemble(abi.GMC_ASM_MATRIX | abi.GMC_ASM_RHS | abi.GMC_ASM_FP32)
        vals, rhs = ctx.system()
        ev, er = rel(vals, vals_ref), rel(rhs, rhs_ref)
        _say(f"  J rel {ev:.2e}  rhs rel {er:.2e}")
        assert ev < J32_TOL and er < 1e-12
        ctx.assemble(abi.GMC_ASM_MATRIX | abi.GMC_ASM_RHS | abi.GMC_ASM_FP32)
        v2, r2 = ctx.system()
        assert np.array_equal(vals, v2) and np.array_equal(rhs, r2)


@pytest.mark.parametrize("cid,scale,scatter,chunks", [(0, 1, "patch", 0), (1, 8, "patch", 0),
                                                      (3, 16, "patch", 0), (0, 1, "ebuf", 0),
                                                      (1, 8, "ebuf", 0), (3, 16, "ebuf", 0),
                                                      (1, 8, "ebuf", 3)])
def test_fp32_assembly_mfma(cid, scale, scatter, chunks, monkeypatch):
    """FP32-tier J with the least-squares term on the matrix cores
    (GMC_FP32_MFMA=1: v_mfma_f32_16x16x4f32, triangles with 2 components),
    patch-local and element-buffer scatter, against the FP64 oracle and
    against the same tier's VALU kernel, for full, matrix-only and
    residual-only assemblies; bitwise reproducible.  chunks: elements through
    the buffer in that many chunks.  Config 3 is the jittered mesh."""
    mesh, P, I, Pst = _case(cid, scale)
    monkeypatch.setenv("GMC_SCATTER", scatter)
    if chunks:
        monkeypatch.setenv("GMC_EBUF_CHUNKS", str(chunks))
    _say(f"fp32 mfma assembly c{cid}/{scale} scatter={scatter} chunks={chunks}")
    A = orc.Assembler(P, mesh)
    vals_ref, rhs_ref = A.assemble(I, Pst)
    full = abi.GMC_ASM_MATRIX | abi.GMC_ASM_RHS | abi.GMC_ASM_FP32
    with Context(mesh, P) as ctx:
        assert ctx.info.fp32_mfma == 0
        ctx.set_state(I, Pst)
        ctx.assemble(full)
        vals_v, _ = ctx.system()
    monkeypatch.setenv("GMC_FP32_MFMA", "1")
    with Context(mesh, P) as ctx:
        assert ctx.info.fp32_mfma == 1 and ctx.info.scatter == (2 if scatter == "patch" else 1)
        ctx.set_state(I, Pst)
        ctx.assemble(full)
        vals, rhs = ctx.system()
        ev, er, evv = rel(vals, vals_ref), rel(rhs, rhs_ref), rel(vals, vals_v)
        _say(f"  J rel {ev:.2e} (VALU kernel {rel(vals_v, vals_ref):.2e}, MFMA vs VALU {evv:.2e})"
             f"  rhs rel {er:.2e}")
        assert ev < J32_TOL and er < 1e-12 and evv < J32_TOL
        assert evv > 0.0  # the products really come from the other kernel
        ctx.assemble(full)
        v2, r2 = ctx.system()
        assert np.array_equal(vals, v2) and np.array_equal(rhs, r2)
        ctx.assemble(abi.GMC_ASM_MATRIX | abi.GMC_ASM_FP32)
        v3, _ = ctx.system()
        assert rel(v3, vals_ref) < J32_TOL, rel(v3, vals_ref)
        ctx.assemble(abi.GMC_ASM_RHS | abi.GMC_ASM_FP32)
        _, r3 = ctx.system()
        assert rel(r3, rhs_ref) < 1e-12, rel(r3, rhs_ref)


def test_fp32_mfma_16_element_patches(monkeypatch):
    """The matrix-core FP32 tier on 16-element patches (one wave per
    workgroup): the star mesh's node of valence 200 leaves no 64-element
    plan.  Against the FP64 oracle."""
    from test_gpu_parity import _star_mesh
    monkeypatch.setenv("GMC_FP32_MFMA", "1")
    mesh = _star_mesh(200)
    P = cases.make_params(2, [cases.MELT, cases.GAS_R461])
    I, Pst = cases.fields(mesh, 2, 5)
    _say("fp32 mfma, star mesh")
    A = orc.Assembler(P, mesh)
    vals_ref, rhs_ref = A.assemble(I, Pst)
    with Context(mesh, P) as ctx:
        assert ctx.info.scatter == 2 and ctx.info.fp32_mfma == 1
        ctx.set_state(I, Pst)
        ctx.assemble(abi.GMC_ASM_MATRIX | abi.GMC_ASM_RHS | abi.GMC_ASM_FP32)
        vals, rhs = ctx.system()
        _say(f"  J rel {rel(vals, vals_ref):.2e}  rhs rel {rel(rhs, rhs_ref):.2e}")
        assert rel(vals, vals_ref) < J32_TOL and rel(rhs, rhs_ref) < 1e-12


def test_jvp_matches_assembled_matrix(c0):
    """Matrix-free FP64 J x equals the assembled FP64 matrix times x and the
    oracle's, and does not depend on the matrix stored in the context."""
    mesh, P, I, Pst, ctx, A, vals_ref, rhs_ref = c0
    _say("jvp c0")
    ctx.set_state(I, Pst)
    x = np.random.default_rng(11).normal(size=I.size) * np.tile([1e4, 1.0, 1.0, 5.0, 0.1], I.size // 5)
    ctx.assemble()
    y_mat = ctx.spmv(x)
    y_free = ctx.jvp(x)
    y_ref = A.spmv(vals_ref, x)
    _say(f"  jvp vs spmv {rel(y_free, y_mat):.2e}, vs oracle {rel(y_free, y_ref):.2e}")
    assert rel(y_free, y_mat) < 1e-12 and rel(y_free, y_ref) < 1e-12
    ctx.assemble(abi.GMC_ASM_MATRIX | abi.GMC_ASM_RHS | abi.GMC_ASM_FP32)
    assert np.array_equal(ctx.jvp(x), y_free)


def test_jvp_3d():
    mesh, P, I, Pst = _case(2, 16)
    _say("jvp c2/16")
    A = orc.Assembler(P, mesh)
    vals_ref, _ = A.assemble(I, Pst)
    x = np.random.default_rng(12).normal(size=I.size)
    with Context(mesh, P) as ctx:
        ctx.set_state(I, Pst)
        assert rel(ctx.jvp(x), A.spmv(vals_ref, x)) < 1e-12


def test_gmres_block_jacobi_c0(c0):
    """GMRES(30) with block-Jacobi right preconditioning matches the oracle's
    run of the same method on the oracle's system."""
    mesh, P, I, Pst, ctx, A, vals_ref, rhs_ref = c0
    _say("gmres block jacobi c0")
    ctx.set_state(I, Pst)
    ctx.assemble()
    info = ctx.solve(restart=30, max_iters=60, rel_tol=0.0, precond=2)
    d_gpu = ctx.solution()
    d_ref, it_ref, rr_ref = A.gmres(vals_ref, rhs_ref, restart=30, max_iters=60, rel_tol=0.0, precond=2)
    _say(f"  it {info.iterations} vs {it_ref}, rel res {info.rel_residual:.3e} vs {rr_ref:.3e}, "
         f"delta diff {rel(d_gpu, d_ref):.2e}")
    assert info.iterations == it_ref == 60
    assert abs(info.rel_residual - rr_ref) <= 1e-6 * rr_ref + 1e-14
    assert rel(d_gpu, d_ref) < 1e-9


def _refine_solver():
    # inner solves to 1e-6: each refinement pass gains ~6 digits
    return dict(restart=120, max_iters=2000, rel_tol=1e-6, precond=2)


def test_refine_recovers_fp64_newton_update(c0):
    """Three refinement passes with the FP32 matrix solve the FP64 Newton
    equation to <= 1e-10 relative (checked with the oracle's own matrix), and
    the update agrees with the FP64 path's (FP64 matrix, GMRES to 1e-12) to
    1e-7 relative.  That bound is the FP64 path's own spread: with the
    oracle, GMRES(1e-12) and a sparse LU solve of the same FP64 system differ
    by 2.7e-9 relative on this case, because J is ill-conditioned.

    The nonlinear residual after the update is reported, not compared: the
    reference's Jacobian freezes tau, nu and the properties and omits the
    source derivative, so R(I + x) reacts to components of x that J barely
    sees; in FP64 alone the two solves above leave residuals 5.6% apart."""
    mesh, P, I, Pst, ctx, A, vals_ref, rhs_ref = c0
    _say("refine c0")
    ctx.set_state(I, Pst)
    ctx.assemble()
    info = ctx.solve(restart=120, max_iters=4000, rel_tol=1e-12, precond=2)
    _say(f"  fp64 gmres it {info.iterations} rel {info.rel_residual:.2e}")
    assert info.rel_residual <= 1e-12
    x64 = ctx.solution()
    ctx.set_state(I, Pst)
    res = ctx.refine(n_refine=3, **_refine_solver())
    hist = res / res[0]
    _say("  refinement history " + " ".join(f"{h:.2e}" for h in hist))
    assert hist[-1] <= 1e-10
    assert all(hist[k + 1] < hist[k] for k in range(len(hist) - 1) if hist[k] > 1e-14)
    x32 = ctx.solution()
    _, rhs = ctx.system()
    assert rel(rhs, rhs_ref) < 1e-12            # -R stays FP64
    lin = rel(A.spmv(vals_ref, x32), rhs_ref)  # FP64 Newton equation, oracle matrix
    dx = rel(x32, x64)
    _say(f"  oracle ||rhs - J64 x32||/||rhs|| {lin:.2e}; update diff vs fp64 {dx:.2e}")
    assert lin <= 1e-10
    assert dx < 1e-7
    ctx.newton_update()
    I_gpu, _ = ctx.get_state()
    assert np.array_equal(I_gpu.reshape(-1), orc.newton_update(I, x32, ctx.ndof, P.dim + 2))
    _say(f"  ||R(I + x32)|| / ||R(I)|| = {ctx.residual() / np.linalg.norm(rhs_ref):.3e}")
    ctx.set_state(I, Pst)


@pytest.mark.parametrize("cid,scatter", [(2, ""), (4, "ebuf32")])
def test_refine_3d(cid, scatter, monkeypatch):
    """Three refinement passes on the 3D boxes at 1/16 resolution (c4 with
    the FP32-entry element buffer it uses at full size) reach the FP64
    Newton equation to 1e-10, checked with the oracle's own matrix."""
    if scatter:
        monkeypatch.setenv("GMC_SCATTER", scatter)
    mesh, P, I, Pst = _case(cid, 16)
    _say(f"refine c{cid}/16 scatter={scatter or 'default'}")
    A = orc.Assembler(P, mesh)
    vals_ref, rhs_ref = A.assemble(I, Pst)
    with Context(mesh, P) as ctx:
        ctx.set_state(I, Pst)
        res = ctx.refine(n_refine=3, **_refine_solver())
        _say("  refinement history " + " ".join(f"{h:.2e}" for h in res / res[0]))
        assert res[-1] / res[0] <= 1e-10
        assert rel(A.spmv(vals_ref, ctx.solution()), rhs_ref) <= 1e-10


def test_refine_argument_errors(c0):
    ctx = c0[4]
    with pytest.raises(Exception):
        ctx.refine(n_refine=0)
    with pytest.raises(Exception):
        ctx.refine(n_refine=1, restart=500)
