"""FP64 parity of the HIP path against the CPU oracle, through the C-ABI.

Bar: FP64 results within 1e-12 relative (norm-wise) of the oracle on the same
seeded inputs; integer / index work (pattern, colouring coverage) exact;
repeated runs bitwise identical.
"""
import sys

import numpy as np
import pytest

import orc
from helpers import entry_rel, rel

from gales_mc import abi, cases
from gales_mc.lib import Context, GmcError

pytestmark = pytest.mark.gpu
TOL = 1e-12
J32_TOL = 1e-6  # FP32 element matrices of the precision tier


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def c0():
    mesh, P, I, Pst = cases.config(0)
    ctx = Context(mesh, P)
    ctx.set_state(I, Pst)
    yield mesh, P, I, Pst, ctx
    ctx.close()


def _elements(mesh):
    bd = sorted(set(mesh["side_elem"].tolist()))
    E = mesh["conn"].shape[0]
    rng = np.random.default_rng(1)
    return sorted(set([0, 1, E - 1] + bd[:4] + bd[-4:] + rng.integers(0, E, 12).tolist()))


def test_element_interior_c0(c0):
    mesh, P, I, Pst, ctx = c0
    _say("element interior c0")
    u = orc.uref_inv(P)
    for e in _elements(mesh):
        xe, Ie, Pe = orc.element_data(mesh, I, Pst, e)
        m_ref, r_ref = orc.element_interior(P, xe, Ie, Pe, u)
        m, r = ctx.element(e)
        assert rel(m, m_ref) < TOL, (e, rel(m, m_ref))
        assert rel(r, r_ref) < TOL, (e, rel(r, r_ref))
        assert entry_rel(m, m_ref) < TOL


def test_element_boundary_c0(c0):
    mesh, P, I, Pst, ctx = c0
    _say("element boundary c0")
    bd = sorted(set(mesh["side_elem"].tolist()))
    for e in bd[:6] + bd[-6:] + bd[len(bd) // 2:len(bd) // 2 + 4]:
        xe, Ie, Pe = orc.element_data(mesh, I, Pst, e)
        sl, sf = orc.element_sides(mesh, e)
        m_ref, r_ref = orc.element_boundary(P, xe, Ie, sl, sf)
        m, r = ctx.element(e, boundary=True)
        assert rel(m, m_ref) < TOL, (e, rel(m, m_ref))
        assert rel(r, r_ref) < TOL, (e, rel(r, r_ref))
    inner = next(e for e in range(100, mesh["conn"].shape[0]) if e not in set(bd))
    m, r = ctx.element(inner, boundary=True)
    assert not m.any() and not r.any()


def _assembly_parity(mesh, P, I, Pst, ctx, tol=TOL):
    A = orc.Assembler(P, mesh)
    vals_ref, rhs_ref = A.assemble(I, Pst)
    rowptr, colidx = ctx.pattern()
    assert np.array_equal(rowptr, A.rowptr) and np.array_equal(colidx, A.colidx)
    ctx.assemble()
    vals, rhs = ctx.system()
    ev, er = rel(vals, vals_ref), rel(rhs, rhs_ref)
    _say(f"assembly parity: J rel {ev:.3e}, rhs rel {er:.3e}")
    assert ev < tol and er < tol
    return A, vals_ref, rhs_ref, vals, rhs


def test_assembly_c0(c0):
    mesh, P, I, Pst, ctx = c0
    _say("assembly c0")
    assert ctx.info.scatter == 2  # patch-local assembly is the 2D FP64 default
    _assembly_parity(mesh, P, I, Pst, ctx)
    ctx.assemble(abi.GMC_ASM_MATRIX | abi.GMC_ASM_RHS | abi.GMC_ASM_FP32)
    assert ctx.timing().gather_s > 0.0


def test_assembly_c0_element_buffer(monkeypatch):
    """The element-buffer scatter (block-sorted slots + gather) gives the same
    system as the patch-local one, to rounding (the sums run in another
    order)."""
    monkeypatch.setenv("GMC_SCATTER", "ebuf")
    mesh, P, I, Pst = cases.config(0)
    with Context(mesh, P) as ctx:
        assert ctx.info.scatter == 1
        ctx.set_state(I, Pst)
        _, _, _, v1, r1 = _assembly_parity(mesh, P, I, Pst, ctx)
    monkeypatch.setenv("GMC_SCATTER", "patch")
    with Context(mesh, P) as ctx:
        assert ctx.info.scatter == 2
        ctx.set_state(I, Pst)
        ctx.assemble()
        v2, r2 = ctx.system()
    assert rel(v2, v1) < 1e-14 and rel(r2, r1) < 1e-14


@pytest.mark.parametrize("cid,scale,chunks", [(0, 1, 3), (2, 16, 2), (4, 16, 3)])
def test_assembly_element_buffer_in_chunks(cid, scale, chunks, monkeypatch):
    """The element buffer taken in element chunks (the default when the whole
    buffer does not fit; forced here): chunk by chunk the interior and
    boundary elements and a gather that adds onto the earlier chunks' values.
    Both tiers against the oracle, against the one-buffer system to rounding,
    and bitwise run to run."""
    mesh, P, I, Pst = cases.config(cid, scale=scale)
    f32 = abi.GMC_ASM_MATRIX | abi.GMC_ASM_RHS | abi.GMC_ASM_FP32
    monkeypatch.setenv("GMC_SCATTER", "ebuf")
    with Context(mesh, P) as ctx:
        assert ctx.info.scatter == 1 and ctx.info.ebuf_chunks == 1
        ctx.set_state(I, Pst)
        ctx.assemble()
        v1, r1 = ctx.system()
        ctx.assemble(f32)
        w1, s1 = ctx.system()
    monkeypatch.setenv("GMC_EBUF_CHUNKS", str(chunks))
    _say(f"element buffer in {chunks} chunks, c{cid}/{scale}")
    with Context(mesh, P) as ctx:
        assert ctx.info.scatter == 1 and ctx.info.ebuf_chunks == chunks
        ctx.set_state(I, Pst)
        _, vref, rref, v2, r2 = _assembly_parity(mesh, P, I, Pst, ctx)
        assert ctx.timing().interior_launches == chunks
        assert rel(v2, v1) < 1e-13 and rel(r2, r1) < 1e-13
        ctx.assemble()
        v3, r3 = ctx.system()
        assert np.array_equal(v2, v3) and np.array_equal(r2, r3)
        ctx.assemble(f32)
        w2, s2 = ctx.system()
        assert rel(w2, vref) < J32_TOL and rel(s2, rref) < 1e-12
        assert rel(w2, w1) < 1e-6 and np.array_equal(s2, s1)


def test_assembly_c0_fp32_only_element_buffer(monkeypatch):
    """An element buffer sized for FP32 entries only (chosen when the FP64 one
    does not fit): FP32-tier assemblies go through it, bitwise equal to the
    full element buffer's; FP64 assemblies take the colour scatter."""
    mesh, P, I, Pst = cases.config(0)
    f32 = abi.GMC_ASM_MATRIX | abi.GMC_ASM_RHS | abi.GMC_ASM_FP32
    monkeypatch.setenv("GMC_SCATTER", "ebuf")
    with Context(mesh, P) as ctx:
        ctx.set_state(I, Pst)
        ctx.assemble(f32)
        v1, r1 = ctx.system()
    monkeypatch.setenv("GMC_SCATTER", "ebuf32")
    with Context(mesh, P) as ctx:
        assert ctx.info.scatter == 0  # the FP64 scatter
        ctx.set_state(I, Pst)
        ctx.assemble(f32)
        assert ctx.timing().interior_launches == 1  # one element-buffer launch
        v2, r2 = ctx.system()
        _assembly_parity(mesh, P, I, Pst, ctx)
        assert ctx.timing().interior_launches > 1  # colour by colour
    assert np.array_equal(v1, v2) and np.array_equal(r1, r2)


def test_assembly_c0_colour_scatter(monkeypatch):
    """The colour-ordered read-modify-write scatter (used when the element
    buffer does not fit) gives the same system."""
    monkeypatch.setenv("GMC_SCATTER", "colour")
    mesh, P, I, Pst = cases.config(0)
    with Context(mesh, P) as ctx:
        assert ctx.info.scatter == 0
        ctx.set_state(I, Pst)
        _assembly_parity(mesh, P, I, Pst, ctx)
        v1, r1 = ctx.system()
        ctx.assemble()
        v2, r2 = ctx.system()
        assert np.array_equal(v1, v2) and np.array_equal(r1, r2)


def test_assembly_deterministic_c0(c0):
    mesh, P, I, Pst, ctx = c0
    ctx.assemble()
    v1, r1 = ctx.system()
    ctx.assemble()
    v2, r2 = ctx.system()
    assert np.array_equal(v1, v2) and np.array_equal(r1, r2)


def test_residual_only_c0(c0):
    mesh, P, I, Pst, ctx = c0
    ctx.assemble()
    _, rhs_full = ctx.system()
    nrm = ctx.residual()
    _, rhs = ctx.system()
    assert rel(rhs, rhs_full) < 1e-14
    assert abs(nrm - np.linalg.norm(rhs_full)) <= 1e-13 * np.linalg.norm(rhs_full)


def test_spmv_c0(c0):
    mesh, P, I, Pst, ctx = c0
    ctx.assemble()
    vals, _ = ctx.system()
    A = orc.Assembler(P, mesh)
    x = np.random.default_rng(3).normal(size=I.size)
    assert rel(ctx.spmv(x), A.spmv(vals, x)) < 1e-14


def test_gmres_newton_c0(c0):
    """One GMRES(30)+Jacobi Newton step (config 0 workload).

    - Solver parity: the oracle's GMRES in the device's summation order
      (orc_gmres_ordered), run on the GPU's own assembled system, gives the
      GPU's update to 1e-12.
    - The GPU update solves the oracle's system as well as the GPU reports.
    - End to end (the oracle's own system and GMRES, in either summation
      order, against the GPU): within 1e-12 (measured ~1e-14).
    - The update I + delta with the Y_1 trim is bitwise the oracle's."""
    mesh, P, I, Pst, ctx = c0
    _say("gmres c0")
    ctx.set_state(I, Pst)
    A, vals_ref, rhs_ref, vals, rhs = _assembly_parity(mesh, P, I, Pst, ctx)
    info = ctx.solve(restart=30, max_iters=30, rel_tol=0.0, precond=1)
    d_gpu = ctx.solution()
    d_ord, it_ord, rr_ord = A.gmres(vals, rhs, restart=30, max_iters=30, rel_tol=0.0, ordered=True)
    d_ref, it_ref, rr_ref = A.gmres(vals_ref, rhs_ref, restart=30, max_iters=30, rel_tol=0.0)
    d_ref_ord, _, _ = A.gmres(vals_ref, rhs_ref, restart=30, max_iters=30, rel_tol=0.0, ordered=True)
    res_on_ref = np.linalg.norm(A.spmv(vals_ref, d_gpu) - rhs_ref) / np.linalg.norm(rhs_ref)
    _say(f"  it {info.iterations} vs {it_ord}; rel res {info.rel_residual:.6e} vs {rr_ord:.6e}; "
         f"delta vs ordered oracle GMRES on the GPU system {rel(d_gpu, d_ord):.3e} "
         f"(bitwise {np.array_equal(d_gpu, d_ord)})")
    _say(f"  end to end: delta vs oracle system + ordered GMRES {rel(d_gpu, d_ref_ord):.3e}, "
         f"+ plain-order GMRES {rel(d_gpu, d_ref):.3e}; GPU delta on the oracle system: rel res "
         f"{res_on_ref:.6e} (GPU reports {info.rel_residual:.6e})")
    assert info.iterations == it_ord == it_ref == 30
    assert rel(d_gpu, d_ord) < 1e-12
    assert abs(info.rel_residual - rr_ord) <= 1e-12 * rr_ord
    assert abs(res_on_ref - info.rel_residual) <= 1e-6 * info.rel_residual
    assert rel(d_gpu, d_ref_ord) < 1e-12 and rel(d_gpu, d_ref) < 1e-12
    ctx.newton_update()
    I_gpu, _ = ctx.get_state()
    I_ref = orc.newton_update(I, d_gpu, ctx.ndof, P.dim + 2)
    assert np.array_equal(I_gpu.reshape(-1), I_ref)
    ctx.set_state(I, Pst)


@pytest.mark.parametrize("scale", [8, 4])
def test_gmres_device_order_c1(scale):
    """The c1 family at 1/8 and 1/4 resolution (165,765 and 659,205 unknowns:
    162 and 644 dot-product blocks, so the final fold adds several block sums
    per lane): GPU GMRES(50) update = the ordered oracle GMRES on the GPU
    system, to 1e-12."""
    mesh, P, I, Pst = cases.config(1, scale=scale)
    _say(f"gmres device order c1/{scale}")
    with Context(mesh, P) as ctx:
        ctx.set_state(I, Pst)
        ctx.assemble()
        vals, rhs = ctx.system()
        info = ctx.solve(restart=50, max_iters=50, rel_tol=0.0, precond=1)
        d_gpu = ctx.solution()
    A = orc.Assembler(P, mesh)
    d_ord, it_ord, rr_ord = A.gmres(vals, rhs, restart=50, max_iters=50, rel_tol=0.0, ordered=True)
    _say(f"  delta vs ordered oracle GMRES {rel(d_gpu, d_ord):.3e} (bitwise {np.array_equal(d_gpu, d_ord)}), "
         f"rel res {info.rel_residual:.6e} vs {rr_ord:.6e}")
    assert info.iterations == it_ord == 50
    assert rel(d_gpu, d_ord) < 1e-12


def test_predictor_c0(c0):
    mesh, P, I, Pst, ctx = c0
    ctx.set_state(I, Pst)
    ctx.predictor()
    I2, P2 = ctx.get_state()
    assert np.array_equal(P2, I2) and np.array_equal(I2.reshape(-1), I.reshape(-1))
    ctx.set_state(I, Pst)


def test_assembly_c1_scaled():
    """c1 family (jittered 2D channel) at 256x128 quads: full-system parity."""
    mesh, P, I, Pst = cases.config(1, scale=8)
    _say("assembly c1/8")
    with Context(mesh, P) as ctx:
        ctx.set_state(I, Pst)
        _assembly_parity(mesh, P, I, Pst, ctx)


def test_assembly_c1_16_element_patches(monkeypatch):
    """GMC_PATCH_ELEMS=16: one-wave workgroups (the persistent FP64 kernel
    then runs four per CU) on the c1 family at 256x128 quads, against the
    oracle; the same matrix as the 64-element plan to rounding."""
    mesh, P, I, Pst = cases.config(1, scale=8)
    _say("assembly c1/8, 16-element patches")
    with Context(mesh, P) as ctx:
        ctx.set_state(I, Pst)
        ctx.assemble()
        v64, r64 = ctx.system()
    monkeypatch.setenv("GMC_PATCH_ELEMS", "16")
    with Context(mesh, P) as ctx:
        assert ctx.info.scatter == 2
        ctx.set_state(I, Pst)
        _assembly_parity(mesh, P, I, Pst, ctx)
        vals, rhs = ctx.system()
        assert rel(vals, v64) < 1e-13 and rel(rhs, r64) < 1e-13


def test_assembly_c1_ragged_patches():
    """c1 family at 682x341 quads (465,124 triangles): the last 64-element
    patch is partial and the persistent FP64 patch kernel loops over several
    patches per workgroup; both tiers against the oracle."""
    mesh, P, I, Pst = cases.config(1, scale=3)
    assert mesh["conn"].shape[0] % 64 != 0
    _say("assembly c1/3 (ragged last patch)")
    with Context(mesh, P) as ctx:
        assert ctx.info.scatter == 2
        ctx.set_state(I, Pst)
        A, vref, rref, _, _ = _assembly_parity(mesh, P, I, Pst, ctx)
        ctx.assemble(abi.GMC_ASM_MATRIX | abi.GMC_ASM_RHS | abi.GMC_ASM_FP32)
        v32, r32 = ctx.system()
    assert rel(r32, rref) < TOL and rel(v32, vref) < 1e-5


def _row_spot_check(mesh, P, I, Pst, ctx, nodes):
    """Block rows of the assembled system recomputed from oracle element
    integrals with the scatter rule (size-independent parity check)."""
    n = ctx.ndof
    nen = P.dim + 1
    conn = mesh["conn"]
    rowptr, colidx = ctx.pattern()
    vals, rhs = ctx.system()
    dmask = orc.dmask_of(P, mesh)
    uref = orc.uref_inv(P)
    bdset = set(mesh["side_elem"].tolist())
    worst_v = worst_r = 0.0
    for node in nodes:
        elems = np.nonzero((conn == node).any(1))[0]
        row = {int(c): np.zeros((n, n)) for c in colidx[rowptr[node]:rowptr[node + 1]]}
        rr = np.zeros(n)
        for e in elems:
            en = conn[e]
            xe, Ie, Pe = orc.element_data(mesh, I, Pst, e)
            parts = [orc.element_interior(P, xe, Ie, Pe, uref)]
            if e in bdset:
                sl, sf = orc.element_sides(mesh, e)
                parts.append(orc.element_boundary(P, xe, Ie, sl, sf))
            b = int(np.nonzero(en == node)[0][0])
            for m, r in parts:
                m = m.copy()
                r = r.copy()
                N = nen * n
                avg = (1.0 + np.trace(m)) / N
                for i in range(N):
                    a_, k = divmod(i, n)
                    if dmask[en[a_]] >> k & 1:
                        m[i, :] = 0.0
                        m[i, i] = avg
                        r[i] = 0.0
                rr += r[b * n:(b + 1) * n]
                for a_ in range(nen):
                    row[int(en[a_])] += m[b * n:(b + 1) * n, a_ * n:(a_ + 1) * n]
        got = np.concatenate([vals[t] for t in range(rowptr[node], rowptr[node + 1])])
        ref = np.concatenate([row[int(c)] for c in colidx[rowptr[node]:rowptr[node + 1]]])
        worst_v = max(worst_v, rel(got, ref))
        worst_r = max(worst_r, rel(rhs[node * n:(node + 1) * n], rr))
    return worst_v, worst_r


def test_assembly_c1_full_rows():
    """Full-size c1 (2048x1024 quads, 4.19M triangles): sampled block rows of
    the assembled J and rhs against oracle element integrals + scatter rule,
    plus run-to-run bitwise determinism."""
    mesh, P, I, Pst = cases.config(1)
    _say("assembly c1 full")
    with Context(mesh, P) as ctx:
        ctx.set_state(I, Pst)
        ctx.assemble()
        rng = np.random.default_rng(7)
        nn = mesh["coords"].shape[0]
        flags = mesh["node_flag"]
        nodes = list(rng.integers(0, nn, 12)) + list(np.nonzero(flags == 5)[0][:3]) + \
            list(np.nonzero(flags == 2)[0][:2]) + list(np.nonzero(flags == 3)[0][:2])
        wv, wr = _row_spot_check(mesh, P, I, Pst, ctx, nodes)
        _say(f"c1 rows: J {wv:.3e} rhs {wr:.3e}")
        assert wv < TOL and wr < TOL
        v1, r1 = ctx.system()
        ctx.assemble()
        v2, r2 = ctx.system()
        assert np.array_equal(v1, v2) and np.array_equal(r1, r2)


@pytest.mark.parametrize("cid", [2, 4])
def test_element_3d(cid):
    mesh, P, I, Pst = cases.config(cid, scale=16)
    _say(f"element 3d c{cid}")
    with Context(mesh, P) as ctx:
        ctx.set_state(I, Pst)
        u = orc.uref_inv(P)
        bd = sorted(set(mesh["side_elem"].tolist()))
        for e in [0, 5, 17, 1000] + bd[:2] + bd[-2:]:
            xe, Ie, Pe = orc.element_data(mesh, I, Pst, e)
            m_ref, r_ref = orc.element_interior(P, xe, Ie, Pe, u)
            m, r = ctx.element(e)
            assert rel(m, m_ref) < TOL and rel(r, r_ref) < TOL, (e, rel(m, m_ref), rel(r, r_ref))
            if e in bd:
                sl, sf = orc.element_sides(mesh, e)
                mb_ref, rb_ref = orc.element_boundary(P, xe, Ie, sl, sf)
                mb, rb = ctx.element(e, boundary=True)
                assert rel(mb, mb_ref) < TOL and rel(rb, rb_ref) < TOL


@pytest.mark.parametrize("cid", [2, 4])
def test_assembly_3d_scaled(cid):
    mesh, P, I, Pst = cases.config(cid, scale=16)
    _say(f"assembly 3d c{cid}/16")
    with Context(mesh, P) as ctx:
        assert ctx.info.scatter == 1  # element buffer is the 3D default
        ctx.set_state(I, Pst)
        _assembly_parity(mesh, P, I, Pst, ctx)


@pytest.mark.parametrize("cid", [2, 4])
def test_assembly_3d_colour_scatter(monkeypatch, cid):
    """The colour scatter on tetrahedra (the FP64 path when the element
    buffer does not fit, as on the 21.2M-tet 4-component box): FP64
    read-modify-write into the BSR values one colour at a time, against the
    oracle, and bitwise the same on a second assembly.  With 7 dofs its
    element kernel runs the rolled column-node loop, with 8 the unrolled one."""
    monkeypatch.setenv("GMC_SCATTER", "colour")
    mesh, P, I, Pst = cases.config(cid, scale=16)
    _say(f"assembly 3d c{cid}/16, colour scatter")
    with Context(mesh, P) as ctx:
        assert ctx.info.scatter == 0
        ctx.set_state(I, Pst)
        _assembly_parity(mesh, P, I, Pst, ctx)
        assert ctx.timing().interior_launches > 1  # colour by colour
        v1, r1 = ctx.system()
        ctx.assemble()
        v2, r2 = ctx.system()
    assert np.array_equal(v1, v2) and np.array_equal(r1, r2)


@pytest.mark.gpu
def test_assembly_3d_patch(monkeypatch):
    """The patch-local scatter on tetrahedra (16 per workgroup, 6 phases),
    selectable in 3D though not the default there."""
    monkeypatch.setenv("GMC_SCATTER", "patch")
    mesh, P, I, Pst = cases.config(2, scale=16)
    _say("assembly 3d c2/16, patch scatter")
    with Context(mesh, P) as ctx:
        assert ctx.info.scatter == 2
        ctx.set_state(I, Pst)
        _assembly_parity(mesh, P, I, Pst, ctx)
        ctx.assemble(abi.GMC_ASM_MATRIX | abi.GMC_ASM_RHS | abi.GMC_ASM_FP32)
        v32, r32 = ctx.system()
    A = orc.Assembler(P, mesh)
    vref, rref = A.assemble(I, Pst)
    assert rel(r32, rref) < 1e-12 and rel(v32, vref) < 1e-5


# ---------------------------------------------------------------------------
# flag variants (fluid_2d.hpp:627-758, 895-1021, 1053-1057; setup.txt:24-31)
# ---------------------------------------------------------------------------
def _variant_params(P, name):
    if name.startswith("tau"):
        P.tau_variant = int(name[3:])
    elif name == "steady":
        P.steady_state = 1
    elif name.startswith("dc_sharp"):
        P.dc_sharp = float(name[len("dc_sharp"):])
    elif name == "mass_flux":
        # the outlet sides (flag 3) prescribe the mass flux along x
        P.neumann[cases.OUTLET].mflux_mask = 1
        P.neumann[cases.OUTLET].mflux[0] = 30.0
    return P


VARIANTS_2D = ["tau2001", "tau2007", "tau2014", "steady", "dc_sharp1", "dc_sharp1.5", "mass_flux"]


@pytest.mark.parametrize("name", VARIANTS_2D)
def test_flag_variant_c0(name):
    """Every interior and boundary element of a sample, and the assembled
    system, against the oracle at 1e-12 (config 0 with the variant set)."""
    mesh, P, I, Pst = cases.config(0)
    P = _variant_params(P, name)
    _say(f"variant {name} c0")
    u = orc.uref_inv(P)
    with Context(mesh, P) as ctx:
        ctx.set_state(I, Pst)
        for e in _elements(mesh):
            xe, Ie, Pe = orc.element_data(mesh, I, Pst, e)
            m_ref, r_ref = orc.element_interior(P, xe, Ie, Pe, u)
            m, r = ctx.element(e)
            assert rel(m, m_ref) < TOL and rel(r, r_ref) < TOL, (e, rel(m, m_ref), rel(r, r_ref))
        bd = sorted(set(mesh["side_elem"].tolist()))
        for e in bd[:4] + bd[-4:]:
            xe, Ie, _ = orc.element_data(mesh, I, Pst, e)
            sl, sf = orc.element_sides(mesh, e)
            m_ref, r_ref = orc.element_boundary(P, xe, Ie, sl, sf)
            m, r = ctx.element(e, boundary=True)
            assert rel(m, m_ref) < TOL and rel(r, r_ref) < TOL, (e, rel(m, m_ref), rel(r, r_ref))
        _assembly_parity(mesh, P, I, Pst, ctx)


@pytest.mark.parametrize("name", ["tau2001", "tau2007", "tau2014", "mass_flux"])
def test_flag_variant_3d(name):
    """The 3D box (3 components) at 1/16 with the variant set: assembly
    against the oracle at 1e-12 (mass flux on the outlet faces along x)."""
    mesh, P, I, Pst = cases.config(2, scale=16)
    P = _variant_params(P, name)
    _say(f"variant {name} c2/16")
    with Context(mesh, P) as ctx:
        ctx.set_state(I, Pst)
        _assembly_parity(mesh, P, I, Pst, ctx)


def test_singular_dudy_reports_numeric():
    """A pure incompressible-melt state (Y_1 = 1 at every node, v != 0): DUDY's
    pressure column is zero (rho beta = 0, de/dp = 0), the LU of tau meets a
    zero pivot, and the assembly returns GMC_ERR_NUMERIC, as the oracle does
    (the reference calls Error() there, algorithm.hpp:407-410).  The context
    stays usable: a regular state afterwards assembles to the oracle's."""
    mesh, P, I, Pst = cases.config(0)
    pure = I.copy()
    pure[:, 4] = 1.0
    pure[:, 1] = np.where(mesh["node_flag"] == cases.WALL, 0.0, 0.5)
    pure[:, 2] = 0.0
    A = orc.Assembler(P, mesh)
    with pytest.raises(ArithmeticError):
        A.assemble(pure, pure)
    with Context(mesh, P) as ctx:
        ctx.set_state(pure, pure)
        with pytest.raises(GmcError) as ei:
            ctx.assemble()
        assert ei.value.code == abi.GMC_ERR_NUMERIC and "singular" in str(ei.value)
        with pytest.raises(GmcError) as ei:
            ctx.residual()
        assert ei.value.code == abi.GMC_ERR_NUMERIC
        ctx.set_state(I, Pst)
        _assembly_parity(mesh, P, I, Pst, ctx)


def _star_mesh(n):
    """n triangles around one centre node (valence n), rim on a natural side."""
    th = 2.0 * np.pi * np.arange(n) / n
    coords = np.concatenate([[[0.0, 0.0]], np.stack([np.cos(th), np.sin(th)], 1)]).astype(np.float64)
    k = np.arange(n)
    conn = np.stack([np.zeros(n, int), 1 + k, 1 + (k + 1) % n], 1).astype(np.int32)
    flag = np.full(n + 1, cases.OUTLET, np.int32)
    flag[0] = 0
    return dict(dim=2, coords=coords, conn=conn, node_flag=flag, side_elem=k.astype(np.int32),
                side_local=np.zeros(n, np.int32), side_flag=np.full(n, cases.OUTLET, np.int32),
                shape=None, extent=(2.0, 2.0))


@pytest.mark.parametrize("scatter", ["", "ebuf", "colour"])
def test_high_valence_node(scatter, monkeypatch):
    """A node shared by 200 elements: every scatter builds (the colour scatter
    with 200 colours) and assembles the oracle's system."""
    if scatter:
        monkeypatch.setenv("GMC_SCATTER", scatter)
    mesh = _star_mesh(200)
    P = cases.make_params(2, [cases.MELT, cases.GAS_R461])
    I, Pst = cases.fields(mesh, 2, 5)
    _say(f"star mesh, scatter={scatter or 'default'}")
    with Context(mesh, P) as ctx:
        _say(f"  scatter {ctx.info.scatter}, interior colours {ctx.info.n_colors}, "
             f"boundary colours {ctx.info.n_bd_colors}")
        assert ctx.info.scatter == {"": 2, "ebuf": 1, "colour": 0}[scatter]
        assert ctx.info.n_colors == (200 if scatter == "colour" else 0)
        ctx.set_state(I, Pst)
        _assembly_parity(mesh, P, I, Pst, ctx)
