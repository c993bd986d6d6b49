"""Golden vectors (tests/golden/, written by tests/golden/make_golden.py from
the FP64 oracle): the oracle still reproduces them (CPU), and the HIP path
matches them through the C-ABI (GPU) at the FP64 parity bar, 1e-12 relative."""
import os

import numpy as np
import pytest

import orc
from helpers import rel

from gales_mc import cases

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-12


def _load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def _element_cases():
    return [("elements_c0.npz", 0, 1), ("elements_c2.npz", 2, 16)]


@pytest.mark.parametrize("name,cid,scale", _element_cases())
def test_oracle_reproduces_element_vectors(name, cid, scale):
    g = _load(name)
    mesh, P, I, Pst = cases.config(cid, scale=scale)
    u = orc.uref_inv(P)
    for k, e in enumerate(g["interior"]):
        xe, Ie, Pe = orc.element_data(mesh, I, Pst, int(e))
        m, r = orc.element_interior(P, xe, Ie, Pe, u)
        assert np.array_equal(m, g["m_interior"][k]) and np.array_equal(r, g["r_interior"][k])
    for k, e in enumerate(g["boundary"]):
        xe, Ie, Pe = orc.element_data(mesh, I, Pst, int(e))
        sl, sf = orc.element_sides(mesh, int(e))
        m, r = orc.element_boundary(P, xe, Ie, sl, sf)
        assert np.array_equal(m, g["m_boundary"][k]) and np.array_equal(r, g["r_boundary"][k])


def test_oracle_reproduces_assembly_vectors():
    g = _load("assembly_c0.npz")
    mesh, P, I, Pst = cases.config(0)
    A = orc.Assembler(P, mesh)
    vals, rhs = A.assemble(I, Pst)
    vals = vals.reshape(-1, 5, 5)
    assert np.array_equal(A.rowptr, g["rowptr"]) and np.array_equal(A.colidx, g["colidx"])
    assert np.array_equal(rhs, g["rhs"])
    assert np.array_equal(np.linalg.norm(vals, axis=(1, 2)), g["block_norms"])
    assert np.array_equal(vals[g["pick"]], g["blocks"])


@pytest.mark.gpu
@pytest.mark.parametrize("name,cid,scale", _element_cases())
def test_gpu_matches_element_vectors(name, cid, scale):
    from gales_mc.lib import Context
    g = _load(name)
    mesh, P, I, Pst = cases.config(cid, scale=scale)
    with Context(mesh, P) as ctx:
        ctx.set_state(I, Pst)
        for k, e in enumerate(g["interior"]):
            m, r = ctx.element(int(e))
            assert rel(m, g["m_interior"][k]) < TOL and rel(r, g["r_interior"][k]) < TOL, e
        for k, e in enumerate(g["boundary"]):
            m, r = ctx.element(int(e), boundary=True)
            assert rel(m, g["m_boundary"][k]) < TOL and rel(r, g["r_boundary"][k]) < TOL, e


@pytest.mark.gpu
def test_gpu_matches_assembly_vectors():
    from gales_mc.lib import Context
    g = _load("assembly_c0.npz")
    mesh, P, I, Pst = cases.config(0)
    with Context(mesh, P) as ctx:
        ctx.set_state(I, Pst)
        ctx.assemble()
        vals, rhs = ctx.system()
        rowptr, colidx = ctx.pattern()
    assert np.array_equal(rowptr, g["rowptr"]) and np.array_equal(colidx, g["colidx"])
    assert rel(rhs, g["rhs"]) < TOL
    assert rel(np.linalg.norm(vals, axis=(1, 2)), g["block_norms"]) < TOL
    assert rel(vals[g["pick"]], g["blocks"]) < TOL
