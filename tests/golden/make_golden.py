"""Writes the golden vectors of tests/golden/ from the FP64 oracle.

The reference ships no tests, vectors or fixtures for this path (and cannot be
built here), so these vectors come from this repository's oracle
(oracle/gmc_oracle.c), which the CPU tests pin against closed forms, finite
differences and an independent NumPy restatement.  Committing them freezes
the oracle's outputs: tests/test_golden.py re-derives them on the CPU and
compares the GPU path with them, so a later change to either side shows.

    python tests/golden/make_golden.py      (from the repository root)

Contents (numpy .npz, no pickles):
  elements_c0.npz   element m (15x15) and r of 6 interior and 4 boundary
                    elements of the 64x32 channel
  elements_c2.npz   element m (28x28) and r of 3 interior and 2 boundary
                    tetrahedra of the 8^3 Kuhn box (3 components)
  assembly_c0.npz   the config-0 system: rhs (10725 values), the Frobenius
                    norm of every BSR block (14625), the full 5x5 values of
                    256 seeded blocks and the pattern
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "gales-multi-fluid_amd")]

import orc  # noqa: E402
from gales_mc import cases  # noqa: E402

C0_INTERIOR = [0, 1, 777, 2048, 3001, 4095]
C2_INTERIOR = [0, 1500, 3071]


def element_set(cid, scale, interior, n_bd):
    mesh, P, I, Pst = cases.config(cid, scale=scale)
    bd = sorted(set(mesh["side_elem"].tolist()))
    bsel = bd[:n_bd // 2] + bd[-(n_bd - n_bd // 2):]
    u = orc.uref_inv(P)
    out = {"interior": np.array(interior), "boundary": np.array(bsel)}
    mi, ri, mb, rb = [], [], [], []
    for e in interior:
        xe, Ie, Pe = orc.element_data(mesh, I, Pst, e)
        m, r = orc.element_interior(P, xe, Ie, Pe, u)
        mi.append(m)
        ri.append(r)
    for e in bsel:
        xe, Ie, Pe = orc.element_data(mesh, I, Pst, e)
        sl, sf = orc.element_sides(mesh, e)
        m, r = orc.element_boundary(P, xe, Ie, sl, sf)
        mb.append(m)
        rb.append(r)
    out.update(m_interior=np.array(mi), r_interior=np.array(ri), m_boundary=np.array(mb),
               r_boundary=np.array(rb))
    return out


def main():
    np.savez_compressed(os.path.join(HERE, "elements_c0.npz"), **element_set(0, 1, C0_INTERIOR, 4))
    np.savez_compressed(os.path.join(HERE, "elements_c2.npz"), **element_set(2, 16, C2_INTERIOR, 2))
    mesh, P, I, Pst = cases.config(0)
    A = orc.Assembler(P, mesh)
    vals, rhs = A.assemble(I, Pst)
    vals = vals.reshape(-1, 5, 5)
    pick = np.sort(np.random.default_rng(7).choice(vals.shape[0], 256, replace=False))
    np.savez_compressed(os.path.join(HERE, "assembly_c0.npz"), rowptr=A.rowptr, colidx=A.colidx,
                        rhs=rhs, block_norms=np.linalg.norm(vals, axis=(1, 2)), pick=pick,
                        blocks=vals[pick])
    print("golden vectors written to", HERE)


if __name__ == "__main__":
    main()
