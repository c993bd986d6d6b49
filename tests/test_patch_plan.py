"""The patch-local assembly plan (gales-multi-fluid_amd/csrc/gmc_patch.cpp),
compiled for the CPU and driven through an emulation of the patch kernel's
summation schedule with integer-valued contributions: every element is in one
patch, same-phase contributions never share a slot, every BSR block and rhs
node is written exactly once, and the sums equal a direct scatter.  CPU only."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import orc
from gales_mc import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gales-multi-fluid_amd")


@pytest.fixture(scope="module")
def patch_shim():
    out = os.path.join(ROOT, "tests", "host_shim", "libpatch_host.so")
    src = os.path.join(ROOT, "tests", "host_shim", "patch_host.cpp")
    deps = [src] + [os.path.join(PKG, "csrc", f) for f in ("gmc_patch.cpp", "gmc_patch.hpp")]
    if not os.path.exists(out) or any(os.path.getmtime(d) > os.path.getmtime(out) for d in deps):
        subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", os.path.join(PKG, "csrc"),
                        "-o", out, src], check=True)
    lib = C.CDLL(out)
    lib.patch_check.restype = C.c_int
    return lib


def _blkmap(mesh):
    conn = mesh["conn"]
    n_nodes = mesh["coords"].shape[0]
    rowptr, colidx = orc.pattern(n_nodes, conn)
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    nen = conn.shape[1]
    rows = np.repeat(conn, nen, axis=1).reshape(-1)
    cols = np.tile(conn, (1, nen)).reshape(-1)
    blk = np.empty(rows.size, np.int32)
    for k in range(rows.size):  # small meshes only
        r = rows[k]
        lo, hi = rowptr[r], rowptr[r + 1]
        blk[k] = lo + np.searchsorted(colidx[lo:hi], cols[k])
    return blk, colidx.size, rowptr


@pytest.mark.parametrize("cid,scale,pe", [(0, 1, 16), (1, 32, 16), (2, 32, 16), (0, 1, 64), (1, 16, 64)])
def test_patch_plan_schedule(patch_shim, cid, scale, pe):
    mesh, P, I, Pst = cases.config(cid, scale=scale)
    blk, nnzb, rowptr = _blkmap(mesh)
    coords = np.ascontiguousarray(mesh["coords"], np.float64)
    conn = np.ascontiguousarray(mesh["conn"], np.int32)
    stats = np.zeros(8, np.int64)
    # the boundary elements' integrals go through halo slots of their own
    bd = np.ascontiguousarray(np.unique(np.asarray(mesh["side_elem"])), np.int32)
    st = patch_shim.patch_check(int(mesh["dim"]), coords.shape[0], C.c_longlong(conn.shape[0]),
                                coords.ctypes.data_as(C.c_void_p), conn.ctypes.data_as(C.c_void_p),
                                blk.ctypes.data_as(C.c_void_p), C.c_longlong(nnzb), pe,
                                bd.ctypes.data_as(C.c_void_p), int(bd.size),
                                stats.ctypes.data_as(C.c_void_p))
    assert st == 0, st
    n_patches, max_slots, max_nodes, max_phase, max_nphase, owned, halo_slots, shared = stats
    print(f"c{cid}/{scale}, {pe} per patch: {n_patches} patches, max slots {max_slots}, max nodes {max_nodes}, "
          f"phases {max_phase + 1}/{max_nphase + 1}, owned blocks {owned / nnzb:.2%}, "
          f"halo slots {halo_slots} for {shared} shared blocks")
    assert owned + shared == nnzb
    if pe == 64:
        # the 2D kernel's workgroup of 64 elements fits the 64 KB LDS budget
        ndof = I.shape[1]
        lds = (max_slots * ndof * ndof + max_nodes * ndof) * 8 + (max_slots + max_nodes) * 4
        assert lds <= 64 * 1024, lds
