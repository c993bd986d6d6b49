"""Case files for the C++ driver (host/fluid_mc_driver.cpp).

Layout: b"GMCCASE1", int32 [dim, n_nodes, n_elems, n_sides, sizeof(gmc_params)],
the gmc_params struct as laid out by the C compiler (abi.Params mirrors it),
then coords f64 [N*dim], conn i32 [E*nen], node_flag i32 [N], side_elem,
side_local, side_flag i32 [S], I f64 [N*ndof], P f64 [N*ndof].
"""
import ctypes as C

import numpy as np

from . import abi

MAGIC = b"GMCCASE1"


def write_case(path, mesh, params, I, P):
    dim = int(mesh["dim"])
    coords = np.ascontiguousarray(mesh["coords"], np.float64)
    conn = np.ascontiguousarray(mesh["conn"], np.int32)
    n_nodes, n_elems = coords.shape[0], conn.shape[0]
    n_sides = int(np.asarray(mesh["side_elem"]).size)
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(np.array([dim, n_nodes, n_elems, n_sides, C.sizeof(abi.Params)], np.int32).tobytes())
        f.write(bytes(C.string_at(C.byref(params), C.sizeof(abi.Params))))
        f.write(coords.tobytes())
        f.write(conn.tobytes())
        for k in ("node_flag", "side_elem", "side_local", "side_flag"):
            f.write(np.ascontiguousarray(mesh[k], np.int32).tobytes())
        f.write(np.ascontiguousarray(I, np.float64).tobytes())
        f.write(np.ascontiguousarray(P, np.float64).tobytes())


def read_state(path, n_nodes, ndof):
    return np.fromfile(path, dtype=np.float64).reshape(n_nodes, ndof)
