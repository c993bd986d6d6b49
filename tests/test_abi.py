"""The C-ABI library: it loads, exports every symbol include/gales_mc.h
declares, its ctypes mirror matches the C layouts, and the host-side argument
checks of gmc_create (which run before any device call) reject bad input with
the documented status codes.  CPU only."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from gales_mc import abi, cases, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gales_mc.h")


def header_functions():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(gmc_\w+)\(", txt, re.M)))


def test_exports_every_declared_symbol():
    L = lib.load()
    names = header_functions()
    assert len(names) >= 20
    assert sorted(names) == sorted(abi.EXPORTS)
    for n in names:
        assert hasattr(L, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    for n in names:
        assert re.search(rf"\bT {n}\b", out), n


def test_struct_layouts_match_c(tmp_path):
    src = tmp_path / "sz.c"
    structs = {"gmc_magma": abi.Magma, "gmc_component": abi.Component, "gmc_neumann": abi.Neumann,
               "gmc_params": abi.Params, "gmc_mesh": abi.Mesh, "gmc_info": abi.Info,
               "gmc_solver": abi.Solver, "gmc_solve_info": abi.SolveInfo,
               "gmc_timing": abi.Timing}
    body = "".join(f'printf("%zu\\n", sizeof({k}));' for k in structs)
    body += 'printf("%zu\\n", offsetof(gmc_params, precision));'
    body += 'printf("%zu\\n", offsetof(gmc_params, neumann));'
    src.write_text(f'#include <stdio.h>\n#include <stddef.h>\n#include "{HEADER}"\n'
                   f"int main(void){{{body}return 0;}}\n")
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    vals = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                           check=True).stdout.split()]
    for (k, cls), v in zip(structs.items(), vals):
        assert C.sizeof(cls) == v, k
    assert abi.Params.precision.offset == vals[-2]
    assert abi.Params.neumann.offset == vals[-1]


def _create(mesh, P):
    L = lib.load()
    m = cases.mesh_struct(mesh)
    h = C.c_void_p()
    st = L.gmc_create(C.byref(m), C.byref(P), C.byref(h))
    return st, L.gmc_last_error(None).decode()


def test_create_rejects_clockwise_element():
    mesh = cases.channel2d(4, 2)
    P = cases.make_params(2, [cases.MELT, cases.GAS_R461])
    mesh["conn"] = np.ascontiguousarray(mesh["conn"][:, [0, 2, 1]])
    st, msg = _create(mesh, P)
    assert st == abi.GMC_ERR_MESH and "Jacobian" in msg


def test_create_rejects_bad_arguments():
    mesh = cases.channel2d(4, 2)
    P = cases.make_params(2, [cases.MELT, cases.GAS_R461])
    P.tau_variant = 1999
    assert _create(mesh, P)[0] == abi.GMC_ERR_ARG
    P = cases.make_params(2, [cases.MELT, cases.GAS_R461])
    P.tau_variant = abi.GMC_TAU_NON_DIAG_COMP_2001
    P.steady_state = 1                                  # tau = 1/sqrt(0): infinite
    st, msg = _create(mesh, P)
    assert st == abi.GMC_ERR_UNSUPPORTED and "infinite" in msg
    P = cases.make_params(2, [cases.MELT, cases.GAS_R461])
    P.dc_sharp = 3.0
    assert _create(mesh, P)[0] == abi.GMC_ERR_ARG
    P = cases.make_params(2, [cases.MELT, cases.GAS_R461, cases.H2O, cases.CRYSTAL])
    assert _create(mesh, P)[0] == abi.GMC_ERR_UNSUPPORTED
    P = cases.make_params(2, [cases.MELT, cases.GAS_R461])
    bad = dict(mesh)
    bad["node_flag"] = mesh["node_flag"].copy()
    bad["node_flag"][0] = 99
    assert _create(bad, P)[0] == abi.GMC_ERR_ARG
    bad = dict(mesh)
    bad["conn"] = mesh["conn"].copy()
    bad["conn"][0, 0] = 10 ** 6
    assert _create(bad, P)[0] == abi.GMC_ERR_ARG
    P = cases.make_params(3, [cases.MELT, cases.CRYSTAL, cases.H2O])
    assert _create(mesh, P)[0] == abi.GMC_ERR_ARG       # dim mismatch


def test_no_fallback_when_library_missing(monkeypatch, tmp_path):
    monkeypatch.setattr(lib, "_lib", None)
    monkeypatch.setattr(lib, "LIB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(RuntimeError):
        lib.load()
