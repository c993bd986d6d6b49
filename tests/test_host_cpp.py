"""The C++ host layer (gales-multi-fluid_amd/host/gales_mc.hpp): the driver
that runs the reference's time / Newton loop (sim/fluid_mc/P0/main.cpp:206-267)
over it builds and links against the C-ABI library, rejects bad arguments
before touching a device, and on a GPU reproduces the Python binding's
Newton iterations bit for bit (the same library calls in the same order),
whose first update is itself checked against the oracle in
test_gpu_parity.py."""
import os
import subprocess

import numpy as np
import pytest

from gales_mc import abi, cases, casefile
from gales_mc.lib import Context

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gales-multi-fluid_amd")
DRIVER = os.path.join(PKG, "host", "fluid_mc_driver")


def _driver():
    if not os.path.exists(DRIVER):
        pytest.fail("host/fluid_mc_driver is not built (run __graft_entry__.build())")
    return DRIVER


def test_driver_builds_and_links():
    out = subprocess.run(["nm", "-u", _driver()], capture_output=True, text=True, check=True).stdout
    for sym in ("gmc_create", "gmc_assemble", "gmc_rhs_norm", "gmc_solve", "gmc_corrector",
                "gmc_dof_trimmer", "gmc_predictor"):
        assert sym in out, sym


def test_driver_usage_and_bad_case(tmp_path):
    r = subprocess.run([_driver()], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
    bad = tmp_path / "bad.case"
    bad.write_bytes(b"NOTACASE" + b"\0" * 64)
    r = subprocess.run([_driver(), str(bad), "1", "1", "30", "30", "1e-8", "1", str(tmp_path / "o")],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "not a case file" in r.stderr


def test_case_file_round_trip(tmp_path):
    mesh, P, I, Pst = cases.config(0)
    path = tmp_path / "c0.case"
    casefile.write_case(path, mesh, P, I, Pst)
    raw = path.read_bytes()
    assert raw[:8] == casefile.MAGIC
    hdr = np.frombuffer(raw[8:28], np.int32)
    assert list(hdr[:4]) == [2, mesh["coords"].shape[0], mesh["conn"].shape[0], mesh["side_elem"].size]
    n, nd = mesh["coords"].shape[0], 5
    tail = np.frombuffer(raw[len(raw) - 2 * n * nd * 8:], np.float64)
    np.testing.assert_array_equal(tail[: n * nd], I.reshape(-1))
    np.testing.assert_array_equal(tail[n * nd:], Pst.reshape(-1))


@pytest.mark.gpu
def test_driver_newton_loop_matches_binding(tmp_path):
    mesh, P, I, Pst = cases.config(0)
    path = tmp_path / "c0.case"
    casefile.write_case(path, mesh, P, I, Pst)
    out = tmp_path / "I.bin"
    # block-Jacobi GMRES solved tightly, so every Newton update is a sound
    # step (point-Jacobi GMRES(30) capped at 30 iterations is too weak for
    # this stiff system: its second update drives DUDY singular, which the
    # driver reports as status -4)
    steps, its = 2, 2
    for keep_p in (False, True):
        args = [_driver(), str(path), str(steps), str(its), "100", "1000", "1e-10", "2", str(out)]
        if keep_p:
            args.append("1")
        r = subprocess.run(args, capture_output=True, text=True, timeout=120)
        print(r.stdout, r.stderr, flush=True)
        assert r.returncode == 0
        assert r.stdout.count("FLUID  it:") == steps * its
        got = casefile.read_state(out, mesh["coords"].shape[0], 5)
        # the same loop through the Python binding: predictor() starts every
        # step (main.cpp:212), the first one too unless the case's P is kept
        with Context(mesh, P) as ctx:
            ctx.set_state(I, Pst)
            for step in range(steps):
                if step > 0 or not keep_p:
                    ctx.predictor()
                for it in range(its):
                    ctx.assemble()
                    ctx.rhs_norm()
                    ctx.solve(restart=100, max_iters=1000, rel_tol=1e-10, precond=2)
                    ctx.corrector()
                    ctx.dof_trimmer(4, 0.0, 1.0)
            ref, _ = ctx.get_state()
        np.testing.assert_array_equal(got, ref)
        assert not np.array_equal(got, I)


@pytest.mark.gpu
def test_corrector_trimmer_and_rhs_norm():
    """corrector + dof_trimmer(first Y, 0, 1) equals gmc_newton_update, and
    gmc_rhs_norm is ||rhs||_2 of the assembled rhs."""
    mesh, P, I, Pst = cases.config(0)
    with Context(mesh, P) as a, Context(mesh, P) as b:
        for c in (a, b):
            c.set_state(I, Pst)
            c.assemble()
        _, rhs = a.system()
        assert abs(a.rhs_norm() - np.linalg.norm(rhs)) <= 1e-13 * np.linalg.norm(rhs)
        for c in (a, b):
            c.solve(restart=30, max_iters=30, rel_tol=1e-8, precond=1)
        a.newton_update()
        b.corrector()
        b.dof_trimmer(4, 0.0, 1.0)
        np.testing.assert_array_equal(a.get_state()[0], b.get_state()[0])
