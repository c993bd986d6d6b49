"""gmc_rhs_norm_noalloc against gmc_rhs_norm, through the C-ABI.

The allocation-free variant keeps its reduction's partial sums in a fixed
device array shared by every call on the GPU; two calls in a row, the second
on a smaller system (fewer partial sums, stale ones left in the array above
them), must each give gmc_rhs_norm's result bit for bit.
"""
import sys

import numpy as np
import pytest

from gales_mc import cases
from gales_mc.lib import Context

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_rhs_norm_noalloc_matches_rhs_norm():
    # c1 family at 1/4 (659,205 unknowns: 644 partial sums), then config 0
    # (10,725 unknowns: 11 partial sums)
    for cid, scale in ((1, 4), (0, 1)):
        mesh, P, I, Pst = cases.config(cid, scale=scale)
        _say(f"rhs norm noalloc c{cid}/{scale}")
        with Context(mesh, P) as ctx:
            ctx.set_state(I, Pst)
            with pytest.raises(Exception):
                ctx.rhs_norm_noalloc()            # no rhs assembled yet
            ctx.assemble()
            a = ctx.rhs_norm_noalloc()
            b = ctx.rhs_norm()
            c = ctx.rhs_norm_noalloc()
            _, rhs = ctx.system()
            ref = np.linalg.norm(rhs)
            _say(f"  noalloc {a:.17e} rhs_norm {b:.17e} numpy {ref:.17e}")
            assert a == b == c
            assert abs(a - ref) <= 1e-14 * ref
