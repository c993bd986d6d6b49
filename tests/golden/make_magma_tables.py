"""Writes tests/golden/magma_tables.npz: the solubility tables the reference's
shipped fluid_mc case reads (sim/fluid_mc/P0/input/lookuptable/: pressure.txt,
wf_l_{h2o,co2}_onliq_{shoshonite,phonolite}.txt), read with the package's own
reader of that format.  These are data files of the reference, kept as a
fixture so the tests run without /root/reference.

    python tests/golden/make_magma_tables.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "gales-multi-fluid_amd"))
from gales_mc.meshio import read_lookup_tables  # noqa: E402

SRC = "/root/reference/sim/fluid_mc/P0/input/lookuptable"


def main():
    out = {}
    for name in ("shoshonite", "phonolite"):
        t = read_lookup_tables(SRC, name)
        assert t["T"] is None
        out["p"] = t["p"]
        out[name + "_h2o"] = t["wf_h2o_l"]
        out[name + "_co2"] = t["wf_co2_l"]
    np.savez_compressed(os.path.join(HERE, "magma_tables.npz"), **out)
    print({k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
