import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


def load_golden(name):
    """(npz dict, System, Params, OrcParams) of a golden case written by make_golden.py."""
    from bd_box_amd import workloads as wl
    from oracle import pyoracle as po
    g = dict(np.load(os.path.join(GOLDEN, name + ".npz")))
    prm = {k[4:]: str(v) for k, v in g.items() if k.startswith("prm.")}
    sysm = wl.System(g["in_coord"], g["in_Q"], g["in_LJ"], g["in_bonds"].astype(np.int32),
                     g["in_bond_parms"])
    P = wl.parse_params(prm)
    return g, sysm, P, po.params_from(P)


@pytest.fixture
def golden():
    return load_golden
