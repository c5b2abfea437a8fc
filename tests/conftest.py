"""Shared fixtures.  `gpu` marks tests that need an MI355X (run with -m gpu)."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (MI355X)")


def golden(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz")))


def load_case(name):
    """Input case of a golden fixture, in the reference's conventions."""
    from dgswem_amd import fortio, mesh
    if name == "c1":
        return mesh.structured_basin()
    if name == "c1var":
        return mesh.structured_basin(nx=24, ny=12, slopeflag=5, cori=1.0e-4, nolibf=0,
                                     manning=None, flow_side=True, flow_amp=2.0, depth=8.0)
    case = fortio.read_case(os.path.join(GOLDEN, name))
    if name == "wetdry":
        case = fortio.barriers_as_land(case)
    return case


# (fixture case, steps) pairs holding dgswem_serial FP64 states
GOLDEN_STATES = [("c1", 1), ("c1", 10), ("c1", 1000), ("c1", 8640), ("c1var", 200), ("delaunay4k", 100),
                 ("delaunay4k", 1000), ("quarter_annular", 100), ("quarter_annular", 2000),
                 ("wetdry", 50), ("wetdry", 500), ("mass_conservation", 1000), ("storm3k", 100),
                 ("storm3k", 500), ("delaunay4k_nolim", 300), ("tipot2k", 300), ("tipotsal2k", 200),
                 ("c1eseries", 1000), ("c1fseries", 400)]
