"""C-ABI library checks that need no GPU: it loads, exports every declared symbol, and its
host-side CREATE_EDGE_DATA equals the oracle's (which equals the reference's numbering)."""
import os
import re

import numpy as np
import pytest

from conftest import ROOT, load_case
from dgswem_amd import gpu
from oracle.oracle import Oracle


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "dgswem_gpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dg_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = gpu.load_library()
    syms = declared_symbols()
    assert len(syms) >= 14
    for s in syms:
        assert hasattr(L, s), s
    assert sorted(gpu.EXPORTED) == syms


@pytest.mark.parametrize("name", ["c1", "c1var", "delaunay4k", "quarter_annular", "wetdry",
                                  "mass_conservation"])
def test_edge_tables_match_oracle(name):
    case = load_case(name)
    ed = gpu.create_edge_data(case)
    oe = Oracle(case).edges()
    for k in ("nedno", "nedel", "nedsd", "niedn", "nledn", "needn", "nfedn"):
        np.testing.assert_array_equal(ed[k], oe[k], err_msg=k)


def test_edge_data_rejects_bad_mesh():
    case = load_case("c1")
    case.nm = case.nm.copy()
    case.nm[5] = case.nm[4]            # duplicate element
    with pytest.raises(gpu.GpuError, match="duplicate"):
        gpu.create_edge_data(case)
