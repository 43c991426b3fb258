"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on the CPU."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def host_tables():
    from scream_amd import p3x
    return p3x.tables_load()


@pytest.fixture(scope="session")
def oracle_tables(host_tables):
    """The product's host tables as the numpy dict the oracle takes."""
    t = host_tables
    a = lambda x, sh: np.ctypeslib.as_array(x).reshape(sh).copy()
    return dict(ice=a(t.ice_table, (5, 4, 50, 12)), coll=a(t.collect_table, (5, 4, 50, 30, 2)),
                vn=a(t.vn_table, (300, 10)), vm=a(t.vm_table, (300, 10)),
                revap=a(t.revap_table, (300, 10)), mu_r=a(t.mu_r_table, (150,)),
                dnu=a(t.dnu, (16,)))
