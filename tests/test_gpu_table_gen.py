"""The device ice lookup-table generator (p3x_ice_table_generate): the whole
table must format to the file the reference's own generator wrote, character for
character, and its doubles must equal the CPU restatement's (portable math) bit
for bit on sampled cells."""
import os
import sys
import tempfile

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import oracle  # noqa: E402
from table_gen_rows import TABLE, mismatches, table_lines  # noqa: E402

pytestmark = pytest.mark.gpu

CELLS = [0, 49, 130, 257, 318, 444, 521, 603, 688, 777, 850, 999]


@pytest.fixture(scope="module")
def generated():
    from scream_amd import p3x
    print("generating the ice table on the device", flush=True)
    ice, coll, ms = p3x.ice_table_generate()
    print(f"generated in {ms:.0f} ms", flush=True)
    return ice, coll


def test_device_table_equals_reference_file(generated):
    ice, coll = generated
    bad = mismatches(ice, coll, range(1000), table_lines())
    assert not bad, f"{len(bad)} rows differ, first: {bad[:2]}"


def test_device_table_bitwise_equals_oracle(generated):
    ice, coll = generated
    want_ice, want_coll = oracle.table_gen("portable", CELLS)
    np.testing.assert_array_equal(ice[CELLS], want_ice[CELLS])
    np.testing.assert_array_equal(coll[CELLS], want_coll[CELLS])


def test_generated_file_loads_as_the_reference_tables(generated):
    from scream_amd import p3x
    ice, coll = generated
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "gen.dat")
        p3x.ice_table_write(path, ice, coll)
        got = p3x.tables_load(path)
    want = p3x.tables_load(TABLE)
    for name in ("ice_table", "collect_table", "vn_table", "vm_table", "revap_table", "mu_r_table", "dnu"):
        np.testing.assert_array_equal(np.ctypeslib.as_array(getattr(got, name)),
                                      np.ctypeslib.as_array(getattr(want, name)), err_msg=name)
