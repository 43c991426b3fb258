"""The ice lookup-table generator (create_p3_lookupTable_1.f90-v4.1) restated in
oracle/p3_table_gen.c, checked against the table the reference's own generator
wrote (scream_amd/data/p3_lookup_table_1.dat-v4.1.1, produced by
scream_amd/data/make_ice_table.sh): every row of the sampled cells must format to
the same characters.  Also checks the library's table writer on the committed
values.  (The device generator is checked in test_gpu_table_gen.py.)"""
import os
import sys
import tempfile

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import oracle  # noqa: E402
from table_gen_rows import TABLE, mismatches, parse, table_lines  # noqa: E402

# 3 masses in each of the 20 (rime density, rime fraction) pairs, spread over
# the normalized-mass range; the whole table is compared on the GPU.
CELLS = sorted({rf * 50 + q for rf in range(20) for q in ((rf * 7) % 50, (rf * 7 + 19) % 50, 49 - rf)})


@pytest.fixture(scope="module")
def lines():
    return table_lines()


@pytest.mark.parametrize("math", ["glibc", "portable"])
def test_oracle_reproduces_reference_table(lines, math):
    ice, coll = oracle.table_gen(math, CELLS)
    bad = mismatches(ice, coll, CELLS, lines)
    assert not bad, f"{len(bad)} rows differ, first: {bad[:2]}"


def test_table_writer_round_trip(lines):
    """p3x_ice_table_write formats the committed values back to the same file."""
    from scream_amd import p3x
    ice, coll = parse(lines)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "t.dat")
        p3x.ice_table_write(path, ice, coll)
        with open(path) as fh:
            got = fh.read()
    with open(TABLE) as fh:
        want = fh.read()
    assert got == want


def test_generator_invariants():
    """Properties the generator's definitions fix: the searched slope reproduces the
    cell's mass (relative error below the 0.13 % candidate spacing), number-weighted
    fall speeds below mass-weighted ones, and the lambda-limiter numbers ordered."""
    cells = CELLS[::6]
    ice, coll = oracle.table_gen("glibc", cells)
    for c in cells:
        q, uns, ums, nlarge, nsmall, lam = ice[c, 0], ice[c, 2], ice[c, 3], ice[c, 8], ice[c, 9], ice[c, 10]
        assert q == pytest.approx(261.7 ** ((c % 50 + 11) * 0.1) * 1e-18, rel=1e-15)
        assert 0 < uns <= ums
        assert nlarge >= nsmall > 0
        assert 100.0 <= lam <= 7.0 / 2e-6
        # the 30 rain sizes: lamrs = 1 / (1.24^i * 10 um), collection rates positive
        np.testing.assert_allclose(coll[c, :, 2], 1.0 / (1.24 ** np.arange(1, 31) * 10e-6), rtol=1e-13)
        assert (coll[c, :, 3:] >= 0).all()
