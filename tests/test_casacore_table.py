"""The casacore table reader (bullseye_amd.casacore_table) against the reference's own table.

Fixture (tests/golden/make_evla_fixture.py): the EVLA antenna table of the reference
(misc/EVLA, the table misc/gen_data.sh:3 hands to makems) and the 27 antenna coordinates
bullseye/mo/benchmarks/benchmark.cpp:63-93 lists as taken from a Measurement Set.  The reader
must return those coordinates to the last printed decimal, and the other columns must read as
the VLA's (a different row order, a misplaced column offset, a wrong string length or a wrong
byte order fails one of these).
"""
import json
import os
import re

import numpy as np
import pytest

from bullseye_amd.casacore_table import AipsIO, Table

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EVLA = os.path.join(GOLDEN, "evla_antenna")


@pytest.fixture(scope="module")
def evla():
    return Table(EVLA)


def test_description(evla):
    assert evla.nrows == 27
    assert evla.colnames() == ["OFFSET", "POSITION", "TYPE", "DISH_DIAMETER", "FLAG_ROW",
                               "MOUNT", "NAME", "STATION"]
    assert evla.managers == [("StandardStMan", 0)]
    assert evla.columns["POSITION"].shape == (3,)
    assert evla.getcolkeyword("POSITION", "MEASINFO") == {"type": "position", "Ref": "ITRF"}
    assert list(evla.getcolkeyword("POSITION", "QuantumUnits")) == ["m", "m", "m"]


def test_positions_equal_the_benchmark_coordinates(evla):
    want = np.array(json.load(open(os.path.join(GOLDEN, "evla_benchmark_positions.json")))
                    ["positions_m"])
    got = evla.getcol("POSITION")
    assert got.shape == (27, 3) and got.dtype == np.float64
    # benchmark.cpp prints the MS doubles with six decimals
    assert np.abs(got - want).max() <= 5e-7, np.abs(got - want).max()


def test_positions_are_the_vla_on_the_earth(evla):
    xyz = evla.getcol("POSITION")
    r = np.linalg.norm(xyz, axis=1)
    assert np.all(np.abs(r - 6.3735e6) < 5e3)              # on the geoid near 2.1 km altitude
    lon = np.degrees(np.arctan2(xyz[:, 1], xyz[:, 0]))
    lat = np.degrees(np.arctan2(xyz[:, 2], np.hypot(xyz[:, 0], xyz[:, 1])))
    assert np.all(np.abs(lon + 107.62) < 0.1) and np.all(np.abs(lat - 33.9) < 0.2)
    # the D configuration: all antennas within 1 km of each other
    assert np.linalg.norm(xyz[:, None] - xyz[None], axis=-1).max() < 1100


def test_scalar_and_string_columns(evla):
    names = list(evla.getcol("NAME"))
    assert names == ["ea%02d" % i for i in range(1, 29) if i != 14]
    assert all(re.fullmatch(r"[NEW]\d\d", s) for s in evla.getcol("STATION"))
    assert len(set(evla.getcol("STATION"))) == 27
    # 12 characters: stored in the string bucket, not inline
    assert list(evla.getcol("TYPE")) == ["GROUND-BASED"] * 27
    assert list(evla.getcol("MOUNT")) == ["ALT-AZ"] * 27
    np.testing.assert_array_equal(evla.getcol("DISH_DIAMETER"), np.full(27, 25.0))
    assert evla.getcol("FLAG_ROW").dtype == bool and not evla.getcol("FLAG_ROW").any()
    off = evla.getcol("OFFSET")
    assert off.shape == (27, 3) and np.abs(off).max() < 1e-9


def test_row_ranges_and_cells(evla):
    full = evla.getcol("POSITION")
    np.testing.assert_array_equal(evla.getcol("POSITION", 5, 7), full[5:12])
    np.testing.assert_array_equal(evla.getcell("POSITION", 26), full[26])
    assert evla.getcell("NAME", 13) == "ea15"
    assert list(evla.getcol("STATION", 20, 3)) == list(evla.getcol("STATION")[20:23])


def test_other_storage_managers_are_refused(evla):
    t = Table(EVLA)
    t.managers = [("TiledShapeStMan", 0)]
    with pytest.raises(NotImplementedError, match="TiledShapeStMan"):
        t.getcol("POSITION")


def test_truncated_or_foreign_files_are_rejected(tmp_path):
    raw = open(os.path.join(EVLA, "table.dat"), "rb").read()
    with pytest.raises(ValueError, match="AipsIO"):
        AipsIO(b"\0" * 16).getstart("Table", outer=True)
    d = tmp_path / "t"
    d.mkdir()
    (d / "table.dat").write_bytes(raw[:400])
    with pytest.raises(Exception):
        Table(str(d))
