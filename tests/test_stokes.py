"""Stokes formation and compaction (bullseye/data/stokes.py:74-172,
bullseye/bin/bullseye_pipeliner.py:406-416).

CPU: the oracle's NumPy restatement (oracle.pipeliner_stokes) and the host-side slot logic
(bullseye_amd.stokes) pinned by closed forms — Stokes I = (XX + YY)/2, V = (RR - LL)/2, the
compaction positions for P = 1, 2, 4, and the reference's behaviour with Jones corrections (no
combination).  GPU: bullseye_form_stokes against the restatement, bit for bit, host and device
buffers.
"""
import numpy as np
import pytest

from bullseye_amd import stokes
from oracle import oracle

LIN = [9, 10, 11, 12]
CIRC = [5, 6, 7, 8]


def grids(F, C, P, n=6, seed=0, dtype=np.complex64):
    rng = np.random.default_rng(seed)
    return (rng.normal(size=(F, C, P, n, n)) + 1j * rng.normal(size=(F, C, P, n, n))).astype(dtype)


def test_stokes_i_linear_closed_form():
    g = grids(2, 1, 2, seed=1)
    out = oracle.pipeliner_stokes(g, False, LIN, [9, 12], "linear", "I")
    np.testing.assert_array_equal(out[:, 0, 0], (g[:, 0, 0] + g[:, 0, 1]) / np.float32(2))
    np.testing.assert_array_equal(out[:, 0, 1], g[:, 0, 1])


def test_stokes_v_circular_closed_form():
    g = grids(1, 1, 2, seed=2)
    out = oracle.pipeliner_stokes(g, False, CIRC, [5, 8], "circular", "V")
    np.testing.assert_array_equal(out[0, 0, 0], (g[0, 0, 0] - g[0, 0, 1]) / np.float32(2))


@pytest.mark.parametrize("P,targets", [(2, [0, 1, 2, 3]), (4, [0, 3, 6, 9]), (3, [0, 2, 4, 6])])
def test_compaction_positions(P, targets):
    """Channel c's slot 0 lands at flat slot (P-1)*c of its facet block."""
    C = 4
    g = grids(2, C, P, seed=3)
    out = oracle.pipeliner_stokes(g, False, LIN, LIN[:P], None, "I")
    flat_out = out.reshape(2, C * P, 6, 6)
    for c, t in enumerate(targets):
        np.testing.assert_array_equal(flat_out[:, t], g[:, c, 0])


def test_compaction_single_correlation_keeps_last_channel():
    g = grids(1, 3, 1, seed=4)
    out = oracle.pipeliner_stokes(g, False, [1], [1], "stokes", "I")
    np.testing.assert_array_equal(out[0, 0, 0], g[0, 2, 0])
    np.testing.assert_array_equal(out[0, 1:], g[0, 1:])


def test_jones_corrections_leave_slot0():
    """With Jones corrections the reference leaves the feed type undetermined, so no
    combination happens (stokes.py:77-78, 91-131)."""
    grid_c, feeds = stokes.correlations_to_grid(True, "I", LIN)
    assert grid_c == LIN and feeds is None
    assert stokes.stokes_slots(True, "I", feeds, LIN, grid_c) == (-1, -1, 0)
    g = grids(1, 1, 4, seed=5)
    out = oracle.pipeliner_stokes(g, True, LIN, grid_c, feeds, "I")
    np.testing.assert_array_equal(out, g)


@pytest.mark.parametrize("wanted,data,expect", [
    ("I", LIN, ([9, 12], "linear", (0, 1, 1))), ("V", CIRC, ([5, 8], "circular", (0, 1, -1))),
    ("U", LIN, ([10, 11], "linear", (0, 1, 1))), ("Q", [1, 2, 3, 4], ([2], "stokes", (-1, -1, 0))),
    ("XY", LIN, ([10], "linear", (-1, -1, 0)))])
def test_slot_selection(wanted, data, expect):
    grid_c, feeds = stokes.correlations_to_grid(False, wanted, data)
    assert (grid_c, feeds) == expect[:2]
    assert stokes.stokes_slots(False, wanted, feeds, data, grid_c) == expect[2]
