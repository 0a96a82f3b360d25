"""Stokes terms from gridded correlations, as bullseye's pipeliner forms them, on the device.

bullseye grids one to four correlation products per facet and channel, normalises them, then
(on host arrays) combines two of them into the wanted Stokes term and compacts the per-channel
grids before finalize (bullseye/data/stokes.py:74-172, bullseye/bin/bullseye_pipeliner.py:
389-416).  `form_stokes` does the same through the library's bullseye_form_stokes, on the
output buffer where it already lives (device memory with device_resident).

The polarisation codes are casacore's Stokes enumeration (the values stokes.py:40 uses).
"""
import ctypes

POL_CODES = {"I": 1, "Q": 2, "U": 3, "V": 4, "RR": 5, "RL": 6, "LR": 7, "LL": 8,
             "XX": 9, "XY": 10, "YX": 11, "YY": 12}

# stokes.py:45-72: per wanted term, the correlation sets it can be derived from and the feed
# type of each set
DERIVATIONS = {
    "I": [(["I"], "stokes"), (["RR", "LL"], "circular"), (["XX", "YY"], "linear")],
    "V": [(["V"], "stokes"), (["RR", "LL"], "circular"), (["XY", "YX"], "linear")],
    "U": [(["U"], "stokes"), (["RL", "LR"], "circular"), (["XY", "YX"], "linear")],
    "Q": [(["Q"], "stokes"), (["RL", "LR"], "circular"), (["XX", "YY"], "linear")],
}
for _c in ("RR", "RL", "LR", "LL"):
    DERIVATIONS[_c] = [([_c], "circular")]
for _c in ("XX", "XY", "YX", "YY"):
    DERIVATIONS[_c] = [([_c], "linear")]

# stokes.py:90-172: (first, second, sign) of the combination per (feeds, wanted term)
COMBINATIONS = {
    ("circular", "I"): ("RR", "LL", +1), ("circular", "V"): ("RR", "LL", -1),
    ("circular", "Q"): ("RL", "LR", +1), ("circular", "U"): ("RL", "LR", -1),
    ("linear", "I"): ("XX", "YY", +1), ("linear", "Q"): ("XX", "YY", -1),
    ("linear", "U"): ("XY", "YX", +1), ("linear", "V"): ("XY", "YX", -1),
}


def correlations_to_grid(jones_corrections, wanted, data_correlations):
    """stokes.py:74-84: the correlation codes to grid and the feed type in use.  With Jones
    corrections every correlation of the data is gridded (feeds stay undetermined, as in the
    reference); otherwise the last derivation whose correlations the data holds."""
    data = [int(c) for c in data_correlations]
    if jones_corrections:
        return data, None
    chosen, feeds = None, None
    for names, feed in DERIVATIONS[wanted]:
        codes = [POL_CODES[n] for n in names]
        if set(codes).issubset(data):
            chosen, feeds = codes, feed
    return chosen, feeds


def stokes_slots(jones_corrections, wanted, feeds, data_correlations, gridded_correlations):
    """(first, second, sign) of the grid slots to combine, or (-1, -1, 0) when the wanted term
    is stored as gridded.  With Jones corrections the slots index the data's correlations
    (stokes.py:91-131), otherwise the gridded list (stokes.py:132-172)."""
    key = (feeds, wanted)
    if key not in COMBINATIONS:
        return -1, -1, 0
    a, b, sign = COMBINATIONS[key]
    order = [int(c) for c in (data_correlations if jones_corrections else gridded_correlations)]
    return order.index(POL_CODES[a]), order.index(POL_CODES[b]), sign


def form_stokes(lib, binding, first, second, sign, wanted, contiguous=False):
    """Run bullseye_form_stokes on the binding's output buffer: the combination, then the
    pipeliner's compaction when the wanted term is I, Q, U or V (bullseye_pipeliner.py:410)."""
    compact = 1 if wanted in ("I", "Q", "U", "V") else 0
    rc = lib.bullseye_form_stokes(binding.byref(), ctypes.c_int(first), ctypes.c_int(second),
                                  ctypes.c_int(sign), ctypes.c_int(compact),
                                  ctypes.c_int(1 if contiguous else 0))
    if rc != 0:
        raise RuntimeError(lib.bullseye_last_error().decode())
