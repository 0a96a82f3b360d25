"""The C-ABI boundary without a GPU: struct layout, exported symbols, loader behaviour."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from bullseye_amd import INCLUDE_DIR, PRECISIONS
from bullseye_amd.gridding_parameters import gridding_parameters_class
from bullseye_amd.library_loader import (bullseye_batch_facet, bullseye_batch_tuning,
                                         bullseye_options, bullseye_stats, library_path,
                                         load_library)
from oracle import oracle

HEADER = os.path.join(INCLUDE_DIR, "bullseye_amd.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(\w+)\s*\([^;]*\);", text, re.M)))


def test_header_declares_the_reference_entry_points():
    names = header_functions()
    reference = ["get_gridding_walltime", "get_inversion_walltime", "gridding_barrier",
                 "initLibrary", "releaseLibrary", "weight_uniformly", "normalize",
                 "repack_input_data", "finalize", "finalize_psf", "grid_single_pol",
                 "facet_single_pol", "grid_duel_pol", "facet_duel_pol", "grid_4_cor",
                 "facet_4_cor", "facet_4_cor_corrections", "grid_sampling_function",
                 "facet_sampling_function"]
    for r in reference:
        assert r in names


@pytest.mark.parametrize("precision", PRECISIONS)
def test_library_exports_every_header_symbol(precision):
    path = library_path(precision)
    assert os.path.exists(path), "build the library first (__graft_entry__.build())"
    lib = ctypes.CDLL(path)   # loading needs no GPU
    for name in header_functions():
        assert hasattr(lib, name), name


@pytest.mark.parametrize("precision", ["single", "double"])
def test_struct_layout_matches_c(precision):
    """ctypes field offsets == offsetof in the C header (compiled into the oracle)."""
    lib = oracle.load(precision)
    lib.oracle_struct_offsets.restype = ctypes.c_size_t
    n = lib.oracle_struct_offsets(None)
    buf = (ctypes.c_size_t * n)()
    lib.oracle_struct_offsets(buf)
    cls = gridding_parameters_class(precision)
    offs = [getattr(cls, f[0]).offset for f in cls._fields_]
    assert offs == list(buf)[:-1]
    assert ctypes.sizeof(cls) == buf[-1]


def test_mixed_layout_equals_double():
    a, b = gridding_parameters_class("mixed"), gridding_parameters_class("double")
    assert [getattr(a, f[0]).offset for f in a._fields_] == \
        [getattr(b, f[0]).offset for f in b._fields_]


def test_options_and_stats_sizes():
    assert ctypes.sizeof(bullseye_options) == 64
    assert ctypes.sizeof(bullseye_stats) == 48


def test_options_and_stats_layout_match_header(tmp_path):
    """The ctypes mirrors of the extension structs against the C header, compiled with gcc."""
    fields = {"bullseye_options": bullseye_options, "bullseye_stats": bullseye_stats,
              "bullseye_batch_facet": bullseye_batch_facet,
              "bullseye_batch_tuning": bullseye_batch_tuning}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "bullseye_amd.h"',
             'int main(void) {']
    for name, cls in fields.items():
        lines.append('printf("%%zu\\n", sizeof(%s));' % name)
        for f in cls._fields_:
            lines.append('printf("%%zu\\n", offsetof(%s, %s));' % (name, f[0]))
    lines.append('return 0; }')
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-DBULLSEYE_DOUBLE", "-I", INCLUDE_DIR, str(src), "-o", str(exe)],
                   check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True,
                                          text=True).stdout.split()]
    want = []
    for cls in fields.values():
        want.append(ctypes.sizeof(cls))
        want += [getattr(cls, f[0]).offset for f in cls._fields_]
    assert got == want


def test_loader_rejects_bad_arguments_like_the_reference():
    with pytest.raises(Exception, match="architecture"):
        load_library("CPU", "single", share_torch_runtime=False)
    with pytest.raises(Exception, match="precision"):
        load_library("GPU", "quad", share_torch_runtime=False)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_calls_without_init_fail_cleanly(precision):
    """Entry points return a status and set bullseye_last_error instead of crashing."""
    lib = load_library("GPU", precision, share_torch_runtime=False)
    p = gridding_parameters_class(precision)()
    assert lib.grid_single_pol(ctypes.byref(p)) != 0
    assert b"initLibrary" in lib.bullseye_last_error()
    o = bullseye_options(struct_size=3)
    assert lib.bullseye_set_options(ctypes.byref(o)) != 0
    facets = (bullseye_batch_facet * 1)()
    assert lib.bullseye_batch_plan(ctypes.byref(p), 1, facets, 1, 0) != 0
    assert b"initLibrary" in lib.bullseye_last_error()
    assert lib.bullseye_batch_execute() != 0
    assert lib.bullseye_batch_kernel_set(0, None, 2, 3, 4, None, ctypes.c_float(1.0)) != 0
    t = bullseye_batch_tuning(struct_size=ctypes.sizeof(bullseye_batch_tuning))
    t.chunk[0] = 3
    assert lib.bullseye_batch_set_tuning(ctypes.byref(t)) != 0
