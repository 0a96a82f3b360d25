"""Fixture for the casacore table reader: the reference's EVLA antenna table (misc/EVLA, the
table its simulation script hands to makems) copied as data, and the antenna coordinates
bullseye/mo/benchmarks/benchmark.cpp:63-93 lists as "taken from an actual measurement set".

    python tests/golden/make_evla_fixture.py /root/reference

writes tests/golden/evla_antenna/{table.dat,table.f0,table.info} and
tests/golden/evla_benchmark_positions.json (27 x 3 metres, the decimal text of benchmark.cpp).
"""
import json
import os
import re
import shutil
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
here = os.path.dirname(os.path.abspath(__file__))
dst = os.path.join(here, "evla_antenna")
os.makedirs(dst, exist_ok=True)
for f in ("table.dat", "table.f0", "table.info"):
    shutil.copyfile(os.path.join(ref, "misc", "EVLA", f), os.path.join(dst, f))
src = open(os.path.join(ref, "bullseye", "mo", "benchmarks", "benchmark.cpp")).read()
block = src[src.index("antenna_coords[] = {"):src.index("};", src.index("antenna_coords[] = {"))]
triples = re.findall(r"\{\s*([-\d.]+)f?\s*,\s*([-\d.]+)f?\s*,\s*([-\d.]+)f?\s*\}", block)
pos = [[float(a), float(b), float(c)] for a, b, c in triples]
assert len(pos) == 27, len(pos)
json.dump({"source": "bullseye/mo/benchmarks/benchmark.cpp:63-93", "positions_m": pos},
          open(os.path.join(here, "evla_benchmark_positions.json"), "w"), indent=1)
print("wrote", dst, "and", len(pos), "positions")
