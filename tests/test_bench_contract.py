"""bench.py's host-side contract, on the CPU: more than one GPU is refused before anything
touches a device, and every profile summary the benchmark reads its roofline traffic and flush
atomics from is committed and parses."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "only --gpus 1" in r.stderr
    assert r.stdout == ""


@pytest.mark.parametrize("profile", ["romein_w_facet_4_cor_corrections_fp32comp_pmc.md",
                                     "wproj_grid_single_pol_fp32comp_pmc.md",
                                     "batch_facets_fp32comp_pmc.md"])
def test_profiles_the_benchmark_reads(profile):
    traffic = bench.profiled_traffic(profile)
    atomics = bench.profiled_value(profile, "atomics_per_launch")
    assert traffic is not None and traffic > 1e9
    assert atomics is not None and atomics > 1e6


def test_missing_profile_reads_as_none():
    assert bench.profiled_traffic("no_such_profile.md") is None
    assert bench.profiled_value("romein_w_facet_4_cor_corrections_fp32comp_pmc.md",
                                "no_such_key") is None
