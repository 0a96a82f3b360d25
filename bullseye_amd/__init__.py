"""bullseye_amd: MI355X (gfx950) imaging hot path behind bullseye's C ABI.

The package mirrors bullseye's host-side boundary (bullseye/mo/base_types.py,
bullseye/mo/gridding_parameters.py, bullseye/mo/library_loader.py) and ships three builds of
a HIP library exporting the 19 entry points of bullseye/mo/cpu_gpu_common/wrapper.h:41-61.
"""
import os

PACKAGE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PACKAGE_DIR, "lib")
INCLUDE_DIR = os.path.join(os.path.dirname(PACKAGE_DIR), "include")

PRECISIONS = ("single", "double", "mixed")
LIBRARY_NAMES = {
    "single": "libmi355x_imaging32.so",
    "double": "libmi355x_imaging64.so",
    "mixed": "libmi355x_imaging_mixed.so",
}
