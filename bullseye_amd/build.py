"""Build the three ABI variants of the gfx950 imaging library in-tree with hipcc.

    python -m bullseye_amd.build            # all builds
    python -m bullseye_amd.build mixed      # one build

Objects go to bullseye_amd/_build/<precision>/, libraries to bullseye_amd/lib/.
Packed-FP32 instructions are disabled for all device code.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

from . import LIB_DIR, LIBRARY_NAMES, PACKAGE_DIR, PRECISIONS

CSRC = os.path.join(PACKAGE_DIR, "csrc")
SOURCES = ("romein.hip", "romein_w.hip", "wproj.hip", "jones.hip", "ingest.hip",
           "finalize.hip", "degrid.hip", "dft.hip", "batch.hip", "calibrate.hip", "wrapper.hip")
DEFINES = {"single": "BULLSEYE_SINGLE", "double": "BULLSEYE_DOUBLE", "mixed": "BULLSEYE_MIXED"}
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
CFLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall",
          "-Wno-unused-function", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build(precisions=PRECISIONS, jobs=None):
    os.makedirs(LIB_DIR, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    compile_cmds = []
    for prec in precisions:
        odir = os.path.join(PACKAGE_DIR, "_build", prec)
        os.makedirs(odir, exist_ok=True)
        for src in SOURCES:
            obj = os.path.join(odir, src.replace(".hip", ".o"))
            compile_cmds.append([HIPCC] + CFLAGS + ["-D" + DEFINES[prec], "-c",
                                                     os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(jobs) as ex:
        for out in ex.map(_run, compile_cmds):
            if out.strip():
                print(out)
    for prec in precisions:
        odir = os.path.join(PACKAGE_DIR, "_build", prec)
        objs = [os.path.join(odir, s.replace(".hip", ".o")) for s in SOURCES]
        lib = os.path.join(LIB_DIR, LIBRARY_NAMES[prec])
        _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", lib + ".tmp"] + objs +
             ["-L/opt/rocm/lib", "-lhipfft", "-Wl,-rpath,/opt/rocm/lib", "-Wl,-Bsymbolic"])
        os.replace(lib + ".tmp", lib)
    return [os.path.join(LIB_DIR, LIBRARY_NAMES[p]) for p in precisions]


if __name__ == "__main__":
    precs = tuple(sys.argv[1:]) or PRECISIONS
    for path in build(precs):
        print(path)
