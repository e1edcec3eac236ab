"""Build script for the MFM pass libraries (runs on a CPU-only host too).

  lib/libgzm.so         HIP kernels + C ABI (include/gzm.h), gfx950
  lib/libgizmo_host.so  C host that keeps GIZMO's density() /
                        hydro_gradient_calc() / hydro_force() over P[]/SphP[]
                        (include/gizmo_mfm_host.h), linked against libgzm.so
  lib/libgizmo_io.so    the HDF5 IC / snapshot layout over the host's P[]/SphP[]
                        (include/gizmo_io.h), against the HDF5 C library

Usage: python gizmo-public_amd/build.py
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib")
INC = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GZM_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: the FP64 pass keeps the reference's rounding (no fused
# multiply-add), so neighbour predicates and sums match the CPU pass exactly.
# The packed-FP32 feature is disabled for every HIP translation unit.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", f"-I{INC}",
]

HIP_SOURCES = ["csrc/gzm_api.hip", "csrc/gzm_kernels.hip", "csrc/gzm_kernels_f32.hip", "csrc/gzm_tree.hip",
               "csrc/gzm_elements.hip", "csrc/gzm_mfv.hip"]
# per-file flags after HIP_FLAGS.  The FP32-bulk kernels stay at
# -ffp-contract=off too: fused multiply-adds gained nothing measurable there,
# and contraction decided per inlining site made the pair-list and scanning
# force loops round differently.
FILE_FLAGS = {}
HOST_SOURCES = ["host/gizmo_host.c", "host/gizmo_integrate.c"]
IO_SOURCE = "host/gizmo_io.c"
# the HDF5 C library of this image (1.10); GIZMO_HDF5_PREFIX selects another
HDF5_PREFIX = os.environ.get("GIZMO_HDF5_PREFIX", "/opt/conda")


def run(cmd, cwd=HERE):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    lines = [ln for ln in r.stdout.splitlines() if "packed-fp32-ops" not in ln]
    if lines:
        print("\n".join(lines[-60:]), flush=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}")


def build_gzm():
    os.makedirs(LIB, exist_ok=True)
    objs = []
    procs = []
    for src in HIP_SOURCES:
        obj = os.path.join(LIB, os.path.basename(src).replace(".hip", ".o"))
        cmd = [HIPCC, *HIP_FLAGS, *FILE_FLAGS.get(src, []), "-c", src, "-o", obj]
        print("+", " ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, cwd=HERE, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    for cmd, p in procs:
        out, _ = p.communicate()
        lines = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln]
        if lines:
            print("\n".join(lines[-60:]), flush=True)
        if p.returncode != 0:
            raise RuntimeError(f"command failed ({p.returncode}): {' '.join(cmd)}")
    tmp = os.path.join(LIB, "libgzm.so.tmp")
    run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, os.path.join(LIB, "libgzm.so"))
    for o in objs:
        os.remove(o)


def build_host():
    srcs = [os.path.join(HERE, s) for s in HOST_SOURCES]
    tmp = os.path.join(LIB, "libgizmo_host.so.tmp")
    run(["gcc", "-O2", "-fPIC", "-shared", "-std=gnu11", "-Wall", f"-I{INC}", *srcs, "-o", tmp,
         f"-L{LIB}", "-lgzm", "-lm", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, os.path.join(LIB, "libgizmo_host.so"))


def build_io():
    inc = os.path.join(HDF5_PREFIX, "include")
    lib = os.path.join(HDF5_PREFIX, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        print(f"skipping libgizmo_io.so: no hdf5.h under {inc}", flush=True)
        return
    tmp = os.path.join(LIB, "libgizmo_io.so.tmp")
    run(["gcc", "-O2", "-fPIC", "-shared", "-std=gnu11", "-Wall", f"-I{INC}", f"-I{inc}",
         os.path.join(HERE, IO_SOURCE), "-o", tmp, f"-L{LIB}", "-lgizmo_host", f"-L{lib}", "-lhdf5", "-lm",
         "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{lib}"])
    os.replace(tmp, os.path.join(LIB, "libgizmo_io.so"))


def build_all():
    build_gzm()
    build_host()
    build_io()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
