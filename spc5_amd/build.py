"""Build recipe of the native libraries (in-tree, gfx950 only).

* ``spc5_amd/libspc5_hip.so``   -- HIP kernels + the C ABI (include/spc5_hip.h)
* ``spc5_amd/libspc5_synth.so`` -- seeded workload generators (host C++/OpenMP)

Run ``python -m spc5_amd.build`` or ``__graft_entry__.build()``.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
HIP_SO = os.path.join(PKG, "libspc5_hip.so")
SYNTH_SO = os.path.join(PKG, "libspc5_synth.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    # packed FP32 (v_pk_*_f32) is kept off for every kernel of this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"build failed ({p.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build_hip(verbose=False, out=HIP_SO, defines=()):
    """Compile every csrc/*.hip and link `out`.  `defines` are extra -D
    options (tuning constants of spc5_mat.hpp / spc5_common.hpp)."""
    bdir = BUILD if not defines else os.path.join(BUILD, "v_" + "_".join(d.replace("=", "") for d in defines))
    os.makedirs(bdir, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    objs = []
    cmds = []
    for s in srcs:
        o = os.path.join(bdir, os.path.basename(s) + ".o")
        objs.append(o)
        cmds.append([HIPCC, *HIP_FLAGS, *["-D" + d for d in defines], "-c", s, "-o", o])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(cmds))) as ex:
        for msg in ex.map(_run, cmds):
            if verbose and msg.strip():
                print(msg)
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,-Bsymbolic", "-o", tmp, *objs])
    os.replace(tmp, out)
    return out


def build_synth(verbose=False):
    src = os.path.join(CSRC, "synth.cpp")
    if not os.path.exists(src):
        return None
    tmp = SYNTH_SO + ".tmp"
    out = _run(["g++", "-O3", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-Wall",
                "-I" + os.path.join(ROOT, "include"), "-o", tmp, src])
    if verbose and out.strip():
        print(out)
    os.replace(tmp, SYNTH_SO)
    return SYNTH_SO


def build_cpp_driver(verbose=False):
    """The reference-style C++ driver of the drop-in header (tests/drivers)."""
    src = os.path.join(ROOT, "tests", "drivers", "spc5_compare.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(ROOT, "tests", "drivers", "spc5_compare")
    msg = _run(["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include"), src,
                "-L" + PKG, "-lspc5_hip", "-Wl,-rpath,$ORIGIN/../../spc5_amd", "-o", out + ".tmp"])
    if verbose and msg.strip():
        print(msg)
    os.replace(out + ".tmp", out)
    return out


def build_all(verbose=False):
    build_hip(verbose)
    build_synth(verbose)
    build_cpp_driver(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print("built", HIP_SO)
