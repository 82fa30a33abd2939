"""Build libqzhip.so in-tree with hipcc for gfx950 (no torch extension machinery).

Inside a QUANTIZER tree (its headers directory given by QZ_REFERENCE_HEADERS,
or found at the default checkout path) also build the drop-in libraries that
export the reference's own entry points: libqzdropin.so (force cycle) and
libqzdropin_resident.so (force cycle + particle updates, particles resident).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libqzhip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -packed-fp32-ops off: the compiler must not form packed FP32 VALU ops.
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wno-unused-result"]

SOURCES = ["qz_sort.hip", "qz_core.hip", "qz_tree.cpp"]
DROPIN_SRC = os.path.join(HERE, "dropin", "qz_dropin.c")
DROPIN_OUT = {"libqzdropin.so": [], "libqzdropin_resident.so": ["-DQZ_DROPIN_RESIDENT"]}


def reference_headers():
    """The QUANTIZER headers the drop-in compiles against, or None."""
    for d in (os.environ.get("QZ_REFERENCE_HEADERS"), "/root/reference/headers"):
        if d and os.path.exists(os.path.join(d, "common.h")):
            return d
    return None


def build_dropin(headers: str, verbose: bool = False) -> list:
    outs = []
    for name, defs in DROPIN_OUT.items():
        out = os.path.join(HERE, name)
        cmd = [os.environ.get("CC", "gcc"), "-O2", "-fPIC", "-shared", "-fopenmp", "-w", *defs, f"-I{headers}",
               "-o", out, DROPIN_SRC, f"-L{HERE}", "-lqzhip", "-Wl,-rpath,$ORIGIN", "-lm"]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        outs.append(out)
    return outs


def _obj(src: str) -> str:
    return os.path.join(CSRC, "_build", os.path.splitext(src)[0] + ".o")


def _run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        sys.stderr.write(p.stdout + p.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd))


def build(force: bool = False, verbose: bool = False) -> str:
    os.makedirs(os.path.join(CSRC, "_build"), exist_ok=True)
    deps = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".h", ".inl", ".hip", ".cpp"))]
    deps += [os.path.join(HERE, "..", "include", f) for f in os.listdir(os.path.join(HERE, "..", "include"))]
    newest_hdr = max(os.path.getmtime(d) for d in deps)
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = _obj(src)
        objs.append(obj)
        stale = force or not os.path.exists(obj) or os.path.getmtime(obj) < os.path.getmtime(path)
        # the sort TU does not include the engine headers; everything else does
        if not stale and src != "qz_sort.hip" and os.path.getmtime(obj) < newest_hdr:
            stale = True
        if stale:
            if src.endswith(".hip"):
                cmd = [HIPCC, *HIP_FLAGS, "-c", path, "-o", obj]
            else:
                cmd = [HIPCC, "-O3", "-fPIC", "-std=c++17", "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-c", path,
                       "-o", obj]
            if verbose:
                print(" ".join(cmd))
            _run(cmd)
    tmp = OUT + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs])
    os.replace(tmp, OUT)
    hdrs = reference_headers()
    if hdrs:
        build_dropin(hdrs, verbose)
    return OUT


if __name__ == "__main__":
    print(build(force="--force" in sys.argv, verbose=True))
