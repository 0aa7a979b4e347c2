"""Build recipe of the gfx950 shared library ``libdiscoeb_hip.so`` (in-tree).

hipcc cross-compiles for gfx950 without a GPU.  Packed FP32 ops are disabled
for all device code.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "discoeb_amd", "libdiscoeb_hip.so")
SOURCES = ["deb_kernels.hip", "deb_adjoint.hip", "deb_observables.hip"]
# Both units at -O3.  The reverse step's row passes are outlined device calls
# (DEB_OUTLINE in deb_adjoint.h): builds with everything inlined into one
# kernel body (thousands of spilled registers) gave build-dependent wrong
# reverse steps at -O3 and at -O1.
OPT = {"deb_kernels.hip": "-O3", "deb_adjoint.hip": "-O3", "deb_observables.hip": "-O3"}
HEADERS = ["deb_device.h", "deb_mode.h", "deb_solver.h", "deb_forward.h", "deb_host.h", "deb_adjoint.h"]
FLAGS = ["--offload-arch=gfx950", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    """Compile every translation unit in parallel, then link the shared library."""
    inc = ["-I", os.path.join(HERE, "..", "include")]
    if os.environ.get("DEB_NS_MASK"):  # development builds: fewer slot counts
        inc += ["-DDEB_NS_MASK=" + os.environ["DEB_NS_MASK"]]
    objs, procs = [], []
    for src in SOURCES:
        obj = os.path.join(CSRC, "_" + src.replace(".hip", ".o"))
        cmd = [hipcc(), OPT[src]] + FLAGS + inc + ["-c", "-o", obj, os.path.join(CSRC, src)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    failed = False
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed = True
    if failed:
        raise RuntimeError("hipcc failed")
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", OUT + ".tmp"] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc link failed")
    for o in objs:
        os.remove(o)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
