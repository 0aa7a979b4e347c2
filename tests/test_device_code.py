"""Checks on the gfx950 code object of the built library (no GPU needed).

Outlined device functions (real calls, not kernels) must contain no
long-branch sequence (s_getpc_b64): in functions large enough to need them
the compiler was seen to build them from the return-address registers
s[30:31] without saving those, so the function's return jumped to a stale
address.  The reverse-mode row passes are kept small for this reason; this
test fails if one of them (or a new outlined function) grows into it.
"""
import os
import re
import shutil
import subprocess

import pytest

from discoeb_amd import _lib

LLVM = "/opt/rocm/llvm/bin"


def _code_objects(tmp_path):
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libdiscoeb_hip.so is not built")
    objdump = os.path.join(LLVM, "llvm-objdump")
    if not os.path.exists(objdump):
        pytest.skip("llvm-objdump not available")
    lib = tmp_path / "lib.so"
    shutil.copy(_lib.LIB_PATH, lib)
    subprocess.run([objdump, "--offloading", str(lib)], cwd=tmp_path, check=True, capture_output=True)
    cos = sorted(p for p in tmp_path.iterdir() if p.name.endswith("gfx950"))
    assert cos, "no gfx950 code object in the library"
    return cos


def _long_branches(body):
    """s_getpc_b64 s[a:b] ... s_setpc_b64 s[a:b] within a few instructions (an indirect jump; a getpc that
    only forms a data address is not one)."""
    ins = [ln.split("//")[0].strip() for ln in body.splitlines() if ln.strip()]
    n = 0
    for i, ln in enumerate(ins):
        m = re.match(r"s_getpc_b64 (s\[\d+:\d+\])", ln)
        if m and any(x.startswith("s_setpc_b64 " + m.group(1)) for x in ins[i + 1:i + 6]):
            n += 1
    return n


def test_long_branch_detector():
    jump = "s_getpc_b64 s[30:31]\ns_add_u32 s30, s30, 0x10\ns_addc_u32 s31, s31, 0\ns_setpc_b64 s[30:31]\n"
    addr = "s_getpc_b64 s[2:3]\ns_add_u32 s2, s2, 0x10\ns_addc_u32 s3, s3, 0\ns_load_dwordx2 s[4:5], s[2:3], 0x0\n"
    assert _long_branches(jump) == 1 and _long_branches(addr) == 0


def test_outlined_functions_have_no_long_branches(tmp_path):
    readelf, objdump = os.path.join(LLVM, "llvm-readelf"), os.path.join(LLVM, "llvm-objdump")
    checked = 0
    for co in _code_objects(tmp_path):
        syms = subprocess.run([readelf, "-sW", str(co)], check=True, capture_output=True, text=True).stdout
        funcs = set(re.findall(r"\sFUNC\s+\S+\s+\S+\s+\S+\s+(\S+)$", syms, re.M))
        kernels = {n[:-3] for n in re.findall(r"\sOBJECT\s+\S+\s+\S+\s+\S+\s+(\S+\.kd)$", syms, re.M)}
        outlined = funcs - kernels
        dis = subprocess.run([objdump, "-d", "--no-show-raw-insn", str(co)], check=True, capture_output=True,
                             text=True).stdout
        # split the listing into per-symbol bodies
        parts = re.split(r"^[0-9a-f]+ <([^>]+)>:$", dis, flags=re.M)
        bodies = dict(zip(parts[1::2], parts[2::2]))
        for f in outlined:
            assert f in bodies, f
            assert not _long_branches(bodies[f]), f"outlined function {f} contains long branches"
            checked += 1
    assert checked > 0
