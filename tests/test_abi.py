"""The C-ABI library loads and exports every function include/*.h declares
(no compute calls: runs on the CPU); host-side argument checks fail cleanly."""
import ctypes
import os
import re

import pytest

from discoeb_amd import _lib, packing

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            src = open(os.path.join(ROOT, "include", fn)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names += re.findall(r"\b(deb_[a-z0-9_]+)\s*\(", src)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libdiscoeb_hip.so not built")
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert "deb_solve_forward" in names and len(names) >= 6
    for n in names:
        assert hasattr(L, n), n


def test_host_argument_checks_without_gpu():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libdiscoeb_hip.so not built")
    L = _lib.lib()
    assert L.deb_abi_version() == 3
    cfg = packing.make_config(lmaxg=32, lmaxgp=32, lmaxr=32, lmaxnu=17, nqmax=5)
    assert L.deb_nvar(ctypes.byref(cfg)) == packing.nvar(32, 32, 32, 17, 5) == 198
    b = ctypes.c_size_t(0)
    assert L.deb_workspace_bytes(ctypes.byref(cfg), 1, 1, 1, ctypes.byref(b)) == 0 and b.value >= 4
    # null arrays are rejected before any HIP call
    rc = L.deb_solve_forward(ctypes.byref(cfg), 1, None, None, 0, None, None, None, 1, None, 1, None, None, None,
                             None, None, 0, None)
    assert rc == -1
    assert b"null" in L.deb_last_error()
    bad = packing.make_config(lmaxg=2)
    assert L.deb_solve_forward(ctypes.byref(bad), 1, 8, 8, 0, 8, 8, 8, 1, 8, 1, 8, 8, 8, 8, 8, 256, None) < 0


def test_noalloc_entry_rejects_sizes_beyond_its_array_without_gpu():
    """deb_solve_forward_batched_noalloc checks the workspace size on the host
    before any HIP call: 1 x 32768 k at lmax 32 needs 128 MiB > 64 MiB."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libdiscoeb_hip.so not built")
    import numpy as np
    L = _lib.lib()
    cfg = packing.make_config(lmaxg=32, lmaxgp=32, lmaxr=32, lmaxnu=17, nqmax=5)
    host = np.zeros(64)
    p = host.ctypes.data_as(ctypes.c_void_p)
    desc = np.zeros(14, dtype=np.int32)
    rc = L.deb_solve_forward_batched_noalloc(ctypes.byref(cfg), 1, p, p, 0, desc.ctypes.data_as(ctypes.c_void_p),
                                             p, p, 32768, p, 1, p, 8, p, p, None, None)
    assert rc == -1
    assert b"allocation-free" in L.deb_last_error()
    ok = packing.make_config(lmaxg=32, lmaxgp=32, lmaxr=32, lmaxnu=17, nqmax=5)
    b = ctypes.c_size_t(0)
    assert L.deb_batched_workspace_bytes(ctypes.byref(ok), 64, 256, ctypes.byref(b)) == 0
    assert b.value == 64 << 20  # the headline's 64 cosmologies x 256 k fit exactly
