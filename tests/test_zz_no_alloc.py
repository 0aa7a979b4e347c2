"""The allocation-free entry point deb_solve_forward_batched_noalloc (its
workspace is a device array inside the library) against the entry point it
varies, deb_solve_forward_batched, through the C ABI: bit-identical outputs,
status and step counts on two calls in a row, the second smaller than the
first.  (Its rejection of sizes beyond the 64 MiB array happens on the host
before any HIP call and is tested on the CPU, tests/test_abi.py.)"""
import ctypes

import numpy as np
import pytest

from discoeb_amd import _lib, packing
from discoeb_amd.background import evolve_background
from discoeb_amd.cosmology import param_from_theta, sample_thetas
from discoeb_amd.perturbations import DeviceInputs

pytestmark = pytest.mark.gpu


def _run(noalloc, inp, k, tau, batch):
    import torch
    L = _lib.lib()
    cfg = inp.cfg
    n_cosmo, nk = k.shape
    nout = tau.shape[1]
    y = torch.full((n_cosmo, nk, nout, 20), float("nan"), dtype=torch.float64, device=inp.device)
    st = torch.full((n_cosmo, nk), -7, dtype=torch.int32, device=inp.device)
    ns = torch.zeros((n_cosmo, nk), dtype=torch.int32, device=inp.device)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    common = (ctypes.byref(cfg), n_cosmo, _lib.ptr(inp.cosmo), _lib.ptr(inp.tables), inp.stride, _lib.ptr(inp.desc),
              _lib.ptr(inp.q), _lib.ptr(inp.w), nk, _lib.ptr(k), nout, _lib.ptr(tau), batch)
    if noalloc:
        rc = L.deb_solve_forward_batched_noalloc(*common, _lib.ptr(y), _lib.ptr(st), _lib.ptr(ns), stream)
        _lib.check(rc, "deb_solve_forward_batched_noalloc")
    else:
        wsb = ctypes.c_size_t(0)
        _lib.check(L.deb_batched_workspace_bytes(ctypes.byref(cfg), n_cosmo, nk, ctypes.byref(wsb)), "ws bytes")
        ws = torch.empty(int(wsb.value), dtype=torch.uint8, device=inp.device)
        _lib.check(L.deb_solve_forward_batched(*common, _lib.ptr(y), _lib.ptr(st), _lib.ptr(ns), _lib.ptr(ws),
                                               wsb.value, stream), "deb_solve_forward_batched")
    torch.cuda.synchronize()
    return y.cpu().numpy(), st.cpu().numpy(), ns.cpu().numpy()


def test_noalloc_batched_matches_workspace_batched(gpu):
    import torch
    th = sample_thetas(3)
    with np.errstate(all="ignore"):
        param = evolve_background(param=param_from_theta(th))
    cfg = packing.make_config(lmaxg=11, lmaxgp=11, lmaxr=11, lmaxnu=8, nqmax=3, rtol=1e-5, atol=1e-8, max_steps=20000)
    inp = DeviceInputs(param, cfg, gpu)
    tau = torch.from_numpy(packing.tau_of_aexp(param, [0.5, 1.0]).real).to(gpu)
    for (nc, nk, batch) in ((3, 24, 8), (2, 8, 4)):   # the second call is smaller than the first
        sub = DeviceInputs.__new__(DeviceInputs)
        sub.__dict__.update(inp.__dict__)
        sub.cosmo, sub.tables = inp.cosmo[:nc], inp.tables[:nc]
        k = torch.from_numpy(np.ascontiguousarray(np.broadcast_to(np.geomspace(1e-3, 0.5, nk), (nc, nk)))).to(gpu)
        ya, sa, na = _run(False, sub, k, tau[:nc].contiguous(), batch)
        yb, sb, nb = _run(True, sub, k, tau[:nc].contiguous(), batch)
        assert np.all(sa == 0) and np.all(np.isfinite(ya))
        np.testing.assert_array_equal(yb, ya)
        np.testing.assert_array_equal(sb, sa)
        np.testing.assert_array_equal(nb, na)
