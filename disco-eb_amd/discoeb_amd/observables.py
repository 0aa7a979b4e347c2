"""Host mirror of DISCO-EB's P(k) observables on the gfx950 kernels.

Same keyword-only signatures and return values as the reference:

* ``get_power_smoothed``  src/discoeb/perturbations.py:1126-1159
* ``power_Kaiser``        src/discoeb/perturbations.py:1162-1198
* ``power_multipoles``    src/discoeb/perturbations.py:1202-1224
* ``get_xi_from_P``       src/discoeb/perturbations.py:1065-1098

``y`` is the solution at one output time, ``(nk, 20)`` as in the reference
or ``(n_cosmo, nk, 20)`` for a batch of cosmologies (``k`` then ``(nk,)`` or
``(n_cosmo, nk)``; ``param`` entries scalars or per-cosmology arrays).  The
arithmetic runs in deb_observables.hip; results are torch FP64 tensors on the
GPU with the batch dimension only if the input had one.  There is no CPU
fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise _lib.DebError("the observables need a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


def _stream():
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev(x, dev):
    torch = _torch()
    if isinstance(x, torch.Tensor):
        return x.to(dev, torch.float64).contiguous()
    return torch.from_numpy(np.ascontiguousarray(np.real(np.asarray(x)), dtype=np.float64)).to(dev)


def _batch(y, k):
    """(y [B, nk, F], k [B, nk], batched?) on the GPU."""
    torch = _torch()
    dev = torch.device("cuda")
    yt = _dev(y, dev)
    batched = yt.dim() == 3
    if not batched:
        yt = yt[None]
    B, nk = yt.shape[0], yt.shape[1]
    kt = _dev(k, dev)
    if kt.dim() == 1:
        kt = kt[None].expand(B, nk).contiguous()
    if tuple(kt.shape) != (B, nk):
        raise _lib.DebError(f"k has shape {tuple(kt.shape)}, expected ({nk},) or ({B}, {nk})")
    return yt.contiguous(), kt, batched


def _per_cosmo(v, B, dev):
    a = np.real(np.asarray(v)).astype(np.float64).ravel()
    if a.size == 1:
        a = np.full(B, a[0])
    if a.size != B:
        raise _lib.DebError(f"parameter has {a.size} entries for {B} cosmologies")
    return _dev(a, dev)


def _prim(param, B, dev):
    return (_per_cosmo(param["A_s"], B, dev), _per_cosmo(param["n_s"], B, dev),
            _per_cosmo(param.get("k_p", 0.05), B, dev))


def _window_length(k: np.ndarray, dlogk: float) -> int:
    """perturbations.py:1150-1151: round half to even, then make odd."""
    wl = int(np.round(dlogk / (np.log(k[1]) - np.log(k[0]))))
    return wl + (wl + 1) % 2


def _windows(kt, dlogk):
    kh = kt.cpu().numpy()
    wls = [_window_length(kh[b], dlogk) for b in range(kh.shape[0])]
    if len(set(wls)) != 1:
        raise _lib.DebError("the cosmologies' k grids give different smoothing windows; call them separately")
    return wls[0]


def _unbatch(t, batched):
    return t if batched else t[0]


def get_power_smoothed(*, k, y, dlogk: float, idx: int, param: dict):
    """Savitzky-Golay (polyorder 3) smoothed P(k) (perturbations.py:1126-1159)."""
    torch = _torch()
    L = _lib.lib()
    yt, kt, batched = _batch(y, k)
    B, nk, F = yt.shape
    A_s, n_s, k_p = _prim(param, B, yt.device)
    wl = _windows(kt, dlogk)
    nbytes = L.deb_observables_scratch_bytes(B, wl, 3)
    scratch = torch.empty(max(nbytes // 8, 1), dtype=torch.float64, device=yt.device)
    out = torch.empty((B, nk), dtype=torch.float64, device=yt.device)
    _lib.check(L.deb_power_smoothed(B, nk, _lib.ptr(kt), _lib.ptr(yt), F, idx, _lib.ptr(A_s), _lib.ptr(n_s),
                                    _lib.ptr(k_p), wl, 3, _lib.ptr(scratch), nbytes, _lib.ptr(out), _stream()),
               "deb_power_smoothed")
    return _unbatch(out, batched)


def power_Kaiser(*, y, kmodes, bias, mu_sampling: bool = True, smooth_dlogk: float = None, nmu: int, param: dict):
    """(P(k, mu), mu) with P = (b delta_m - mu^2 theta_m)^2 (perturbations.py:1162-1198)."""
    torch = _torch()
    L = _lib.lib()
    yt, kt, batched = _batch(y, kmodes)
    B, nk, F = yt.shape
    A_s, n_s, k_p = _prim(param, B, yt.device)
    b = _per_cosmo(bias, B, yt.device)
    if mu_sampling:
        mu = np.linspace(-1, 1, nmu)
    else:
        mu = np.cos(np.linspace(0, np.pi, nmu))
    mut = _dev(mu, yt.device)
    wl = 0 if smooth_dlogk is None else _windows(kt, smooth_dlogk)
    nbytes = L.deb_observables_scratch_bytes(B, wl, 3) if wl else 0
    scratch = torch.empty(max(nbytes // 8, 1), dtype=torch.float64, device=yt.device)
    out = torch.empty((B, nk, nmu), dtype=torch.float64, device=yt.device)
    _lib.check(L.deb_power_kaiser(B, nk, nmu, _lib.ptr(kt), _lib.ptr(yt), F, _lib.ptr(A_s), _lib.ptr(n_s),
                                  _lib.ptr(k_p), _lib.ptr(b), _lib.ptr(mut), wl, 3, _lib.ptr(scratch), nbytes,
                                  _lib.ptr(out), _stream()), "deb_power_kaiser")
    return _unbatch(out, batched), mut


def power_multipoles(*, y, kmodes, b, param):
    """(P0, P2, P4) of the Kaiser spectrum (perturbations.py:1202-1224)."""
    torch = _torch()
    L = _lib.lib()
    yt, kt, batched = _batch(y, kmodes)
    B, nk, F = yt.shape
    A_s, n_s, k_p = _prim(param, B, yt.device)
    bt = _per_cosmo(b, B, yt.device)
    P = [torch.empty((B, nk), dtype=torch.float64, device=yt.device) for _ in range(3)]
    _lib.check(L.deb_power_multipoles(B, nk, _lib.ptr(kt), _lib.ptr(yt), F, _lib.ptr(A_s), _lib.ptr(n_s),
                                      _lib.ptr(k_p), _lib.ptr(bt), _lib.ptr(P[0]), _lib.ptr(P[1]), _lib.ptr(P[2]),
                                      _stream()), "deb_power_multipoles")
    return tuple(_unbatch(p, batched) for p in P)


def get_xi_from_P(*, k, Pk, N: int, ell: int = 0):
    """(xi(r), r) by FFTlog (perturbations.py:1065-1098); N is taken from k as
    in the reference, which recomputes it."""
    torch = _torch()
    L = _lib.lib()
    dev = torch.device("cuda")
    Pt = _dev(Pk, dev).contiguous()
    batched = Pt.dim() == 2
    if not batched:
        Pt = Pt[None].contiguous()
    B, n = Pt.shape
    kt = _dev(k, dev)
    if kt.dim() == 1:
        kt = kt[None].expand(B, n).contiguous()
    xi = torch.empty((B, n), dtype=torch.float64, device=dev)
    r = torch.empty((B, n), dtype=torch.float64, device=dev)
    _lib.check(L.deb_xi_from_power(B, n, _lib.ptr(kt), _lib.ptr(Pt), ell, _lib.ptr(xi), _lib.ptr(r),
                                   _stream()), "deb_xi_from_power")
    return _unbatch(xi, batched), _unbatch(r, batched)
