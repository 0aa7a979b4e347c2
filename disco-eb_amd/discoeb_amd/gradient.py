"""Gradient path: taped forward solve plus one batched reverse launch.

The reference obtains dP(k)/dtheta with jax.jacfwd through
``evolve_perturbations`` (nb_minimal_example.ipynb:297,
nb_discoeb_autodiff.ipynb:274).  Here the derivative of the same discrete
solution is computed in reverse mode on the GPU for every (cosmology, k)
mode at once (``deb_solve_forward_tape`` + ``deb_solve_adjoint``), projected
on the input tangents of the theta directions that enter the solve
(:mod:`discoeb_amd.tangents`).  The step-size controller's decisions carry no
derivative (accepted steps keep their fractions of [t0, t1]).

``power_spectrum`` is a ``torch.autograd.Function``: P(k) of a batch of
theta vectors, with ``backward`` running the reverse launch.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib, packing, tangents
from .cosmology import param_from_theta
from .perturbations import DeviceInputs, _torch

FIELD_DELTA_M = 4


CKPT_EVERY = 64            # accepted steps per checkpoint segment
TAPE_BUDGET = 64 * 2**30   # default device bytes for the tape and the reverse launch's rings
STATES_PER_MODE = 4096     # first estimate of the accepted states of a mode


class Tape:
    """Device checkpoint tape of one taped forward solve."""

    def __init__(self, buf, nbytes, ckpt_every):
        self.buf = buf
        self.nbytes = int(nbytes)
        self.ckpt_every = int(ckpt_every)


def tape_bytes(cfg, n_cosmo: int, nk: int, n_states: int, ckpt_every: int = CKPT_EVERY) -> int:
    b = ctypes.c_size_t(0)
    _lib.check(_lib.lib().deb_tape_bytes(ctypes.byref(cfg), n_cosmo, nk, int(n_states), int(ckpt_every),
                                         ctypes.byref(b)), "deb_tape_bytes")
    return int(b.value)


def adjoint_workspace_bytes(cfg, n_cosmo: int, nk: int, ckpt_every: int = CKPT_EVERY) -> int:
    b = ctypes.c_size_t(0)
    _lib.check(_lib.lib().deb_adjoint_workspace_bytes(ctypes.byref(cfg), int(n_cosmo), int(nk), int(ckpt_every),
                                                      ctypes.byref(b)), "deb_adjoint_workspace_bytes")
    return int(b.value)


def solve_forward_tape(inp: DeviceInputs, kmodes, tau_out, *, n_states: int | None = None,
                       ckpt_every: int = CKPT_EVERY, budget_bytes: int = TAPE_BUDGET, stream=None):
    """Taped forward solve: every mode's accepted states in segments of
    ``ckpt_every`` steps (their first state and every step time are kept).
    ``n_states``: tape size in states (default STATES_PER_MODE per mode; on
    overflow the solve is repeated with the exact count).  The tape plus the
    reverse launch's workspace must fit ``budget_bytes``: DebError otherwise."""
    torch = _torch()
    L = _lib.lib()
    cfg = inp.cfg
    kmodes = kmodes.to(inp.device, torch.float64).contiguous()
    tau_out = tau_out.to(inp.device, torch.float64).contiguous()
    n_cosmo, nk = kmodes.shape
    nout = tau_out.shape[1]
    st = stream if stream is not None else torch.cuda.current_stream(inp.device)
    if n_states is None:
        n_states = n_cosmo * nk * STATES_PER_MODE
    ws = adjoint_workspace_bytes(cfg, n_cosmo, nk, ckpt_every)
    for attempt in range(2):
        nb = tape_bytes(cfg, n_cosmo, nk, n_states, ckpt_every)
        if nb + ws > budget_bytes:
            raise _lib.DebError(f"checkpoint tape ({nb / 2**30:.1f} GiB for {n_states} states) and reverse-launch "
                                f"workspace ({ws / 2**30:.1f} GiB) exceed the budget of {budget_bytes / 2**30:.1f} "
                                f"GiB; raise ckpt_every ({ckpt_every}) or the budget")
        buf = torch.empty(nb, dtype=torch.uint8, device=inp.device)
        y = torch.empty((n_cosmo, nk, nout, packing.N_OUT_FIELDS), dtype=torch.float64, device=inp.device)
        status = torch.full((n_cosmo, nk), -1, dtype=torch.int32, device=inp.device)
        nsteps = torch.zeros((n_cosmo, nk), dtype=torch.int32, device=inp.device)
        _lib.check(L.deb_solve_forward_tape(ctypes.byref(cfg), n_cosmo, _lib.ptr(inp.cosmo), _lib.ptr(inp.tables),
                                            inp.stride, _lib.ptr(inp.desc), _lib.ptr(inp.q), _lib.ptr(inp.w), nk,
                                            _lib.ptr(kmodes), nout, _lib.ptr(tau_out), _lib.ptr(y),
                                            _lib.ptr(status), _lib.ptr(nsteps), int(ckpt_every), _lib.ptr(buf), nb,
                                            ctypes.c_void_p(st.cuda_stream)), "deb_solve_forward_tape")
        full = bool((status == packing.STATUS_TAPE_FULL).any())
        if not full:
            return y, status, nsteps, Tape(buf, nb, ckpt_every)
        # nsteps counts attempted steps: an upper bound of the accepted states
        n_states = int(nsteps.sum().item()) + n_cosmo * nk
        del buf
    raise _lib.DebError("checkpoint tape still too small")


class Tangents:
    """Input tangents of n_cosmo cosmologies on the device."""

    def __init__(self, dcos, dtab, dtau, device):
        torch = _torch()
        self.ndir = dcos.shape[1]
        self.dcos = torch.from_numpy(np.ascontiguousarray(dcos)).to(device)
        self.dtab = torch.from_numpy(np.ascontiguousarray(dtab)).to(device)
        self.dtau = torch.from_numpy(np.ascontiguousarray(dtau)).to(device)


def solve_adjoint(inp: DeviceInputs, tan: Tangents, kmodes, tau_out, ybar, status, tape: Tape, *, stream=None):
    """grad[c, k, j] = sum over outputs of ybar . d(y_out)/d(direction j)."""
    torch = _torch()
    L = _lib.lib()
    cfg = inp.cfg
    kmodes = kmodes.to(inp.device, torch.float64).contiguous()
    tau_out = tau_out.to(inp.device, torch.float64).contiguous()
    ybar = ybar.to(inp.device, torch.float64).contiguous()
    n_cosmo, nk = kmodes.shape
    nout = tau_out.shape[1]
    grad = torch.empty((n_cosmo, nk, tan.ndir), dtype=torch.float64, device=inp.device)
    wsb = adjoint_workspace_bytes(cfg, n_cosmo, nk, tape.ckpt_every)
    ws = torch.empty(wsb, dtype=torch.uint8, device=inp.device)
    st = stream if stream is not None else torch.cuda.current_stream(inp.device)
    _lib.check(L.deb_solve_adjoint(ctypes.byref(cfg), n_cosmo, _lib.ptr(inp.cosmo), _lib.ptr(inp.tables), inp.stride,
                                   _lib.ptr(inp.desc), _lib.ptr(inp.q), _lib.ptr(inp.w), nk, _lib.ptr(kmodes), nout,
                                   _lib.ptr(tau_out), tan.ndir, _lib.ptr(tan.dcos), _lib.ptr(tan.dtab),
                                   _lib.ptr(tan.dtau), _lib.ptr(ybar), _lib.ptr(status), _lib.ptr(grad),
                                   tape.ckpt_every, _lib.ptr(tape.buf), tape.nbytes, _lib.ptr(ws), wsb,
                                   ctypes.c_void_p(st.cuda_stream)), "deb_solve_adjoint")
    return grad


def _prefactor(theta, kk):
    torch = _torch()
    A_s = torch.exp(theta[:, 4:5]) / 1e10
    n_s = theta[:, 3:4]
    return 2 * np.pi**2 * A_s * (kk / 0.05) ** (n_s - 1) * kk ** (-3)


class PowerSpectrumFn:
    """P(k) at one scale factor and its gradient for a batch of theta."""

    @staticmethod
    def prepare(theta_np, aexp, cfg, device, massive_nu=True):
        param, dcos, dtab, dtau = tangents.input_tangents(theta_np, [aexp], massive_nu=massive_nu)
        inp = DeviceInputs(param, cfg, device)
        tau = packing.tau_of_aexp(param, [aexp]).real
        return inp, Tangents(dcos, dtab, dtau, device), tau


def power_and_gradient(theta, kmodes, *, aexp: float = 1.0, cfg=None, massive_nu: bool = True, device=None,
                       n_states: int | None = None, prepared=None):
    """P(k) [B, nk] and dP/dtheta [B, nk, 8] (theta order of
    :data:`discoeb_amd.cosmology.THETA_NAMES`) for theta [B, 8] and kmodes [nk].

    Six directions come from the reverse launch, n_s and ln10As analytically.
    """
    torch = _torch()
    theta = np.atleast_2d(np.asarray(theta, dtype=np.float64))
    device = torch.device(device if device is not None else "cuda")
    if cfg is None:
        cfg = packing.make_config(lmaxg=32, lmaxgp=32, lmaxr=32, lmaxnu=17, nqmax=5, rtol=1e-5, atol=1e-8,
                                  max_steps=32768)
    inp, tan, tau = prepared if prepared is not None else PowerSpectrumFn.prepare(theta, aexp, cfg, device,
                                                                                 massive_nu)
    B = theta.shape[0]
    k = torch.as_tensor(np.broadcast_to(np.asarray(kmodes, dtype=np.float64), (B, len(kmodes))).copy(),
                        device=device)
    tau_t = torch.from_numpy(tau).to(device)
    y, status, nsteps, tape = solve_forward_tape(inp, k, tau_t, n_states=n_states)
    bad = status != packing.STATUS_OK
    if bool(bad.any()):
        raise _lib.DebError(f"{int(bad.sum())} mode(s) failed in the taped solve")
    th = torch.from_numpy(theta).to(device)
    pref = _prefactor(th, k)
    dm = y[:, :, 0, FIELD_DELTA_M]
    P = pref * dm**2
    ybar = torch.zeros_like(y)
    ybar[:, :, 0, FIELD_DELTA_M] = 2.0 * pref * dm
    g = solve_adjoint(inp, tan, k, tau_t, ybar, status, tape)
    G = torch.empty((B, k.shape[1], 8), dtype=torch.float64, device=device)
    for j, d in enumerate(tangents.SOLVE_DIRS):
        G[:, :, d] = g[:, :, j]
    G[:, :, 3] = torch.log(k / 0.05) * P
    G[:, :, 4] = P
    return P, G, dict(status=status, nsteps=nsteps, tape_bytes=tape.nbytes)


def power_spectrum(theta, kmodes, *, aexp: float = 1.0, cfg=None, massive_nu: bool = True):
    """Differentiable P(k) [B, nk] of theta [B, 8] (a torch FP64 tensor on the
    GPU).  The forward pass runs the taped solve; ``backward`` seeds the
    reverse launch with the incoming cotangent of P."""
    torch = _torch()
    if cfg is None:
        cfg = packing.make_config(lmaxg=32, lmaxgp=32, lmaxr=32, lmaxnu=17, nqmax=5, rtol=1e-5, atol=1e-8,
                                  max_steps=32768)

    class _Fn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, th):
            theta_np = th.detach().cpu().numpy()
            inp, tan, tau = PowerSpectrumFn.prepare(theta_np, aexp, cfg, th.device, massive_nu)
            B = theta_np.shape[0]
            k = torch.as_tensor(np.broadcast_to(np.asarray(kmodes, dtype=np.float64), (B, len(kmodes))).copy(),
                                device=th.device)
            tau_t = torch.from_numpy(tau).to(th.device)
            y, status, _, tape = solve_forward_tape(inp, k, tau_t)
            if bool((status != packing.STATUS_OK).any()):
                raise _lib.DebError("a mode failed in the taped solve")
            pref = _prefactor(th.detach(), k)
            dm = y[:, :, 0, FIELD_DELTA_M]
            P = pref * dm**2
            ctx.state = (inp, tan, k, tau_t, y, status, tape, pref, dm, P)
            return P

        @staticmethod
        def backward(ctx, gout):
            inp, tan, k, tau_t, y, status, tape, pref, dm, P = ctx.state
            ybar = torch.zeros_like(y)
            ybar[:, :, 0, FIELD_DELTA_M] = gout * 2.0 * pref * dm
            g = solve_adjoint(inp, tan, k, tau_t, ybar, status, tape)
            gth = torch.zeros((k.shape[0], 8), dtype=torch.float64, device=k.device)
            gs = g.sum(dim=1)
            for j, d in enumerate(tangents.SOLVE_DIRS):
                gth[:, d] = gs[:, j]
            gth[:, 3] = (gout * torch.log(k / 0.05) * P).sum(dim=1)
            gth[:, 4] = (gout * P).sum(dim=1)
            ctx.state = None
            return gth

    return _Fn.apply(theta)
