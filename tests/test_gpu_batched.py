"""Batched variant (evolve_perturbations_batched, perturbations.py:1000-1061):
groups of batch_size consecutive k share one step sequence.  The GPU path
(deb_solve_forward_batched) against the oracle's restatement of
evolve_modes_batched, and against the per-mode solve for groups of one."""
import numpy as np
import pytest

import oracle
from discoeb_amd import packing
from discoeb_amd.perturbations import DeviceInputs, evolve_perturbations_batched, solve_forward

pytestmark = pytest.mark.gpu

SMALL = dict(lmaxg=11, lmaxgp=11, lmaxr=11, lmaxnu=8, nqmax=3)


def _gpu(param, cfg, k, tau, batch):
    import torch
    inp = DeviceInputs(param, cfg)
    y, st, ns = solve_forward(inp, torch.from_numpy(k[None, :]), torch.from_numpy(tau[None, :]), throw=False,
                              batch_size=batch)
    return y[0].cpu().numpy(), st[0].cpu().numpy(), ns[0].cpu().numpy()


def test_batched_tight_tolerance_matches_oracle(gpu, test_param):
    cfg = packing.make_config(rtol=1e-10, atol=1e-14, max_steps=400000, **SMALL)
    k = np.geomspace(1e-3, 0.3, 8)
    tau = packing.tau_of_aexp(test_param, [1.0])[0]
    ref = oracle.Problem(test_param, cfg).evolve_batched(k, tau, 4, nthreads=2)
    y, st, ns = _gpu(test_param, cfg, k, tau, 4)
    assert np.all(st == 0) and np.all(ref["status"] == 0)
    for g in range(2):
        assert len(set(ns[4 * g:4 * g + 4].tolist())) == 1
    for f in (4, 6, 8, 10):
        rel = np.abs(y[:, -1, f] - ref["y"][:, -1, f]) / np.abs(ref["y"][:, -1, f])
        assert rel.max() < 1e-9, (f, rel.max())


def test_batched_reference_defaults(gpu, test_param):
    """rtol = atol = 1e-4, batch_size 16 (the reference's defaults): P(k)-level agreement."""
    cfg = packing.make_config(rtol=1e-4, atol=1e-4, max_steps=2048, **SMALL)
    k = np.geomspace(1e-4, 1.0, 32)
    tau = packing.tau_of_aexp(test_param, [0.5, 1.0])[0]
    ref = oracle.Problem(test_param, cfg).evolve_batched(k, tau, 16, nthreads=2)
    y, st, ns = _gpu(test_param, cfg, k, tau, 16)
    assert np.all(st == 0) and np.all(ref["status"] == 0)
    assert np.all(np.abs(ns - ref["nsteps"]) <= 0.05 * ref["nsteps"] + 2), (ns, ref["nsteps"])
    rel = np.abs(y[:, -1, 4] ** 2 - ref["y"][:, -1, 4] ** 2) / ref["y"][:, -1, 4] ** 2
    assert rel.max() < 1e-3, rel.max()


def test_batch_of_one_is_the_per_mode_solve(gpu, test_param):
    """Groups of one follow the per-mode algorithm; the two kernels round
    differently, which can flip an accept/reject decision (the error estimate
    is a cancelling sum), so they agree at the tolerance level: final outputs
    (step ends) within 30 rtol, step counts within 2 % (1.3 % measured at
    ~3800 steps)."""
    rtol = 1e-5
    cfg = packing.make_config(rtol=rtol, atol=1e-8, max_steps=20000, **SMALL)
    k = np.geomspace(1e-3, 0.5, 6)
    tau = packing.tau_of_aexp(test_param, [0.1, 1.0])[0]
    yb, sb, nb = _gpu(test_param, cfg, k, tau, 1)
    ys, ss, nss = _gpu(test_param, cfg, k, tau, 0)
    assert np.all(sb == 0) and np.all(ss == 0)
    for f in (4, 6, 8, 10):
        rel = np.abs(yb[:, -1, f] - ys[:, -1, f]) / np.abs(ys[:, -1, f])
        assert rel.max() < 30 * rtol, (f, rel.max())
    assert np.all(np.abs(nb - nss) <= 0.02 * nss + 2), (nb, nss)


def test_batched_api(gpu, test_param):
    p = dict(test_param)
    y, k = evolve_perturbations_batched(param=p, aexp_out=[1.0], kmin=1e-3, kmax=0.5, num_k=16, batch_size=8)
    assert tuple(y.shape) == (16, 1, 20) and np.all(np.isfinite(y.cpu().numpy()))
    with pytest.raises(ValueError):
        evolve_perturbations_batched(param=p, aexp_out=[1.0], kmin=1e-3, kmax=0.5, num_k=10, batch_size=4)
