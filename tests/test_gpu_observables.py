"""GPU parity of the P(k) observables (deb_observables.hip through the C ABI,
via the discoeb_amd.observables mirror) against the oracle's restatement
(oracle/observables.py), on batches of cosmologies."""
import numpy as np
import pytest

import observables as ob

pytestmark = pytest.mark.gpu


def _inputs(B=3, nk=256, seed=7):
    rng = np.random.default_rng(seed)
    k = np.geomspace(1e-4, 1.0, nk)
    y = np.zeros((B, nk, 20))
    for b in range(B):
        amp = rng.uniform(0.5, 2.0)
        # delta_m ~ transfer-function shape with baryon wiggles, theta_m/(aH) ~ -f delta_m
        T = 1.0 / (1.0 + (k / rng.uniform(0.01, 0.03)) ** 2) * (1 + 0.05 * np.sin(k / 0.01))
        y[b, :, 4] = amp * k**2 * T * 1e4
        y[b, :, 5] = -rng.uniform(0.4, 0.9) * y[b, :, 4]
        y[b, :, 6:] = rng.standard_normal((nk, 14))
    param = dict(A_s=rng.uniform(1.8e-9, 2.4e-9, B), n_s=rng.uniform(0.92, 1.0, B), k_p=np.full(B, 0.05))
    return k, y, param


def _p(param, b):
    return param["A_s"][b], param["n_s"][b], param["k_p"][b]


def test_smoothed_power_matches_oracle(gpu):
    from discoeb_amd import observables as gob
    k, y, param = _inputs()
    for dlogk in (0.3, 1.0):
        got = gob.get_power_smoothed(k=k, y=y, dlogk=dlogk, idx=4, param=param).cpu().numpy()
        for b in range(y.shape[0]):
            want = ob.power_smoothed(k, y[b], dlogk, 4, *_p(param, b))
            np.testing.assert_allclose(got[b], want, rtol=1e-12)
    # one cosmology, reference shapes (nk, 20) -> (nk,)
    got1 = gob.get_power_smoothed(k=k, y=y[1], dlogk=0.5, idx=5, param=dict(A_s=param["A_s"][1],
                                                                               n_s=param["n_s"][1]))
    assert tuple(got1.shape) == (k.size,)
    np.testing.assert_allclose(got1.cpu().numpy(), ob.power_smoothed(k, y[1], 0.5, 5, *_p(param, 1)), rtol=1e-12)


@pytest.mark.parametrize("smooth", [None, 0.5])
@pytest.mark.parametrize("mu_sampling", [True, False])
def test_kaiser_matches_oracle(gpu, smooth, mu_sampling):
    from discoeb_amd import observables as gob
    k, y, param = _inputs()
    bias = np.array([1.0, 1.5, 2.2])
    got, mu = gob.power_Kaiser(y=y, kmodes=k, bias=bias, mu_sampling=mu_sampling, smooth_dlogk=smooth, nmu=11,
                               param=param)
    got = got.cpu().numpy()
    for b in range(3):
        want, mu_ref = ob.power_kaiser(y[b], k, bias[b], 11, *_p(param, b), mu_sampling=mu_sampling,
                                       smooth_dlogk=smooth)
        np.testing.assert_allclose(mu.cpu().numpy(), mu_ref, rtol=0, atol=1e-15)
        np.testing.assert_allclose(got[b], want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


def test_multipoles_match_oracle(gpu):
    from discoeb_amd import observables as gob
    k, y, param = _inputs()
    P0, P2, P4 = gob.power_multipoles(y=y, kmodes=k, b=1.3, param=param)
    for b in range(3):
        want = ob.power_multipoles(y[b], k, 1.3, *_p(param, b))
        for got, w in zip((P0, P2, P4), want):
            np.testing.assert_allclose(got[b].cpu().numpy(), w, rtol=1e-13, atol=1e-13 * np.abs(w).max())


@pytest.mark.parametrize("ell", [0, 2, 4, 1])
def test_xi_matches_oracle(gpu, ell):
    """The kernel's direct transforms against numpy's FFT in the oracle."""
    from discoeb_amd import observables as gob
    B, N = 3, 512
    k = np.geomspace(1e-4, 10.0, N)
    rng = np.random.default_rng(3)
    Pk = np.stack([k ** rng.uniform(0.5, 1.5) * np.exp(-(k * rng.uniform(3, 10)) ** 2) for _ in range(B)])
    xi, r = gob.get_xi_from_P(k=k, Pk=Pk, N=N, ell=ell)
    xi, r = xi.cpu().numpy(), r.cpu().numpy()
    for b in range(B):
        want, r_ref = ob.xi_from_P(k, Pk[b], ell)
        np.testing.assert_allclose(r[b], r_ref, rtol=1e-15)
        np.testing.assert_allclose(xi[b], want, rtol=0, atol=1e-11 * np.abs(want).max() + 1e-300)


def test_observable_argument_errors(gpu):
    import torch
    from discoeb_amd import _lib
    from discoeb_amd import observables as gob
    k = np.geomspace(1e-4, 1.0, 15)
    with pytest.raises(_lib.DebError):  # odd N
        gob.get_xi_from_P(k=k, Pk=k, N=15)
    y = np.ones((15, 20))
    with pytest.raises(_lib.DebError):  # window longer than the k grid
        gob.get_power_smoothed(k=k, y=y, dlogk=20.0, idx=4, param=dict(A_s=2e-9, n_s=0.96))
    torch.cuda.synchronize()
