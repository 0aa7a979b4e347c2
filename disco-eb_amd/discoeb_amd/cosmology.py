"""Cosmological parameter vectors and the benchmark workloads.

The 8 differentiable parameters theta = (omega_c, omega_b, h, n_s,
ln(1e10 A_s), m_nu [eV], w0, wa) map onto the reference's ``param`` dict as
(tests/test_perturbations.py:12-57 for the conventions):

  Omegam = (omega_c + omega_b) / h^2   (CDM + baryons; massive nu separate)
  Omegab = omega_b / h^2,  H0 = 100 h,  A_s = exp(ln10As) / 1e10
  one massive nu: Nmnu = 1, Neff = 3.046 - 1;  none: Nmnu = 0, Neff = 3.046
  Tcmb = 2.7255 K, YHe = 0.248, cs2_DE = 1, Omegak = 0, k_p = 0.05 / Mpc
"""
from __future__ import annotations

import numpy as np

THETA_NAMES = ("omega_c", "omega_b", "h", "n_s", "ln10As", "m_nu", "w0", "wa")
# sampling box of the multi-cosmology workloads
THETA_LO = np.array([0.10, 0.020, 0.60, 0.92, 2.8, 0.02, -1.2, -0.5])
THETA_HI = np.array([0.14, 0.024, 0.75, 1.00, 3.2, 0.30, -0.8, 0.5])
SEED = 20261202


def param_from_theta(theta, massive_nu: bool = True) -> dict:
    """Reference ``param`` dict (arrays of shape (B,)) from theta[B, 8].

    Works for complex theta (complex-step derivatives)."""
    th = np.atleast_2d(np.asarray(theta))
    oc, ob, h, ns, lnAs, mnu, w0, wa = (th[:, i] for i in range(8))
    B = th.shape[0]
    one = np.ones(B)
    p = dict(
        Omegam=(oc + ob) / h**2,
        Omegab=ob / h**2,
        H0=100.0 * h,
        A_s=np.exp(lnAs) / 1e10,
        n_s=ns,
        w_DE_0=w0,
        w_DE_a=wa,
        cs2_DE=1.0 * one,
        Omegak=0.0 * one,
        Tcmb=2.7255 * one,
        YHe=0.248 * one,
        k_p=0.05 * one,
    )
    if massive_nu:
        p.update(Neff=(3.046 - 1.0) * one, Nmnu=1.0 * one, mnu=mnu)
    else:
        p.update(Neff=3.046 * one, Nmnu=0.0 * one, mnu=0.0 * one)
    return p


def base_theta(w0: float = -0.9, wa: float = 0.1, m_nu: float = 0.06) -> np.ndarray:
    """h = 0.67, Omega_m = 0.31, Omega_b = 0.049, A_s = 2.1e-9, n_s = 0.965."""
    h = 0.67
    ob = 0.049 * h**2
    oc = (0.31 - 0.049) * h**2
    return np.array([oc, ob, h, 0.965, np.log(21.0), m_nu, w0, wa])


def sample_thetas(n: int, seed: int = SEED) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.uniform(THETA_LO, THETA_HI, size=(n, 8))


def loss_weights(n_k: int, n_cosmo: int, seed: int = SEED) -> np.ndarray:
    """w(k) ~ U[0, 1] of the batched-adjoint loss sum_k w(k) ln P(k): drawn from
    the same generator as the cosmologies, after the n_cosmo draws."""
    rng = np.random.default_rng(seed)
    rng.uniform(THETA_LO, THETA_HI, size=(n_cosmo, 8))
    return rng.uniform(0.0, 1.0, size=n_k)


def test_cosmology_param() -> dict:
    """The reference test suite's cosmology (tests/test_perturbations.py:12-57)."""
    Tnu = (4 / 11) ** (1 / 3)
    return dict(Omegam=0.3099, Omegab=0.0488911, w_DE_0=-0.99, w_DE_a=0.0, cs2_DE=1.0, Omegak=0.0,
                A_s=2.1064e-09, n_s=0.96822, H0=67.742, Tcmb=2.7255, YHe=0.248,
                Neff=3.046 - 1 * (Tnu / ((4 / 11) ** (1 / 3))) ** 4, Nmnu=1, mnu=0.06, k_p=0.05)


# workloads: hierarchy sizes, k-grid, output scale factors, tolerances
WORKLOADS = {
    "plain-lcdm-64k": dict(lmaxg=16, lmaxgp=16, lmaxr=16, lmaxnu=8, nqmax=3, kmin=1e-4, kmax=1.0, num_k=64,
                           aexp_out=[1.0], rtol=1e-5, atol=1e-8, massive_nu=False, w0=-1.0, wa=0.0),
    "plain-mnu-256k-16out": dict(lmaxg=32, lmaxgp=32, lmaxr=32, lmaxnu=17, nqmax=5, kmin=1e-5, kmax=10.0,
                                 num_k=256, aexp_out=list(np.geomspace(0.01, 1.0, 16)), rtol=1e-5, atol=1e-8,
                                 massive_nu=True, w0=-0.9, wa=0.1),
    "batched-64cosmo-256k": dict(lmaxg=32, lmaxgp=32, lmaxr=32, lmaxnu=17, nqmax=5, kmin=1e-5, kmax=10.0,
                                 num_k=256, aexp_out=[1.0], rtol=1e-5, atol=1e-8, massive_nu=True, n_cosmo=64),
    "grad-base-256k": dict(lmaxg=32, lmaxgp=32, lmaxr=32, lmaxnu=17, nqmax=5, kmin=1e-5, kmax=10.0, num_k=256,
                           aexp_out=[1.0], rtol=1e-5, atol=1e-8, massive_nu=True, w0=-0.9, wa=0.1),
    "batched-adjoint-256cosmo-512k": dict(lmaxg=50, lmaxgp=50, lmaxr=50, lmaxnu=17, nqmax=5, kmin=1e-5,
                                          kmax=20.0, num_k=512, aexp_out=[1.0], rtol=1e-5, atol=1e-8,
                                          massive_nu=True, n_cosmo=256),
}
