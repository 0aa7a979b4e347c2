"""Pins of the oracle's P(k) observables (oracle/observables.py) against
independent results: scipy's Savitzky-Golay coefficients and log-gamma, the
Legendre projection of the Kaiser spectrum, and the Hankel transform the
FFTlog approximates (CPU only)."""
import numpy as np
import pytest
from scipy import integrate, signal, special

import observables as ob


@pytest.mark.parametrize("W,p", [(5, 2), (7, 3), (11, 3), (21, 3), (51, 3), (9, 4), (8, 3)])
def test_savgol_coefficients_match_scipy(W, p):
    """The least-squares filter of util.py:426-442 is scipy's savgol_coeffs
    (deriv 0, convolution order; symmetric for odd W)."""
    c = ob.savgol_coeffs(W, p)
    np.testing.assert_allclose(c, signal.savgol_coeffs(W, p, use="conv"), rtol=0, atol=1e-13)


def test_savgol_reproduces_polynomials_in_the_interior():
    """A degree-polyorder polynomial is a fixed point of the filter away from
    the zero-padded ends."""
    x = np.linspace(-2, 3, 200)
    y = 0.3 - 1.2 * x + 0.7 * x**2 - 0.25 * x**3
    s = ob.savgol_filter(y, 15, 3)
    np.testing.assert_allclose(s[7:-7], y[7:-7], rtol=0, atol=1e-11)


def test_smoothing_keeps_a_power_law_and_replaces_the_ends():
    """log P linear in log k passes the cubic filter unchanged; the first
    W//2 and last ceil(W/2) points are the unsmoothed P (perturbations.py:1153-1157)."""
    k = np.geomspace(1e-4, 1.0, 300)
    y = np.zeros((300, 20))
    y[:, 4] = 3.0 * k**0.8
    P = ob.power(k, y, 4, 2.1e-9, 0.96)
    Ps = ob.power_smoothed(k, y, 0.5, 4, 2.1e-9, 0.96)
    np.testing.assert_allclose(Ps, P, rtol=1e-12)
    # a wiggle is damped in the interior, untouched at the ends
    y[:, 4] *= 1 + 0.05 * np.sin(60 * np.log(k))
    P = ob.power(k, y, 4, 2.1e-9, 0.96)
    Ps = ob.power_smoothed(k, y, 0.5, 4, 2.1e-9, 0.96)
    W = ob.window_length_of(k, 0.5)
    assert W % 2 == 1
    np.testing.assert_array_equal(Ps[: W // 2], P[: W // 2])
    np.testing.assert_array_equal(Ps[-(W + 1) // 2:], P[-(W + 1) // 2:])
    mid = slice(W, -W)
    assert np.abs(np.log(Ps[mid] / P[mid]) - np.log(1 / 1)).max() < 0.5 * np.abs(np.log(P[mid])).std()


def test_window_length_rounds_half_to_even_then_odd():
    k = np.array([1.0, np.e, np.e**2])  # log-spacing exactly 1
    assert ob.window_length_of(k, 2.5) == 3   # 2.5 -> 2 -> 3
    assert ob.window_length_of(k, 3.5) == 5   # 3.5 -> 4 -> 5
    assert ob.window_length_of(k, 4.5) == 5   # 4.5 -> 4 -> 5
    assert ob.window_length_of(k, 6.2) == 7


def test_multipoles_are_the_legendre_projection_of_kaiser():
    """P_l(k) = (2l+1)/2 int_{-1}^{1} P(k, mu) L_l(mu) dmu for the Kaiser
    spectrum (b delta - mu^2 theta)^2 (a degree-4 polynomial in mu: 8-point
    Gauss-Legendre is exact)."""
    rng = np.random.default_rng(1)
    k = np.geomspace(1e-3, 0.5, 40)
    y = rng.standard_normal((40, 20))
    b = 1.7
    P0, P2, P4 = ob.power_multipoles(y, k, b, 2.1e-9, 0.97)
    mu, w = np.polynomial.legendre.leggauss(8)
    fac = 2 * np.pi**2 * 2.1e-9
    f = np.sqrt(fac * (k / 0.05) ** (0.97 - 1) * k ** (-3))
    Pkmu = (b * f[:, None] * y[:, 4:5] - mu[None, :] ** 2 * f[:, None] * y[:, 5:6]) ** 2
    for ell, Pl in ((0, P0), (2, P2), (4, P4)):
        L = special.eval_legendre(ell, mu)
        proj = (2 * ell + 1) / 2 * (Pkmu * L[None, :] * w[None, :]).sum(1)
        np.testing.assert_allclose(Pl, proj, rtol=1e-12, atol=1e-12 * np.abs(Pkmu).max())
    # and power_Kaiser samples the same P(k, mu)
    Pk, mus = ob.power_kaiser(y, k, b, 9, 2.1e-9, 0.97)
    ref = (b * f[:, None] * y[:, 4:5] - mus[None, :] ** 2 * f[:, None] * y[:, 5:6]) ** 2
    np.testing.assert_allclose(Pk, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
    _, mus2 = ob.power_kaiser(y, k, b, 9, 2.1e-9, 0.97, mu_sampling=False)
    np.testing.assert_allclose(mus2, np.cos(np.linspace(0, np.pi, 9)))


@pytest.mark.parametrize("z", [0.75 + 0.0j, 0.75 + 3.2j, 1.75 + 17.0j, 2.75 + 0.4j, 0.3 + 1.1j, -1.3 + 0.7j,
                               0.75 + 150.0j])
def test_lngamma_matches_scipy(z):
    """Lanczos log-gamma (util.py:12-44): real part equal, phase equal modulo 2 pi."""
    got = ob.lngamma_complex(z)
    want = special.loggamma(z)
    assert abs(got.real - want.real) <= 1e-12 * max(1.0, abs(want.real))
    d = (got.imag - want.imag) / (2 * np.pi)
    assert abs(d - round(d)) < 1e-12 * max(1.0, abs(want.imag))


@pytest.mark.parametrize("ell", [0, 2, 4])
def test_fftlog_converges_to_the_hankel_transform(ell):
    """xi_l(r) -> i^l int k^2 P(k) j_l(kr) dk / (2 pi^2) for a smooth, localised
    P(k) = k^2 exp(-(8k)^2); the reference's FFTlog (output radii r = 2 pi / k)
    is first order in the log-k spacing: the error halves when N doubles."""
    R = 8.0
    errs = []
    for N in (1024, 2048):
        k = np.geomspace(1e-5, 1e3, N)
        xi, r = ob.xi_from_P(k, k**2 * np.exp(-(k * R) ** 2), ell)
        assert np.all(np.diff(r) > 0)
        e = 0.0
        for rr in (4.0, 8.0, 16.0, 32.0):
            i = int(np.argmin(np.abs(r - rr)))
            f = lambda kk: kk**4 * np.exp(-(kk * R) ** 2) * special.spherical_jn(ell, kk * r[i]) / (2 * np.pi**2)
            want = integrate.quad(f, 0, 3.0, limit=1000, epsabs=0, epsrel=1e-13)[0] * (1j**ell).real
            e = max(e, abs(xi[i] / want - 1))
        errs.append(e)
    assert errs[1] < 6e-3  # 0.9e-3, 2.8e-3, 4.4e-3 for l = 0, 2, 4 at N = 2048
    assert 1.6 < errs[0] / errs[1] < 2.5


def test_fftlog_odd_multipole_is_zero():
    """The real part of i^l times a real transform vanishes for odd l."""
    k = np.geomspace(1e-4, 10, 256)
    xi, _ = ob.xi_from_P(k, k * np.exp(-k), 1)
    assert np.all(xi == 0)
