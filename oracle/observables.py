"""CPU restatement of the reference's P(k) observables (TEST INFRASTRUCTURE ONLY).

Only tests/ may import this module.  numpy restatement, one cosmology at a
time, of:

* ``get_power``           src/discoeb/perturbations.py:1101-1123
* ``get_power_smoothed``  src/discoeb/perturbations.py:1126-1159
* ``savgol_filter``       src/discoeb/util.py:407-444
* ``power_Kaiser``        src/discoeb/perturbations.py:1162-1198
* ``power_multipoles``    src/discoeb/perturbations.py:1202-1224
* ``lngamma_complex_e``   src/discoeb/util.py:12-44 (Lanczos, after GSL)
* ``get_xi_from_P``       src/discoeb/perturbations.py:1065-1098 (FFTlog)

Pinned in tests/test_observables_oracle.py against scipy.signal.savgol_coeffs,
scipy.special.loggamma, the Legendre projection of the Kaiser spectrum and
analytic Hankel pairs.
"""
from __future__ import annotations

import numpy as np


def power(k, y, idx, A_s, n_s, k_p=0.05):
    """get_power (perturbations.py:1123)."""
    return 2 * np.pi**2 * A_s * (k / k_p) ** (n_s - 1) * k ** (-3) * y[..., idx] ** 2


def savgol_coeffs(window_length: int, polyorder: int) -> np.ndarray:
    """The filter of savgol_filter (util.py:426-442): least-squares (minimum
    norm) solution of A c = e_0, A[m, j] = x_j^m on the reversed grid."""
    halflen, rem = divmod(window_length, 2)
    pos = halflen - 0.5 if rem == 0 else halflen
    x = np.arange(-pos, window_length - pos, dtype=float)[::-1]
    A = x ** np.arange(polyorder + 1).reshape(-1, 1)
    Y = np.zeros(polyorder + 1)
    Y[0] = 1.0
    return np.linalg.lstsq(A, Y, rcond=None)[0]


def savgol_filter(y, window_length: int, polyorder: int):
    """util.py:444: convolution 'same' with the coefficients."""
    return np.convolve(y, savgol_coeffs(window_length, polyorder), mode="same")


def window_length_of(k, dlogk: float) -> int:
    """perturbations.py:1150-1151 (round half to even, then odd)."""
    wl = int(np.round(dlogk / (np.log(k[1]) - np.log(k[0]))))
    return wl + (wl + 1) % 2


def power_smoothed(k, y, dlogk, idx, A_s, n_s, k_p=0.05):
    """get_power_smoothed (perturbations.py:1150-1159)."""
    wl = window_length_of(k, dlogk)
    Pm = power(k, y, idx, A_s, n_s, k_p)
    Pms = np.exp(savgol_filter(np.log(Pm), wl, 3))
    Pms[: wl // 2] = Pm[: wl // 2]
    Pms[-wl // 2:] = Pm[-wl // 2:]
    return Pms


def power_kaiser(y, kmodes, bias, nmu, A_s, n_s, k_p=0.05, mu_sampling=True, smooth_dlogk=None):
    """power_Kaiser (perturbations.py:1182-1198)."""
    if mu_sampling:
        mu = np.linspace(-1, 1, nmu)
    else:
        mu = np.cos(np.linspace(0, np.pi, nmu))
    if smooth_dlogk is None:
        fac = 2 * np.pi**2 * A_s
        f = np.sqrt(fac * (kmodes / k_p) ** (n_s - 1) * kmodes ** (-3))
        deltam, thetam = f * y[:, 4], f * y[:, 5]
    else:
        deltam = np.sqrt(power_smoothed(kmodes, y, smooth_dlogk, 4, A_s, n_s, k_p))
        thetam = -np.sqrt(power_smoothed(kmodes, y, smooth_dlogk, 5, A_s, n_s, k_p))
    return (bias * deltam[:, None] - mu[None, :] ** 2 * thetam[:, None]) ** 2, mu


def power_multipoles(y, kmodes, b, A_s, n_s, k_p=0.05):
    """power_multipoles (perturbations.py:1215-1224)."""
    fac = 2 * np.pi**2 * A_s
    f = np.sqrt(fac * (kmodes / k_p) ** (n_s - 1) * kmodes ** (-3))
    d, t = f * y[:, 4], f * y[:, 5]
    P0 = b**2 * d**2 - 2 * b / 3 * d * t + 1 / 5 * t**2
    P2 = -4 * b / 3 * d * t + 4 / 7 * t**2
    P4 = 8 / 35 * t**2
    return P0, P2, P4


_LANCZOS7 = np.array([0.99999999999980993227684700473478, 676.520368121885098567009190444019,
                      -1259.13921672240287047156078755283, 771.3234287776530788486528258894,
                      -176.61502916214059906584551354, 12.507343278686904814458936853,
                      -0.13857109526572011689554707, 9.984369578019570859563e-6, 1.50563273514931155834e-7])


def lngamma_complex(z: complex) -> complex:
    """lngamma_complex_e (util.py:12-44): Lanczos series, reflection for Re z <= 0.5."""
    def lanczos(z):
        z = z - 1.0
        i = np.arange(1, 9)
        Ag = _LANCZOS7[0] + np.sum(_LANCZOS7[1:] / np.abs(z + i) ** 2 * np.conj(z + i))
        return (z + 0.5) * np.log(z + 7.5) - (z + 7.5) + 0.9189385332046727418 + np.log(Ag)
    if z.real <= 0.5:
        return 1.14472988584940017414342735135 - np.log(np.sin(np.pi * z)) - lanczos(1.0 - z)
    return lanczos(z)


def xi_from_P(k, Pk, ell=0):
    """get_xi_from_P (perturbations.py:1081-1098)."""
    N = len(k)
    L = np.log(k[N - 1] / k[0])
    fPk = np.fft.rfft(Pk * k**1.5)
    ki = np.pi * np.arange(N // 2 + 1) / L
    zp = (1.5 + ell) / 2 + 1j * ki
    theta = np.array([lngamma_complex(z).imag for z in zp])
    fPk = fPk * np.exp(2j * (theta - np.log(np.pi) * ki))
    r = 2 * np.pi / k
    xi = np.real(1j**ell * np.fft.irfft(fPk) / (2 * np.pi * r) ** 1.5)
    return xi[::-1], r[::-1]
