"""RECFAST-style recombination history (host-side table producer).

This file is derived from RECFAST through DISCO-EB's
thermodynamics_recfast.py and carries the RECFAST copyright notice, which its
licence requires on all copies and modifications:

 # Integrator for Cosmic Recombination of Hydrogen and Helium,
 # developed by Douglas Scott (dscott@astro.ubc.ca)
 # based on calculations in the papers Seager, Sasselov & Scott
 # (ApJ, 523, L1, 1999; ApJS, 128, 407, 2000)
 # and "fudge" updates in Wong, Moss & Scott (2008).
 #
 # Permission to use, copy, modify and distribute without fee or royalty at
 # any tier, this software and its documentation, for any purpose and without
 # fee or royalty is hereby granted, provided that you agree to comply with
 # the following copyright notice and statements, including the disclaimer,
 # and that the same appear on ALL copies of the software and documentation,
 # including modifications that you make for internal use or for distribution:
 #
 # Copyright 1999-2010 by University of British Columbia.  All rights reserved.

The original RECFAST code: https://www.astro.ubc.ca/people/scott/recfast.html

Follows ``src/discoeb/thermodynamics_recfast.py`` of DISCO-EB, itself adapted
from RECFAST (Seager, Sasselov & Scott 1999/2000; Wong, Moss & Scott 2008):

* physical and atomic constants: lines 37-122 (values must agree with the
  reference: they define the tables);
* ``ionization`` (H, He and matter-temperature ODE right-hand side): 124-281;
* ``solve_ionization``: one adaptive GRK4A Rosenbrock integration per segment,
  diffrax ``PIDController(rtol, atol)`` defaults, ``max_steps=128``,
  ``throw=False``: 283-300; GRK4A step ``ode_integrators_stiff.py:117-223``;
* ``Saha_HeII``: 302-321; ``_get_adaptive_sampling``: 324-360;
* ``compute_thermal_history`` (segment loop with Saha shortcuts): 362-452;
* ``evaluate_thermo`` (xe, mu, cs2, tau by Romberg per segment): 455-501.

All arrays carry a leading cosmology axis (B,).  Values may be complex
(complex-step derivatives); step-size control uses real parts only, which
matches diffrax's ``stop_gradient`` on the controller factor.
"""
from __future__ import annotations

import numpy as np

from . import _dual as D

const_G = 6.67430e-11
const_mH = 1.67353284e-27
const_me = 9.1093837015e-31
const_mHe_mH = 3.97146570884
const_c = 2.99792458e08
const_h = 6.62607015e-34
const_kB = 1.380649e-23
const_sigma = 6.6524587321e-29
const_arad = 4 * 5.670374419e-8 / const_c
bigH = 3.2407792902755102e-18
const_dens_fac = 11.223810928601939
const_EionHe12s = 6.314878282674e5

Lambda = 8.2245809
Lambda_He = 51.3
L_H_ion = 1.096787737e7
L_H_alpha = 8.225916453e6
L_He1_ion = 1.98310772e7
L_He2_ion = 4.389088863e7
L_He_2s = 1.66277434e7
L_He_2p = 1.71134891e7
A2P_s = 1.798287e9
A2P_t = 177.58
L_He_2Pt = 1.690871466e7
L_He_2St = 1.5985597526e7
L_He2St_ion = 3.8454693845e6
sigma_He_2Ps = 1.436289e-22
sigma_He_2Pt = 1.484872e-22
AGauss1 = -0.14
AGauss2 = 0.05
zGauss1 = 7.28
zGauss2 = 6.75
wGauss1 = 0.18
wGauss2 = 0.33
H_frac = 1.0e-3

Lalpha = 1.0 / L_H_alpha
Lalpha_He = 1.0 / L_He_2p
DeltaB = const_h * const_c * (L_H_ion - L_H_alpha)
CDB = DeltaB / const_kB
DeltaB_He = const_h * const_c * (L_He1_ion - L_He_2s)
CDB_He = DeltaB_He / const_kB
CB1 = const_h * const_c * L_H_ion / const_kB
CB1_He1 = const_h * const_c * L_He1_ion / const_kB
CB1_He2 = const_h * const_c * L_He2_ion / const_kB
CR = 2.0 * np.pi * (const_me / const_h) * (const_kB / const_h)
CK = Lalpha**3 / (8.0 * np.pi)
CK_He = Lalpha_He**3 / (8.0 * np.pi)
CL = const_c * const_h / (const_kB * Lalpha)
CL_He = const_c * const_h / (const_kB / L_He_2s)
CT = (8.0 / 3.0) * (const_sigma / (const_me * const_c)) * const_arad
Bfact = const_h * const_c * (L_He_2p - L_He_2s) / const_kB
CL_PSt = const_h * const_c * (L_He_2Pt - L_He_2St) / const_kB
const_h_c_L_He_2St_kB = const_h * const_c * L_He_2St / const_kB

a_PPB = 4.309
b_PPB = -0.6166
c_PPB = 0.6703
d_PPB = 0.5300
a_VF = 10.0 ** (-16.744)
b_VF = 0.711
T_0 = 10 ** (0.477121)
T_1 = 10 ** (5.114)
a_trip = 10.0 ** (-16.306)
b_trip = 0.761

# GRK4A coefficients (Kaps & Rentrop 1979), ode_integrators_stiff.py:131-150
_G = dict(gamma=0.395, g21=-0.767672395484, g31=-0.851675323742, g32=0.522967289188,
          g41=0.288463109545, g42=0.880214273381e-1, g43=-0.337389840627,
          a21=0.438, a31=0.796920457938, a32=0.730795420615e-1,
          ch1=0.346325833758, ch2=0.285693175712, ch3=0.367980990530,
          c1=0.199293275701, c2=0.482645235674, c3=0.680614886256e-1, c4=0.25)


def ionization(a, y, p):
    """d(x_H, x_He, T_m)/da; ``a`` (B,) real, ``y`` three (B,) values/Duals."""
    from .background import get_aprimeoa
    z = 1 / a - 1
    H = p["H0"] / 100.0
    HO = H * bigH
    fu = 1.105
    mu_H = 1.0 / (1.0 - p["YHe"])
    fHe = p["YHe"] / (const_mHe_mH * (1.0 - p["YHe"]))
    Nnow = const_dens_fac * H * H * p["Omegab"] / mu_H
    x_H = y[0]
    x_He = y[1]
    x = x_H + fHe * x_He
    Tmat = D.rabs(y[2])
    n = Nnow * (1 + z) ** 3
    n_He = fHe * n
    Trad = p["Tcmb"] * (1 + z)
    z_term = 1 + z
    Hz = (1e-5 * get_aprimeoa(p, a)) / a * const_c * bigH
    Tmat_1e4 = Tmat / 1e4
    CR_Tmat_15 = (CR * Tmat) ** 1.5
    Rdown = 1e-19 * a_PPB * Tmat_1e4**b_PPB / (1 + c_PPB * Tmat_1e4**d_PPB)
    Rup = Rdown * CR_Tmat_15 * D.exp(-CDB / Tmat)
    sq_0 = D.sqrt(Tmat / T_0)
    sq_1 = D.sqrt(Tmat / T_1)
    sq_0_term = 1 + sq_0
    sq_1_term = 1 + sq_1
    Rdown_He = a_VF / (sq_0 * sq_0_term ** (1 - b_VF) * sq_1_term ** (1 + b_VF))
    Rup_He = 4 * Rdown_He * CR_Tmat_15 * D.exp(-CDB_He / Tmat)
    He_Boltz = D.exp(D.minimum(680.0, Bfact / Tmat))
    Rdown_trip = a_trip / (sq_0 * sq_0_term ** (1 - b_trip) * sq_1_term ** (1 + b_trip))
    Rup_trip = Rdown_trip * D.exp(-const_h * const_c * L_He2St_ion / (const_kB * Tmat)) * CR_Tmat_15 * (4 / 3)
    lz = np.log(z_term)
    K_gaussian = (1 + AGauss1 * np.exp(-(((lz - zGauss1) / wGauss1) ** 2))
                  + AGauss2 * np.exp(-(((lz - zGauss2) / wGauss2) ** 2)))
    K = CK / Hz * K_gaussian
    one_minus_x_He = 1 - x_He
    one_minus_x_H = 1 - x_H
    n_He_1mxHe = n_He * one_minus_x_He
    tauHe_s = A2P_s * CK_He * 3 * n_He_1mxHe / Hz
    pHe_s = (1 - D.exp(-tauHe_s)) / tauHe_s
    Doppler_term = D.sqrt(2 * const_kB * Tmat / (const_mH * const_mHe_mH * const_c**2))
    Doppler_2p = const_c * L_He_2p * Doppler_term
    Doppler_2Pt = const_c * L_He_2Pt * Doppler_term
    gamma_2Ps = (3 * A2P_s * fHe * one_minus_x_He * const_c**2) / (
        np.sqrt(np.pi) * sigma_He_2Ps * 8 * np.pi * Doppler_2p * one_minus_x_H * (const_c * L_He_2p) ** 2)
    AHcon = A2P_s / (1 + 0.36 * (gamma_2Ps**0.86))
    xHe_r = np.real(D.val(x_He))
    xH_r = np.real(D.val(x_H))
    cond1 = (xHe_r < 5e-9) | (xHe_r > 0.98)
    cond2 = xH_r < 0.9999999
    K_He_default = CK_He / Hz
    K_He_case1 = 1.0 / ((A2P_s * pHe_s + AHcon) * 3.0 * n_He_1mxHe)
    K_He_case2 = 1.0 / (A2P_s * pHe_s * 3.0 * n_He_1mxHe)
    K_He = D.where(cond1, K_He_default, D.where(cond2, K_He_case1, K_He_case2))
    tauHe_t = A2P_t * n_He_1mxHe * 3 / (8 * np.pi * Hz * L_He_2Pt**3)
    pHe_t = (1 - D.exp(-tauHe_t)) / tauHe_t
    gamma_2Pt = (3 * A2P_t * fHe * one_minus_x_He * const_c**2) / (
        np.sqrt(np.pi) * sigma_He_2Pt * 8 * np.pi * Doppler_2Pt * one_minus_x_H * (const_c * L_He_2Pt) ** 2)
    AHcon_t = A2P_t / (1 + 0.66 * gamma_2Pt**0.9) / 3
    exp_CL_PSt = D.exp(-CL_PSt / Tmat)
    CfHe_t = D.where(xH_r > 0.99999, A2P_t * pHe_t * exp_CL_PSt, (A2P_t * pHe_t + AHcon_t) * exp_CL_PSt)
    CfHe_t = CfHe_t / (Rup_trip + CfHe_t)
    timeTh = (1 / (CT * Trad**4)) * (1 + x + fHe) / x
    timeH = 2 / (3 * HO * z_term**1.5)
    Hz_z_term = Hz * z_term
    rate_diff = x * x_H * n * Rdown - Rup * one_minus_x_H * D.exp(-CL / Tmat)
    KLn = K * Lambda * n * one_minus_x_H
    denom_term = 1 / fu + KLn / fu + K * Rup * n * one_minus_x_H
    f0_case2 = rate_diff / Hz_z_term
    f0_case3 = (rate_diff * (1 + KLn)) / (Hz_z_term * denom_term)
    f0 = D.where(xH_r > 0.99, 0.0 * f0_case2, D.where(xH_r > 0.985, f0_case2, f0_case3))
    rate_diff_He = x * x_He * n * Rdown_He - Rup_He * one_minus_x_He * D.exp(-CL_He / Tmat)
    KHe_L = K_He * Lambda_He * n_He_1mxHe * He_Boltz
    He_denom = 1 + K_He * (Lambda_He + Rup_He) * n_He_1mxHe * He_Boltz
    f1_main = (rate_diff_He * (1 + KHe_L)) / (Hz_z_term * He_denom)
    trip_rate_diff = x * x_He * n * Rdown_trip - one_minus_x_He * 3 * Rup_trip * D.exp(-const_h_c_L_He_2St_kB / Tmat)
    trip_contrib = trip_rate_diff * CfHe_t / Hz_z_term
    trip_cond = (xHe_r < 5e-9) | (xHe_r > 0.98)
    f1 = D.where(xHe_r < 1e-8, 0.0 * f1_main, f1_main + D.where(trip_cond, 0.0 * trip_contrib, trip_contrib))
    Tmr = np.real(D.val(timeTh)) < H_frac * np.real(timeH)
    f2_case1 = Tmat / z_term
    f2_case2 = CT * (Trad**4) * x / (1 + x + fHe) * (Tmat - Trad) / Hz_z_term + 2 * Tmat / z_term
    f2 = D.where(Tmr, f2_case1, f2_case2)
    dzda = -1 / a**2
    return f0 * dzda, f1 * dzda, f2 * dzda


def _grk4a_step(a0, a1, y0, p):
    """One GRK4A step over [a0, a1]; y0 is (3, B). Returns (y1, err)."""
    g = _G
    dt = a1 - a0
    B = y0.shape[1]
    yd = [D.Dual(y0[i], np.eye(3)[:, i][:, None] * np.ones((1, B))) for i in range(3)]
    f = ionization(a0, yd, p)
    fv = np.stack([fi.v for fi in f]) * dt
    # J[b, i, j] = d(f_i dt)/dy_j
    J = np.stack([np.stack([f[i].d[j] for j in range(3)], axis=-1) for i in range(3)], axis=1) * dt[:, None, None]
    A = (np.eye(3)[None] - g["gamma"] * J).real  # stop_gradient(I - gamma J)
    Jm = lambda k: np.einsum("bij,jb->ib", J, k)
    sol = lambda b: np.linalg.solve(A, b.T[..., None])[..., 0].T

    def fval(y):
        return np.stack(ionization(a0, list(y), p)) * dt

    k1 = sol(fv)
    y1 = y0 + g["a21"] * k1
    f1 = fval(y1)
    Jk1 = Jm(k1)
    k2 = sol(f1 + g["g21"] * Jk1)
    y2 = y0 + g["a31"] * k1 + g["a32"] * k2
    f2 = fval(y2)
    Jk2 = Jm(k2)
    k3 = sol(f2 + g["g31"] * Jk1 + g["g32"] * Jk2)
    Jk3 = Jm(k3)
    k4 = sol(f2 + g["g41"] * Jk1 + g["g42"] * Jk2 + g["g43"] * Jk3)
    ynew = y0 + g["c1"] * k1 + g["c2"] * k2 + g["c3"] * k3 + g["c4"] * k4
    err = (g["c1"] - g["ch1"]) * k1 + (g["c2"] - g["ch2"]) * k2 + (g["c3"] - g["ch3"]) * k3 + g["c4"] * k4
    return ynew, err


def solve_ionization(astart, aend, ystart, p, rtol=1e-3, atol=1e-6, max_steps=128):
    """Adaptive GRK4A integration of (x_H, x_He, T_m) from astart to aend.

    diffrax semantics: dt0 = |astart| 1e-3, default PIDController (I-controller,
    safety 0.9, factormin 0.2, factormax 10, error order 3, RMS norm), final
    step clipped to aend within 1e-10.  Returns (y_end (3,B), dy/da (3,B),
    failed (B,) bool) where ``failed`` marks lanes that hit max_steps (the
    reference then keeps its initial inf fill: throw=False).
    """
    B = ystart.shape[1]
    y = ystart.copy()
    tprev = astart.copy()
    tnext = np.minimum(astart + np.abs(astart) * 1e-3, aend)
    done = tprev >= aend
    steps = 0
    while not np.all(done) and steps < max_steps:
        act = ~done
        y1, err = _grk4a_step(tprev, tnext, y, p)
        yr, y1r, er = y.real, y1.real, err.real
        scale = er / (atol + np.maximum(np.abs(yr), np.abs(y1r)) * rtol)
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            nrm = np.sqrt(np.mean(scale**2, axis=0))
            keep = nrm < 1
            inv = 1.0 / nrm
            f1 = inv ** (1.0 / 3.0)
            fmin = np.where(keep, 1.0, 0.2)
            factor = np.clip(0.9 * f1, fmin, 10.0)
        dt = (tnext - tprev) * factor
        nt0 = np.where(keep, tnext, tprev)
        nt1 = nt0 + dt
        clip = nt1 > aend - 1e-10
        nt1 = np.where(clip, np.where(keep, aend, nt0 + 0.5 * (aend - nt0)), nt1)
        upd = act & keep
        y = np.where(upd[None], y1, y)
        tprev = np.where(act, nt0, tprev)
        tnext = np.where(act, nt1, tnext)
        done = tprev >= aend
        steps += 1
    failed = ~done
    dyda = np.stack(ionization(aend, list(y), p))
    return y, dyda, failed


def saha_HeII(a, p):
    """HeII -> HeIII Saha equilibrium x_e contribution (lines 302-321)."""
    const_Hfac = 1 / (1.0e06 * 3.0856775807e13)
    mu_H = 1 / (1 - p["YHe"])
    rho_c = 3 * (p["H0"] * const_Hfac) ** 2 / (8 * np.pi * const_G)
    NHnow = rho_c * p["Omegab"] / (const_mH * mu_H) / a**3
    T = p["Tcmb"] / a
    betaE = const_EionHe12s / T
    A = 1 + p["fHe"]
    Bc = 1 + 2 * p["fHe"]
    R = (2 * np.pi * const_me * const_kB / const_h**2 * T) ** 1.5 / NHnow * D.exp(-betaE)
    asym = p["fHe"] * (1 - Bc / R + (1 + 5 * p["fHe"] + 6 * p["fHe"] ** 2) / R**2)
    full = -(R - A) / 2 + D.sqrt((R - A) ** 2 / 4 + R * Bc) - A
    return D.where(np.real(D.val(R)) > 1e5, asym, full)


def adaptive_sampling(a0: float, a1: float, N: int) -> np.ndarray:
    n_very_early = max(8, int(N * 0.05))
    n_pre = max(8, int(N * 0.10))
    n_recomb = int(N * 0.50)
    n_post = N - n_very_early - n_pre - n_recomb
    ab1 = max(1.0 / 3001.0, a0)
    ab2 = 1.0 / 1401.0
    ab3 = 1.0 / 601.0
    return np.concatenate([
        np.geomspace(a0, ab1, n_very_early, endpoint=False),
        np.geomspace(ab1, ab2, n_pre, endpoint=False),
        np.geomspace(ab2, ab3, n_recomb, endpoint=False),
        np.geomspace(ab3, a1, n_post)])


def compute_thermal_history(*, a0, a1, N, param, rtol=1e-3, atol=1e-6):
    """Segment loop of ``compute_thermal_history`` (lines 362-452).

    Returns y (6, B, N) = (x_HI, x_HeI, T_m, dx_H/da, dx_He/da, dT_m/da) at
    a[1:], the sampled scale factors a[1:], and the per-lane failure count.
    """
    a = adaptive_sampling(a0, a1, N + 1)
    p = param
    B = p["H0"].shape[0]
    H = p["H0"] / 100.0
    HO = H * bigH
    mu_H = 1.0 / (1.0 - p["YHe"])
    Nnow = 3.0 * HO * HO * p["Omegab"] / (8.0 * np.pi * const_G * mu_H * const_mH)
    fHe = p["YHe"] / (const_mHe_mH * (1.0 - p["YHe"]))
    Tcmb = p["Tcmb"]
    Tcmb3 = Tcmb**3
    dtype = np.result_type(*[np.asarray(v).dtype for v in (p["H0"], p["Omegab"], Tcmb, p["YHe"])], np.float64)
    yarr = np.zeros((6, B, N), dtype=dtype)
    nfail = np.zeros(B, dtype=np.int64)
    for i in range(N):
        astart = a[i]
        zstart = 1.0 / astart - 1.0
        aend = a[i + 1]
        zend = 1.0 / aend - 1.0
        dzda = -1.0 / aend**2
        if i > 0:
            yprev = yarr[:, :, i - 1]
        else:
            yprev = np.stack([np.ones(B), np.ones(B), Tcmb * (1 + zstart), np.zeros(B), np.zeros(B),
                              -Tcmb * (1 + zstart)]).astype(dtype)
        if zend > 3500.0:
            yarr[:, :, i] = np.stack([np.ones(B), np.ones(B), Tcmb * (1 + zend), np.zeros(B), np.zeros(B),
                                      -Tcmb * (1 + zend)])
            continue
        c2 = (i > 0) & (yprev[1].real > 0.99)
        c3 = (~c2) & (i > 0) & (yprev[0].real > 0.99)
        c4 = ~(c2 | c3)
        out = np.zeros((6, B), dtype=dtype)
        if np.any(c2):
            rhs = (np.exp(1.5 * np.log(CR * Tcmb / (1.0 + zend)) - CB1_He1 / (Tcmb * (1.0 + zend))) / Nnow) * 4.0
            x_He0 = 0.5 * (np.sqrt((rhs - 1.0) ** 2 + 4.0 * (1.0 + fHe) * rhs) - (rhs - 1.0))
            u = -((CB1_He1 - CR * Tcmb**2) / (Tcmb * (1 + zend)))
            u2 = -((CB1_He1 - CR * Tcmb**2) / (Tcmb + Tcmb * zend))
            sq = np.sqrt(Nnow**2 + (16 * (-CB1_He1 + CR * Tcmb**2) ** 3) / (Tcmb3 * (1 + zend) ** 3)
                         + 8 * (1 + 2 * fHe) * Nnow * u**1.5)
            sq2 = np.sqrt(Nnow**2 + (16 * (-CB1_He1 + CR * Tcmb**2) ** 3) / (Tcmb**3 * (1 + zend) ** 3)
                          + 8 * (1 + 2 * fHe) * Nnow * u2**1.5)
            dxHeIdz = ((-3 * (-(CB1_He1 / Tcmb) + CR * Tcmb) ** 1.5 * (Nnow + 2 * fHe * Nnow + 4 * u**1.5 - sq))
                       / (Nnow * (1 + zend) ** 2.5 * sq2))
            v2 = np.stack([np.ones(B), (x_He0 - 1.0) / fHe, Tcmb * (1.0 + zend), np.zeros(B), dxHeIdz * dzda,
                           -Tcmb * (1.0 + zend)])
            out = np.where(c2[None], v2, out)
        if np.any(c3 | c4):
            ys, dyda, failed = solve_ionization(np.full(B, astart), np.full(B, aend), yprev[:3], p, rtol=rtol, atol=atol)
            nfail += ((c3 | c4) & failed).astype(np.int64)
            v4 = np.concatenate([ys, dyda])
            if np.any(c3):
                rhs = np.exp(1.5 * np.log(CR * Tcmb / (1.0 + zend)) - CB1 / (Tcmb * (1.0 + zend))) / Nnow
                x_H0 = 0.5 * (np.sqrt(rhs**2 + 4.0 * rhs) - rhs)
                s = np.sqrt(-((CB1 - CR * Tcmb**2) / (Tcmb * (1 + zend))))
                dxHdz = ((3 * ((2 * (-((CB1 - CR * Tcmb**2) / (Tcmb * (1 + zend)))) ** 1.5) / (1 + zend)
                               + ((CB1 - CR * Tcmb**2) * (2 * CB1**2 - 4 * CB1 * CR * Tcmb**2 + Tcmb**2
                                                          * (2 * CR**2 * Tcmb**2 + Nnow * (1 + zend) ** 2 * s)))
                               / (Tcmb**1.5 * (1 + zend) ** 2.5 * np.sqrt((-CB1 + CR * Tcmb**2)
                                  * (CB1**2 - 2 * CB1 * CR * Tcmb**2 + Tcmb**2 * (CR**2 * Tcmb**2 + Nnow * (1 + zend) ** 2 * s))))))
                         / (2.0 * Nnow))
                v3 = v4.copy()
                v3[0] = x_H0
                v3[3] = dxHdz * dzda
                out = np.where(c3[None], v3, out)
            out = np.where(c4[None], v4, out)
        yarr[:, :, i] = out
    return yarr, a[1:], nfail


def evaluate_thermo(*, param, num_thermo=256):
    """x_e, c_s^2, T_m and conformal time on the thermal-history grid (lines 455-501)."""
    from .background import romb, dtauda
    p = param
    p["fHe"] = p["YHe"] / (const_mHe_mH * (1.0 - p["YHe"]))
    y, a, nfail = compute_thermal_history(a0=p["amin"], a1=p["amax"], N=num_thermo, param=p)
    p["recfast_failures"] = nfail
    B = p["H0"].shape[0]
    xeHI = y[0]
    xeHeI = y[1]
    aa = np.broadcast_to(a, (B, a.size))
    pb = {k: (v[:, None] if isinstance(v, np.ndarray) and v.ndim == 1 else v) for k, v in p.items()}
    xeHeII = D.val(saha_HeII(aa, pb))
    xe = xeHI + pb["fHe"] * xeHeI + xeHeII
    mu = 1 / (1 + (1 / const_mHe_mH - 1) * pb["YHe"] + (1 - pb["YHe"]) * xe)
    Tm = y[2]
    dTmda = y[5]
    daTmda = Tm + a * dTmda
    cs2 = const_kB / const_mH / const_c**2 / mu * Tm * (4 - daTmda / Tm) / 3
    seg_lo = np.broadcast_to(a[:-1], (B, a.size - 1))
    seg_hi = np.broadcast_to(a[1:], (B, a.size - 1))
    dtau = romb(lambda x: dtauda(x, p), seg_lo, seg_hi)
    tau0 = p["taumin"]
    tau = np.concatenate([tau0[:, None], tau0[:, None] + np.cumsum(dtau, axis=-1)], axis=-1)
    return tau, a, cs2, Tm, mu, xe, xeHI, xeHeI, xeHeII
