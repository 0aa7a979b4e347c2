"""Pins of the oracle functions that the reference's fixtures do not reach.

* determine_starting_time (perturbations.py:630-681, root_find_bisect
  util.py:365-396): against a dense scan of both start conditions on the host
  tables (numpy background, an independent restatement);
* convert_to_output_variables (:374-523): each field against something
  independent -- the raw state entries it reads (index map), the metric
  functionals of the right-hand side (:252-260), the evolved a*h' accumulator,
  the conformal Hubble rate from the a(tau) spline's derivative, and the
  massive-nu weight of delta_m / theta_m from the non-relativistic closed form;
* evolve_modes_batched (:786-922): the group-minimum start time and the group
  error norm (one RMS over every mode's filtered entries, :690-694) from the
  single-mode step hook;
* the step-size controller's treatment of non-finite errors (diffrax: a NaN
  error counts as inf; entries outside rms_norm_filtered's filter are never read).
"""
import numpy as np
import pytest

import oracle
from discoeb_amd import packing
from discoeb_amd.background import evolve_background, get_aprimeoa
from discoeb_amd.cosmology import base_theta, param_from_theta, sample_thetas


def _cfg(**kw):
    base = dict(lmaxg=11, lmaxgp=11, lmaxr=11, lmaxnu=8, nqmax=3, rtol=1e-4, atol=1e-8, max_steps=20000)
    base.update(kw)
    return packing.make_config(**base)


@pytest.fixture(scope="module")
def box_param():
    with np.errstate(all="ignore"):
        return evolve_background(param=param_from_theta(sample_thetas(3)))


# ---------------------------------------------------------------------------
# start time

def _conditions(param, b, k, lt):
    """The two start conditions of determine_starting_time at ln tau = lt,
    evaluated with the host table producer's splines and get_aprimeoa."""
    tau = np.exp(lt)[None, :]
    sl = lambda key: _slice(param[key], b)  # noqa: E731
    a = sl("a_of_tau_spline").evaluate(tau)
    xe = sl("xe_of_tau_spline").evaluate(tau)
    pb = {kk: (v[b:b + 1] if isinstance(v, np.ndarray) and v.ndim == 1 else v) for kk, v in param.items()
          if not hasattr(v, "evaluate")}
    pb["logrhonu_of_loga_spline"] = sl("logrhonu_of_loga_spline")
    aH = get_aprimeoa(pb, a)
    akthom = 2.3048e-9 * (1.0 - pb["YHe"][0]) * pb["Omegab"][0] * pb["H0"][0] ** 2
    opac = xe * akthom / a**2
    large = (1.0 / aH) * k / 0.07 - 1.0
    small = (1.0 / opac) / (1.0 / aH) / 0.0004 - 1.0
    return large[0], small[0]


def _slice(sp, b):
    import copy
    s = copy.copy(sp)
    s.x, s.y, s.S = sp.x[b:b + 1], sp.y[b:b + 1], sp.S[b:b + 1]
    return s


def _bisect_by_scan(f, xl, xr, per=64):
    """Where root_find_bisect's 7 halvings end, from a dense scan: the
    midpoint of the 2^-7 sub-bracket holding the first sign change relative to
    f(xl) (the last sub-bracket when there is none).  None when the change
    lies within one scan point of a sub-bracket edge (too close to call)."""
    n = 128
    g = xl + (xr - xl) * np.arange(n * per + 1) / (n * per)
    v = f(g)
    same = v * v[0] > 0
    if np.all(same[1:]):
        j = n - 1
    else:
        i = int(np.argmin(same[1:])) + 1  # first grid point with the other sign
        if (i - 1) % per in (0, per - 1):
            return None
        j = (i - 1) // per
    return xl + (j + 0.5) * (xr - xl) / n


@pytest.mark.parametrize("which", ["test", "box0", "box2"])
def test_start_time_matches_dense_scan(test_param, box_param, which):
    param, b = (test_param, 0) if which == "test" else (box_param, int(which[3:]))
    prob = oracle.Problem(param, _cfg(), b=b)
    tau0 = float(param["taumin"][b])
    tau1 = float(_slice(param["tau_of_a_spline"], b).evaluate(np.array([[0.1]]))[0, 0])
    xl, xr = np.log(tau0), np.log(tau1)
    checked = 0
    for k in np.geomspace(1e-5, 10.0, 24):
        lt_large = _bisect_by_scan(lambda lt: _conditions(param, b, k, lt)[0], xl, xr)
        lt_small = _bisect_by_scan(lambda lt: _conditions(param, b, k, lt)[1], xl, xr)
        if lt_large is None or lt_small is None:
            continue
        want = 0.99 * np.exp(min(lt_small, lt_large))
        ts, _ = prob.ics(k, 1e30)
        assert abs(ts / want - 1) < 1e-12, (k, ts, want)
        # tau_start = 0.99 min(min tau_out, tau*) (perturbations.py:743)
        ts2, _ = prob.ics(k, 0.5 * want)
        assert ts2 == 0.99 * 0.5 * want
        checked += 1
    assert checked >= 20


# ---------------------------------------------------------------------------
# output fields

LAYOUT = dict(lmaxg=12, lmaxgp=9, lmaxr=14, lmaxnu=7, nqmax=4)


def _full_states(param, aexp, kk, rtol=1e-10, **layout):
    cfg = _cfg(rtol=rtol, atol=1e-14, max_steps=400000, return_full=True, **layout)
    prob = oracle.Problem(param, cfg)
    tau = packing.tau_of_aexp(param, aexp)[0]
    r = prob.evolve(np.asarray(kk, dtype=np.float64), tau, nthreads=len(kk))
    assert np.all(r["status"] == 0)
    return prob, tau, r["y"]


def test_output_fields_read_the_right_state_entries(test_param):
    """Raw fields equal the state entries at their reference indices
    (perturbations.py:115-167 layout) on an uneven hierarchy layout, for a
    random state (a wrong index in the conversion shows as a mismatch)."""
    prob = oracle.Problem(test_param, _cfg(**LAYOUT))
    n = prob.n
    lg, lgp, lr = LAYOUT["lmaxg"], LAYOUT["lmaxgp"], LAYOUT["lmaxr"]
    rng = np.random.default_rng(1)
    y = rng.standard_normal(n)
    y[0] = 0.37
    o = prob.outputs(0.2, y)
    assert o[0] == y[2]              # eta
    assert o[8] == y[3]              # delta_c
    assert o[10] == y[5]             # delta_b
    assert o[12] == y[7]             # delta_gamma
    assert o[14] == y[9 + lg + lgp]  # delta_r
    assert o[18] == y[n - 2]         # delta_Q
    # theta fields are the state entries over the same aH
    aH = o[9] and y[4] / o[9]
    for f, i in ((11, 6), (13, 8), (15, 10 + lg + lgp), (19, n - 1)):
        np.testing.assert_allclose(y[i] / o[f], aH, rtol=1e-14)


@pytest.mark.parametrize("k", [1e-3, 0.05, 1.0])
def test_metric_fields_match_the_rhs_functionals(test_param, k):
    """eta' (field 1), h' (field 2) and alpha (field 3) against the functionals
    the right-hand side evolves with (perturbations.py:252-260): f[2] = eta',
    f[3] = -theta_c - h'/2."""
    prob, tau, ys = _full_states(test_param, [0.01, 0.3, 1.0], [k], rtol=1e-8, **LAYOUT)
    for j, t in enumerate(tau):
        y = ys[0, j]
        o = prob.outputs(k, y)
        f = prob.rhs(k, t, y)
        etap, hp = f[2], -2.0 * (f[3] + y[4])
        assert abs(o[1] - etap) <= 1e-12 * abs(etap), (j, o[1], etap)
        assert abs(o[2] - hp) <= 1e-12 * abs(hp), (j, o[2], hp)
        assert abs(o[3] - (hp + 6 * etap) / (2 * k * k)) <= 1e-11 * abs(o[3])


@pytest.fixture(scope="module")
def lcdm_param():
    with np.errstate(all="ignore"):
        return evolve_background(param=param_from_theta(base_theta(w0=-1.0, wa=0.0)[None], massive_nu=False))


@pytest.mark.parametrize("cosmo,tol", [("test", 2e-7), ("lcdm-nonu", 1e-10)])
@pytest.mark.parametrize("k", [1e-3, 0.02, 0.5])
def test_hprime_matches_the_evolved_accumulator(test_param, lcdm_param, cosmo, tol, k):
    """(a h')' = -(delta rho + 3 delta p) a is the a*h' row of the right-hand
    side (:252); its integral from y[1](t0) = 0 equals a h' - a0 h'(t0).  The
    constraint h' (field 2) must therefore equal y[1] / a up to the start
    offset a0 h'(t0) / a, taken from the field at t0.  Checked at a = 1: y[1]
    is not error-controlled (outside the filtered norm), so earlier its
    integration error can reach 1e-5.  With massive nu the identity holds to
    ~1e-7 only: the background density table integrates the nu momenta with 8
    nodes, the perturbations with nqmax (the reference does the same)."""
    param, lay = (test_param, {}) if cosmo == "test" else (lcdm_param, dict(nqmax=0))
    prob, tau, ys = _full_states(param, [1.0], [k], **lay)
    ts, y0 = prob.ics(k, tau[0])
    off = y0[0] * prob.outputs(k, y0)[2]
    y = ys[0, 0]
    hp = prob.outputs(k, y)[2]
    acc = (y[1] + off) / y[0]
    assert abs(acc / hp - 1) < tol, (acc, hp)


def test_theta_fields_use_the_conformal_hubble_rate(test_param):
    """theta/aH fields with aH = (da/dtau)/a from the derivative of the a(tau)
    spline (an independent route to the Friedmann rate the conversion uses).
    Interior scale factors only: a = 1 is the spline's last knot, where its
    natural end condition (zero curvature) distorts the derivative by ~2%."""
    kk = [0.01, 0.3]
    prob, tau, ys = _full_states(test_param, [1e-3, 0.1, 0.6], kk, rtol=1e-8, **LAYOUT)
    sp = test_param["a_of_tau_spline"]
    lg, lgp = LAYOUT["lmaxg"], LAYOUT["lmaxgp"]
    for i, k in enumerate(kk):
        for j, t in enumerate(tau):
            y = ys[i, j]
            o = prob.outputs(k, y)
            a = sp.evaluate(np.array([[t]]))[0, 0]
            aH = sp.derivative(np.array([[t]]))[0, 0] / a
            for f, idx in ((11, 6), (13, 8), (15, 10 + lg + lgp)):
                assert abs(o[f] * aH / y[idx] - 1) < 2e-4, (k, j, f)


def _rhonu_nonrel(a, amnu):
    """rho_nu of one massive flavour over one massless flavour in the
    non-relativistic limit: (180 zeta(3) / (7 pi^4)) a amnu."""
    zeta3 = 1.2020569031595942
    return 180.0 * zeta3 / (7.0 * np.pi**4) * a * amnu


@pytest.mark.parametrize("k", [0.01, 1.0])
def test_total_matter_fields_weight_neutrinos_by_their_density(test_param, k):
    """delta_m = (1 - f_nu) delta_bc + f_nu delta_nu and
    3 theta_m,raw = (1 - f_nu) theta_cb + f_nu theta_nu, with
    f_nu = rho_nu / (rho_cb + rho_nu) from the non-relativistic closed form of
    rho_nu (a amnu >> 1 at a >= 0.5; the closed form holds to ~1e-4 there)."""
    prob, tau, ys = _full_states(test_param, [0.5, 1.0], [k], rtol=1e-8)
    c = prob.cosmo
    F = {name: c[i] for i, name in enumerate(packing.COSMO_FIELDS)}
    assert F["Nmnu"] == 1.0
    for j in range(len(tau)):
        y = ys[0, j]
        o = prob.outputs(k, y)
        a = y[0]
        rcb = F["grhom"] * F["Omegam"] / a
        rnu = F["grhor"] * F["Nmnu"] * _rhonu_nonrel(a, F["amnu"]) / a**2
        fnu = rnu / (rcb + rnu)
        assert 1e-3 < fnu < 2e-2
        dm = (1 - fnu) * o[6] + fnu * o[16]
        assert abs(o[4] - dm) <= 1e-3 * fnu * abs(o[16] - o[6]) + 1e-13 * abs(o[4]), (j, o[4], dm)
        aH = y[6] / o[11]
        alpha_k2 = o[3] * k * k
        tcb = (F["Omegam"] - F["Omegab"]) / F["Omegam"] * y[4] + F["Omegab"] / F["Omegam"] * y[6]
        tnu = o[17] * aH
        tm_raw = o[5] * aH - alpha_k2
        want = ((1 - fnu) * tcb + fnu * tnu) / 3.0
        assert abs(tm_raw - want) <= 2e-3 * fnu * abs(tnu - tcb) + 1e-10 * abs(tm_raw), (j, tm_raw, want)


def test_total_matter_fields_without_massive_neutrinos():
    """No massive nu (Nmnu = 0): delta_m = delta_bc = (Omega_c delta_c +
    Omega_b delta_b) / Omega_m, theta_m,raw = theta_cb / 3 and the reference's
    theta_bc,raw = a^2 theta_cb / 3 (perturbations.py:490-502), both shifted by
    alpha k^2 (:505-506)."""
    with np.errstate(all="ignore"):
        p = evolve_background(param=param_from_theta(base_theta(w0=-1.0, wa=0.0)[None], massive_nu=False))
    k = 0.1
    prob, tau, ys = _full_states(p, [0.2, 1.0], [k], rtol=1e-8, lmaxg=16, lmaxgp=16, lmaxr=16, lmaxnu=8, nqmax=0)
    F = {name: prob.cosmo[i] for i, name in enumerate(packing.COSMO_FIELDS)}
    oc, ob, om = F["Omegam"] - F["Omegab"], F["Omegab"], F["Omegam"]
    for j in range(len(tau)):
        y = ys[0, j]
        o = prob.outputs(k, y)
        a = y[0]
        dbc = (oc * y[3] + ob * y[5]) / om
        assert abs(o[6] / dbc - 1) < 1e-14 and abs(o[4] / dbc - 1) < 1e-14
        aH = y[6] / o[11]
        tcb = (oc * y[4] + ob * y[6]) / om
        ak2 = o[3] * k * k
        assert abs((o[5] * aH - ak2) - tcb / 3.0) <= 1e-10 * abs(ak2)
        assert abs((o[7] * aH - ak2) - a * a * tcb / 3.0) <= 1e-10 * abs(ak2)
        assert o[18] == 0.0 and o[19] == 0.0  # Lambda: no dark-energy perturbations


# ---------------------------------------------------------------------------
# batched groups

def test_batched_group_start_time_and_norm(test_param):
    """The first attempted step of a group of two modes (evolve_modes_batched):
    it starts at the minimum of the modes' start times
    (determine_start_times_per_batch :786-803) with dt0 = min(t0/4, (t1-t0)/2),
    and the controller sees one RMS over both modes' filtered weighted errors
    (rms_norm_filtered_batched :690-694) -- not the larger norm, not the mean."""
    prob = oracle.Problem(test_param, _cfg())
    kk = np.array([2e-3, 0.4])
    tau = packing.tau_of_aexp(test_param, [1.0])[0]
    ns, tr = prob.batch_trace(kk, tau, max_trace=4)
    assert ns > 4
    starts = [prob.ics(k, tau[0])[0] for k in kk]
    t0 = min(starts)
    assert starts[0] != starts[1]
    assert tr[0, 0] == t0
    assert tr[0, 1] == min(t0 + min(t0 / 4.0, 0.5 * (tau[0] - t0)), tau[0])
    norms = []
    for k in kk:
        y0 = prob.ics_at(k, t0)
        y1, err = prob.rodas5_step(k, tr[0, 0], tr[0, 1], y0)
        norms.append(prob.step_control(k, err, y0, y1)[0])
    norms = np.array(norms)
    rms = np.sqrt(np.mean(norms**2))
    assert abs(tr[0, 2] / rms - 1) < 1e-13, (tr[0, 2], norms)
    assert abs(norms.max() / rms - 1) > 1e-3 and abs(norms.mean() / rms - 1) > 1e-3
    # the second attempt follows the controller's factor for that norm
    nrm, keep, factor, _ = prob.step_control(kk[0], np.zeros(prob.n), np.ones(prob.n), np.ones(prob.n))
    start2 = tr[0, 1] if tr[0, 2] < 1 else tr[0, 0]
    fac = np.clip(0.9 * tr[0, 2] ** (-(0.25 + 0.8) / 5.0), 1.0 if tr[0, 2] < 1 else 0.3, 20.0)
    assert tr[1, 0] == start2
    np.testing.assert_allclose(tr[1, 1] - tr[1, 0], (tr[0, 1] - tr[0, 0]) * fac, rtol=1e-12)


# ---------------------------------------------------------------------------
# non-finite errors and failed modes

def test_step_control_nonfinite_semantics(test_param):
    prob = oracle.Problem(test_param, _cfg(rtol=1e-5, atol=1e-8))
    n = prob.n
    rng = np.random.default_rng(5)
    y0 = rng.standard_normal(n)
    y1 = y0 * 1.01
    err = 1e-7 * rng.standard_normal(n)
    base = prob.step_control(0.1, err, y0, y1)
    assert np.isfinite(base[0]) and base[1]
    # entries outside the filter (a*h', the hierarchies) are never read
    for i in (1, 4, 12, n - 1):
        e2 = err.copy(); e2[i] = np.inf
        assert prob.step_control(0.1, e2, y0, y1) == base
        e2[i] = np.nan
        assert prob.step_control(0.1, e2, y0, y1) == base
        y2 = y1.copy(); y2[i] = np.nan
        assert prob.step_control(0.1, err, y0, y2) == base
    # a non-finite filtered error rejects the step and shrinks it by factormin
    for i in (0, 2, 3, 5, 6, 7):
        for bad in (np.nan, np.inf):
            e2 = err.copy(); e2[i] = bad
            nrm, keep, factor, inv = prob.step_control(0.1, e2, y0, y1)
            assert nrm == np.inf and not keep and factor == 0.3 and inv == 1.0


def test_nonfinite_tables_end_the_mode_with_inf_outputs(test_param):
    """A NaN knot in the opacity table poisons every right-hand side past it:
    the steps are rejected until the mode ends (NONFINITE or MAX_STEPS, both
    failures) and the outputs it never reached hold inf."""
    prob = oracle.Problem(test_param, _cfg(max_steps=3000))
    off, nn = prob.desc[packing.SPLINE_KEYS.index("xe_of_loga_spline")]
    prob.tables[off + nn + nn // 2] = np.nan  # a y knot mid-table
    tau = packing.tau_of_aexp(test_param, [1e-4, 1.0])[0]
    r = prob.evolve(np.array([0.05]), tau)
    assert r["status"][0] in (packing.STATUS_MAX_STEPS, packing.STATUS_NONFINITE)
    assert np.all(np.isinf(r["y"][0, -1]))
