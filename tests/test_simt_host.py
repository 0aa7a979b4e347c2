"""The kernels' per-item device code, run on the CPU under AddressSanitizer and
UndefinedBehaviorSanitizer (tests/simt: 64 host threads per wavefront, every
cross-lane operation through a shared exchange buffer).  Any out-of-bounds
access, use of an invalid pointer or undefined arithmetic in integrate_mode,
the checkpoint tape writer or adjoint_mode aborts the run; a cross-lane
operation reached by part of a wave only would never complete (the run's
time limit catches it).  The results must agree with the oracle at the
tolerance level.  Short integrations only (an emulated step takes about a
second): k = 1e-3 to a = 5e-4 at rtol 1e-4.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
from discoeb_amd import packing
from discoeb_amd.cosmology import base_theta

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SIMT = os.path.join(ROOT, "tests", "simt")
EXE = os.path.join(SIMT, "simt_host")
AOUT = [2e-4, 5e-4]
KK = np.array([1e-3])


@pytest.fixture(scope="module")
def harness():
    r = subprocess.run(["make", "-C", SIMT], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail("building the emulation harness failed:\n" + r.stdout[-3000:] + r.stderr[-3000:])
    return EXE


def _write(path, mode, cfg, prob, k, tau, extra=b""):
    with open(path, "wb") as f:
        f.write(np.int32(mode).tobytes())
        f.write(bytes(cfg))
        f.write(np.ascontiguousarray(prob.cosmo, dtype=np.float64).tobytes())
        f.write(np.int64(prob.tables.size).tobytes())
        f.write(prob.tables.tobytes())
        f.write(np.ascontiguousarray(prob.desc, dtype=np.int32).ravel().tobytes())
        f.write(np.int32(cfg.nqmax).tobytes())
        f.write(prob.q.tobytes())
        f.write(prob.w.tobytes())
        f.write(np.int32(k.size).tobytes())
        f.write(k.tobytes())
        f.write(np.int32(tau.size).tobytes())
        f.write(tau.tobytes())
        f.write(extra)


def _start(exe, tmp, name):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    return subprocess.Popen([exe, str(tmp / f"{name}.in"), str(tmp / f"{name}.out")], env=env,
                            stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)


def _finish(p, timeout):
    try:
        out, err = p.communicate(timeout=timeout)
    except subprocess.TimeoutExpired:
        p.kill()
        p.communicate()
        pytest.fail("emulated run did not finish (a cross-lane operation reached by part of a wave?)")
    assert p.returncode == 0, err[-4000:]


def test_device_code_under_sanitizers_matches_oracle(harness, test_param, grad_inputs, tmp_path):
    runs = {}
    layouts = {"ns2": dict(lmaxg=11, lmaxgp=11, lmaxr=11, lmaxnu=8, nqmax=3),
               "ns4": dict(lmaxg=32, lmaxgp=32, lmaxr=32, lmaxnu=17, nqmax=5)}
    tau = packing.tau_of_aexp(test_param, AOUT)[0]
    for name, lay in layouts.items():
        cfg = packing.make_config(rtol=1e-4, atol=1e-8, max_steps=2000, **lay)
        prob = oracle.Problem(test_param, cfg)
        _write(tmp_path / f"{name}.in", 0, cfg, prob, KK, tau)
        runs[name] = (cfg, prob, _start(harness, tmp_path, name))
    # taped forward + reverse sweep (checkpoints every 5 steps) on the gradient inputs
    g = grad_inputs
    cfg = packing.make_config(rtol=1e-4, atol=1e-8, max_steps=2000, **layouts["ns2"])
    gprob = oracle.Problem(g["param"], cfg)
    gtau = packing.tau_of_aexp(g["param"], AOUT)[0]
    dtau = np.stack([packing.tau_of_aexp(g["param"], AOUT)[0] * 0.0 for _ in range(6)])
    ndir = 6
    ybar = np.zeros((KK.size, len(AOUT), 20))
    ybar[:, :, 4] = 1.0
    extra = (np.int32(5).tobytes() + np.int32(ndir).tobytes() + np.ascontiguousarray(g["dcos"][0]).tobytes()
             + np.ascontiguousarray(g["dtab"][0]).tobytes() + dtau.tobytes() + ybar.tobytes())
    _write(tmp_path / "grad.in", 1, cfg, gprob, KK, gtau, extra)
    pg = _start(harness, tmp_path, "grad")
    for name, (cfg, prob, p) in runs.items():
        _finish(p, 900)
        raw = (tmp_path / f"{name}.out").read_bytes()
        st = np.frombuffer(raw[:4], np.int32)
        ns = np.frombuffer(raw[4:8], np.int32)
        y = np.frombuffer(raw[8:], np.float64).reshape(KK.size, len(AOUT), 20)
        ref = prob.evolve(KK, tau)
        assert st[0] == 0 and ref["status"][0] == 0
        assert abs(int(ns[0]) - int(ref["nsteps"][0])) <= 2
        for f in (4, 6, 8, 10, 12):
            assert abs(y[0, -1, f] / ref["y"][0, -1, f] - 1) < 30 * 1e-4, (name, f)
    _finish(pg, 1500)
    raw = (tmp_path / "grad.out").read_bytes()
    st = np.frombuffer(raw[:4], np.int32)
    y = np.frombuffer(raw[8:8 + 8 * KK.size * len(AOUT) * 20], np.float64).reshape(KK.size, len(AOUT), 20)
    grad = np.frombuffer(raw[8 + 8 * KK.size * len(AOUT) * 20:], np.float64).reshape(KK.size, ndir)
    assert st[0] == 0 and np.all(np.isfinite(grad))
    # the oracle's forward-mode tangent of the same quantity (sum over both
    # outputs of delta_m) along each direction, at the tolerance level
    for j in range(ndir):
        r = gprob.evolve_tangent(KK, gtau, dtau[j], g["dcos"][0, j], g["dtab"][0, j])
        want = r["dy"][0, :, 4].sum()
        scale = np.abs(r["dy"][0, :, 4]).sum() + 1e-30
        assert abs(grad[0, j] - want) <= 3e-3 * scale, (j, grad[0, j], want)


def test_row_coefficient_wave_form_is_bit_identical(harness, test_param, tmp_path):
    """The row coefficients' wave form (row_coef_wave: per-lane selects and the
    fixed-lane rows of slot 0 as wave-uniform values) equals the row switch
    (row_coef) bit for bit on every lane and slot, at seven a in [1e-7, 1],
    with and without the dark-energy fluid, for three layouts (lmax 11 with
    3 bins, lmax 32 with 5 bins, lmax 50)."""
    layouts = {"ns2": dict(lmaxg=11, lmaxgp=11, lmaxr=11, lmaxnu=8, nqmax=3),
               "ns4": dict(lmaxg=32, lmaxgp=32, lmaxr=32, lmaxnu=17, nqmax=5),
               "ns5": dict(lmaxg=50, lmaxgp=50, lmaxr=50, lmaxnu=17, nqmax=5)}
    tau = packing.tau_of_aexp(test_param, AOUT)[0]
    kk = np.array([1e-4, 3e-2, 2.0])
    procs = {}
    for name, lay in layouts.items():
        cfg = packing.make_config(rtol=1e-4, atol=1e-8, max_steps=2000, **lay)
        prob = oracle.Problem(test_param, cfg)
        _write(tmp_path / f"rc_{name}.in", 2, cfg, prob, kk, tau)
        procs[name] = _start(harness, tmp_path, f"rc_{name}")
    for name, p in procs.items():
        _finish(p, 600)
        bad, checked = np.frombuffer((tmp_path / f"rc_{name}.out").read_bytes(), np.int32)
        assert checked > 0 and bad == 0, (name, int(bad), int(checked))
