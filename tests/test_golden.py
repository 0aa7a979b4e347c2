"""The committed golden vectors (tests/golden/golden_<config>.json, written by
the oracle through tests/golden/make_golden.py) stay reproducible: the oracle
recomputes the two lowest-k modes of every cosmology of every configuration at
rtol 1e-5 to rounding, and the stored P(k) follow from the stored delta_m by
get_power (perturbations.py:1101-1123)."""
import json
import os

import numpy as np
import pytest

import oracle
from discoeb_amd import flops, packing
from discoeb_amd.background import evolve_background
from discoeb_amd.cosmology import param_from_theta

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = sorted(f[7:-5] for f in os.listdir(GOLDEN) if f.startswith("golden_") and f.endswith(".json"))


def _load(name):
    with open(os.path.join(GOLDEN, f"golden_{name}.json")) as f:
        return json.load(f)


def test_golden_set_is_complete():
    assert NAMES == ["c1-lcdm-lmax16", "c2-mnu-w0wa-lmax32", "c3-box-lmax32", "c4-grad-base", "c5-box-lmax50"]


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden(name):
    g = _load(name)
    with np.errstate(all="ignore"):
        param = evolve_background(param=param_from_theta(np.array(g["thetas"]), g["massive_nu"]))
    cfg = packing.make_config(rtol=g["loose"]["rtol"], atol=g["loose"]["atol"], max_steps=2_000_000, **g["layout"])
    k = np.asarray(g["k"])
    for b, c in enumerate(g["cosmologies"]):
        prob = oracle.Problem(param, cfg, b=b)
        r = prob.evolve(k[:2], np.asarray(g["tau_out"][b]), nthreads=2)
        ref = c["loose"]
        np.testing.assert_allclose(r["y"], np.asarray(ref["y"])[:2], rtol=1e-12, atol=0)
        assert r["nsteps"].tolist() == ref["nsteps"][:2]
        np.testing.assert_allclose(r["tau_start"], ref["tau_start"][:2], rtol=1e-14)
        th = np.asarray(g["thetas"][b])
        P = 2 * np.pi**2 * np.exp(th[4]) / 1e10 * (k / 0.05) ** (th[3] - 1) * k ** (-3) \
            * np.asarray(ref["y"])[:, -1, 4] ** 2
        np.testing.assert_allclose(P, c["P"], rtol=1e-14)


def test_gradient_fixture_tangent_agrees_with_differences():
    """The c4 fixture's two derivatives of the rtol 1e-10 P(k): the oracle's
    forward-mode tangent and central differences of the whole pipeline (tables
    included, relative step 1e-4).  They agree to 1e-5 below k = 0.01 /Mpc and
    to 5e-3 at k = 0.1 /Mpc (omega_b there 3.0e-3: the differences' own
    truncation through the recombination tables)."""
    g = _load("c4-grad-base")
    c = g["cosmologies"][0]
    k = np.asarray(g["k"])[np.asarray(c["tight_kidx"])]
    tan = np.asarray(c["dP_dtheta_tangent"])
    fd = np.asarray(c["dP_dtheta_fd"])
    assert tan.shape == fd.shape == (k.size, len(c["fd_dirs"]))
    rel = np.abs(tan / fd - 1)
    assert rel[k < 0.01].max() < 1e-5
    assert rel.max() < 5e-3


def test_flop_model_scales_with_the_layout():
    """The roofline numerator grows linearly with the state size at a fixed
    number of cyclic-reduction levels, the reverse step costs several forward
    steps, and the structured step is far below the dense LU step."""
    small = packing.make_config(lmaxg=32, lmaxgp=32, lmaxr=32, lmaxnu=17, nqmax=3)
    big = packing.make_config(lmaxg=32, lmaxgp=32, lmaxr=32, lmaxnu=17, nqmax=5)
    fs, fb = flops.flops_per_step(small), flops.flops_per_step(big)
    ns, nb = flops.n_state(small), flops.n_state(big)
    assert flops.pcr_levels(small) == flops.pcr_levels(big) == 6
    slope = (fb - fs) / (nb - ns)
    assert abs(fs - (slope * ns + (fb - slope * nb))) < 1e-6 * fs
    assert 3.0 < flops.flops_per_reverse_step(big) / fb < 12.0
    assert flops.dense_flops_per_step(big) > 20 * fb
