"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on CPU."""
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("disco-eb_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built libdiscoeb_hip.so")


@pytest.fixture(scope="session")
def test_param():
    """Background tables of the reference test cosmology (one cosmology)."""
    from discoeb_amd.background import evolve_background
    from discoeb_amd.cosmology import test_cosmology_param
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        with np.errstate(all="ignore"):
            return evolve_background(param=test_cosmology_param())


@pytest.fixture(scope="session")
def class_data():
    d = json.load(open(os.path.join(GOLDEN, "CLASS_data.json")))
    return np.array(d["k"]), np.array(d["Pkbc"])


@pytest.fixture(scope="session")
def recfast_data():
    d = json.load(open(os.path.join(GOLDEN, "RECFAST_DISCO_EB_data.json")))
    return np.array(d["a"]), np.array(d["xe"])


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no ROCm GPU visible")
    from discoeb_amd import _lib
    _lib.lib()
    return torch.device("cuda:0")


@pytest.fixture(scope="session")
def grad_inputs():
    """w0waCDM + massive-nu base cosmology with its input tangents along the six
    theta directions that enter the perturbation solve (complex step)."""
    from discoeb_amd import tangents
    from discoeb_amd.cosmology import base_theta
    import warnings
    th = base_theta()[None]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        param, dcos, dtab, dtau = tangents.input_tangents(th, [0.5, 1.0])
    return dict(theta=th, param=param, dcos=dcos, dtab=dtab, dtau=dtau, aexp_out=[0.5, 1.0])
