"""Input tangents of the perturbation solve with respect to theta.

The gradient kernels take, per cosmology, the directional derivatives of
their inputs (the deb_cosmo_t record, every spline knot and the output times)
along each requested theta direction.  They are computed here by the complex
step (Squire & Trapp 1998) through the table producer
(``evolve_background``, the reference's background.py:140-342 and
thermodynamics_recfast.py): one complex evaluation per direction, exact to
rounding because no difference is taken.  In the reference the same
derivatives come from jax.jacfwd through ``evolve_background``
(nb_minimal_example.ipynb:297).

Of theta = (omega_c, omega_b, h, n_s, ln10As, m_nu, w0, wa), n_s and ln10As
enter only the primordial spectrum of ``get_power``
(perturbations.py:1101-1123); their derivatives are analytic and the
perturbation solve is differentiated along the other six.
"""
from __future__ import annotations

import numpy as np

from . import packing
from .background import evolve_background
from .cosmology import param_from_theta

SOLVE_DIRS = (0, 1, 2, 5, 6, 7)
ANALYTIC_DIRS = (3, 4)
CSTEP = 1e-20


def input_tangents(theta, aexp_out, dirs=SOLVE_DIRS, massive_nu: bool = True, param=None):
    """Tables and their theta-tangents for cosmologies theta[B, 8].

    Returns (param, dcosmo[B, nd, 24], dtables[B, nd, stride], dtau_out[B, nd, nout])
    where ``param`` is the real table set (as ``evolve_background`` returns it).
    """
    theta = np.atleast_2d(np.asarray(theta, dtype=np.float64))
    B, nd = theta.shape[0], len(dirs)
    aexp_out = np.atleast_1d(np.asarray(aexp_out, dtype=np.float64))
    with np.errstate(all="ignore"):
        if param is None:
            param = evolve_background(param=param_from_theta(theta, massive_nu))
        thc = np.repeat(theta[:, None, :], nd, axis=1).astype(np.complex128)
        for j, d in enumerate(dirs):
            thc[:, j, d] += 1j * CSTEP
        pc = evolve_background(param=param_from_theta(thc.reshape(B * nd, 8), massive_nu))
    dcos = packing.pack_cosmo(pc, part=np.imag).reshape(B, nd, -1) / CSTEP
    dtab = packing.pack_tables(pc, part=np.imag)[0].reshape(B, nd, -1) / CSTEP
    dtau = packing.tau_of_aexp(pc, aexp_out).imag.reshape(B, nd, aexp_out.size) / CSTEP
    return param, np.ascontiguousarray(dcos), np.ascontiguousarray(dtab), np.ascontiguousarray(dtau)
