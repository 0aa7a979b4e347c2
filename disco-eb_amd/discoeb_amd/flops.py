"""Algorithmic FP64 operation counts of the structured Rodas5 step (the
roofline numerator of the benchmark).

Counts are per mode and per step for the algorithm the kernels execute, one
operation per active state row (the n rows of the reference's state vector,
perturbations.py:739) -- not per lane: padding lanes and the wave-uniform
background arithmetic every lane repeats are not counted.  n is the number of
state entries, L the number of cyclic-reduction levels of the longest chain.

  right-hand side   per row the coefficient formulas (~10), the stencil
                    f = sub y[i-1] + diag y + sup y[i+1] + u_h h' + u_eta eta'
                    + u_pi Pi (11) and the four functionals h', eta', Pi, f_1
                    (8); plus the background at one scale factor (~150)
  factorisation     a dual (value + d/da) right-hand side for J and df/da,
                    the cyclic-reduction sweep (per level and row 2 divisions
                    and 10 multiply-adds), three solves for the Woodbury
                    columns and their 3x3 capacitance
  solve             a-row, the sweep (5 per level and row), the rank-3
                    update and the a*h' row
  stage combination sum_j a_ij k_j and sum_j C_ij k_j of the 8 stages
"""
from __future__ import annotations

import math

F_BACKGROUND = 150.0


def n_state(cfg) -> int:
    return 12 + cfg.lmaxg + cfg.lmaxgp + cfg.lmaxr + cfg.nqmax * (cfg.lmaxnu + 1)


def pcr_levels(cfg) -> int:
    maxlen = max(cfg.lmaxg + 2, cfg.lmaxgp + 1, cfg.lmaxr + 1, cfg.lmaxnu + 1 if cfg.nqmax else 0)
    return int(math.ceil(math.log2(maxlen)))


def parts(cfg) -> dict:
    n, L = n_state(cfg), pcr_levels(cfg)
    rhs = 29.0 * n + F_BACKGROUND
    solve = (5.0 * L + 15.0) * n
    factor = 2.0 * rhs + 14.0 * L * n + 3.0 * solve + 24.0 * n
    comb = 2.0 * 2.0 * sum(range(8)) * n
    return dict(n=n, levels=L, rhs=rhs, solve=solve, factor=factor, comb=comb)


def flops_per_step(cfg) -> float:
    """One attempted step: 8 right-hand sides, the factorisation, 8 solves,
    the stage combinations and the error norm."""
    p = parts(cfg)
    return 8 * p["rhs"] + p["factor"] + 8 * p["solve"] + p["comb"] + 20.0 * 6


def flops_per_reverse_step(cfg) -> float:
    """One accepted step of the reverse launch: the step again from its
    segment's checkpoint (stage vectors kept), the factorisation once more,
    then per stage a transposed solve, the transposed Jacobian at the stage
    point (one right-hand side's work), the first-order parameter contraction
    through the background features (two right-hand sides' worth in
    three-component duals) and the stage cotangent updates; once per step the
    a-column second derivative (a dual transposed right-hand side and a
    nested-dual feature contraction)."""
    p = parts(cfg)
    n = p["n"]
    d3 = 3.0 * 3.0 * p["rhs"]
    per_stage = p["solve"] + p["rhs"] + 2.0 * d3 + 12.0 * n
    return flops_per_step(cfg) + p["factor"] + 8.0 * per_stage + 2.0 * p["rhs"] + 3.0 * d3


def dense_flops_per_step(cfg) -> float:
    """The reference's dense step for comparison: LU of an n x n matrix, 8
    solves, a Jacobian by n directional derivatives and 9 right-hand sides
    (ode_integrators_stiff.py:718-834 with jax.jacfwd)."""
    n = n_state(cfg)
    rhs = 29.0 * n + F_BACKGROUND
    return (2.0 / 3.0) * n**3 + 8 * 2.0 * n**2 + n * rhs + 9 * rhs
