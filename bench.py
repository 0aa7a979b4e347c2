"""Benchmark of the perturbation hot path on one MI355X.

Workload (default): 64 cosmologies drawn with seed 20261202 from the
(omega_c, omega_b, h, n_s, ln10As, m_nu, w0, wa) box x 256 k log-spaced in
[1e-5, 10] /Mpc, photon / polarization / massless-nu lmax 32, one massive nu
(lmax 17, 5 momentum bins), Rodas5 at rtol 1e-5, atol 1e-8, P(k) at a = 1.

A "step" returns P(k) and dP(k)/dtheta for all 64 cosmologies:
  --arm fd      (default) one persistent launch over the base cosmologies and
                their +-1e-4 relative perturbations of omega_c, omega_b, h,
                m_nu, w0, wa (13 solves per cosmology); n_s and ln10As are
                analytic (dP/dn_s = ln(k/k_p) P, dP/dln10As = P)
  --arm forward P(k) only (64 x 256 modes)
Inputs (background tables) are built and copied to HBM before timing.
Prints one JSON line (see the repository's measurement notes).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "disco-eb_amd"))

METRIC = "(P(k), dP/dtheta) per sec, 64 cosmo x 256 k, rtol 1e-5: batched adjoint vs plain"
FD_PARAMS = (0, 1, 2, 5, 6, 7)  # omega_c, omega_b, h, m_nu, w0, wa
FD_REL = 1e-4


def flops_per_step(cfg) -> float:
    """FP64 flops of one attempted Rodas5 step of one mode in the structured
    algorithm as executed (n = state size, L = PCR levels); see DESIGN.md."""
    n = 12 + cfg.lmaxg + cfg.lmaxgp + cfg.lmaxr + cfg.nqmax * (cfg.lmaxnu + 1)
    maxlen = max(cfg.lmaxg + 2, cfg.lmaxgp + 1, cfg.lmaxr + 1, cfg.lmaxnu + 1 if cfg.nqmax else 0)
    L = int(np.ceil(np.log2(maxlen)))
    f_bg = 150.0                       # background at one scale factor (uniform)
    f_rhs = 14.0 * n + f_bg            # per-row stencil + metric terms + 4 functionals
    f_factor = 2 * f_rhs + 3 * n + 14.0 * L * n + 3 * (5.0 * L + 15) * n + 24 * n
    f_solve = (5.0 * L + 15) * n
    f_comb = 2 * 2 * (sum(range(8))) * n  # stage combinations u_i and rhs_i
    return 8 * f_rhs + f_factor + 8 * f_solve + f_comb + 20 * 6


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--arm", choices=["fd", "forward"], default="fd")
    ap.add_argument("--n-cosmo", type=int, default=64)
    ap.add_argument("--num-k", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-modes", type=int, default=16)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    from discoeb_amd import packing
    from discoeb_amd.background import evolve_background
    from discoeb_amd.cosmology import param_from_theta, sample_thetas, SEED
    from discoeb_amd.perturbations import DeviceInputs, solve_forward

    dev = torch.device("cuda:0")
    kw = dict(lmaxg=32, lmaxgp=32, lmaxr=32, lmaxnu=17, nqmax=5, rtol=1e-5, atol=1e-8, max_steps=32768)
    cfg = packing.make_config(**kw)
    nc, nk = args.n_cosmo, args.num_k
    th0 = sample_thetas(nc, SEED)
    if args.arm == "fd":
        ths = [th0]
        for j in FD_PARAMS:
            for sgn in (+1.0, -1.0):
                t = th0.copy()
                t[:, j] = t[:, j] * (1.0 + sgn * FD_REL)
                ths.append(t)
        theta = np.concatenate(ths)
    else:
        theta = th0
    t_setup = time.time()
    with np.errstate(all="ignore"):
        param = evolve_background(param=param_from_theta(theta))
    kmodes = np.geomspace(1e-5, 10.0, nk)
    tau_out = packing.tau_of_aexp(param, [1.0]).real
    inp = DeviceInputs(param, cfg, dev)
    B = theta.shape[0]
    k_t = torch.from_numpy(np.ascontiguousarray(np.broadcast_to(kmodes, (B, nk)))).to(dev)
    tau_t = torch.from_numpy(np.ascontiguousarray(tau_out)).to(dev)
    A_s = torch.from_numpy(np.exp(theta[:, 4]) / 1e10).to(dev)[:, None]
    n_s = torch.from_numpy(theta[:, 3]).to(dev)[:, None]
    kk = torch.from_numpy(kmodes).to(dev)[None, :]
    dth = torch.from_numpy(np.stack([th0[:, j] * FD_REL for j in FD_PARAMS], axis=1)).to(dev)
    setup_s = time.time() - t_setup

    kern_ms = []
    last = {}

    def step():
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record()
        y, st, ns = solve_forward(inp, k_t, tau_t, throw=False)
        ev1.record()
        dm = y[:, :, 0, 4]
        P = 2 * np.pi**2 * A_s * (kk / 0.05) ** (n_s - 1) * kk ** (-3) * dm**2
        if args.arm == "fd":
            P0 = P[:nc]
            grads = [(P[nc * (1 + 2 * i):nc * (2 + 2 * i)] - P[nc * (2 + 2 * i):nc * (3 + 2 * i)]) / (2 * dth[:, i:i + 1])
                     for i in range(len(FD_PARAMS))]
            G = torch.stack(grads[:3] + [torch.log(kk / 0.05) * P0, P0] + grads[3:], dim=-1)  # theta order
            last["P"], last["G"] = P0, G
        else:
            last["P"] = P
        last["st"], last["ns"], last["ev"] = st, ns, (ev0, ev1)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    kern_ms.clear()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        torch.cuda.synchronize()
        kern_ms.append(last["ev"][0].elapsed_time(last["ev"][1]))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    st = last["st"].cpu().numpy()
    ns = last["ns"].cpu().numpy()
    n_fail = int((st != 0).sum())
    total_steps = float(ns.sum())
    fl = total_steps * flops_per_step(cfg)
    kmean = float(np.mean(kern_ms))
    achieved = fl / (kmean * 1e-3) / 1e12
    value = nc / (ms_per_step * 1e-3)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "(P,dP/dtheta)/s" if args.arm == "fd" else "P(k)/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"batched-{nc}cosmo-{nk}k", "n_cosmo": nc, "num_k": nk, "lmax": 32, "lmaxnu": 17,
                   "nqmax": 5, "rtol": 1e-5, "atol": 1e-8, "aexp_out": [1.0], "arm": args.arm,
                   "solves_per_cosmology": B // nc, "gradient": "central FD (6 params) + analytic n_s, ln10As"
                   if args.arm == "fd" else None},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": 78.6, "unit": "TFLOP/s",
                     "frac": achieved / 78.6, "traffic": None, "kernel": "deb_forward_kernel",
                     "kernel_ms": kmean, "flops_per_launch": fl},
        "modes": int(ns.size), "mode_steps": total_steps, "failed_modes": n_fail,
        "k_modes_per_s": ns.size / (kmean * 1e-3), "setup_s": setup_s,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(param, cfg, kmodes, args.cpu_sample_modes, B // nc)
    print(json.dumps(out))


def cpu_baseline(param, cfg, kmodes, nsample, solves_per_cosmo):
    """The oracle (C++ restatement of the reference algorithm: dense LU Rodas5)
    timed on this host's cores on a bounded sample of modes of cosmology 0."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from discoeb_amd import packing
    cores = os.cpu_count() or 1
    sel = np.unique(np.linspace(0, kmodes.size - 1, nsample).astype(int))
    prob = oracle.Problem(param, cfg, b=0)
    tau_out = packing.tau_of_aexp(param, [1.0])[0:1].real[0]
    t = time.perf_counter()
    r = prob.evolve(kmodes[sel], tau_out, nthreads=min(cores, sel.size))
    dt = time.perf_counter() - t
    cpu_secs = dt * min(cores, sel.size)  # core-seconds for the sample
    per_cosmo = cpu_secs / sel.size * kmodes.size * solves_per_cosmo / cores
    return {"value": 1.0 / per_cosmo, "unit": "(P,dP/dtheta)/s" if solves_per_cosmo > 1 else "P(k)/s",
            "cores": cores, "kind": "port",
            "sample": f"{sel.size} of {kmodes.size} k (log-spaced subset) of cosmology 0, "
                      f"{dt:.1f}s wall on {min(cores, sel.size)} threads, scaled to {solves_per_cosmo} solves x "
                      f"{kmodes.size} k on {cores} cores", "sample_steps": int(r["nsteps"].sum())}


if __name__ == "__main__":
    main()
