"""Benchmark of the perturbation hot path on one MI355X.

Workload: 64 cosmologies drawn with seed 20261202 from the (omega_c, omega_b,
h, n_s, ln10As, m_nu, w0, wa) box x 256 k log-spaced in [1e-5, 10] /Mpc,
photon / polarization / massless-nu lmax 32, one massive nu (lmax 17, 5
momentum bins), Rodas5 at rtol 1e-5, atol 1e-8, P(k) at a = 1.

A "step" returns P(k) and dP(k)/dtheta (all 8 parameters) for all 64
cosmologies x 256 k:
  --arm adjoint (default) one taped forward launch over all 16384 modes, then
                one reverse launch (batched adjoint); six directions come from
                the reverse launch, n_s and ln10As analytically
  --arm fd      one launch over the 64 base cosmologies and their +-1e-4
                relative perturbations of omega_c, omega_b, h, m_nu, w0, wa
                (13 solves per cosmology)
  --arm forward P(k) only
The "plain" baseline of the metric (one cosmology per launch, forward + central
differences, 13 launches per cosmology) is timed on a sample of cosmologies
beside the adjoint arm and reported under "plain_baseline".
Inputs (background tables and their theta-tangents) are built and copied to
HBM before timing.  Prints one JSON line.
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
PEAK_FP64 = 78.6  # TFLOP/s, MI355X dense FP64 vector


def n_state(cfg) -> int:
    return 12 + cfg.lmaxg + cfg.lmaxgp + cfg.lmaxr + cfg.nqmax * (cfg.lmaxnu + 1)


def pcr_levels(cfg) -> int:
    maxlen = max(cfg.lmaxg + 2, cfg.lmaxgp + 1, cfg.lmaxr + 1, cfg.lmaxnu + 1 if cfg.nqmax else 0)
    return int(np.ceil(np.log2(maxlen)))


def flops_per_step(cfg) -> float:
    """FP64 flops of one attempted Rodas5 step of one mode in the structured
    algorithm as executed (n state entries, L PCR levels): 8 right-hand sides
    and one dual (value + d/da) right-hand side for J, the PCR factorisation
    and Woodbury set-up, 8 structured solves and the stage combinations."""
    n, L = n_state(cfg), pcr_levels(cfg)
    f_bg = 150.0                       # background at one scale factor
    f_rhs = 14.0 * n + f_bg            # per-row stencil + metric terms + 4 functionals
    f_solve = (5.0 * L + 15) * n       # a-row, PCR sweep, Woodbury update, a*h' row
    f_factor = 2 * f_rhs + 3 * n + 14.0 * L * n + 3 * f_solve + 24 * n
    f_comb = 2 * 2 * sum(range(8)) * n
    return 8 * f_rhs + f_factor + 8 * f_solve + f_comb + 20 * 6


def flops_per_reverse_step(cfg) -> float:
    """FP64 flops of the reverse of one accepted step: the recomputed forward
    step, 8 transposed solves, per stage one transposed right-hand side and two
    right-hand-side contractions in 8-component dual arithmetic (stage point and
    J(t0, y0) k_i), and once per step a dual transposed right-hand side and a
    16-component (nested dual) contraction for the a-column of J."""
    n = n_state(cfg)
    f_bg = 150.0
    f_rhs = 14.0 * n + f_bg
    f_solve = (5.0 * pcr_levels(cfg) + 15) * n
    f_d8 = 3.0 * 8 * f_rhs             # dn<7>: ~3 flops per component per rhs flop
    f_d16 = 3.0 * 16 * f_rhs
    per_stage = f_solve + f_rhs + 2 * f_d8 + 12 * n
    return flops_per_step(cfg) + 8 * per_stage + 2 * f_rhs + f_d16


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_inputs(theta, cfg, dev, nk, with_tangents):
    import torch
    from discoeb_amd import packing, tangents
    from discoeb_amd.background import evolve_background
    from discoeb_amd.cosmology import param_from_theta
    from discoeb_amd.gradient import Tangents
    from discoeb_amd.perturbations import DeviceInputs
    with np.errstate(all="ignore"):
        if with_tangents:
            param, dcos, dtab, dtau = tangents.input_tangents(theta, [1.0])
        else:
            param = evolve_background(param=param_from_theta(theta))
    B = theta.shape[0]
    kmodes = np.geomspace(1e-5, 10.0, nk)
    inp = DeviceInputs(param, cfg, dev)
    k_t = torch.from_numpy(np.ascontiguousarray(np.broadcast_to(kmodes, (B, nk)))).to(dev)
    tau_t = torch.from_numpy(np.ascontiguousarray(packing.tau_of_aexp(param, [1.0]).real)).to(dev)
    tan = Tangents(dcos, dtab, dtau, dev) if with_tangents else None
    return param, kmodes, inp, k_t, tau_t, tan


def prefactor(theta_t, kk):
    import torch
    A_s = torch.exp(theta_t[:, 4:5]) / 1e10
    n_s = theta_t[:, 3:4]
    return 2 * np.pi**2 * A_s * (kk / 0.05) ** (n_s - 1) * kk ** (-3)


def timed(fn, stream):
    import torch
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    r = fn()
    e1.record(stream)
    return r, (e0, e1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--arm", choices=["adjoint", "fd", "forward"], default="adjoint")
    ap.add_argument("--n-cosmo", type=int, default=64)
    ap.add_argument("--num-k", type=int, default=256)
    ap.add_argument("--plain-sample", type=int, default=2, help="cosmologies timed one per launch (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-modes", type=int, default=6)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    from discoeb_amd import gradient, packing
    from discoeb_amd.cosmology import sample_thetas, SEED
    from discoeb_amd.perturbations import solve_forward

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    kw = dict(lmaxg=32, lmaxgp=32, lmaxr=32, lmaxnu=17, nqmax=5, rtol=1e-5, atol=1e-8, max_steps=32768)
    cfg = packing.make_config(**kw)
    nc, nk = args.n_cosmo, args.num_k
    th0 = sample_thetas(nc, SEED)
    t_setup = time.time()
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
    param, kmodes, inp, k_t, tau_t, tan = make_inputs(theta, cfg, dev, nk, args.arm == "adjoint")
    B = theta.shape[0]
    th_t = torch.from_numpy(theta).to(dev)
    kk = torch.from_numpy(kmodes).to(dev)[None, :]
    pref = prefactor(th_t, kk)
    dth = torch.from_numpy(np.stack([th0[:, j] * FD_REL for j in FD_PARAMS], axis=1)).to(dev)
    setup_s = time.time() - t_setup
    progress(f"setup done ({setup_s:.1f} s): {B} solves x {nk} k")
    last = {}
    n_states = [nc * nk * 2048]

    def step():
        if args.arm == "adjoint":
            (y, st, ns, tape), ev_f = timed(lambda: gradient.solve_forward_tape(inp, k_t, tau_t, n_states=n_states[0]),
                                            stream)
            dm = y[:, :, 0, 4]
            P = pref * dm**2
            ybar = torch.zeros_like(y)
            ybar[:, :, 0, 4] = 2.0 * pref * dm
            g, ev_a = timed(lambda: gradient.solve_adjoint(inp, tan, k_t, tau_t, ybar, st, tape), stream)
            G = torch.empty((nc, nk, 8), dtype=torch.float64, device=dev)
            for j, d in enumerate(FD_PARAMS):
                G[:, :, d] = g[:, :, j]
            G[:, :, 3] = torch.log(kk / 0.05) * P
            G[:, :, 4] = P
            last.update(P=P, G=G, st=st, ns=ns, ev={"forward_tape": ev_f, "adjoint": ev_a}, tape=tape.nbytes)
        else:
            (y, st, ns), ev = timed(lambda: solve_forward(inp, k_t, tau_t, throw=False), stream)
            dm = y[:, :, 0, 4]
            P = pref * dm**2
            if args.arm == "fd":
                P0 = P[:nc]
                grads = [(P[nc * (1 + 2 * i):nc * (2 + 2 * i)] - P[nc * (2 + 2 * i):nc * (3 + 2 * i)]) /
                         (2 * dth[:, i:i + 1]) for i in range(len(FD_PARAMS))]
                G = torch.stack(grads[:3] + [torch.log(kk / 0.05) * P0, P0] + grads[3:], dim=-1)
                last.update(P=P0, G=G)
            else:
                last.update(P=P)
            last.update(st=st, ns=ns, ev={"forward": ev})

    for i in range(args.warmup):
        tw = time.perf_counter()
        step()
        torch.cuda.synchronize()
        progress(f"warmup step {i + 1}/{args.warmup} done ({time.perf_counter() - tw:.1f} s)")
        if args.arm == "adjoint":
            n_states[0] = int(last["ns"].sum().item()) + nc * nk  # exact bound for the timed steps
    kern = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        torch.cuda.synchronize()
        progress(f"timed step {i + 1}/{args.steps} done")
        for name, (e0, e1) in last["ev"].items():
            kern.setdefault(name, []).append(e0.elapsed_time(e1))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    st = last["st"].cpu().numpy()
    ns = last["ns"].cpu().numpy()
    n_fail = int((st != 0).sum())
    if n_fail:
        # a mode that stopped early did less work than the workload asks for:
        # no throughput is reported for such a run
        vals, cnt = np.unique(st[st != 0], return_counts=True)
        sys.exit(f"bench: {n_fail} of {st.size} modes ended with a nonzero status "
                 f"({dict(zip(vals.tolist(), cnt.tolist()))}); no result reported")
    total_steps = float(ns.sum())
    kern_ms = {k: float(np.mean(v)) for k, v in kern.items()}
    value = nc / (ms_per_step * 1e-3)
    if args.arm == "adjoint":
        # the tape holds every accepted state: the reverse launch processes
        # (states - modes) steps; attempted steps bound the forward work
        dom = "adjoint"
        fl = total_steps * flops_per_reverse_step(cfg)
        kname = "deb_adjoint_kernel"
    else:
        dom = "forward"
        fl = total_steps * flops_per_step(cfg)
        kname = "deb_forward_kernel"
    achieved = fl / (kern_ms[dom] * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "(P,dP/dtheta)/s" if args.arm != "forward" else "P(k)/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"batched-{nc}cosmo-{nk}k", "n_cosmo": nc, "num_k": nk, "lmax": 32, "lmaxnu": 17,
                   "nqmax": 5, "rtol": 1e-5, "atol": 1e-8, "aexp_out": [1.0], "arm": args.arm,
                   "gradient": {"adjoint": "batched reverse launch (6 directions) + analytic n_s, ln10As",
                                "fd": "central differences, 13 solves per cosmology in one launch",
                                "forward": None}[args.arm]},
        "roofline": {"bound": "fp64-vector", "achieved": achieved, "peak": PEAK_FP64, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64, "traffic": None, "kernel": kname,
                     "kernel_ms": kern_ms[dom], "flops_per_launch": fl},
        "kernel_ms": kern_ms, "modes": int(ns.size), "mode_steps": total_steps, "failed_modes": n_fail,
        "k_modes_per_s": ns.size / (sum(kern_ms.values()) * 1e-3), "setup_s": setup_s,
    }
    if args.arm == "adjoint":
        out["tape_bytes"] = last["tape"]
        if args.plain_sample > 0:
            progress("timing the plain baseline")
            out["plain_baseline"] = plain_baseline(cfg, th0[:args.plain_sample], nk, dev, value)
    if not args.no_cpu_baseline:
        progress("timing the CPU baseline")
        out["cpu_baseline"] = cpu_baseline(cfg, th0[:1], nk, args.cpu_sample_modes, args.arm)
    print(json.dumps(out))


def plain_baseline(cfg, thetas, nk, dev, adjoint_value):
    """One cosmology per launch: forward + central differences of the six
    solve parameters (13 launches per cosmology), timed on a sample."""
    import torch
    from discoeb_amd.perturbations import solve_forward
    per = []
    for th in thetas:
        ths = [th[None]]
        for j in FD_PARAMS:
            for sgn in (+1.0, -1.0):
                t = th[None].copy()
                t[0, j] *= (1.0 + sgn * FD_REL)
                ths.append(t)
        prepared = [make_inputs(t, cfg, dev, nk, False) for t in ths]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        outs = [solve_forward(inp, k_t, tau_t, throw=False) for (_, _, inp, k_t, tau_t, _) in prepared]
        torch.cuda.synchronize()
        per.append(time.perf_counter() - t0)
        n_fail = sum(int((st != 0).sum().item()) for (_, st, _) in outs)
        if n_fail:
            sys.exit(f"bench: plain baseline: {n_fail} modes ended with a nonzero status; no result reported")
    sec = float(np.mean(per))
    return {"value": 1.0 / sec, "unit": "(P,dP/dtheta)/s", "launches_per_cosmology": 13,
            "sample": f"{len(thetas)} cosmologies, each 13 launches of 1 cosmology x {nk} k",
            "batched_adjoint_over_plain": adjoint_value * sec}


def cpu_baseline(cfg, theta, nk, nsample, arm):
    """The oracle (C++ restatement of the reference algorithm: dense LU Rodas5;
    gradients by forward-mode tangents through the frozen steps, like
    jax.jacfwd) on this host's cores, on a bounded sample of modes of
    cosmology 0.  The six tangent directions cost the same; one is timed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from discoeb_amd import packing, tangents
    cores = os.cpu_count() or 1
    kmodes = np.geomspace(1e-5, 10.0, nk)
    sel = np.unique(np.linspace(0, nk - 1, nsample).astype(int))
    with np.errstate(all="ignore"):
        param, dcos, dtab, dtau = tangents.input_tangents(theta, [1.0])
    prob = oracle.Problem(param, cfg, b=0)
    tau_out = packing.tau_of_aexp(param, [1.0])[0].real
    thr = min(cores, sel.size)
    t = time.perf_counter()
    prob.evolve(kmodes[sel], tau_out, nthreads=thr)
    t_fwd = time.perf_counter() - t
    if arm == "forward":
        wall, what = t_fwd, "forward solve"
    else:
        t = time.perf_counter()
        prob.evolve_tangent(kmodes[sel], tau_out, dtau[0, 0], dcos[0, 0], dtab[0, 0], nthreads=thr)
        t_tan = time.perf_counter() - t  # a forward solve plus one tangent replay
        wall = t_fwd + len(tangents.SOLVE_DIRS) * max(t_tan - t_fwd, 0.0)
        what = f"forward {t_fwd:.1f}s + 6 x tangent replay {max(t_tan - t_fwd, 0.0):.1f}s"
    per_cosmo = wall * thr / sel.size * nk / cores
    return {"value": 1.0 / per_cosmo, "unit": "(P,dP/dtheta)/s" if arm != "forward" else "P(k)/s",
            "cores": cores, "kind": "port",
            "sample": f"{sel.size} of {nk} k (log-spaced subset) of cosmology 0 on {thr} threads: {what}; "
                      f"scaled to {nk} k on {cores} cores"}


if __name__ == "__main__":
    main()
