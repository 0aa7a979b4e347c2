"""Benchmark of the perturbation hot path on one MI355X.

Default workload (batched-64cosmo-256k, the headline metric): 64 cosmologies
drawn with seed 20261202 from the (omega_c, omega_b, h, n_s, ln10As, m_nu, w0,
wa) box x 256 k log-spaced in [1e-5, 10] /Mpc, photon / polarization /
massless-nu lmax 32, one massive nu (lmax 17, 5 momentum bins), Rodas5 at
rtol 1e-5, atol 1e-8, P(k) at a = 1.

A "step" returns P(k) and dP(k)/dtheta (all 8 parameters) for all 64
cosmologies x 256 k:
  --arm adjoint (default) one taped forward launch over all 16384 modes
                (checkpoints every 64 accepted steps), then one reverse launch
                (batched adjoint); six directions come from the reverse
                launch, n_s and ln10As analytically
  --arm fd      one launch over the 64 base cosmologies and their +-1e-4
                relative perturbations of omega_c, omega_b, h, m_nu, w0, wa
                (13 solves per cosmology)
  --arm forward P(k) only
The "plain" baseline of the metric (one cosmology per launch, forward + central
differences, 13 launches per cosmology) is timed on a sample of cosmologies
beside the adjoint arm and reported under "plain_baseline".

Other workloads (one JSON line each, measured on request):
  --workload batched-adjoint-256cosmo-512k  256 seeded cosmologies x 512 k in
        [1e-5, 20], lmax 50: loss sum_k w(k) ln P(k) with seeded w(k) ~ U[0, 1]
        and its gradient (checkpoints every 64 steps, 64 GiB device budget)
  --workload plain-lcdm-64k          one flat LCDM cosmology (Lambda, no
        massive nu), lmax 16, 64 k in [1e-4, 1]: P(k)
  --workload plain-mnu-256k-16out    the base cosmology with one massive nu,
        w0 -0.9, wa 0.1, lmax 32, 256 k, all fields at 16 a in [0.01, 1]

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
PEAK_FP64 = 78.6  # TFLOP/s, MI355X dense FP64 (vector and matrix)
LMAX32 = dict(lmaxg=32, lmaxgp=32, lmaxr=32, lmaxnu=17, nqmax=5)
WORKLOADS = ("batched-64cosmo-256k", "batched-adjoint-256cosmo-512k", "plain-lcdm-64k", "plain-mnu-256k-16out")


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_inputs(theta, cfg, dev, kmodes, aexp=(1.0,), with_tangents=False, massive_nu=True):
    import torch
    from discoeb_amd import packing, tangents
    from discoeb_amd.background import evolve_background
    from discoeb_amd.cosmology import param_from_theta
    from discoeb_amd.gradient import Tangents
    from discoeb_amd.perturbations import DeviceInputs
    with np.errstate(all="ignore"):
        if with_tangents:
            param, dcos, dtab, dtau = tangents.input_tangents(theta, list(aexp), massive_nu=massive_nu)
        else:
            param = evolve_background(param=param_from_theta(theta, massive_nu))
    B = theta.shape[0]
    inp = DeviceInputs(param, cfg, dev)
    k_t = torch.from_numpy(np.ascontiguousarray(np.broadcast_to(kmodes, (B, kmodes.size)))).to(dev)
    tau_t = torch.from_numpy(np.ascontiguousarray(packing.tau_of_aexp(param, list(aexp)).real)).to(dev)
    tan = Tangents(dcos, dtab, dtau, dev) if with_tangents else None
    return param, inp, k_t, tau_t, tan


def prefactor(theta_t, kk):
    import torch
    A_s = torch.exp(theta_t[:, 4:5]) / 1e10
    n_s = theta_t[:, 3:4]
    return 2 * np.pi**2 * A_s * (kk / 0.05) ** (n_s - 1) * kk ** (-3)


class Timer:
    """HIP events on the launch stream around one launch."""

    def __init__(self, stream):
        import torch
        self.stream = stream
        self.e0 = torch.cuda.Event(enable_timing=True)
        self.e1 = torch.cuda.Event(enable_timing=True)

    def __enter__(self):
        self.e0.record(self.stream)
        return self

    def __exit__(self, *a):
        self.e1.record(self.stream)

    def ms(self):
        return self.e0.elapsed_time(self.e1)


def traffic_record(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE /
    WRITE_SIZE summary (profiles/traffic_<kernel>_<workload>.json), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=WORKLOADS, default=WORKLOADS[0])
    ap.add_argument("--arm", choices=["adjoint", "fd", "forward"], default="adjoint")
    ap.add_argument("--n-cosmo", type=int, default=None)
    ap.add_argument("--num-k", type=int, default=None)
    ap.add_argument("--ckpt-every", type=int, default=64)
    ap.add_argument("--plain-sample", type=int, default=2, help="cosmologies timed one per launch (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-modes", type=int, default=16)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 2 GiB device-to-device (the reference kernel for HBM byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload == "batched-64cosmo-256k":
        run_batched(args)
    elif args.workload == "batched-adjoint-256cosmo-512k":
        run_loss_adjoint(args)
    else:
        run_plain(args)


def copy_calibration():
    """A plain 2 GiB device-to-device copy (2^28 doubles read, 2^28 written):
    what FETCH_SIZE and WRITE_SIZE read for it scales their readings of the
    kernels to bytes."""
    import torch
    x = torch.ones(1 << 28, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_each_way": (1 << 28) * 8}))


def _accepted_per_mode(tape, n_modes):
    """Accepted steps of every mode from the tape's per-mode state counts."""
    import torch
    off = 256 + 8 * n_modes
    cnt = tape.buf[off:off + 4 * n_modes].view(torch.int32)
    return cnt.cpu().numpy().astype(np.int64) - 1


def _accepted(tape, n_modes):
    return int(_accepted_per_mode(tape, n_modes).sum())


def tail_model(cost, n_cosmo, nk, slots, longest_first=False):
    """The launch's tail: items are dispatched in workgroup order to whichever
    of `slots` resident waves frees up first -- item i is k index
    nk-1-i//n_cosmo of cosmology i%n_cosmo (largest k first), or, with
    longest_first (the reverse launch's order), the items sorted by their step
    counts.  With each item's cost proportional to its step count, the
    makespan of that list schedule against the ideal (total / slots) gives the
    fraction of wave-time idle at the end."""
    import heapq
    order = [nk - 1 - i // n_cosmo for i in range(n_cosmo * nk)]
    cosmo = [i % n_cosmo for i in range(n_cosmo * nk)]
    c = np.asarray(cost, dtype=np.float64).reshape(n_cosmo, nk)
    heap = [0.0] * min(slots, n_cosmo * nk)
    heapq.heapify(heap)
    for ki, ci in zip(order, cosmo):
        t = heapq.heappop(heap)
        heapq.heappush(heap, t + c[ci, ki])
    makespan = max(heap)
    ideal = c.sum() / slots
    # the same dispatch with the items sorted by their step counts (longest first)
    heap = [0.0] * min(slots, n_cosmo * nk)
    for v in np.sort(c.ravel())[::-1]:
        t = heapq.heappop(heap)
        heapq.heappush(heap, t + v)
    sorted_makespan = max(heap)
    if longest_first:
        return {"slots": slots, "makespan_steps": float(sorted_makespan), "ideal_steps": float(ideal),
                "idle_fraction": float(1.0 - ideal / sorted_makespan), "longest_item_steps": float(c.max()),
                "largest_k_first_makespan_steps": float(makespan),
                "model": "greedy dispatch of per-mode step counts, longest first"}
    return {"slots": slots, "makespan_steps": float(makespan), "ideal_steps": float(ideal),
            "idle_fraction": float(1.0 - ideal / makespan), "longest_item_steps": float(c.max()),
            "sorted_makespan_steps": float(sorted_makespan),
            "model": "greedy dispatch of per-mode step counts in workgroup order, largest k first"}


def run_batched(args):
    import torch
    from discoeb_amd import flops, gradient, packing
    from discoeb_amd.cosmology import SEED, sample_thetas
    from discoeb_amd.perturbations import solve_forward

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    cfg = packing.make_config(rtol=1e-5, atol=1e-8, max_steps=32768, **LMAX32)
    nc = args.n_cosmo or 64
    nk = args.num_k or 256
    th0 = sample_thetas(nc, SEED)
    kmodes = np.geomspace(1e-5, 10.0, nk)
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
    _, inp, k_t, tau_t, tan = make_inputs(theta, cfg, dev, kmodes, with_tangents=args.arm == "adjoint")
    th_t = torch.from_numpy(theta).to(dev)
    kk = torch.from_numpy(kmodes).to(dev)[None, :]
    pref = prefactor(th_t, kk)
    dth = torch.from_numpy(np.stack([th0[:, j] * FD_REL for j in FD_PARAMS], axis=1)).to(dev)
    setup_s = time.time() - t_setup
    progress(f"setup done ({setup_s:.1f} s): {theta.shape[0]} solves x {nk} k")
    last = {}
    n_states = [None]

    def step():
        if args.arm == "adjoint":
            with Timer(stream) as tf:
                y, st, ns, tape = gradient.solve_forward_tape(inp, k_t, tau_t, n_states=n_states[0],
                                                              ckpt_every=args.ckpt_every)
            dm = y[:, :, 0, 4]
            P = pref * dm**2
            ybar = torch.zeros_like(y)
            ybar[:, :, 0, 4] = 2.0 * pref * dm
            with Timer(stream) as ta:
                g = gradient.solve_adjoint(inp, tan, k_t, tau_t, ybar, st, tape)
            G = torch.empty((nc, nk, 8), dtype=torch.float64, device=dev)
            for j, d in enumerate(FD_PARAMS):
                G[:, :, d] = g[:, :, j]
            G[:, :, 3] = torch.log(kk / 0.05) * P
            G[:, :, 4] = P
            last.update(P=P, G=G, st=st, ns=ns, ev={"forward_tape": tf, "adjoint": ta}, tape=tape)
        else:
            with Timer(stream) as tf:
                y, st, ns = solve_forward(inp, k_t, tau_t, throw=False)
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
            last.update(st=st, ns=ns, ev={"forward": tf})

    for i in range(args.warmup):
        tw = time.perf_counter()
        step()
        torch.cuda.synchronize()
        progress(f"warmup step {i + 1}/{args.warmup} done ({time.perf_counter() - tw:.1f} s)")
        if args.arm == "adjoint":
            n_states[0] = int(last["ns"].sum().item()) + nc * nk  # a sufficient bound for the timed steps
    kern = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        torch.cuda.synchronize()
        progress(f"timed step {i + 1}/{args.steps} done")
        for name, tm in last["ev"].items():
            kern.setdefault(name, []).append(tm.ms())
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
    attempted = float(ns.sum())
    kern_ms = {k: float(np.mean(v)) for k, v in kern.items()}
    value = nc / (ms_per_step * 1e-3)
    if args.arm == "adjoint":
        accepted = _accepted(last["tape"], nc * nk)
        dom, kname, units = "adjoint", "deb_adjoint_kernel", accepted
        per_unit = flops.flops_per_reverse_step(cfg)
        unit_name = "accepted steps (each recomputed from its checkpoint and reversed)"
    else:
        accepted = None
        dom, kname, units = "forward", "deb_forward_kernel", attempted
        per_unit = flops.flops_per_step(cfg)
        unit_name = "attempted steps"
    fl = units * per_unit
    achieved = fl / (kern_ms[dom] * 1e-3) / 1e12
    workload = f"batched-{nc}cosmo-{nk}k"
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "(P,dP/dtheta)/s" if args.arm != "forward" else "P(k)/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": workload, "n_cosmo": nc, "num_k": nk, "lmax": 32, "lmaxnu": 17, "nqmax": 5,
                   "rtol": 1e-5, "atol": 1e-8, "aexp_out": [1.0], "arm": args.arm,
                   "ckpt_every": args.ckpt_every if args.arm == "adjoint" else None,
                   "gradient": {"adjoint": "batched reverse launch (6 directions) + analytic n_s, ln10As",
                                "fd": "central differences, 13 solves per cosmology in one launch",
                                "forward": None}[args.arm]},
        "roofline": roofline(kname, workload, fl, kern_ms[dom], per_unit, units, unit_name),
        "kernel_ms": kern_ms, "modes": int(ns.size), "attempted_steps": attempted, "accepted_steps": accepted,
        "failed_modes": n_fail, "k_modes_per_s": ns.size / (ms_per_step * 1e-3), "setup_s": setup_s,
    }
    # one wave per SIMD at lmax 32 (512 registers): 4 resident waves per CU
    slots = 4 * torch.cuda.get_device_properties(dev).multi_processor_count
    # (the fd arm launches 13 solves per cosmology)
    out["tail"] = {"forward": tail_model(ns, ns.size // nk, nk, slots)}
    if args.arm == "adjoint":
        out["tail"]["adjoint"] = tail_model(_accepted_per_mode(last["tape"], nc * nk), nc, nk, slots,
                                            longest_first=True)
        out["tape_bytes"] = last["tape"].nbytes
        if args.plain_sample > 0:
            progress("timing the plain baseline")
            out["plain_baseline"] = plain_baseline(cfg, th0[:args.plain_sample], kmodes, dev, value)
    if not args.no_cpu_baseline:
        progress("timing the CPU baseline")
        out["cpu_baseline"] = cpu_baseline(cfg, th0[:1], kmodes, args.cpu_sample_modes, args.arm)
    print(json.dumps(out))


def roofline(kname, workload, fl, kernel_ms, per_unit, units, unit_name):
    achieved = fl / (kernel_ms * 1e-3) / 1e12
    r = {"bound": "mfma", "pipe": "fp64 (vector ALU; the MI355X FP64 peak is the same for the matrix cores)",
         "achieved": achieved, "peak": PEAK_FP64, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64, "traffic": None,
         "kernel": kname, "kernel_ms": kernel_ms, "flops_per_launch": fl, "flops_per_unit": per_unit,
         "units_per_launch": units, "work_unit": unit_name}
    tr = traffic_record(kname, workload)
    if tr is not None:
        r["traffic"] = tr["bytes_per_launch"]
        r["traffic_source"] = tr["source"]
    return r


def run_loss_adjoint(args):
    """Config 5: loss = sum_k w(k) ln P(k) per cosmology and its gradient."""
    import torch
    from discoeb_amd import flops, gradient, packing
    from discoeb_amd.cosmology import SEED, loss_weights, sample_thetas

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    cfg = packing.make_config(lmaxg=50, lmaxgp=50, lmaxr=50, lmaxnu=17, nqmax=5, rtol=1e-5, atol=1e-8,
                              max_steps=65536)
    nc = args.n_cosmo or 256
    nk = args.num_k or 512
    th0 = sample_thetas(256, SEED)[:nc]
    wk = torch.from_numpy(loss_weights(512, 256, SEED)[:nk]).to(dev)
    kmodes = np.geomspace(1e-5, 20.0, nk)
    t_setup = time.time()
    _, inp, k_t, tau_t, tan = make_inputs(th0, cfg, dev, kmodes, with_tangents=True)
    th_t = torch.from_numpy(th0).to(dev)
    kk = torch.from_numpy(kmodes).to(dev)[None, :]
    pref = prefactor(th_t, kk)
    setup_s = time.time() - t_setup
    progress(f"setup done ({setup_s:.1f} s): {nc} cosmologies x {nk} k, lmax 50")
    last = {}
    n_states = [None]

    def step():
        with Timer(stream) as tf:
            y, st, ns, tape = gradient.solve_forward_tape(inp, k_t, tau_t, n_states=n_states[0],
                                                          ckpt_every=args.ckpt_every)
        dm = y[:, :, 0, 4]
        loss = (wk[None, :] * torch.log(pref * dm**2)).sum(dim=1)
        ybar = torch.zeros_like(y)
        ybar[:, :, 0, 4] = wk[None, :] * 2.0 / dm  # d(w ln(pref dm^2)) / d dm
        with Timer(stream) as ta:
            g = gradient.solve_adjoint(inp, tan, k_t, tau_t, ybar, st, tape)
        G = torch.empty((nc, 8), dtype=torch.float64, device=dev)
        gs = g.sum(dim=1)
        for j, d in enumerate(FD_PARAMS):
            G[:, d] = gs[:, j]
        G[:, 3] = (wk[None, :] * torch.log(kk / 0.05)).sum(dim=1)
        G[:, 4] = wk.sum()
        last.update(loss=loss, G=G, st=st, ns=ns, tape=tape, ev={"forward_tape": tf, "adjoint": ta})

    for i in range(args.warmup):
        tw = time.perf_counter()
        step()
        torch.cuda.synchronize()
        progress(f"warmup step {i + 1}/{args.warmup} done ({time.perf_counter() - tw:.1f} s)")
        n_states[0] = int(last["ns"].sum().item()) + nc * nk
    kern = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        torch.cuda.synchronize()
        progress(f"timed step {i + 1}/{args.steps} done")
        for name, tm in last["ev"].items():
            kern.setdefault(name, []).append(tm.ms())
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    st = last["st"].cpu().numpy()
    n_fail = int((st != 0).sum())
    if n_fail:
        vals, cnt = np.unique(st[st != 0], return_counts=True)
        sys.exit(f"bench: {n_fail} of {st.size} modes ended with a nonzero status "
                 f"({dict(zip(vals.tolist(), cnt.tolist()))}); no result reported")
    kern_ms = {k: float(np.mean(v)) for k, v in kern.items()}
    accepted = _accepted(last["tape"], nc * nk)
    fl = accepted * flops.flops_per_reverse_step(cfg)
    workload = f"batched-adjoint-{nc}cosmo-{nk}k"
    out = {
        "metric": "(loss, dloss/dtheta) per sec, loss = sum_k w(k) ln P(k), lmax 50, rtol 1e-5",
        "value": nc / (ms_per_step * 1e-3), "unit": "(loss,grad)/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": workload, "n_cosmo": nc, "num_k": nk, "kmax": 20.0, "lmax": 50, "lmaxnu": 17,
                   "nqmax": 5, "rtol": 1e-5, "atol": 1e-8, "ckpt_every": args.ckpt_every,
                   "device_budget_gib": gradient.TAPE_BUDGET / 2**30},
        "roofline": roofline("deb_adjoint_kernel", workload, fl, kern_ms["adjoint"],
                             flops.flops_per_reverse_step(cfg), accepted,
                             "accepted steps (each recomputed from its checkpoint and reversed)"),
        "kernel_ms": kern_ms, "modes": int(st.size), "attempted_steps": float(last["ns"].sum().item()),
        "accepted_steps": accepted, "tape_bytes": last["tape"].nbytes,
        "adjoint_workspace_bytes": gradient.adjoint_workspace_bytes(cfg, nc, nk, args.ckpt_every), "setup_s": setup_s,
        "loss_mean": float(last["loss"].mean().item()),
    }
    print(json.dumps(out))


def run_plain(args):
    """Configs 1 and 2: one cosmology, forward solve."""
    import torch
    from discoeb_amd import flops, packing
    from discoeb_amd.cosmology import base_theta
    from discoeb_amd.perturbations import solve_forward

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    if args.workload == "plain-lcdm-64k":
        cfg = packing.make_config(lmaxg=16, lmaxgp=16, lmaxr=16, lmaxnu=8, nqmax=3, rtol=1e-5, atol=1e-8,
                                  max_steps=32768)
        th = base_theta(w0=-1.0, wa=0.0)[None]
        kmodes, aexp, massive = np.geomspace(1e-4, 1.0, args.num_k or 64), (1.0,), False
    else:
        cfg = packing.make_config(rtol=1e-5, atol=1e-8, max_steps=32768, **LMAX32)
        th = base_theta()[None]
        kmodes, aexp, massive = np.geomspace(1e-5, 10.0, args.num_k or 256), tuple(np.geomspace(0.01, 1, 16)), True
    _, inp, k_t, tau_t, _ = make_inputs(th, cfg, dev, kmodes, aexp=aexp, massive_nu=massive)
    res = {}
    for i in range(args.warmup + args.steps):
        with Timer(stream) as tm:
            y, st, ns = solve_forward(inp, k_t, tau_t, throw=False)
        torch.cuda.synchronize()
        if i >= args.warmup:
            res.setdefault("ms", []).append(tm.ms())
    st = st.cpu().numpy()
    if (st != 0).any():
        sys.exit(f"bench: {int((st != 0).sum())} modes ended with a nonzero status; no result reported")
    ms = float(np.mean(res["ms"]))
    attempted = float(ns.sum().item())
    fl = attempted * flops.flops_per_step(cfg)
    out = {"metric": f"k-modes per sec, {args.workload}", "value": kmodes.size / (ms * 1e-3), "unit": "k-modes/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": args.workload, "num_k": int(kmodes.size), "nout": len(aexp),
                      "lmax": cfg.lmaxg, "nqmax": cfg.nqmax, "massive_nu": massive, "rtol": 1e-5, "atol": 1e-8},
           "p_k_per_s": 1.0 / (ms * 1e-3),
           "roofline": roofline("deb_forward_kernel", args.workload, fl, ms, flops.flops_per_step(cfg), attempted,
                                "attempted steps"),
           "attempted_steps": attempted}
    print(json.dumps(out))


def plain_baseline(cfg, thetas, kmodes, dev, adjoint_value):
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
        prepared = [make_inputs(t, cfg, dev, kmodes) for t in ths]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        outs = [solve_forward(inp, k_t, tau_t, throw=False) for (_, inp, k_t, tau_t, _) in prepared]
        torch.cuda.synchronize()
        per.append(time.perf_counter() - t0)
        n_fail = sum(int((st != 0).sum().item()) for (_, st, _) in outs)
        if n_fail:
            sys.exit(f"bench: plain baseline: {n_fail} modes ended with a nonzero status; no result reported")
    sec = float(np.mean(per))
    return {"value": 1.0 / sec, "unit": "(P,dP/dtheta)/s", "launches_per_cosmology": 13,
            "sample": f"{len(thetas)} cosmologies, each 13 launches of 1 cosmology x {kmodes.size} k",
            "batched_adjoint_over_plain": adjoint_value * sec}


def cpu_baseline(cfg, theta, kmodes, nsample, arm):
    """The oracle (C++ restatement of the reference algorithm: dense LU Rodas5;
    gradients by forward-mode tangents through the frozen steps, like
    jax.jacfwd) on this host's cores, on a bounded sample: nsample of the k,
    evenly spaced in index over the log grid, of cosmology 0, one mode per
    thread.  Each mode's forward solve and its tangent replay are timed in its
    own thread; the six tangent directions cost the same, one is timed.  The
    cosmology's cost is the sample's summed mode time scaled to all k; value =
    cores / cost (and 1 / cost for one thread)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from discoeb_amd import packing, tangents
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    nk = kmodes.size
    sel = np.unique(np.linspace(0, nk - 1, nsample).round().astype(int))
    with np.errstate(all="ignore"):
        param, dcos, dtab, dtau = tangents.input_tangents(theta, [1.0])
    prob = oracle.Problem(param, cfg, b=0)
    tau_out = packing.tau_of_aexp(param, [1.0])[0].real
    thr = min(cores, sel.size)
    t = time.perf_counter()
    rf = prob.evolve(kmodes[sel], tau_out, nthreads=thr)
    wall = time.perf_counter() - t
    fwd = rf["seconds"]
    if arm == "forward":
        per_mode = fwd
        what = f"forward solves: {fwd.sum():.1f} mode-s"
    else:
        t = time.perf_counter()
        rt = prob.evolve_tangent(kmodes[sel], tau_out, dtau[0, 0], dcos[0, 0], dtab[0, 0], nthreads=thr)
        wall += time.perf_counter() - t
        rep = np.maximum(rt["seconds"] - fwd, 0.0)  # a tangent solve is a forward solve plus the replay
        per_mode = fwd + len(tangents.SOLVE_DIRS) * rep
        what = f"forward {fwd.sum():.1f} mode-s + 6 x tangent replay {rep.sum():.1f} mode-s"
    cost = per_mode.sum() * nk / sel.size  # seconds of one core for the whole cosmology
    return {"value": cores / cost, "unit": "(P,dP/dtheta)/s" if arm != "forward" else "P(k)/s",
            "cores": cores, "kind": "port", "value_1_thread": 1.0 / cost,
            "sample": f"{sel.size} of {nk} k (every {max(nk // sel.size, 1)}th of the log grid) of cosmology 0, "
                      f"one mode per thread on {thr} threads ({wall:.1f} s wall): {what}; "
                      f"scaled to {nk} k"}


if __name__ == "__main__":
    main()
