#!/usr/bin/env python3
"""Benchmark of the MI355X MFM hydro pass: particle-neighbour interactions/s.

Workload (one GPU): 128^3 = 2,097,152 particles, jittered lattice in a periodic
unit box, rho = 1, v = 0, u = 1e-6, total energy 1 deposited kernel-weighted
into the 64 central particles (blast wave); gamma 5/3, 32 neighbours, every
particle active.  One step = one full plain-FP64 MFM pass on data resident in
HBM: hsml iteration from the stored initial guess (density), gradients, force.

Interactions per pass I = sum over density sweeps of |{j : r_ij < h_i}| +
|pair set| of the gradient loop + |pair set| of the force loop, counted by the
kernels (equal to the CPU oracle's count: the neighbour sets are bit-exact).

Prints one JSON line.  Also reported: the roofline of the dominant kernel
(algorithmic bytes / its mean HIP-event duration vs 8 TB/s) and the FP64 CPU
oracle (OpenMP, this host's cores) timed on the same workload, whose outputs
double as a parity check of the timed pass.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gizmo-public_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published

# algorithmic bytes per particle (FP64 records), read + written:
#   density sweep: Pos, VelPred, Mass, Hsml read (64) + h, NumNgb, Dhsml, DivVel,
#                  rho, P, NV_T(9), CN, FCE written (136)
#   gradients:     Pos, h, m, rho, P, v of j (80) + NV_T(6) of i (48) read, 15 gradients written (120)
#   force:         Pos, v, h, m, rho, P, CN, FCE, Dhsml, NV_T(6), gradients(15) read (272),
#                  dv/dt(3), du/dt, v_sig written (40)
BYTES_PER_PARTICLE = {"density": 64 + 136, "gradients": 128 + 120, "force": 272 + 40}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=128, help="lattice side (default 128 -> 2,097,152 particles)")
    ap.add_argument("--cpu-repeats", type=int, default=3, help="oracle passes for the CPU baseline (0 = skip)")
    ap.add_argument("--cell-size", type=float, default=0.0)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="instead of the pass, copy a buffer of this many MiB three times and exit: the "
                         "known-size reference that calibrates the FETCH_SIZE / WRITE_SIZE counters")
    return ap.parse_args()


def copy_probe(mib):
    """Device-to-device copies of exactly `mib` MiB (read mib, write mib each)."""
    import torch
    n = mib * (1 << 20) // 8
    src = torch.arange(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    ok = bool(torch.equal(dst[-16:], src[-16:]))
    print(json.dumps({"copy_probe_mib": mib, "bytes_read_per_copy": n * 8, "bytes_written_per_copy": n * 8,
                      "copies": 3, "ok": ok}))


def cpu_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        n = min(n, int(env))
    return max(1, n)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe_mib > 0:
        copy_probe(args.copy_probe_mib)
        return

    import gzm
    import ics

    ic = ics.blast(args.n)
    N = ic["pos"].shape[0]
    params = gzm.make_params(box=tuple(ic["box"]), cell_size=args.cell_size)
    ctx = gzm.MFMContext(params)
    ctx.upload(ic)

    def one_pass():
        ctx.restore_initial_state()
        s1 = ctx.density()
        s2 = ctx.gradients()
        s3 = ctx.hydro_force()
        return s1, s2, s3

    for _ in range(args.warmup):
        one_pass()
    ctx.timing_reset()
    interactions = []
    sweeps = []
    updates = []
    t0 = time.perf_counter()
    for _ in range(args.steps):  # each stage call returns after its kernels finished
        s1, s2, s3 = one_pass()
        interactions.append(s1.density_interactions + s2.gradient_interactions + s3.force_interactions)
        sweeps.append(s1.density_sweeps)
        updates.append(s1.density_updates)
    t1 = time.perf_counter()
    elapsed = t1 - t0
    tm = ctx.timing()
    out = ctx.download()
    last = (s1, s2, s3)
    ctx.close()

    ms_per_step = 1e3 * elapsed / args.steps
    i_pass = interactions[-1]
    value = float(np.sum(interactions)) / elapsed

    # roofline of the dominant kernel (largest total device time)
    kern = {
        "density": (tm["density_ms"], tm["density_launches"], float(np.mean(updates))),
        "gradients": (tm["gradients_ms"], tm["gradients_launches"], float(N)),
        "force": (tm["force_ms"], tm["force_launches"], float(N)),
    }
    dom = max(kern, key=lambda k: kern[k][0])
    tot_ms, launches, units = kern[dom]
    mean_ms = tot_ms / max(launches, 1)
    bytes_per_launch = BYTES_PER_PARTICLE[dom] * units / max(launches / args.steps, 1)
    achieved = bytes_per_launch / (mean_ms * 1e-3) / 1e9

    result = {
        "metric": "particle-neighbor interactions/s, MFM pass; FP32-batched vs FP64 at stated error",
        "value": value,
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"blast wave, {args.n}^3 jittered lattice, plain FP64 MFM pass (density+gradients+force)",
            "n_particles": N,
            "active": "all",
            "gamma": 5.0 / 3.0,
            "des_num_ngb": 32,
            "interactions_per_pass": int(i_pass),
            "density_sweeps": int(sweeps[-1]),
            "parallelism": "single GPU",
        },
        "kernels_ms_per_pass": {
            "density_sweeps": tm["density_ms"] / args.steps,
            "gradients": tm["gradients_ms"] / args.steps,
            "force": tm["force_ms"] / args.steps,
            "other": tm["other_ms"] / args.steps,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "bytes_per_launch": bytes_per_launch,
            "mean_launch_ms": mean_ms,
        },
    }

    # CPU baseline: the FP64 oracle (OpenMP) on this host, same workload
    if args.cpu_repeats > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle as O
        from parity import compare_pass

        cores = cpu_cores()
        prm = O.make_params(box=tuple(ic["box"]), nthreads=cores)
        times = []
        ref = st = None
        for _ in range(args.cpu_repeats):
            t = time.perf_counter()
            ref, st, _ = O.run_pass(ic, params=prm)
            times.append(time.perf_counter() - t)
        t_med = statistics.median(times)
        cpu_int = st.dens_interactions + st.grad_interactions + st.force_interactions
        result["cpu_baseline"] = {
            "value": cpu_int / t_med,
            "unit": "interactions/s",
            "cores": cores,
            "kind": "port",
            "sample": f"full {args.n}^3 blast pass, median of {args.cpu_repeats} (s: "
                      + ", ".join(f"{x:.2f}" for x in times) + ")",
        }
        errs = compare_pass(out, ref)
        result["parity_vs_oracle"] = {k: float(f"{v:.3e}") for k, v in errs.items()}
        result["parity_vs_oracle"]["interactions_equal"] = bool(cpu_int == i_pass)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
