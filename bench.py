#!/usr/bin/env python3
"""Benchmark of the MI355X MFM hydro pass: particle-neighbour interactions/s.

Default workload (one GPU), the batched configuration: 512 independent
periodic unit boxes, box b a jittered n_b^3 lattice with n_b^3 log-uniform in
[4096, 262144] (31,542,962 particles in all), Gaussian velocities with
sigma ~ U(0.1, 1) and u ~ log-uniform [0.1, 10] per box, gamma 5/3, 32
neighbours, every particle active.  One step = one full MFM pass over all
boxes (hsml iteration from the stored initial guess, gradients, force) on
data resident in HBM; the K timed steps are K passes captured in one HIP graph
and launched once (gzm_run_passes), timed with HIP events around the launch.
`value` is the FP32-bulk pass (precision 1: FP32 gradient and force pair
arithmetic; FP64 hsml iteration, NV_T inversion, limiter, exact Riemann solver
and flux sums), with the FP64 pass's interaction count; precision 2 (FP32
density sums too) and the plain FP64 pass on the same boxes are reported
beside it, with the max / rms error of every FP32 output against the FP64
oracle run on all boxes, and dv/dt and du/dt on identical inputs.

Interactions per pass I = sum over density sweeps of |{j : r_ij < h_i}| +
|pair set| of the gradient loop + |pair set| of the force loop, counted by the
kernels (for FP64 equal to the CPU oracle's count: the neighbour sets are
bit-exact).

Also reported: the roofline of the dominant kernel (algorithmic bytes over its
mean HIP-event duration in the eager warm-up passes of the same workload, vs
8 TB/s, and the HBM bytes the committed counter summary measured for it), the
FP64 CPU oracle (OpenMP, this host's cores) timed on the same workload, and
the 128^3 blast-wave case (--workload blast makes it the headline).
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
# FP32-bulk mode, stored widths: density sweep reads offsets+m (16), velocity (16),
# hsml (8) and writes the same FP64 outputs (136); gradients read the 48-byte
# FP32 record of j (rho and P as high and low words) plus NV_T(6) of i (48) and
# write 15 FP64 gradients (120); force reads the 160-byte FP32 record and writes
# dv/dt, du/dt, v_sig (40)
BYTES_PER_PARTICLE_F32 = {"density": 40 + 136, "gradients": 96 + 120, "force": 160 + 40}


# meshless finite-volume pass: every particle on the integer timestep 2^12
# of a 2^-20 time unit (dt = 1/256), which sets the exchanged masses
MFV_TIMEBASE = 2.0 ** -20
MFV_TI_STEP = 2 ** 12

_T0 = time.perf_counter()


def log(msg):
    """Progress line on stderr (the JSON result is the only stdout line)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


def fp64_kernel(precision, kernel):
    """The stage runs the FP64 kernel: precision 0, and precision 1's density stage."""
    return precision == 0 or (precision == 1 and kernel == "density")


def bytes_per_particle(precision, kernel):
    return (BYTES_PER_PARTICLE if fp64_kernel(precision, kernel) else BYTES_PER_PARTICLE_F32)[kernel]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed passes (default: 100 batched, 10 blast)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=("batch", "blast", "shear", "clustered"), default="batch",
                    help="batch: 512 boxes (default); blast: 128^3 blast wave; shear: 16.8M-particle shear "
                         "layers; clustered: Plummer spheres in a uniform background, tree search")
    ap.add_argument("--cluster-n", type=int, default=8388608, help="particles of the clustered workload")
    ap.add_argument("--spheres", type=int, default=1000, help="Plummer spheres of the clustered workload")
    ap.add_argument("--cpu-sample", type=int, default=262144,
                    help="clustered workload: particles of the CPU-baseline sample")
    ap.add_argument("--list-cap", type=int, default=None,
                    help="pair-list capacity (default 64: a particle with more force pairs is scanned on its "
                         "own, the lists keep the exact-solver pairs out of the loop)")
    ap.add_argument("--boxes", type=int, default=512, help="boxes of the batched workload")
    ap.add_argument("--n", type=int, default=128, help="lattice side of the blast (default 128 -> 2,097,152 particles)")
    ap.add_argument("--cpu-repeats", type=int, default=3,
                    help="timed oracle passes for the CPU baseline, after one warm-up; the median is reported "
                         "(0 = skip the CPU leg)")
    ap.add_argument("--cpu-boxes", type=int, default=48,
                    help="batched workload: boxes of the CPU-baseline sample (the first ones)")
    ap.add_argument("--cell-size", type=float, default=0.0)
    ap.add_argument("--fp32", type=int, default=1, help="blast workload: also time the FP32-bulk mode")
    ap.add_argument("--plain", type=int, default=1,
                    help="batch workload: also time the plain form (FP64, one box per pass, eager API)")
    ap.add_argument("--fp32-density", type=int, default=1,
                    help="batch workload: also time precision 2 (FP32 density pair sums)")
    ap.add_argument("--blast", type=int, default=1, help="batch workload: also time the 128^3 blast case")
    ap.add_argument("--mfv", type=int, default=1,
                    help="batch workload: also time the meshless finite-volume pass (FP64) on the same boxes")
    ap.add_argument("--autotune", type=int, default=1,
                    help="batch workload: pick cell size, pair-list capacity, workgroup sizes and the density "
                         "load-balance thresholds per mode on a sample of boxes")
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


def cpu_model():
    """Model name of the host CPU and the number of logical CPUs the OS shows."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline(O, samples, cores, repeats, what):
    """The FP64 oracle with its octree search (node hsml maxima, the role of
    ngb_treefind_variable/pairs_threads) over the particle sets `samples`
    (one pass each, one after another), on `cores` OpenMP threads: one
    untimed warm-up, then the median of `repeats` timed runs.  Returns the
    cpu_baseline record and the interactions of one run."""
    prm = [O.make_params(box=tuple(s["box"]), nthreads=cores, tree=1) for s in samples]

    def one():
        inter = 0
        for s, p in zip(samples, prm):
            _, st, _ = O.run_pass(s, params=p)
            inter += st.dens_interactions + st.grad_interactions + st.force_interactions
        return inter

    inter = one()
    times = []
    for _ in range(repeats):
        t = time.perf_counter()
        one()
        times.append(time.perf_counter() - t)
    t_med = statistics.median(times)
    model, nproc = cpu_model()
    n = sum(s["pos"].shape[0] for s in samples)
    return {
        "value": inter / t_med,
        "unit": "interactions/s",
        "cores": cores,
        "kind": "port",
        "search": "octree with node hsml maxima (the reference's ngb_treefind_* walk)",
        "sample": f"{what} ({n:,} particles, {inter:,} interactions per pass): median of {repeats} timed "
                  f"passes after one warm-up, s: " + ", ".join(f"{x:.2f}" for x in times),
        "cpu_model": model,
        "nproc": nproc,
    }, inter


def run_mode(gzm, ic, args, precision, search=0, active=None):
    """Time `args.steps` passes (after `args.warmup`) of one precision mode on
    `ic`, resident on the GPU, over all particles or the `active` subset.
    Returns the measurements and the last outputs."""
    N = ic["pos"].shape[0]
    n_act = N if active is None else len(active)
    params = gzm.make_params(box=tuple(ic["box"]), cell_size=args.cell_size, precision=precision, search=search,
                             pair_list_cap=getattr(args, "list_cap", None) or 0)
    ctx = gzm.MFMContext(params)
    ctx.upload(ic)  # first upload allocates
    t_up = time.perf_counter()
    ctx.upload(ic)  # host arrays -> sorted device state (PCIe + sort + grid)
    t_up = time.perf_counter() - t_up

    def one_pass():
        ctx.restore_initial_state()
        s1 = ctx.density(active)
        s2 = ctx.gradients(active)
        s3 = ctx.hydro_force(active)
        return s1, s2, s3

    for _ in range(args.warmup):
        one_pass()
    ctx.timing_reset()
    interactions, sweeps, updates = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):  # each stage call returns after its kernels finished
        s1, s2, s3 = one_pass()
        interactions.append(s1.density_interactions + s2.gradient_interactions + s3.force_interactions)
        sweeps.append(s1.density_sweeps)
        updates.append(s1.density_updates)
    elapsed = time.perf_counter() - t0
    tm = ctx.timing()
    t_down = time.perf_counter()
    out = ctx.download()
    t_down = time.perf_counter() - t_down
    ctx.close()
    # roofline of the dominant kernel (largest total device time)
    kern = {
        "density": (tm["density_ms"], tm["density_launches"], float(np.mean(updates))),
        "gradients": (tm["gradients_ms"], tm["gradients_launches"], float(n_act)),
        "force": (tm["force_ms"], tm["force_launches"], float(n_act)),
    }
    dom = max(kern, key=lambda k: kern[k][0])
    tot_ms, launches, units = kern[dom]
    mean_ms = tot_ms / max(launches, 1)
    bytes_per_launch = bytes_per_particle(precision, dom) * units / max(launches / args.steps, 1)
    return {
        "value": float(np.sum(interactions)) / elapsed,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "interactions_per_pass": int(interactions[-1]),
        "density_sweeps": int(sweeps[-1]),
        "kernels_ms_per_pass": {
            "density_sweeps": tm["density_ms"] / args.steps,
            "density_wave_per_particle": tm["density_wave_ms"] / args.steps,
            "gradients": tm["gradients_ms"] / args.steps,
            "force": tm["force_ms"] / args.steps,
            "force_deferred_pairs": tm["force_deferred_ms"] / args.steps,
            "other": tm["other_ms"] / args.steps,
        },
        "dominant": dom,
        "mean_launch_ms": mean_ms,
        "bytes_per_launch": bytes_per_launch,
        "achieved_gbs": bytes_per_launch / (mean_ms * 1e-3) / 1e9,
        "out": out,
        "upload_ms": 1e3 * t_up,
        "download_ms": 1e3 * t_down,
        # one pass with the host arrays handed over and the results copied back
        "pcie_inclusive_value": float(interactions[-1]) / (t_up + elapsed / args.steps + t_down),
    }


def measured_traffic(kernel, precision, n):
    """HBM bytes per launch of `kernel` from the committed counter summary of
    this workload (FETCH_SIZE and WRITE_SIZE, calibrated on a copy of known
    size), or None when no summary covers it."""
    name = "hbm_traffic_blast128.json" if n == 2097152 else "hbm_traffic_batch512.json"
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            tab = json.load(f)
    except (OSError, ValueError):
        return None
    rec = tab.get("kernels", {}).get(f"{kernel}:{'fp64' if fp64_kernel(precision, kernel) else 'fp32'}")
    if rec is None or tab.get("n_particles") != n:
        return None
    return rec["fetch_bytes"] + rec["write_bytes"]


def measured_valu(kernel, precision, n):
    """VALU busy share and FLOP rates of `kernel` from the committed rocprof
    counter summary of the 512-box workload, or None when it does not cover it."""
    if n != 31542962:
        return None
    try:
        with open(os.path.join(ROOT, "profiles", "valu_batch512.json")) as f:
            tab = json.load(f)
    except (OSError, ValueError):
        return None
    rec = tab.get("kernels", {}).get(f"{kernel}:{'fp64' if fp64_kernel(precision, kernel) else 'fp32'}")
    if rec is None:
        return None
    return {"kernel": kernel, "valu_busy": rec["valu_busy"], "tflops_fp64": rec["tflops_fp64"],
            "tflops_fp32": rec["tflops_fp32"], "peak_tflops": tab["peaks_tflops"],
            "source": "profiles/valu_batch512.json"}


def run_plain_per_box(gzm, ic, tuned=None):
    """The plain form: FP64, one box per pass through the eager API (density,
    gradients, hydro_force, each returning when its kernels finished), as the
    drop-in runs one particle set per call.  Each box gets its own context,
    upload and one untimed warm-up pass; the timed pass starts from the
    restored initial state with the inputs resident in HBM."""
    import ics
    params = dict(precision=0)
    if tuned:
        params.update(tuned["params"])
    nb = len(ic["box_offsets"]) - 1
    total_s, inter = 0.0, 0
    for b in range(nb):
        sub = ics.box_slice(ic, b)
        with gzm.MFMContext(gzm.make_params(box=tuple(sub["box"]), **params)) as ctx:
            ctx.upload(sub)
            for timed in (False, True):
                ctx.restore_initial_state()
                t0 = time.perf_counter()
                s1 = ctx.density()
                s2 = ctx.gradients()
                s3 = ctx.hydro_force()
                dt = time.perf_counter() - t0
            total_s += dt
            inter += s1.density_interactions + s2.gradient_interactions + s3.force_interactions
        if (b + 1) % 128 == 0:
            log(f"plain FP64 per box: {b + 1}/{nb} boxes, {1e3 * total_s:.1f} ms so far")
    return {"value": inter / total_s, "ms_per_pass": 1e3 * total_s, "interactions_per_pass": int(inter),
            "boxes": nb}


def run_graph_mode(gzm, ic, args, precision, steps, tuned=None, method=0):
    """Warm-up eager passes (per-kernel HIP events), then `steps` passes in one
    HIP graph (timed with events around the launch).  method 1: the
    meshless finite-volume pass."""
    N = ic["pos"].shape[0]
    mk = dict(precision=precision, method=method, timebase_interval=MFV_TIMEBASE)
    if tuned:
        params = gzm.make_params(**mk, **tuned["params"])
    else:
        params = gzm.make_params(cell_size=args.cell_size, **mk)
    ctx = gzm.MFMContext(params)
    ctx.upload(ic)
    ctx.timing_reset()
    updates = []
    for _ in range(max(args.warmup, 1)):
        ctx.restore_initial_state()
        s1 = ctx.density()
        ctx.gradients()
        ctx.hydro_force()
        updates.append(s1.density_updates)
    tw = ctx.timing()
    w = max(args.warmup, 1)
    ctx.timing_reset()
    st = ctx.run_passes(steps)
    tm = ctx.timing()
    out = ctx.download()
    ctx.close()
    graph_s = tm["graph_ms"] * 1e-3
    inter = st.density_interactions + st.gradient_interactions + st.force_interactions
    kern = {
        "density": (tw["density_ms"], tw["density_launches"], float(np.mean(updates))),
        "gradients": (tw["gradients_ms"], tw["gradients_launches"], float(N)),
        "force": (tw["force_ms"], tw["force_launches"], float(N)),
    }
    dom = max(kern, key=lambda k: kern[k][0])
    tot_ms, launches, units = kern[dom]
    mean_ms = tot_ms / max(launches, 1)
    bytes_per_launch = bytes_per_particle(precision, dom) * units / max(launches / w, 1)
    return {
        "value": inter / graph_s,
        "ms_per_step": 1e3 * graph_s / steps,
        "interactions_per_pass": inter // steps,
        "density_sweeps": int(st.density_sweeps),
        "kernels_ms_per_pass_eager": {
            "density_sweeps": tw["density_ms"] / w,
            "density_wave_per_particle": tw["density_wave_ms"] / w,
            "gradients": tw["gradients_ms"] / w,
            "force": tw["force_ms"] / w,
            "force_deferred_pairs": tw["force_deferred_ms"] / w,
            "other": tw["other_ms"] / w,
        },
        "dominant": dom,
        "mean_launch_ms": mean_ms,
        "bytes_per_launch": bytes_per_launch,
        "achieved_gbs": bytes_per_launch / (mean_ms * 1e-3) / 1e9,
        "out": out,
        "density_interactions": st.density_interactions // steps,
        "grad_interactions": st.gradient_interactions // steps,
        "force_interactions": st.force_interactions // steps,
        "exact_solver_calls": st.exact_solver_calls // steps,
        "fp64_force_particles": st.fp64_force_particles // steps,
        "density_fp64_redo": int(s1.density_fp64_redo),
        "riemann_retries": st.riemann_retries // steps,
        "precision": precision,
    }


def oracle_boxes(O, ic, cores, method=0):
    """FP64 oracle pass box by box; returns concatenated outputs, interactions, seconds."""
    import ics
    mk = dict(nthreads=cores, method=method, timebase_interval=MFV_TIMEBASE)
    if "box_offsets" not in ic:
        t0 = time.perf_counter()
        ref, st, _ = O.run_pass(ic, params=O.make_params(box=tuple(ic["box"]), **mk))
        return ref, st.dens_interactions + st.grad_interactions + st.force_interactions, time.perf_counter() - t0
    offs = ic["box_offsets"]
    outs, inter, t0 = [], 0, time.perf_counter()
    prm = O.make_params(**mk)
    for b in range(len(offs) - 1):
        one = ics.box_slice(ic, b)
        ref, st, _ = O.run_pass(one, params=prm)
        outs.append(ref)
        inter += st.dens_interactions + st.grad_interactions + st.force_interactions
    secs = time.perf_counter() - t0
    ref = {k: np.concatenate([o[k] for o in outs]) for k in outs[0] if isinstance(outs[0][k], np.ndarray)}
    return ref, inter, secs


IDENTICAL_KEYS = ("hsml", "num_ngb", "dhsml", "divvel", "density", "pressure", "nvt", "cn", "fce",
                  "grad_rho", "grad_p", "grad_v", "vsig")


def force_on_identical_inputs(gzm, O, ic, state_out, cores):
    """The FP32-bulk force stage (GPU) and the oracle's force stage on the FP64
    pass's own density and gradient outputs, so that every P/rho tie of the
    entropic switch resolves alike: dv/dt and du/dt errors over every
    particle, ties included, relative to the sum of |pair terms|."""
    import ics
    from parity import TOL32_FLUX, l1_errs
    st = dict(ic)
    for k in IDENTICAL_KEYS:
        st[k] = state_out[k]
    offs = st.get("box_offsets")
    prm = O.make_params(nthreads=cores)
    parts = []
    rng = [(0, st["pos"].shape[0])] if offs is None else [(int(offs[b]), int(offs[b + 1])) for b in range(len(offs) - 1)]
    for b, (lo, hi) in enumerate(rng):
        one = dict(st) if offs is None else ics.box_slice(st, b)
        for k in IDENTICAL_KEYS:
            one[k] = np.asarray(st[k])[lo:hi]
        r, _, _ = O.run_pass(one, params=prm if offs is not None else
                             O.make_params(box=tuple(st["box"]), nthreads=cores), stages=("force",))
        parts.append(r)
    ref = {k: np.concatenate([p[k] for p in parts]) for k in ("acc", "dtu", "acc_l1", "dtu_l1", "tie_flag")}
    out, gst = gzm.run_pass(st, params=gzm.make_params(precision=1, box=tuple(st.get("box", (1.0, 1.0, 1.0)))),
                            stages=("force",))
    e_dtu = l1_errs(out["dtu"], ref["dtu"], ref["dtu_l1"])
    e_acc = l1_errs(out["acc"], ref["acc"], ref["acc_l1"]).max(axis=1)
    tied = ref["tie_flag"].astype(bool)
    stat = lambda e: {"max": float(f"{e.max():.3e}"), "rms": float(f"{np.sqrt(np.mean(e * e)):.3e}"),
                      "n_over": int((e > TOL32_FLUX).sum())}
    return {"dtu": stat(e_dtu), "dtu_tied_only": stat(e_dtu[tied]) if tied.any() else None, "acc": stat(e_acc),
            "n_tied": int(tied.sum()), "n_particles": int(len(e_dtu)),
            "fp64_force_particles": int(gst["fp64_force_particles"]),
            "inputs": "density and gradients of the FP64 GPU pass, given unchanged to both force stages"}


def main_batch(args, gzm, ics):
    steps = args.steps or 100
    if args.workload == "shear":
        ic = ics.shear_layers(224)
        nb = 1
        what = (f"two shear layers, density 2:1 by lattice spacing (282^2x141 + 224^2x112 lattices, "
                f"{ic['pos'].shape[0]:,} particles), vx = +-0.5, P = 2.5")
    else:
        ic = ics.batch_boxes(args.boxes)
        nb = len(ic["box_offsets"]) - 1
        what = f"{nb} periodic boxes, jittered lattices of 16^3..64^3 particles ({ic['pos'].shape[0]:,} in all)"
    N = ic["pos"].shape[0]
    log(f"workload built: {N:,} particles")
    tuned = {}
    if args.autotune:
        import autotune
        t0 = time.perf_counter()
        for prec in (1, 0):
            tuned[prec] = autotune.tune(ic, prec)
            log(f"autotune precision {prec}: {tuned[prec]['setting']}")
        tune_s = time.perf_counter() - t0
    r32 = run_graph_mode(gzm, ic, args, 1, steps, tuned.get(1))
    log(f"precision 1: {r32['ms_per_step']:.2f} ms/pass")
    r32d = run_graph_mode(gzm, ic, args, 2, steps, tuned.get(1)) if args.fp32_density else None
    if r32d is not None:
        log(f"precision 2: {r32d['ms_per_step']:.2f} ms/pass")
    r64 = run_graph_mode(gzm, ic, args, 0, steps, tuned.get(0))
    log(f"precision 0: {r64['ms_per_step']:.2f} ms/pass")
    ic_mfv = None
    rmfv = None
    if args.mfv:
        ic_mfv = dict(ic, ti_step=np.full(N, MFV_TI_STEP, dtype=np.int64))
        rmfv = run_graph_mode(gzm, ic_mfv, args, 0, steps, tuned.get(0), method=1)
        log(f"meshless finite volume (FP64): {rmfv['ms_per_step']:.2f} ms/pass")
    plain = run_plain_per_box(gzm, ic, tuned.get(0)) if (args.plain and "box_offsets" in ic) else None
    if plain is not None:
        log(f"plain FP64, one box per pass: {plain['ms_per_pass']:.1f} ms for all boxes")
    # The headline is precision 1 (FP64 density sweeps, FP32 gradient and
    # force pairs): its inputs to every threshold decision are the FP64
    # pass's, so its FP64 redo of near-threshold pairs recovers the FP64
    # decisions.  Precision 2 (FP32 density sums) is reported beside it.
    # Every mode's rate is the FP64 pass's interaction count (the oracle's:
    # the neighbour sets are bit-exact) over that mode's time per pass.
    head = r32
    inter = r64["interactions_per_pass"]
    for r in (r32, r32d, r64):
        if r is not None:
            r["value"] = inter / (r["ms_per_step"] * 1e-3)
    traffic = measured_traffic(head["dominant"], head["precision"], N)
    result = {
        "metric": "particle-neighbor interactions/s, MFM pass; FP32-batched vs FP64 at stated error",
        "value": head["value"],
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{what}, FP32-bulk MFM pass (density+gradients+force), {steps} passes in one HIP graph",
            "n_particles": N,
            "n_boxes": nb,
            "active": "all",
            "gamma": 5.0 / 3.0,
            "des_num_ngb": 32,
            "passes_per_graph": steps,
            "interactions_per_pass": int(inter),
            "density_sweeps": head["density_sweeps"],
            "parallelism": "single GPU",
        },
        "fp64_same_workload": {
            "value": r64["value"],
            "unit": "interactions/s",
            "ms_per_step": r64["ms_per_step"],
            "interactions_per_pass": int(r64["interactions_per_pass"]),
            "density_sweeps": r64["density_sweeps"],
            "kernels_ms_per_pass_eager": r64["kernels_ms_per_pass_eager"],
        },
        # precision alone: FP32-bulk vs FP64, both batched in one graph
        "fp32_speedup_vs_fp64": head["value"] / r64["value"],
        # the plain form (FP64, one box per pass through the eager API) and the
        # chosen form's speed-up over it (FP32-bulk and batching together)
        "fp64_plain_per_box": plain,
        "fp32_batched_speedup_vs_plain_fp64": (None if plain is None else head["value"] / plain["value"]),
        "plain_interactions_equal_batched": (None if plain is None else
                                             bool(plain["interactions_per_pass"] == r64["interactions_per_pass"])),
        # pair sets: the gradient and force interactions of each FP32 mode
        # equal the FP64 pass's (density sweeps of precision 2 are its own)
        "fp32_pair_sets_equal_fp64": bool(
            head["grad_interactions"] == r64["grad_interactions"] and
            head["force_interactions"] == r64["force_interactions"]),
        "fp32_interactions_equal_fp64": bool(r32["interactions_per_pass"] == r64["interactions_per_pass"]),
        "fp32_density_too": (None if r32d is None else {
            "precision": 2,
            "value": r32d["value"],
            "ms_per_step": r32d["ms_per_step"],
            "speedup_vs_fp64": r32d["value"] / r64["value"],
            "interactions_per_pass": int(r32d["interactions_per_pass"]),
            "pair_sets_equal_fp64": bool(r32d["grad_interactions"] == r64["grad_interactions"] and
                                         r32d["force_interactions"] == r64["force_interactions"]),
            "fp64_density_redo_particles_per_pass": r32d["density_fp64_redo"],
            "density_sweeps": r32d["density_sweeps"],
            "kernels_ms_per_pass_eager": r32d["kernels_ms_per_pass_eager"],
        }),
        # the meshless finite-volume variant (gzm_params.method 1), FP64, same
        # boxes and neighbour sets, mass fluxes through every face
        "mfv_same_workload": (None if rmfv is None else {
            "value": rmfv["value"],
            "unit": "interactions/s",
            "ms_per_step": rmfv["ms_per_step"],
            "interactions_per_pass": int(rmfv["interactions_per_pass"]),
            "interactions_equal_mfm": bool(rmfv["interactions_per_pass"] == r64["interactions_per_pass"]),
            "exact_solver_calls_per_pass": int(rmfv["exact_solver_calls"]),
            "riemann_retries_per_pass": int(rmfv["riemann_retries"]),
            "vs_mfm_fp64": rmfv["value"] / r64["value"],
            "kernels_ms_per_pass_eager": rmfv["kernels_ms_per_pass_eager"],
            "timesteps": f"every particle on integer step {MFV_TI_STEP} of {MFV_TIMEBASE:g} time units",
        }),
        "exact_solver_calls_per_pass": r64["exact_solver_calls"],
        # FP32 force loops redone in FP64: a pair's branch decision within FP32 rounding of its threshold
        "fp64_force_particles_per_pass": head["fp64_force_particles"],
        "autotune": ({"fp32": tuned[1], "fp64": tuned[0], "seconds": round(tune_s, 1)} if tuned else None),
        "kernels_ms_per_pass_eager": head["kernels_ms_per_pass_eager"],
        "roofline": {
            "bound": "hbm",
            "kernel": head["dominant"],
            "achieved": head["achieved_gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": head["achieved_gbs"] / HBM_PEAK_GBS,
            "traffic": traffic,
            "bytes_per_launch": head["bytes_per_launch"],
            "mean_launch_ms": head["mean_launch_ms"],
        },
        # rocprof VALU counters of the dominant kernel on this workload (committed summary)
        "valu": measured_valu(head["dominant"], head["precision"], N),
    }
    if args.cpu_repeats > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle as O
        from parity import compare_mfv, compare_pass, error_stats

        cores = cpu_cores()
        if "box_offsets" in ic:
            nbs = min(args.cpu_boxes, nb)
            samples = [ics.box_slice(ic, b) for b in range(nbs)]
            what = f"the first {nbs} of the {nb} boxes"
        else:
            samples, what = [ic], "the whole workload"
        log(f"CPU baseline: octree oracle on {what}, {cores} cores")
        result["cpu_baseline"], _ = cpu_baseline(O, samples, cores, args.cpu_repeats, what)
        log(f"CPU baseline: {result['cpu_baseline']['sample']}")
        # the reference outputs of the error statistics: the oracle (cell-grid
        # search, the same sets) over every box
        log(f"CPU oracle on all {nb} box(es), {cores} cores")
        ref, cpu_int, secs = oracle_boxes(O, ic, cores)
        log(f"CPU oracle: {secs:.1f} s")
        result["cpu_full_pass"] = {"value": cpu_int / secs, "unit": "interactions/s", "seconds": round(secs, 2),
                                   "search": "cell grid", "sample": f"one pass over all {nb} box(es) ({N:,} particles)"}
        def fmt(es):
            return {k: (v if isinstance(v, int) else {q: (float(f"{x:.3e}") if isinstance(x, float) else x)
                                                      for q, x in v.items()})
                    for k, v in es.items()}
        result["error_vs_fp64_oracle"] = fmt(error_stats(head["out"], ref))
        if r32d is not None:
            result["fp32_density_too"]["error_vs_fp64_oracle"] = fmt(error_stats(r32d["out"], ref))
        # du/dt with P/rho ties included: both force stages on the FP64
        # pass's own density and gradient outputs (identical inputs)
        log("force stage on identical inputs: oracle and FP32-bulk GPU")
        result["force_on_identical_inputs"] = force_on_identical_inputs(gzm, O, ic, r64["out"], cores)
        errs = compare_pass(r64["out"], ref)
        result["fp64_same_workload"]["parity_vs_oracle"] = {k: float(f"{v:.3e}") for k, v in errs.items()}
        result["fp64_same_workload"]["interactions_equal"] = bool(cpu_int == r64["interactions_per_pass"])
        if rmfv is not None:
            log(f"CPU oracle, meshless finite volume, on all {nb} box(es)")
            ref_v, int_v, secs_v = oracle_boxes(O, ic_mfv, cores, method=1)
            log(f"CPU oracle (MFV): {secs_v:.1f} s")
            ev = compare_mfv(rmfv["out"], ref_v)
            result["mfv_same_workload"]["parity_vs_oracle"] = {k: float(f"{v:.3e}") for k, v in ev.items()}
            result["mfv_same_workload"]["interactions_equal_oracle"] = bool(int_v == rmfv["interactions_per_pass"])
            result["mfv_same_workload"]["cpu_oracle_value"] = int_v / secs_v
    if args.blast and args.workload == "batch":
        b_args = argparse.Namespace(**vars(args))
        b_args.steps = 5
        b_args.warmup = 1
        blast_ic = ics.blast(args.n)
        b64 = run_mode(gzm, blast_ic, b_args, 0)
        b32 = run_mode(gzm, blast_ic, b_args, 1)
        log("blast done")
        result["blast_128"] = {
            "n_particles": blast_ic["pos"].shape[0],
            "fp64": {"value": b64["value"], "ms_per_step": b64["ms_per_step"],
                     "kernels_ms_per_pass": b64["kernels_ms_per_pass"],
                     "upload_ms": b64["upload_ms"], "download_ms": b64["download_ms"],
                     "pcie_inclusive_value": b64["pcie_inclusive_value"]},
            "fp32_bulk": {"value": b32["value"], "ms_per_step": b32["ms_per_step"],
                          "kernels_ms_per_pass": b32["kernels_ms_per_pass"]},
            "fp32_speedup_vs_fp64": b32["value"] / b64["value"],
        }
    print(json.dumps(result))


STATE_KEYS = ("hsml", "num_ngb", "dhsml", "divvel", "density", "pressure", "nvt", "cn", "fce",
              "grad_rho", "grad_p", "grad_v", "acc", "dtu", "vsig")


def main_clustered(args, gzm, ics):
    """Plummer spheres in a uniform background, FP64 pass with the tree search:
    every particle active, then 10 % active on the state the full pass left
    (the other 90 % keep their hsml, density, gradients)."""
    args.steps = steps = args.steps or 5
    if args.list_cap is None:
        args.list_cap = 64
    t0 = time.perf_counter()
    ic = ics.plummer_clusters(args.cluster_n, args.spheres)
    ic_s = time.perf_counter() - t0
    N = ic["pos"].shape[0]
    log(f"clustered set built: {N:,} particles, {ic_s:.1f} s")
    full = run_mode(gzm, ic, args, 0, search=1)
    log(f"full pass: {full['ms_per_step']:.1f} ms")
    state = dict(ic)
    for k in STATE_KEYS:
        state[k] = full["out"][k]
    active = ics.random_subset(N, 0.1)
    part = run_mode(gzm, state, args, 0, search=1, active=active)
    log(f"10% active pass: {part['ms_per_step']:.1f} ms")
    # FP32-bulk (precision 1: FP64 density stage, FP32 gradient and force pairs)
    full32 = run_mode(gzm, ic, args, 1, search=1)
    log(f"FP32-bulk full pass: {full32['ms_per_step']:.1f} ms")
    part32 = run_mode(gzm, state, args, 1, search=1, active=active)
    log(f"FP32-bulk 10% active pass: {part32['ms_per_step']:.1f} ms")
    h = ic["hsml"]
    result = {
        "metric": "particle-neighbor interactions/s, MFM pass; FP32-batched vs FP64 at stated error",
        "value": full["value"],
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": full["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.spheres} Plummer spheres + 20% uniform background ({N:,} particles), "
                        "plain FP64 MFM pass (density+gradients+force), tree neighbour search",
            "n_particles": N,
            "active": "all",
            "hsml_range": [float(h.min()), float(h.max())],
            "interactions_per_pass": full["interactions_per_pass"],
            "density_sweeps": full["density_sweeps"],
            "ic_seconds": round(ic_s, 1),
            "pair_list_cap": args.list_cap,
            "parallelism": "single GPU",
        },
        "kernels_ms_per_pass": full["kernels_ms_per_pass"],
        "roofline": {
            "bound": "hbm",
            "kernel": full["dominant"],
            "achieved": full["achieved_gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": full["achieved_gbs"] / HBM_PEAK_GBS,
            "traffic": None,
            "bytes_per_launch": full["bytes_per_launch"],
            "mean_launch_ms": full["mean_launch_ms"],
        },
        "active_10pct": {
            "n_active": int(len(active)),
            "value": part["value"],
            "unit": "interactions/s",
            "ms_per_step": part["ms_per_step"],
            "interactions_per_pass": part["interactions_per_pass"],
            "density_sweeps": part["density_sweeps"],
            "kernels_ms_per_pass": part["kernels_ms_per_pass"],
        },
        "fp32_same_workload": {
            "precision": 1,
            "value": full32["value"],
            "unit": "interactions/s",
            "ms_per_step": full32["ms_per_step"],
            "speedup_vs_fp64": full["ms_per_step"] / full32["ms_per_step"],
            "interactions_equal_fp64": bool(full32["interactions_per_pass"] == full["interactions_per_pass"]),
            "kernels_ms_per_pass": full32["kernels_ms_per_pass"],
            "active_10pct": {
                "value": part32["value"],
                "ms_per_step": part32["ms_per_step"],
                "speedup_vs_fp64": part["ms_per_step"] / part32["ms_per_step"],
                "interactions_equal_fp64": bool(part32["interactions_per_pass"] == part["interactions_per_pass"]),
            },
        },
    }
    if args.cpu_repeats > 0:
        # CPU baseline on a bounded sample: a clustered set of --cpu-sample
        # particles with proportionally fewer spheres (same generator), FP64
        # oracle on the host cores; the GPU pass of the same sample is held
        # to the FP64 parity bar against it
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle as O
        from parity import compare_pass

        n_s = min(args.cpu_sample, N)
        sph_s = max(1, round(args.spheres * n_s / N))
        sample = ics.plummer_clusters(n_s, sph_s) if n_s < N else ic
        cores = cpu_cores()
        log(f"CPU baseline: octree oracle on a {n_s:,}-particle sample, {cores} cores")
        result["cpu_baseline"], cpu_int = cpu_baseline(
            O, [sample], cores, args.cpu_repeats,
            f"one full pass of a {n_s:,}-particle clustered set ({sph_s} spheres, same generator)")
        log(f"CPU baseline: {result['cpu_baseline']['sample']}")
        ref, _, _ = O.run_pass(sample, params=O.make_params(box=tuple(sample["box"]), nthreads=cores))
        out_s, st_s = gzm.run_pass(sample, params=gzm.make_params(box=tuple(sample["box"]), search=1,
                                                                   pair_list_cap=args.list_cap))
        errs = compare_pass(out_s, ref)
        result["parity_vs_oracle_on_sample"] = {k: float(f"{v:.3e}") for k, v in errs.items()}
        result["parity_vs_oracle_on_sample"]["interactions_equal"] = bool(
            cpu_int == st_s["density_interactions"] + st_s["gradient_interactions"] + st_s["force_interactions"])
        # every particle: the oracle (octree) over the whole set, full pass and
        # the 10 %-active pass on the state the FP64 full pass left
        from parity import error_stats

        def fmt(es):
            return {k: (v if isinstance(v, int) else {q: (float(f"{x:.3e}") if isinstance(x, float) else x)
                                                      for q, x in v.items()})
                    for k, v in es.items()}
        t = time.perf_counter()
        prm_o = O.make_params(box=tuple(ic["box"]), nthreads=cores, tree=1)
        ref_f, ost_f, _ = O.run_pass(ic, params=prm_o)
        ref_p, ost_p, _ = O.run_pass(state, active=active, params=prm_o)
        log(f"CPU oracle on all {N:,} particles (full and 10 % active): {time.perf_counter() - t:.1f} s")
        int_f = ost_f.dens_interactions + ost_f.grad_interactions + ost_f.force_interactions
        int_p = ost_p.dens_interactions + ost_p.grad_interactions + ost_p.force_interactions
        for rec, out, ref_o, n_int, act in ((result, full, ref_f, int_f, None),
                                            (result["active_10pct"], part, ref_p, int_p, active)):
            errs = compare_pass(out["out"], ref_o, active=act)
            rec["parity_vs_oracle"] = {k: float(f"{v:.3e}") for k, v in errs.items()}
            rec["parity_vs_oracle"]["interactions_equal"] = bool(n_int == out["interactions_per_pass"])
        result["fp32_same_workload"]["error_vs_fp64_oracle"] = fmt(error_stats(full32["out"], ref_f))
        result["fp32_same_workload"]["active_10pct"]["error_vs_fp64_oracle"] = fmt(
            error_stats(part32["out"], ref_p, active=active))
    print(json.dumps(result))


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

    if args.workload in ("batch", "shear"):
        main_batch(args, gzm, ics)
        return
    if args.workload == "clustered":
        main_clustered(args, gzm, ics)
        return
    args.steps = args.steps or 10
    ic = ics.blast(args.n)
    N = ic["pos"].shape[0]
    r64 = run_mode(gzm, ic, args, 0)
    r32 = run_mode(gzm, ic, args, 1) if args.fp32 else None

    traffic = measured_traffic(r64["dominant"], 0, N)
    result = {
        "metric": "particle-neighbor interactions/s, MFM pass; FP32-batched vs FP64 at stated error",
        "value": r64["value"],
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["ms_per_step"],
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
            "interactions_per_pass": r64["interactions_per_pass"],
            "density_sweeps": r64["density_sweeps"],
            "parallelism": "single GPU",
        },
        "kernels_ms_per_pass": r64["kernels_ms_per_pass"],
        "roofline": {
            "bound": "hbm",
            "kernel": r64["dominant"],
            "achieved": r64["achieved_gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": r64["achieved_gbs"] / HBM_PEAK_GBS,
            "traffic": traffic,
            "bytes_per_launch": r64["bytes_per_launch"],
            "mean_launch_ms": r64["mean_launch_ms"],
        },
    }
    if r32 is not None:
        result["fp32_bulk"] = {
            "value": r32["value"],
            "unit": "interactions/s",
            "ms_per_step": r32["ms_per_step"],
            "speedup_vs_fp64": r32["value"] / r64["value"],
            "interactions_per_pass": r32["interactions_per_pass"],
            "density_sweeps": r32["density_sweeps"],
            "kernels_ms_per_pass": r32["kernels_ms_per_pass"],
            "roofline": {"kernel": r32["dominant"], "achieved": r32["achieved_gbs"], "unit": "GB/s",
                         "frac": r32["achieved_gbs"] / HBM_PEAK_GBS,
                         "traffic": measured_traffic(r32["dominant"], 1, N)},
        }

    # CPU baseline: the FP64 oracle (OpenMP) on this host, same workload
    if args.cpu_repeats > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle as O
        from parity import compare_pass, error_stats

        cores = cpu_cores()
        log(f"CPU baseline: octree oracle on the {args.n}^3 blast, {cores} cores")
        result["cpu_baseline"], cpu_int = cpu_baseline(O, [ic], cores, args.cpu_repeats,
                                                       f"one full {args.n}^3 blast pass")
        ref, _, _ = O.run_pass(ic, params=O.make_params(box=tuple(ic["box"]), nthreads=cores))
        errs = compare_pass(r64["out"], ref)
        result["parity_vs_oracle"] = {k: float(f"{v:.3e}") for k, v in errs.items()}
        result["parity_vs_oracle"]["interactions_equal"] = bool(cpu_int == r64["interactions_per_pass"])
        if r32 is not None:
            es = error_stats(r32["out"], ref)
            result["fp32_bulk"]["error_vs_fp64_oracle"] = {
                k: (v if isinstance(v, int) else {q: (float(f"{x:.3e}") if isinstance(x, float) else x) for q, x in v.items()})
                for k, v in es.items()}
    print(json.dumps(result))


if __name__ == "__main__":
    main()
