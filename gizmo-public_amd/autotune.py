"""Autotuner of the MFM pass's launch parameters on the GPU.

For a particle set (one box or a batch) it times eager passes through the C
ABI for candidate settings of each axis in turn (coordinate descent from the
library's defaults) and returns the fastest setting it met:

  cell_factor     grid cell side as a factor of each box's mean initial hsml
                  (gzm_params.cell_size < 0): candidate tests per neighbour
                  against cells walked
  pair_list_cap   neighbours per particle kept in the gradient-to-force pair
                  list (longer pair sets take the wavefront fallback)
  block_size      workgroup size of the density-sweep, gradient and force
                  loops (64, 128, 256, 512), one axis each (threads per
                  particle are fixed at one, so this trades scheduling
                  granularity against tail effects)
  density_budget  candidates a density-sweep thread tests before its particle
                  moves to a wavefront of its own
  wave_sweep_max  density sweeps over at most this many particles give every
                  particle a wavefront (neighbours per wavefront rather than
                  per thread)

A batch is tuned on a sample of its boxes spread over the size range, so that
tuning costs a small fraction of a full pass.  None of these settings changes
a neighbour set; they change summation order only, within the parity bar.
"""
from __future__ import annotations

import time

import numpy as np

import gzm

CELL_FACTORS = (1.0, 0.75, 0.6, 0.5, 0.42, 0.35)
LIST_CAPS = (64, 48)
BLOCK_SIZES = (256, 128, 64, 512)
DENSITY_BUDGETS = (512, 256, 1024)
WAVE_SWEEP_MAX = (4096, 1024, 16384)
DEFAULTS = dict(cell_factor=1.0, pair_list_cap=64, block_density=256, block_gradients=256, block_force=256,
                density_budget=512, wave_sweep_max=4096)
# axes in the order they are tuned, each with its candidates (the first is
# the default the descent starts from)
AXES = (("cell_factor", CELL_FACTORS), ("pair_list_cap", LIST_CAPS), ("block_density", BLOCK_SIZES),
        ("block_gradients", BLOCK_SIZES), ("block_force", BLOCK_SIZES), ("density_budget", DENSITY_BUDGETS),
        ("wave_sweep_max", WAVE_SWEEP_MAX))
# A later candidate replaces the best so far only when it is faster by more
# than this fraction: timings of one setting on the sample vary by ~1 % from
# run to run, and a choice made on noise would differ between otherwise
# identical runs.
MIN_GAIN = 0.02


def make_params_kwargs(setting: dict) -> dict:
    """gzm.make_params keyword arguments of an autotuner setting."""
    return dict(cell_size=-float(setting["cell_factor"]), pair_list_cap=int(setting["pair_list_cap"]),
                block_size=(int(setting["block_density"]), int(setting["block_gradients"]),
                            int(setting["block_force"])),
                density_budget=int(setting["density_budget"]), wave_sweep_max=int(setting["wave_sweep_max"]))


def sample_boxes(ic: dict, n_sample: int = 32) -> dict:
    """Up to n_sample boxes of a batch, evenly spaced in box size."""
    offs = np.asarray(ic["box_offsets"])
    nb = len(offs) - 1
    if nb <= n_sample:
        return ic
    sizes = np.diff(offs)
    order = np.argsort(sizes, kind="stable")
    pick = np.sort(order[np.linspace(0, nb - 1, n_sample).round().astype(int)])
    rows = np.concatenate([np.arange(offs[b], offs[b + 1]) for b in pick])
    out = {k: v[rows] for k, v in ic.items()
           if isinstance(v, np.ndarray) and v.shape[:1] == ic["pos"].shape[:1]}
    new_offs = np.zeros(len(pick) + 1, dtype=np.int64)
    new_offs[1:] = np.cumsum(sizes[pick])
    out["box_offsets"] = new_offs
    out["box_sizes"] = np.asarray(ic["box_sizes"])[pick]
    out["name"] = f"{ic.get('name', 'batch')}[sample {len(pick)}]"
    return out


def time_setting(ic: dict, precision: int, setting: dict, passes: int = 2) -> float:
    """Mean wall time (s) of `passes` eager passes after one warm-up pass."""
    box = tuple(ic["box"]) if "box" in ic and np.ndim(ic["box"]) == 1 else (1.0, 1.0, 1.0)
    params = gzm.make_params(box=box, precision=precision, **make_params_kwargs(setting))
    with gzm.MFMContext(params) as ctx:
        ctx.upload(ic)

        def one():
            ctx.restore_initial_state()
            ctx.density()
            ctx.gradients()
            ctx.hydro_force()

        one()
        t0 = time.perf_counter()
        for _ in range(passes):
            one()
        return (time.perf_counter() - t0) / passes


def _label(setting: dict) -> str:
    return " ".join(f"{k}={setting[k]}" for k, _ in AXES)


def tune(ic: dict, precision: int, axes=AXES, n_sample: int = 32, min_gain: float = MIN_GAIN) -> dict:
    """Fastest setting on a sample of `ic` by coordinate descent over `axes`,
    and the table of every timing: {"setting": {...}, "params": make_params
    kwargs, "cell_factor", "pair_list_cap", "table": {label: ms},
    "axes": {axis: {candidate: ms}}}.  Along an axis the candidates are tried
    in the order given, the others held at the best setting so far; one
    replaces the best only when it is faster by more than `min_gain`."""
    sample = sample_boxes(ic, n_sample) if "box_offsets" in ic else ic
    best = dict(DEFAULTS)
    timed = {}

    def measure(setting):
        key = _label(setting)
        if key not in timed:
            timed[key] = time_setting(sample, precision, setting)
        return timed[key]

    best_t = measure(best)
    per_axis = {}
    for name, candidates in axes:
        per_axis[name] = {}
        for v in candidates:
            trial = dict(best, **{name: v})
            t = measure(trial)
            per_axis[name][str(v)] = round(t * 1e3, 3)
            if t < (1 - min_gain) * best_t:
                best, best_t = trial, t
    return {"setting": best, "params": make_params_kwargs(best), "cell_factor": best["cell_factor"],
            "pair_list_cap": best["pair_list_cap"], "ms": round(best_t * 1e3, 3),
            "table": {k: round(v * 1e3, 3) for k, v in timed.items()}, "axes": per_axis,
            "sample_particles": int(sample["pos"].shape[0])}
