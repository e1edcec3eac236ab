"""Autotuner of the MFM pass's launch parameters on the GPU.

For a particle set (one box or a batch) it times eager passes through the C
ABI for each candidate setting and returns the fastest: the grid cell side
as a factor of each box's mean initial hsml (gzm_params.cell_size < 0), which
trades candidate tests per neighbour against cells walked, and the capacity
of the gradient-to-force pair list (neighbours per particle kept in its row;
particles with more pairs fall back to the grid scan).  A batch is tuned on a
sample of its boxes spread over the size range, so that tuning costs a small
fraction of a full pass.
"""
from __future__ import annotations

import time

import numpy as np

import gzm

CELL_FACTORS = (1.0, 0.75, 0.6, 0.5)
LIST_CAPS = (64, 48)
# A later candidate replaces the best so far only when it is faster by more
# than this fraction: timings of one setting on the sample vary by ~1 % from
# run to run, and a choice made on noise would differ between otherwise
# identical runs.
MIN_GAIN = 0.02


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


def time_setting(ic: dict, precision: int, cell_factor: float, list_cap: int, passes: int = 2) -> float:
    """Mean wall time (s) of `passes` eager passes after one warm-up pass."""
    box = tuple(ic["box"]) if "box" in ic and np.ndim(ic["box"]) == 1 else (1.0, 1.0, 1.0)
    params = gzm.make_params(box=box, precision=precision, cell_size=-float(cell_factor),
                             pair_list_cap=int(list_cap))
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


def tune(ic: dict, precision: int, cell_factors=CELL_FACTORS, list_caps=LIST_CAPS, n_sample: int = 32,
         min_gain: float = MIN_GAIN) -> dict:
    """Fastest (cell factor, list capacity) on a sample of `ic`, and the table
    of all timings: {"cell_factor": f, "pair_list_cap": c, "table": {...}}.
    Candidates are tried in the order given; one replaces the best so far
    only when it is faster by more than `min_gain`."""
    sample = sample_boxes(ic, n_sample) if "box_offsets" in ic else ic
    table = {}
    best_f = None
    for f in cell_factors:
        table[(f, list_caps[0])] = time_setting(sample, precision, f, list_caps[0])
        if best_f is None or table[(f, list_caps[0])] < (1 - min_gain) * table[(best_f, list_caps[0])]:
            best_f = f
    best = (best_f, list_caps[0])
    for c in list_caps[1:]:
        table[(best_f, c)] = time_setting(sample, precision, best_f, c)
        if table[(best_f, c)] < (1 - min_gain) * table[best]:
            best = (best_f, c)
    return {"cell_factor": best[0], "pair_list_cap": best[1],
            "table": {f"cell {k[0]} cap {k[1]}": round(v * 1e3, 3) for k, v in table.items()},
            "sample_particles": int(sample["pos"].shape[0])}
