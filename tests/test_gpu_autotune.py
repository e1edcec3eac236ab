"""The autotuner's axes on the GPU: workgroup sizes of the three loops, the
density load-balance thresholds (candidates per thread before a particle gets
a wavefront, sweep size below which every particle gets one), cell size and
pair-list capacity change the order of summation only.  A tuned setting, and
settings at the ends of every axis, keep the default's interaction counts and
meet the parity bar against the oracle."""
import numpy as np
import pytest

import autotune
import gzm
import ics
import oracle as O
from parity import TOL32_FLUX, check_pass, error_stats

pytestmark = pytest.mark.gpu


def small_batch():
    return ics.batch_boxes(8, n_min=512, n_max=8192)


def oracle_batch(ic):
    offs = ic["box_offsets"]
    parts = []
    for b in range(len(offs) - 1):
        r, _, _ = O.run_pass(ics.box_slice(ic, b))
        parts.append(r)
    return {k: np.concatenate([p[k] for p in parts]) for k in parts[0] if isinstance(parts[0][k], np.ndarray)}


def counts(st):
    return (st["density_interactions"], st["gradient_interactions"], st["force_interactions"],
            st["exact_solver_calls"])


SETTINGS = {
    "blocks_64": dict(block_size=(64, 64, 64)),
    "blocks_128": dict(block_size=(128, 128, 128)),
    "mixed_blocks": dict(block_size=(64, 256, 128)),
    "blocks_512": dict(block_size=(512, 512, 512)),
    "no_wave_sweeps": dict(wave_sweep_max=-1, density_budget=-1),
    "wave_every_sweep": dict(wave_sweep_max=1 << 20),
    "small_budget": dict(density_budget=16, wave_sweep_max=16),
    "fine_cells": dict(cell_size=-0.35, pair_list_cap=48),
}


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name", list(SETTINGS))
def test_launch_settings_keep_sets_and_parity(name, precision):
    ic = small_batch()
    ref = oracle_batch(ic)
    print(f"launch setting {name} precision {precision}", flush=True)
    base, bst = gzm.run_pass(ic, params=gzm.make_params(precision=precision))
    out, st = gzm.run_pass(ic, params=gzm.make_params(precision=precision, **SETTINGS[name]))
    assert counts(st)[1:] == counts(bst)[1:]
    if "wave" not in name and "budget" not in name:
        assert st["density_interactions"] == bst["density_interactions"]
    if precision == 0:
        errs, bad = check_pass(out, ref)
        assert not bad, bad
    else:
        es = error_stats(out, ref)
        for k in ("acc", "dtu"):
            assert es[k]["max"] <= TOL32_FLUX, (k, es[k])


def test_tuned_setting_reproduces_the_default():
    ic = small_batch()
    print("autotune on 8 boxes, precision 1", flush=True)
    r = autotune.tune(ic, 1, n_sample=8)
    for name, cands in autotune.AXES:
        assert set(r["axes"][name]) == {str(v) for v in cands}
        assert all(v > 0 for v in r["axes"][name].values())
    ref = oracle_batch(ic)
    base, bst = gzm.run_pass(ic, params=gzm.make_params(precision=1))
    out, st = gzm.run_pass(ic, params=gzm.make_params(precision=1, **r["params"]))
    assert counts(st)[1:] == counts(bst)[1:]
    es = error_stats(out, ref)
    for k in ("acc", "dtu"):
        assert es[k]["max"] <= TOL32_FLUX, (k, es[k])


def test_bad_block_size_is_rejected():
    with pytest.raises(gzm.GzmError, match="block_size"):
        gzm.MFMContext(gzm.make_params(block_size=(96, 256, 256)))
