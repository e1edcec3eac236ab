"""Selection rule of the launch-parameter autotuner, with the GPU timings
replaced by a model of the setting (runs on the CPU)."""
import numpy as np

import autotune


def _tune(monkeypatch, cost):
    calls = []

    def fake(sample, precision, setting):
        calls.append(dict(setting))
        return cost(setting)

    monkeypatch.setattr(autotune, "time_setting", fake)
    ic = {"pos": np.zeros((8, 3))}
    return autotune.tune(ic, precision=1), calls


def _model(cell=None, cap48=1.0, blocks=None, budget=None, wave=None):
    """time = product of per-axis factors (1.0 = the default's)"""
    cell = cell or {}
    blocks = blocks or {}
    budget = budget or {}
    wave = wave or {}

    def cost(s):
        t = 10.0 * cell.get(s["cell_factor"], 1.0) * (cap48 if s["pair_list_cap"] == 48 else 1.0)
        for ax in ("block_density", "block_gradients", "block_force"):
            t *= blocks.get((ax, s[ax]), 1.0)
        return t * budget.get(s["density_budget"], 1.0) * wave.get(s["wave_sweep_max"], 1.0)
    return cost


def test_coordinate_descent_takes_each_clear_gain(monkeypatch):
    cost = _model(cell={0.75: 0.9, 0.6: 0.85, 0.5: 0.8, 0.42: 0.78, 0.35: 0.83}, cap48=0.97,
                  blocks={("block_force", 128): 0.95, ("block_density", 64): 1.2},
                  budget={1024: 0.9}, wave={16384: 1.1})
    r, calls = _tune(monkeypatch, cost)
    s = r["setting"]
    assert s["cell_factor"] == 0.42          # interior optimum of the cell axis
    assert s["pair_list_cap"] == 48
    assert (s["block_density"], s["block_gradients"], s["block_force"]) == (256, 256, 128)
    assert s["density_budget"] == 1024 and s["wave_sweep_max"] == 4096
    assert r["params"]["cell_size"] == -0.42 and r["params"]["block_size"] == (256, 256, 128)
    # every candidate of every axis appears in the per-axis table, none timed twice
    for name, cands in autotune.AXES:
        assert set(r["axes"][name]) == {str(v) for v in cands}
    labels = [autotune._label(c) for c in calls]
    assert len(labels) == len(set(labels))


def test_gain_within_noise_keeps_the_earlier_setting(monkeypatch):
    # every alternative within 2 % of the default: nothing changes
    cost = _model(cell={0.75: 0.99, 0.5: 0.985}, cap48=0.99, blocks={("block_gradients", 64): 0.985})
    r, _ = _tune(monkeypatch, cost)
    assert r["setting"] == autotune.DEFAULTS
    assert r["params"]["block_size"] == (256, 256, 256)
