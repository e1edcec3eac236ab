"""Selection rule of the launch-parameter autotuner, with the GPU timings
replaced by a fixed table (runs on the CPU)."""
import numpy as np

import autotune


def _tune(monkeypatch, times):
    monkeypatch.setattr(autotune, "time_setting", lambda sample, precision, f, c: times[(f, c)])
    ic = {"pos": np.zeros((8, 3))}
    return autotune.tune(ic, precision=1)


def _table(cell_ms, cap48_ms):
    t = {(f, 64): ms for f, ms in zip(autotune.CELL_FACTORS, cell_ms)}
    for f in autotune.CELL_FACTORS:
        t[(f, 48)] = cap48_ms
    return t


def test_clear_gain_is_taken(monkeypatch):
    r = _tune(monkeypatch, _table([10.4, 12.7, 13.2, 8.2], 7.9))
    assert (r["cell_factor"], r["pair_list_cap"]) == (0.5, 48)
    assert r["table"]["cell 0.5 cap 48"] == 7.9e3


def test_gain_within_noise_keeps_the_earlier_setting(monkeypatch):
    # cap 48 is 0.9 % faster than cap 64 at cell 0.5: within run-to-run noise
    r = _tune(monkeypatch, _table([10.4, 12.7, 13.2, 8.155], 8.083))
    assert (r["cell_factor"], r["pair_list_cap"]) == (0.5, 64)
    # a cell factor 1 % faster than the first one does not replace it either
    r = _tune(monkeypatch, _table([10.0, 9.9, 10.2, 10.1], 9.0))
    assert (r["cell_factor"], r["pair_list_cap"]) == (1.0, 48)
