"""The shear-layer (C3) and clustered (C5) configurations at reduced size:
FP64 pass against the oracle to the parity bar, FP32-bulk pass within its
error bounds, and the clustered case's 10 %-active pass that reuses the state
of the other 90 %."""
import numpy as np
import pytest

import gzm
import ics
import oracle as O
from parity import TOL32_DENSITY, TOL32_FLUX, check_pass, error_stats

pytestmark = pytest.mark.gpu

STATE_KEYS = ("hsml", "num_ngb", "dhsml", "divvel", "density", "pressure", "nvt", "cn", "fce",
              "grad_rho", "grad_p", "grad_v", "acc", "dtu", "vsig")

CASES = {
    "shear_layers_24": lambda: ics.shear_layers(24),
    "plummer_clusters_32k": lambda: ics.plummer_clusters(32768, 50),
}


@pytest.mark.parametrize("name", list(CASES))
def test_fp64_pass_matches_oracle(name):
    ic = CASES[name]()
    ref, st, _ = O.run_pass(ic)
    out, gst = gzm.run_pass(ic)
    assert gst["density_sweeps"] == st.iterations
    assert gst["density_interactions"] == st.dens_interactions
    assert gst["gradient_interactions"] == st.grad_interactions
    assert gst["force_interactions"] == st.force_interactions
    errs, bad = check_pass(out, ref)
    print(name, {k: f"{v:.2e}" for k, v in errs.items()})
    assert not bad, bad


@pytest.mark.parametrize("name", list(CASES))
def test_fp32_pass_within_bounds(name):
    ic = CASES[name]()
    ref, st, _ = O.run_pass(ic)
    out, gst = gzm.run_pass(ic, params=gzm.make_params(precision=1))
    assert gst["density_interactions"] == st.dens_interactions
    assert gst["gradient_interactions"] == st.grad_interactions
    assert gst["force_interactions"] == st.force_interactions
    es = error_stats(out, ref)
    print(name, {k: v for k, v in es.items()})
    for k in ("density", "hsml", "pressure"):
        assert es[k]["max"] <= TOL32_DENSITY, (k, es[k])
    for k in ("acc", "dtu"):
        assert es[k]["max"] <= TOL32_FLUX, (k, es[k])


def test_clustered_ten_percent_active_after_full_pass():
    ic = ics.plummer_clusters(32768, 50)
    N = ic["pos"].shape[0]
    full, _, _ = O.run_pass(ic)
    state = dict(ic)
    for k in STATE_KEYS:
        state[k] = full[k]
    active = ics.random_subset(N, 0.1)
    ref, st, _ = O.run_pass(state, active=active)
    out, gst = gzm.run_pass(state, active=active)
    assert gst["density_interactions"] == st.dens_interactions
    assert gst["gradient_interactions"] == st.grad_interactions
    assert gst["force_interactions"] == st.force_interactions
    errs, bad = check_pass(out, ref, active=active)
    assert not bad, bad
    np.testing.assert_array_equal(out["wakeup"], ref["wakeup"])
    inactive = np.setdiff1d(np.arange(N), active)
    for k in ("hsml", "density", "grad_v", "acc", "dtu"):
        np.testing.assert_array_equal(out[k][inactive], state[k][inactive])
