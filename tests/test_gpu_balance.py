"""Load-balanced density sweeps (gzm_params.density_budget).

A sweep thread that passes its candidate budget stops and lists its particle
for a second launch that gives each listed particle a whole wavefront (FP64
sums reduced across the lanes).  The budget moves work between the two
launches, never changes a neighbour set: in FP64 interaction counts and
sweep counts must not depend on it, and the outputs stay within the FP64
parity bar of the oracle (a listed particle's sums are added in another
order).  With FP32 density sums (precision 2) a listed particle is summed in
FP64, so its hsml may differ at the FP32 level: the FP32 bounds apply.  A graph
replay equals the eager pass bit for bit.
"""
import numpy as np
import pytest

import gzm
import ics
import oracle as O
from parity import TOL32_DENSITY, TOL32_FLUX, check_pass, error_stats

pytestmark = pytest.mark.gpu

OUT_KEYS = ("hsml", "density", "pressure", "num_ngb", "dhsml", "divvel", "cn", "fce", "nvt",
            "grad_rho", "grad_p", "grad_v", "acc", "dtu", "vsig")

CASES = {
    "blast_16": (lambda: ics.blast(16), {}),
    "plummer_clusters_32k_tree": (lambda: ics.plummer_clusters(32768, 50), {"search": 1}),
    "plummer_clusters_32k_grid": (lambda: ics.plummer_clusters(32768, 50), {}),
    "random_state_14_fp32_density": (lambda: ics.random_state(14), {"precision": 2}),
}


@pytest.mark.parametrize("name", list(CASES))
def test_density_budget_keeps_sets_and_parity(name):
    make, kw = CASES[name]
    ic = make()
    box = tuple(ic["box"])
    ref, rst, _ = O.run_pass(ic)
    base, bst = gzm.run_pass(ic, params=gzm.make_params(box=box, density_budget=-1, **kw))
    for budget in (0, 8, 40):
        out, st = gzm.run_pass(ic, params=gzm.make_params(box=box, density_budget=budget, **kw))
        if kw.get("precision", 0) == 0:
            for k in ("density_sweeps", "density_interactions", "gradient_interactions", "force_interactions"):
                assert st[k] == bst[k], (budget, k)
            assert st["density_interactions"] == rst.dens_interactions
            errs, bad = check_pass(out, ref)
            assert not bad, (budget, bad)
        else:
            es = error_stats(out, ref)
            for k in ("density", "hsml", "pressure"):
                assert es[k]["max"] <= TOL32_DENSITY, (budget, k, es[k])
            for k in ("acc", "dtu"):
                assert es[k]["max"] <= TOL32_FLUX, (budget, k, es[k])


def test_budgeted_graph_passes_equal_eager():
    ic = ics.batch_boxes(8, n_min=512, n_max=4096)
    with gzm.MFMContext(gzm.make_params(density_budget=16)) as ctx:
        ctx.upload(ic)
        ctx.restore_initial_state()
        s1 = ctx.density()
        ctx.gradients()
        ctx.hydro_force()
        eager = ctx.download()
        for rep in range(2):
            st = ctx.run_passes(2)
            assert st.density_interactions == 2 * s1.density_interactions
            out = ctx.download()
            for k in OUT_KEYS:
                np.testing.assert_array_equal(out[k], eager[k], err_msg=f"replay {rep} {k}")
