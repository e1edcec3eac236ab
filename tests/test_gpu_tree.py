"""Tree neighbour search (gzm_params.search = 1) against the cell grid and the
FP64 oracle.

The tree changes which candidates a particle tests, never which pairs pass
the exact predicates, so neighbour counts and interaction totals must equal
the grid's and the oracle's exactly; the outputs are held to the FP64 parity
bar (the visiting order, hence the summation order, differs from the grid's).
A replayed graph pass must equal the eager tree pass bit for bit.
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
    "blast_16": lambda: ics.blast(16),
    "random_state_12": lambda: ics.random_state(12),
    "lattice_gaussian_16": lambda: ics.lattice_gaussian(16),
    "perfect_lattice_12": lambda: ics.perfect_lattice(12, velocity=(0.2, 0.0, -0.1)),
    "plummer_clusters_32k": lambda: ics.plummer_clusters(32768, 50),
}


def params_for(ic, **kw):
    return gzm.make_params(box=tuple(ic["box"]) if "box" in ic else (1.0, 1.0, 1.0), **kw)


@pytest.mark.parametrize("name", ["blast_16", "random_state_12", "plummer_clusters_32k"])
def test_tree_neighbour_counts_equal_grid(name):
    ic = CASES[name]()
    for mode in (0, 1):
        with gzm.MFMContext(params_for(ic)) as g:
            g.upload(ic)
            cg, tg = g.neighbor_counts(mode)
        with gzm.MFMContext(params_for(ic, search=1)) as t:
            t.upload(ic)
            ct, tt = t.neighbor_counts(mode)
        print(name, mode, tg, tt)
        assert tt == tg
        np.testing.assert_array_equal(ct, cg)


def test_tree_deep_equal_key_cluster_counts_equal_grid():
    """A tree 73 levels deep: 600 of 1728 particles moved into a cube of side
    2^-23 at (0.5, 0.5, 0.5), below the 2^-21 resolution of the Morton keys,
    so that they share one key and are split by slot; and 63 particles at
    0.5 + 0.75 * 2^-k along one axis at a time (k = 1..21), each branching off
    at its own key bit above the cluster.  Every node on that chain holds the
    cluster, so the bounds pass completes them one combine launch after
    another; the tree's neighbour counts still equal the grid's exactly."""
    ic = dict(ics.random_state(12))
    rng = np.random.default_rng(11)
    pos = ic["pos"].copy()
    pos[:600] = 0.5 + rng.uniform(0.0, 2.0 ** -23, size=(600, 3))
    q = 600
    for k in range(1, 22):
        for ax in range(3):
            pos[q] = 0.5
            pos[q, ax] += 0.75 * 2.0 ** -k
            q += 1
    ic["pos"] = pos
    for mode in (0, 1):
        with gzm.MFMContext(params_for(ic)) as g:
            g.upload(ic)
            cg, tg = g.neighbor_counts(mode)
        with gzm.MFMContext(params_for(ic, search=1)) as t:
            t.upload(ic)
            ct, tt = t.neighbor_counts(mode)
        assert tt == tg
        np.testing.assert_array_equal(ct, cg)
        assert cg[:600].min() >= 599


@pytest.mark.parametrize("name", list(CASES))
def test_tree_pass_matches_oracle(name):
    ic = CASES[name]()
    ref, st, _ = O.run_pass(ic)
    out, gst = gzm.run_pass(ic, params=params_for(ic, search=1))
    print(name, {k: v for k, v in gst.items() if k != "timing"})
    assert gst["density_sweeps"] == st.iterations
    assert gst["density_interactions"] == st.dens_interactions
    assert gst["gradient_interactions"] == st.grad_interactions
    assert gst["force_interactions"] == st.force_interactions
    assert gst["exact_solver_calls"] == st.exact_solver_calls
    errs, bad = check_pass(out, ref)
    print(name, {k: f"{v:.2e}" for k, v in errs.items()})
    assert not bad, bad


def test_tree_partial_active_pass():
    ic = ics.random_state(14)
    N = ic["pos"].shape[0]
    full, _, _ = O.run_pass(ic)
    state = dict(ic)
    for k in OUT_KEYS:
        state[k] = full[k]
    active = ics.random_subset(N, 0.1)
    ref, st, _ = O.run_pass(state, active=active)
    out, gst = gzm.run_pass(state, active=active, params=params_for(ic, search=1))
    assert gst["density_interactions"] == st.dens_interactions
    assert gst["gradient_interactions"] == st.grad_interactions
    assert gst["force_interactions"] == st.force_interactions
    errs, bad = check_pass(out, ref, active=active)
    assert not bad, bad
    np.testing.assert_array_equal(out["wakeup"], ref["wakeup"])


def test_tree_batch_matches_oracle_and_graph_replays():
    ic = ics.batch_boxes(8, n_min=512, n_max=4096)
    params = gzm.make_params(search=1)
    with gzm.MFMContext(params) as ctx:
        ctx.upload(ic)
        ctx.restore_initial_state()
        ctx.density()
        ctx.gradients()
        ctx.hydro_force()
        eager = ctx.download()
        for rep in range(2):
            ctx.run_passes(2)
            out = ctx.download()
            for k in OUT_KEYS:
                np.testing.assert_array_equal(out[k], eager[k], err_msg=f"replay {rep} {k}")
    offs = ic["box_offsets"]
    for b in range(len(offs) - 1):
        ref, _, _ = O.run_pass(ics.box_slice(ic, b))
        sl = slice(int(offs[b]), int(offs[b + 1]))
        errs, bad = check_pass({k: eager[k][sl] for k in OUT_KEYS}, ref)
        assert not bad, (b, bad)


def test_tree_fp32_pass_within_bounds():
    """FP32-bulk with the tree: neighbour sets equal the FP64 oracle's, outputs
    within the FP32 bounds, on the clustered set where grid-anchored FP32
    offsets would be too coarse (every tree pair takes the FP64 geometry)."""
    ic = CASES["plummer_clusters_32k"]()
    ref, st, _ = O.run_pass(ic)
    out, gst = gzm.run_pass(ic, params=params_for(ic, search=1, precision=1))
    assert gst["density_interactions"] == st.dens_interactions
    assert gst["gradient_interactions"] == st.grad_interactions
    assert gst["force_interactions"] == st.force_interactions
    es = error_stats(out, ref)
    print({k: v for k, v in es.items()})
    for k in ("density", "hsml", "pressure"):
        assert es[k]["max"] <= TOL32_DENSITY, (k, es[k])
    for k in ("acc", "dtu"):
        assert es[k]["max"] <= TOL32_FLUX, (k, es[k])


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name", ["plummer_clusters_32k", "blast_16"])
def test_tree_candidate_rows_give_the_walk_bits(name, precision):
    """A tree pass with FP64 density sums (precisions 0 and 1), all particles
    active, takes each gradient pair set from the density stage's candidate
    rows merged with the symmetric pair rows (no tree walk).  Gradients run on
    the density outputs uploaded into a fresh context have no rows and walk
    the tree; both must visit the same pairs in the same (ascending slot)
    order, so gradients, the pair lists and every force output agree bit for
    bit."""
    ic = CASES[name]()
    prm = params_for(ic, search=1, precision=precision)
    rows, st_rows = gzm.run_pass(ic, params=prm)
    state = dict(ic)
    for k in ("hsml", "num_ngb", "dhsml", "divvel", "density", "pressure", "nvt", "cn", "fce"):
        state[k] = rows[k]
    walk, st_walk = gzm.run_pass(state, params=prm, stages=("gradients", "force"))
    assert st_walk["gradient_interactions"] == st_rows["gradient_interactions"]
    assert st_walk["force_interactions"] == st_rows["force_interactions"]
    for k in ("grad_rho", "grad_p", "grad_v", "acc", "dtu", "vsig"):
        np.testing.assert_array_equal(walk[k], rows[k], err_msg=k)
