"""Edge cases of the C ABI on the GPU: empty active sets, empty and tiny boxes
in a batch, massless particles, and the argument checks that turn bad input
into a status instead of a kernel launch."""
import numpy as np
import pytest

import gzm
import ics
import oracle as O

pytestmark = pytest.mark.gpu

OUT_KEYS = ("hsml", "density", "pressure", "num_ngb", "dhsml", "divvel", "cn", "fce", "nvt",
            "grad_rho", "grad_p", "grad_v", "acc", "dtu", "vsig")


def test_empty_active_list_changes_nothing():
    ic = ics.blast(12)
    full, _ = gzm.run_pass(ic, params=gzm.make_params(box=tuple(ic["box"])))
    state = dict(ic)
    for k in OUT_KEYS:
        state[k] = full[k]
    empty = np.zeros(0, dtype=np.int32)
    for search in (0, 1):
        out, st = gzm.run_pass(state, active=empty, params=gzm.make_params(box=tuple(ic["box"]), search=search))
        assert st["density_interactions"] == 0 and st["force_interactions"] == 0
        for k in ("hsml", "density", "grad_v", "acc", "dtu"):
            np.testing.assert_array_equal(out[k], state[k])


@pytest.mark.parametrize("search", [0, 1])
def test_batch_with_empty_and_single_particle_boxes(search):
    """Boxes of 0 and 1 particles between ordinary ones: the ordinary boxes
    give their own results, the lone particle sees only itself (as in the
    oracle).  A lone particle's neighbour number does not grow with hsml, so
    its hsml grows until MaxHsml stops it (with the default 1e56 that takes
    more than MAXITER sweeps: endrun 1155, in the reference too); MaxHsml is
    0.4 here."""
    base = ics.batch_boxes(4, n_min=512, n_max=1000)
    offs = np.asarray(base["box_offsets"])
    sizes = np.asarray(base["box_sizes"], dtype=np.float64).reshape(-1, 3)
    parts = [ics.box_slice(base, b) for b in range(len(offs) - 1)]
    lone = {k: v[:1].copy() for k, v in parts[0].items() if isinstance(v, np.ndarray) and v.ndim >= 1
            and v.shape[:1] == parts[0]["pos"].shape[:1]}
    # order: box0, empty, box1, lone, box2, box3
    order = [parts[0], None, parts[1], lone, parts[2], parts[3]]
    box_sizes = [sizes[0], np.ones(3), sizes[1], np.ones(3), sizes[2], sizes[3]]
    keys = [k for k in parts[0] if isinstance(parts[0][k], np.ndarray) and parts[0][k].shape[:1] == parts[0]["pos"].shape[:1]]
    ic = {k: np.concatenate([p[k] for p in order if p is not None]) for k in keys}
    counts = [0 if p is None else p["pos"].shape[0] for p in order]
    ic["box_offsets"] = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    ic["box_sizes"] = np.asarray(box_sizes, dtype=np.float64)
    out, st = gzm.run_pass(ic, params=gzm.make_params(search=search, max_hsml=0.4))
    o = ic["box_offsets"]
    for b, p in enumerate(order):
        if p is None or p is lone:
            continue
        alone, _ = gzm.run_pass(p, params=gzm.make_params(box=tuple(p["box"]), search=search, max_hsml=0.4))
        sl = slice(int(o[b]), int(o[b + 1]))
        for k in ("hsml", "density", "num_ngb"):
            np.testing.assert_allclose(out[k][sl], alone[k], rtol=1e-12, atol=0, err_msg=f"box {b} {k}")
    k = int(o[3])
    lone_box = dict(lone, box=np.ones(3))
    ref, _, _ = O.run_pass(lone_box, params=O.make_params(box=(1.0, 1.0, 1.0), max_hsml=0.4))
    for key in ("hsml", "num_ngb", "density"):
        assert out[key][k] == pytest.approx(ref[key][0], rel=1e-12), key
    assert np.all(np.isfinite(out["acc"])) and np.all(np.isfinite(out["dtu"]))


def test_massless_particles_are_invisible():
    """Particles with m = 0 are neither updated nor seen as neighbours
    (density.c:85, hydro_evaluate.h:149-150): adding them at random positions
    leaves every other particle's outputs unchanged, bit for bit."""
    ic = ics.random_state(12)
    N = ic["pos"].shape[0]
    rng = np.random.default_rng(9)
    extra = 500
    ghost = {k: v.copy() for k, v in ic.items() if isinstance(v, np.ndarray) and v.shape[:1] == (N,)}
    ghost["pos"] = rng.uniform(0, 1, size=(extra, 3))
    ghost["vel"] = rng.normal(size=(extra, 3))
    ghost["mass"] = np.zeros(extra)
    ghost["u"] = np.ones(extra)
    ghost["hsml"] = np.full(extra, float(np.mean(ic["hsml"])))
    mixed = dict(ic)
    for k in ghost:
        mixed[k] = np.concatenate([ic[k], ghost[k][:extra] if ghost[k].shape[0] >= extra else ghost[k]])
    # a fixed cell side (not derived from the mean hsml of all particles) and no
    # candidate budget (the extra candidates must not move a particle to the
    # one-wavefront-per-particle launch, which sums in another order)
    prm = dict(box=tuple(ic["box"]), cell_size=float(np.mean(ic["hsml"])), density_budget=-1)
    a, sa = gzm.run_pass(ic, params=gzm.make_params(**prm))
    b, sb = gzm.run_pass(mixed, params=gzm.make_params(**prm))
    assert sa["density_interactions"] == sb["density_interactions"]
    assert sa["force_interactions"] == sb["force_interactions"]
    for k in OUT_KEYS:
        np.testing.assert_array_equal(b[k][:N], a[k], err_msg=k)


def test_bad_input_is_a_status_not_a_launch():
    ic = ics.blast(8)
    with gzm.MFMContext(gzm.make_params(box=tuple(ic["box"]))) as ctx:
        bad = dict(ic)
        bad["pos"] = ic["pos"].copy()
        bad["pos"][3, 1] = 1.5  # outside the unit box
        with pytest.raises(gzm.GzmError) as e:
            ctx.upload(bad)
        assert e.value.status == 1
        ctx.upload(ic)
        with pytest.raises(gzm.GzmError) as e:
            ctx.density(np.array([0, ic["pos"].shape[0]], dtype=np.int32))  # index out of range
        assert e.value.status == 1
        st = ctx.density()  # the context still works
        assert st.density_interactions > 0
    with pytest.raises(gzm.GzmError):
        gzm.MFMContext(gzm.make_params(precision=3))
    with pytest.raises(gzm.GzmError):
        gzm.MFMContext(gzm.make_params(search=2))
