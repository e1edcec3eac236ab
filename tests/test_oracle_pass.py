"""Pins the oracle's three stages by brute force and by invariants that a
plausible mistake (wrong sign, transposed index, dropped term) would break."""
import math

import numpy as np
import pytest

import ics
import oracle as O

NORM_COEFF = 4.188790204786  # allvars.h:1443


def brute_neighbors(pos, hsml, mode, box=1.0):
    """Independent O(N^2) neighbour sets with the reference's FP64 predicate."""
    N = pos.shape[0]
    out = []
    for i in range(N):
        d = pos[i] - pos
        d = np.where(d > 0.5 * box, d - box, np.where(d < -0.5 * box, d + box, d))
        r2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
        if mode == 0:
            sel = r2 < hsml[i] * hsml[i]
        else:
            sel = (r2 > 0) & ~((r2 >= hsml[i] * hsml[i]) & (r2 >= hsml * hsml))
        out.append(np.nonzero(sel)[0])
    return out


@pytest.mark.parametrize("mode", [0, 1])
def test_grid_neighbor_sets_equal_brute_force(mode):
    rng = np.random.default_rng(5)
    N = 1500
    pos = rng.uniform(0, 1, size=(N, 3))
    pos[:40] = rng.uniform(0, 0.02, size=(40, 3))          # dense clump at the corner seam
    pos[40:60] = np.mod(pos[40:60] * 1e-3 + 0.999, 1.0)     # just below the far faces
    pos[60] = pos[61]                                       # coincident pair
    hsml = rng.uniform(0.03, 0.12, size=N)
    ref = brute_neighbors(pos, hsml, mode)
    for brute, tree in ((0, 0), (1, 0), (0, 1)):
        offs, idx = O.neighbor_lists(pos, hsml, mode, params=O.make_params(brute_force=brute, tree=tree))
        for i in range(N):
            np.testing.assert_array_equal(idx[offs[i]:offs[i + 1]], ref[i])


@pytest.mark.parametrize("maker", [lambda: ics.random_state(12), lambda: ics.plummer_clusters(6000, 6)])
def test_octree_search_pass_equals_grid_pass(maker):
    """The octree walk (the reference's search, CPU baseline mode) finds the
    same sets as the grid, so the pass takes the same sweeps and interaction
    counts; outputs agree to summation order."""
    from parity import check_pass

    ic = maker()
    ref, st, _ = O.run_pass(ic)
    out, st2, _ = O.run_pass(ic, params=O.make_params(tree=1))
    assert (st2.iterations, st2.dens_interactions, st2.grad_interactions, st2.force_interactions) == \
           (st.iterations, st.dens_interactions, st.grad_interactions, st.force_interactions)
    errs, bad = check_pass(out, ref)
    assert not bad, bad


def _pass(ic, **kw):
    out, st, _ = O.run_pass(ic, **kw)
    return out, st


def test_hsml_converges_to_neighbour_target():
    for ic in (ics.lattice_gaussian(12), ics.random_state(10), ics.blast(12)):
        out, st = _pass(ic, stages=("density",))
        neff = out["num_ngb"] ** 3
        # density.c:659-661 with the first-step tolerance 0.05 (or a closed bracket)
        assert np.all(np.abs(neff - 32.0) <= 0.05 + 1e-9), (ic["name"], neff.min(), neff.max())
        # final density: own mass times the kernel number density (density.c:968)
        np.testing.assert_allclose(out["density"], ic["mass"] * neff / (NORM_COEFF * out["hsml"] ** 3), rtol=1e-14)
        np.testing.assert_allclose(out["pressure"], (5.0 / 3.0 - 1.0) * ic["u"] * out["density"], rtol=1e-15)


def test_neff_is_the_kernel_sum_at_the_final_hsml():
    ic = ics.random_state(8)
    out, st = _pass(ic, stages=("density",))
    h = out["hsml"]
    nb = brute_neighbors(ic["pos"], h, 0)
    for i in range(0, ic["pos"].shape[0], 7):
        d = ic["pos"][i] - ic["pos"][nb[i]]
        d = np.where(d > 0.5, d - 1, np.where(d < -0.5, d + 1, d))
        q = np.sqrt((d * d).sum(1)) / h[i]
        w = np.where(q < 0.5, 1 + 6 * (q - 1) * q * q, 2 * (1 - q) ** 3) * (8 / math.pi) / h[i] ** 3
        assert out["num_ngb"][i] ** 3 == pytest.approx(NORM_COEFF * h[i] ** 3 * w.sum(), rel=1e-12)


def test_perfect_lattice_is_in_equilibrium():
    ic = ics.perfect_lattice(10, velocity=(0.3, -0.2, 0.1))
    out, st = _pass(ic)
    h = out["hsml"][0]
    P = out["pressure"][0]
    m = ic["mass"][0]
    # every particle sees the same neighbourhood
    assert np.ptp(out["hsml"]) == 0 and np.ptp(out["density"]) <= 1e-14 * out["density"][0]
    nvt = out["nvt"].reshape(-1, 3, 3)
    off = nvt - np.einsum("ij,ni->nij", np.eye(3), np.einsum("nii->ni", nvt))
    assert np.abs(off).max() <= 1e-12 * np.abs(nvt).max()
    np.testing.assert_allclose(np.einsum("nii->ni", nvt), nvt[0, 0, 0], rtol=1e-12)
    assert np.abs(out["fce"]).max() < 1e-12
    # uniform state moving with a bulk velocity: no force, no heating (Galilean invariance)
    scale = P * h * h / m
    assert np.abs(out["acc"]).max() < 1e-11 * scale
    assert np.abs(out["dtu"]).max() < 1e-11 * scale * 0.4
    assert np.abs(out["grad_p"]).max() < 1e-11 * P / h


def test_linear_fields_have_exact_mls_gradients():
    ic = ics.lattice_gaussian(14)
    out, st = _pass(ic, stages=("density",))
    g_p = np.array([0.3, -0.7, 0.2])
    G_v = np.array([[0.1, 0.2, -0.3], [0.0, -0.4, 0.5], [0.25, 0.05, 0.0]])
    x = ic["pos"]
    state = dict(ic)
    state.update({k: out[k] for k in ("hsml", "num_ngb", "dhsml", "nvt", "cn", "fce")})
    state["density"] = out["density"]
    state["pressure"] = 10.0 + x @ g_p
    state["vel"] = 5.0 + x @ G_v.T
    gout, _ = _pass(state, stages=("gradients",))
    # interior particles: the kernel does not straddle the periodic seam of the field
    hmax = out["hsml"].max()
    inner = np.all((x > 1.05 * hmax) & (x < 1 - 1.05 * hmax), axis=1)
    assert inner.sum() > 100
    np.testing.assert_allclose(gout["grad_p"][inner], np.tile(g_p, (inner.sum(), 1)), rtol=0, atol=1e-10)
    np.testing.assert_allclose(gout["grad_v"][inner], np.tile(G_v, (inner.sum(), 1, 1)), rtol=0, atol=1e-10)


@pytest.mark.parametrize("maker", [lambda: ics.random_state(9), lambda: ics.blast(10), lambda: ics.lattice_gaussian(10)])
def test_momentum_and_energy_conservation(maker):
    ic = maker()
    out, st = _pass(ic)
    m = ic["mass"]
    p = (m[:, None] * out["acc"]).sum(0)
    e = (m * (out["dtu"] + (ic["vel"] * out["acc"]).sum(1))).sum()
    assert np.abs(p).max() <= 1e-13 * np.abs(m[:, None] * out["acc"]).sum()
    assert abs(e) <= 1e-13 * (np.abs(m * out["dtu"]).sum() + np.abs(m[:, None] * ic["vel"] * out["acc"]).sum())


def test_pair_fluxes_are_antisymmetric():
    ic = ics.random_state(7)
    out, st = _pass(ic, stages=("density", "gradients"))
    state = dict(ic)
    state.update(out)
    S = O.OracleState(state)
    offs, idx = O.neighbor_lists(S.pos, S.hsml, 1)
    for i in range(0, S.n, 5):
        for j in idx[offs[i]:offs[i + 1]]:
            a, _ = O.force_pair(S, i, j)
            b, _ = O.force_pair(S, j, i)
            scale = np.abs(a[:3]).max() + 1e-300
            assert np.abs(a[:3] + b[:3]).max() <= 1e-12 * scale
            assert abs(a[3] + b[3]) <= 1e-12 * (abs(a[3]) + np.abs(a[:3]).max() * np.abs(ic["vel"]).max() + 1e-300)
            assert a[4] == pytest.approx(b[4], rel=1e-15)


def test_partial_active_set_only_updates_active_particles():
    ic = ics.random_state(8)
    full, _ = _pass(ic)
    state = dict(ic)
    state.update({k: full[k] for k in full if not k.endswith("_l1") and k != "wakeup"})
    rng = np.random.default_rng(2)
    active = np.sort(rng.choice(ic["pos"].shape[0], 50, replace=False)).astype(np.int32)
    state["hsml"] = full["hsml"] * (1 + 0.05 * np.sin(np.arange(ic["pos"].shape[0])))
    part, _ = _pass(state, active=active)
    inactive = np.setdiff1d(np.arange(ic["pos"].shape[0]), active)
    for k in ("hsml", "density", "grad_p", "acc", "dtu"):
        np.testing.assert_array_equal(part[k][inactive], state[k][inactive])
    assert part["wakeup"][active].sum() == 0


def test_maxiter_failure_is_reported():
    ic = ics.random_state(8)
    ic = dict(ic)
    ic["hsml"] = ic["hsml"] * 3.0
    with pytest.raises(RuntimeError, match="1155"):
        O.run_pass(ic, params=O.make_params(maxiter=1), stages=("density",))


def test_generators_of_the_larger_configs():
    """Shapes, masses and densities of the shear-layer, batched and clustered sets."""
    c = ics.shear_layers(32)
    n_in, n_out = 40 * 20 * 40, 32 * 16 * 32
    assert c["pos"].shape == (n_in + n_out, 3)
    m = c["mass"][0]
    assert np.all(c["mass"] == m)
    assert abs(n_out * m / 0.5 - 1.0) < 1e-12                     # rho_out = 1
    inside = np.abs(c["pos"][:, 1] - 0.5) < 0.25
    assert inside.sum() == n_in
    assert np.all(c["vel"][inside, 0] == 0.5) and np.all(c["vel"][~inside, 0] == -0.5)
    P = (2.0 / 3.0) * c["u"] * np.where(inside, n_in * m / 0.5, 1.0)
    assert np.allclose(P, 2.5, rtol=1e-14)
    assert ics.shear_layers.__defaults__[0] == 224                 # full size 16,832,596
    assert 282 * 141 * 282 + 224 * 112 * 224 == 16832596

    b = ics.batch_boxes(16, n_min=512, n_max=4096)
    offs = b["box_offsets"]
    assert offs[0] == 0 and offs[-1] == b["pos"].shape[0] and np.all(np.diff(offs) > 0)
    for k in range(16):
        sl = slice(offs[k], offs[k + 1])
        nb = offs[k + 1] - offs[k]
        assert round(nb ** (1 / 3)) ** 3 == nb and 512 <= nb <= 4096
        assert np.allclose(b["mass"][sl], 1.0 / nb) and np.ptp(b["u"][sl]) == 0
        one = ics.box_slice(b, k)
        assert one["pos"].shape[0] == nb and np.all(one["box"] == 1.0)
    assert np.all((b["pos"] >= 0) & (b["pos"] < 1))

    p = ics.plummer_clusters(4096, 20)
    assert p["pos"].shape == (4096, 3) and np.all((p["pos"] >= 0) & (p["pos"] < 1))
    # hsml = midpoint of the 32nd and 33rd neighbour distances (brute force)
    for i in (7, 1000, 4095):
        d = p["pos"] - p["pos"][i]
        d -= np.rint(d)
        r = np.sort(np.sqrt((d * d).sum(1)))[1:]
        assert abs(0.5 * (r[31] + r[32]) - p["hsml"][i]) < 1e-12
    assert p["hsml"].max() / p["hsml"].min() > 10
    act = ics.random_subset(1000, 0.1)
    assert act.size == 100 and np.all(np.diff(act) > 0)
