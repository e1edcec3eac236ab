"""Parity of the HIP meshless finite-volume (MFV) pass (gzm_params.method = 1)
with the FP64 oracle's MFV restatement.

Bar: interaction, retry and exact-solver counts equal; density, hsml,
pressure, signal speed within 1e-12 relative; ParticleVel, gradients, dv/dt,
du/dt, DtMass, dMass and GravWorkTerm within 1e-10 of the sum of |terms|
(tests/parity.py).  MFV has no entropic switch, so no particle is exempt.
"""
import numpy as np
import pytest

import gzm
import ics
import oracle as O
from parity import check_mfv

pytestmark = pytest.mark.gpu

STEPS = 2 ** 12
TIMEBASE = 2.0 ** -20  # dt = 2^-8 for the equal-step sets


def with_steps(ic, ti=None):
    ic = dict(ic)
    n = ic["pos"].shape[0]
    ic["ti_step"] = np.full(n, STEPS, dtype=np.int64) if ti is None else np.asarray(ti, dtype=np.int64)
    return ic


def mixed_steps(n, seed=9):
    """Integer timesteps 2^10, 2^11 or 2^12 (a hierarchy of three time bins)."""
    return (2 ** (10 + np.floor(3 * ics.uniform(seed, 3, n)))).astype(np.int64)


CONFIGS = {
    "lattice_gaussian_16": lambda: with_steps(ics.lattice_gaussian(16)),
    "random_state_12": lambda: with_steps(ics.random_state(12)),
    "blast_16": lambda: with_steps(ics.blast(16)),
    "random_state_14_mixed_steps": lambda: with_steps(ics.random_state(14), mixed_steps(14 ** 3)),
    "perfect_lattice_12_moving": lambda: with_steps(ics.perfect_lattice(12, velocity=(0.2, 0.0, -0.1))),
    # supersonic random flows: the first-order retry and the exact solver with
    # its face sampling (~90 and ~1150 exact-solver pairs)
    "random_state_12_supersonic": lambda: with_steps(ics.random_state(12, sigma_v=3.0)),
    "random_state_12_cold_supersonic": lambda: with_steps(_cold(ics.random_state(12, sigma_v=3.0))),
}


def _cold(ic, factor=0.01):
    ic = dict(ic)
    ic["u"] = ic["u"] * factor
    return ic

MFV_KEYS = ("particle_vel", "dt_mass", "d_mass", "grav_work")


def oracle_params(ic, **kw):
    return O.make_params(box=tuple(ic["box"]), method=O.METHOD_MFV, timebase_interval=TIMEBASE, **kw)


def gpu_params(ic=None, **kw):
    box = tuple(ic["box"]) if ic is not None and "box" in ic else (1.0, 1.0, 1.0)
    return gzm.make_params(box=box, method=gzm.METHOD_MFV, timebase_interval=TIMEBASE, **kw)


def check_counts(gst, st):
    assert gst["density_sweeps"] == st.iterations
    assert gst["density_interactions"] == st.dens_interactions
    assert gst["gradient_interactions"] == st.grad_interactions
    assert gst["force_interactions"] == st.force_interactions
    assert gst["exact_solver_calls"] == st.exact_solver_calls
    assert gst["riemann_retries"] == st.riemann_retries


@pytest.mark.parametrize("name", list(CONFIGS))
def test_mfv_pass_matches_oracle(name):
    ic = CONFIGS[name]()
    print(f"mfv pass {name}", flush=True)
    ref, st, _ = O.run_pass(ic, oracle_params(ic))
    out, gst = gzm.run_pass(ic, params=gpu_params(ic))
    print(name, {k: v for k, v in gst.items() if k != "timing"})
    check_counts(gst, st)
    errs, bad = check_mfv(out, ref)
    print(name, {k: f"{v:.2e}" for k, v in errs.items()})
    assert not bad, bad
    total = np.abs(out["d_mass"]).sum()
    assert abs(out["d_mass"].sum()) <= 1e-12 * max(total, 1e-300)  # the exchange conserves mass
    if "supersonic" in name:
        assert gst["exact_solver_calls"] > 0 and gst["riemann_retries"] > 0


def test_mfv_partial_active_set_with_mixed_steps():
    """10 % active: inactive neighbours enter with their stored state and
    receive their share of dMass on top of the value they carry."""
    ic = with_steps(ics.random_state(14), mixed_steps(14 ** 3, seed=4))
    N = ic["pos"].shape[0]
    full, _, _ = O.run_pass(ic, oracle_params(ic))
    state = dict(ic)
    for k in ("hsml", "density", "pressure", "num_ngb", "dhsml", "divvel", "cn", "fce", "nvt", "grad_rho",
              "grad_p", "grad_v", "acc", "dtu", "vsig", "particle_vel", "dt_mass", "grav_work"):
        state[k] = full[k]
    state["d_mass"] = 1e-3 * full["d_mass"]  # a carried value that inactive particles keep
    active = ics.random_subset(N, 0.1)
    print("mfv partial active", flush=True)
    ref, st, _ = O.run_pass(state, oracle_params(ic), active=active)
    out, gst = gzm.run_pass(state, active=active, params=gpu_params(ic))
    check_counts(gst, st)
    errs, bad = check_mfv(out, ref, active=active)
    print({k: f"{v:.2e}" for k, v in errs.items()})
    assert not bad, bad
    inactive = np.setdiff1d(np.arange(N), active)
    np.testing.assert_array_equal(out["particle_vel"][inactive], state["particle_vel"][inactive])
    np.testing.assert_array_equal(out["wakeup"], ref["wakeup"])
    assert np.abs(out["d_mass"][inactive] - state["d_mass"][inactive]).max() > 0  # shares received


@pytest.mark.parametrize("variant", ["tree", "no_list", "short_list"])
def test_mfv_search_and_list_variants(variant):
    """Tree search, no pair list (every particle scans), and a pair list too
    short for most pair sets (those particles scan in the same kernel)."""
    ic = with_steps(ics.plummer_clusters(16384, 30) if variant == "tree" else ics.random_state(12))
    kw = {"tree": dict(search=1), "no_list": dict(pair_list_cap=-1), "short_list": dict(pair_list_cap=16)}[variant]
    print(f"mfv variant {variant}", flush=True)
    ref, st, _ = O.run_pass(ic, oracle_params(ic))
    out, gst = gzm.run_pass(ic, params=gpu_params(ic, **kw))
    check_counts(gst, st)
    errs, bad = check_mfv(out, ref)
    assert not bad, bad


def test_mfv_batch_and_graph_passes():
    """Four boxes in one context; passes in one HIP graph repeat the eager
    pass (dMass to its summation-order rounding, the rest bit for bit)."""
    ic = ics.batch_boxes(4, n_min=1024, n_max=4096)
    ic["ti_step"] = np.full(ic["pos"].shape[0], STEPS, dtype=np.int64)
    keys = ("hsml", "density", "acc", "dtu", "vsig", "grad_v", "particle_vel", "dt_mass", "grav_work")
    print("mfv batch + graph", flush=True)
    with gzm.MFMContext(gpu_params()) as ctx:
        ctx.upload(ic)
        ctx.restore_initial_state()
        ctx.density()
        ctx.gradients()
        ctx.hydro_force()
        eager = ctx.download()
        for rep in range(2):
            ctx.run_passes(2)
            out = ctx.download()
            for k in keys:
                np.testing.assert_array_equal(out[k], eager[k], err_msg=f"replay {rep} {k}")
            scale = np.abs(eager["d_mass"]).max()
            assert np.abs(out["d_mass"] - eager["d_mass"]).max() <= 1e-13 * scale
    offs = ic["box_offsets"]
    for b in range(len(offs) - 1):
        box = ics.box_slice(ic, b)
        ref, _, _ = O.run_pass(box, oracle_params(box))
        sl = slice(int(offs[b]), int(offs[b + 1]))
        errs, bad = check_mfv({k: np.asarray(v)[sl] for k, v in eager.items()}, ref)
        assert not bad, (b, bad)


def test_mfv_needs_fp64():
    with pytest.raises(gzm.GzmError, match="FP64"):
        gzm.MFMContext(gzm.make_params(method=gzm.METHOD_MFV, precision=1))


def _element_states(seed, n):
    """Random face states: mild, supersonic and strongly separating flows,
    with random unit normals."""
    rng = np.random.default_rng(seed)
    L = np.column_stack([rng.uniform(0.1, 3.0, n), rng.uniform(0.05, 5.0, n), rng.normal(0, 2.0, (n, 3))])
    R = np.column_stack([rng.uniform(0.1, 3.0, n), rng.uniform(0.05, 5.0, n), rng.normal(0, 2.0, (n, 3))])
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    return L, R, nrm


def _vacuum_states():
    """One side empty (left or right, gas moving either way), and gas flying
    apart fast enough to open a vacuum inside the fan."""
    gas = [(1.0, 1.0, u, 0.3, -0.2) for u in (-3.5, -1.0, 0.0, 0.5, 1.5, 4.0)]
    vac = (0.0, 0.0, 0.0, 0.0, 0.0)
    L = [g for g in gas] + [vac] * len(gas) + [(1.0, 0.5, -5.0, 0.1, 0.0), (0.5, 1.0, -2.0, 0.0, 0.4)]
    R = [vac] * len(gas) + [(g[0], g[1], -g[2], g[3], g[4]) for g in gas] + [(1.0, 0.5, 5.0, 0.0, 0.2),
                                                                           (0.8, 0.3, 6.0, 0.2, 0.0)]
    n = len(L)
    nrm = np.tile([1.0, 0.0, 0.0], (n, 1))
    return np.asarray(L, dtype=np.float64), np.asarray(R, dtype=np.float64), nrm


def _oracle_elements(L, R, nrm, mode, limiter=None):
    out = np.zeros((L.shape[0], 7))
    mask = np.zeros(L.shape[0], dtype=np.int32)
    for k in range(L.shape[0]):
        if mode == 0:
            r = O.riemann_mfv(L[k], R[k], nrm[k], float(limiter[k]))
            mask[k] = r["mask"]
        elif mode == 1:
            r = O.riemann_kt_mfv(L[k], R[k], nrm[k])
            mask[k] = 2
        else:
            r = O.riemann_exact_mfv(L[k], R[k], nrm[k])
            mask[k] = -r["mask"]
        out[k] = [r["P_M"], r["S_M"], r["Frho"], *r["Fv"], r["Fp"]]
    return out, mask


def _close(a, b, tol=1e-12):
    scale = np.maximum(1.0, np.abs(b))
    both_nan = np.isnan(a) & np.isnan(b)
    return np.where(both_nan, True, np.abs(a - b) <= tol * scale)


@pytest.mark.parametrize("case", ["hllc", "exact_past_limiter", "kt", "exact_alone", "vacuum_solver",
                                  "vacuum_exact"])
def test_device_mfv_riemann_elements_match_oracle(case):
    """The MFV Riemann solver of the force loop, one face state per thread,
    against the oracle's: every stage, including the exact solver's vacuum
    fans and the vacuum inside the fan, which particle sets rarely reach."""
    if case.startswith("vacuum"):
        L, R, nrm = _vacuum_states()
    else:
        L, R, nrm = _element_states({"hllc": 1, "exact_past_limiter": 2, "kt": 3, "exact_alone": 4}[case], 2000)
    n = L.shape[0]
    mode = {"hllc": 0, "exact_past_limiter": 0, "kt": 1, "exact_alone": 2, "vacuum_solver": 0, "vacuum_exact": 2}[case]
    lim = np.full(n, 1e300) if case == "hllc" else np.zeros(n)
    print(f"mfv riemann elements {case}: {n} states, mode {mode}", flush=True)
    out, mask = gzm.riemann_states_mfv(L, R, nrm, lim if mode == 0 else None, mode=mode)
    ref, rmask = _oracle_elements(L, R, nrm, mode, lim)
    np.testing.assert_array_equal(mask, rmask)
    ok = _close(out, ref)
    assert ok.all(), (np.argwhere(~ok)[:5], out[~ok.all(axis=1)][:3], ref[~ok.all(axis=1)][:3])
    if case == "exact_past_limiter":
        assert np.all(mask & 4)
    if case == "hllc":
        assert np.all(mask & 1)
