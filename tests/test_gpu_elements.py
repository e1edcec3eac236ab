"""The device Riemann solver (the code the force kernels inline) on
independent face states through gzm_riemann_states, against the oracle's
Riemann_solver and KT restatements, which tests/test_oracle_pins.py and
tests/test_oracle_elements.py hold to closed forms and Toro's tables.
Covers the branches that particle sets rarely reach: the Kurganov-Tadmor
flux (reimann.h:484-519, taken only on a NaN or negative HLLC pressure), the
Roe and PVRS estimates (reimann.h:548-590), the exact solver and its
zero-density exit (reimann.h:685-724)."""
import numpy as np
import pytest

import gzm
import oracle as O

pytestmark = pytest.mark.gpu

GAMMA = 5.0 / 3.0


def random_states(n, seed, vscale=0.5):
    rng = np.random.default_rng(seed)
    L = np.column_stack([rng.uniform(0.1, 3, n), rng.uniform(0.05, 5, n), vscale * rng.normal(size=(n, 3))])
    R = np.column_stack([rng.uniform(0.1, 3, n), rng.uniform(0.05, 5, n), vscale * rng.normal(size=(n, 3))])
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    return L, R, nrm


def oracle_rows(L, R, nrm, lim):
    out = np.zeros((len(L), 6))
    mask = np.zeros(len(L), dtype=np.int32)
    for k in range(len(L)):
        o = O.riemann(L[k], R[k], n=nrm[k], press_tot_limiter=lim[k])
        out[k] = [o["P_M"], o["S_M"], *o["Fv"], o["Fp"]]
        mask[k] = o["mask"]
    return out, mask


def test_kt_flux_matches_oracle_bit_for_bit():
    L, R, nrm = random_states(4000, 1, vscale=1.0)
    out, mask = gzm.riemann_states(L, R, nrm, mode=1)
    assert np.all(mask == 2)
    for k in range(len(L)):
        o = O.riemann_kt(L[k], R[k], n=nrm[k])
        np.testing.assert_array_equal(out[k], [o["P_M"], o["S_M"], *o["Fv"], o["Fp"]], err_msg=str(k))


def test_kt_flux_fp32_within_rounding():
    L, R, nrm = random_states(4000, 2, vscale=1.0)
    out, _ = gzm.riemann_states(L, R, nrm, mode=1, precision=1)
    ref = np.array([[o["P_M"], o["S_M"], *o["Fv"], o["Fp"]] for o in
                    (O.riemann_kt(L[k], R[k], n=nrm[k]) for k in range(len(L)))])
    scale = np.abs(ref[:, 0:1]) * (1 + np.abs(ref[:, 1:2]))  # P_M (1 + |S_M|): size of every output
    assert np.max(np.abs(out - ref) / scale) < 2e-5


@pytest.mark.parametrize("which", ["hllc", "roe", "pvrs", "exact"])
def test_riemann_solver_branches_match_oracle(which):
    """The limiter steers each state to one stage of the cascade: above the
    HLLC pressure (HLLC stands), between the HLLC and Roe pressures (Roe
    stands), below both with PVRS inside it (PVRS stands), or zero (every
    estimate fails it and the exact solver runs)."""
    L, R, nrm = random_states(3000, {"hllc": 3, "roe": 4, "pvrs": 5, "exact": 6}[which])
    lim = np.zeros(len(L))
    keep = np.zeros(len(L), bool)
    for k in range(len(L)):
        w = O.wavespeeds(L[k], R[k], n=nrm[k])
        if w["which"] == 0:
            continue
        e = [O.pstar_estimate(L[k], R[k], q, n=nrm[k]) for q in range(3)]
        ph, pr, pp = e[0]["P_M"], e[1]["P_M"], e[2]["P_M"]
        if which == "hllc":
            lim[k], keep[k] = 2.0 * ph, True
        elif which == "roe" and pr < ph:
            lim[k], keep[k] = 0.5 * (pr + ph), True
        elif which == "pvrs" and pp < min(ph, pr):
            lim[k], keep[k] = 0.5 * (pp + min(ph, pr)), True
        elif which == "exact":
            lim[k], keep[k] = 0.0, True
    L, R, nrm, lim = L[keep], R[keep], nrm[keep], lim[keep]
    assert len(L) > 200
    out, mask = gzm.riemann_states(L, R, nrm, limiter=lim)
    ref, rmask = oracle_rows(L, R, nrm, lim)
    np.testing.assert_array_equal(mask, rmask)
    want = 4 if which == "exact" else 0
    assert np.all((mask & 4) == want)
    if which == "exact":
        # the iterative solver uses pow(): device and host libm may differ by
        # an ulp, well inside the solver's own 1e-6 tolerance
        np.testing.assert_allclose(out[:, :2], ref[:, :2], rtol=1e-12, atol=1e-14)
        np.testing.assert_array_equal(out[:, 2:], ref[:, 2:])  # fluxes stay HLLC's (reimann.h:873-877)
    else:
        np.testing.assert_array_equal(out, ref)


def test_nan_states_reach_kt_and_the_exact_solver_exit():
    """Zero densities on both sides give infinite sound speeds and NaN in every
    estimate: the solver takes the KT flux (NaN as well) and then the exact
    solver's zero-density exit, P_M = S_M = 0 (reimann.h:705-707)."""
    L = np.array([[0.0, 1.0, 0.1, 0.0, 0.0], [0.0, 2.0, -0.3, 0.2, 0.0]])
    R = np.array([[0.0, 0.5, -0.1, 0.0, 0.0], [0.0, 1.0, 0.0, 0.0, 0.4]])
    nrm = np.array([[1.0, 0.0, 0.0], [0.0, 0.6, 0.8]])
    lim = np.full(2, 10.0)
    out, mask = gzm.riemann_states(L, R, nrm, limiter=lim)
    ref, rmask = oracle_rows(L, R, nrm, lim)
    assert np.all(rmask == 7) and np.all(mask == 7)
    np.testing.assert_array_equal(out[:, :2], 0.0)
    np.testing.assert_array_equal(np.isnan(out[:, 2:]), np.isnan(ref[:, 2:]))


def test_exact_solver_on_toro_table():
    """Toro's Table 4.3 star states (gamma = 1.4) from the device exact solver."""
    rows = [((1.0, 1.0, 0.0), (0.125, 0.1, 0.0), 0.30313, 0.92745),
            ((1.0, 0.4, -2.0), (1.0, 0.4, 2.0), 0.00189, 0.0),
            ((1.0, 1000.0, 0.0), (1.0, 0.01, 0.0), 460.894, 19.5975),
            ((1.0, 0.01, 0.0), (1.0, 100.0, 0.0), 46.0950, -6.19633),
            ((5.99924, 460.894, 19.5975), (5.99242, 46.0950, -6.19633), 1691.64, 8.68975)]
    L = np.array([[a[0], a[1], a[2], 0, 0] for a, _, _, _ in rows])
    R = np.array([[b[0], b[1], b[2], 0, 0] for _, b, _, _ in rows])
    nrm = np.tile([1.0, 0.0, 0.0], (len(rows), 1))
    out, mask = gzm.riemann_states(L, R, nrm, limiter=np.zeros(len(rows)), gamma=1.4)
    assert np.all(mask & 4)
    for k, (_, _, p, u) in enumerate(rows):
        assert out[k, 0] == pytest.approx(p, rel=2e-5, abs=1e-5)
        assert out[k, 1] == pytest.approx(u, rel=2e-5, abs=1e-5)
