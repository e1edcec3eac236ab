"""Pins the oracle's meshless finite-volume (MFV) restatement against what the
mathematics fixes, never against a retyping of its own expressions.

Reference functions covered (restated in oracle/mfm_oracle.c):
  HLLC four-wave fluxes                  hydro/reimann.h:603-675
  KT/PWK flux, MFV branch                hydro/reimann.h:484-507
  exact solver with the face sampling    hydro/reimann.h:685-1030, 1146-1166
  Riemann_solver, MFV                    hydro/reimann.h:286-388, 400-425
  MFV face frame, limiter, flux boost    hydro/hydro_core_meshless.h:93-94, 194-196, 285-295
  mass exchange between the pair         hydro/hydro_evaluate.h:379-401
  smoothed mesh velocity (motion 5)      hydro/density.c:291-293, 331-334, 886-894
  MFV final operations                   hydro/hydro_toplevel.c:661-665, 742-752

Closed forms: every solver returns the Euler flux of two equal states; the
exact solver's face state is Toro's self-similar solution (a Sod tube seen
from a moving frame, rarefactions into vacuum, Toro sections 4.4-4.6);
swapping the states and the normal negates every flux.  Invariants of the
whole pass: Galilean invariance, exact mass conservation of the exchange,
a lattice at rest stays at rest, a uniform flow keeps ParticleVel = v, and
ParticleVel against a brute-force kernel sum.
"""
import math

import numpy as np
import pytest

import ics
import oracle as O
from sod_exact import sod_profile, star_state

GAMMA = 5.0 / 3.0


def euler_flux(rho, p, v, n, gamma=GAMMA):
    """Mass, momentum and total-energy flux of one state through a face at rest."""
    v = np.asarray(v, dtype=np.float64)
    n = np.asarray(n, dtype=np.float64)
    vn = float(v @ n)
    E = 0.5 * rho * float(v @ v) + p / (gamma - 1)
    return rho * vn, rho * v * vn + p * n, vn * (E + p)


def check_flux(f, ref, scale, tol=1e-13):
    fr, fv, fp = ref
    assert abs(f["Frho"] - fr) <= tol * scale, (f, ref)
    assert np.abs(f["Fv"] - fv).max() <= tol * scale, (f, ref)
    assert abs(f["Fp"] - fp) <= tol * scale, (f, ref)


EQUAL_STATES = [
    (1.0, 1.0, (0.3, -0.2, 0.1)),     # subsonic
    (2.0, 0.5, (3.0, 2.0, 0.0)),      # supersonic along the normal
    (0.7, 2.0, (-0.5, 0.2, 0.7)),     # subsonic, flow against the normal
    (1.3, 0.2, (-2.0, -1.5, 0.4)),    # supersonic against the normal
    (1.0, 1.0, (0.0, 0.0, 0.0)),      # at rest: contact speed exactly 0
]


@pytest.mark.parametrize("state", EQUAL_STATES)
def test_equal_states_give_the_euler_flux(state):
    rho, p, v = state
    n = np.array([0.6, 0.8, 0.0])
    S = (rho, p, *v)
    ref = euler_flux(rho, p, v, n)
    scale = rho * (1 + np.dot(v, v)) + p
    check_flux(O.riemann_mfv(S, S, n, 1e300), ref, scale)   # HLLC with its fluxes
    check_flux(O.riemann_kt_mfv(S, S, n), ref, scale)
    check_flux(O.riemann_exact_mfv(S, S, n), ref, scale)


@pytest.mark.parametrize("u0", [-2.5, -1.0, -0.5, -0.2, 0.0, 0.3, 0.6, 0.9, 1.2, 2.5])
def test_exact_flux_samples_the_sod_solution_in_a_moving_frame(u0):
    """States moving at u0 along x with different tangential speeds: the face
    at rest samples the Sod solution at x/t = -u0 in the states' frame
    (left fan, both star states, shocked and undisturbed right gas)."""
    rl, pl, rr, pr = 1.0, 1.0, 0.125, 0.1
    rho, u, p = (x[0] for x in sod_profile(np.array([-u0]), rl, pl, rr, pr, GAMMA))
    _, us = star_state(rl, pl, rr, pr, GAMMA)
    vt = 0.3 if -u0 < us else -0.2  # tangential velocity of the side the face sees
    f = O.riemann_exact_mfv((rl, pl, u0, 0.3, 0.0), (rr, pr, u0, -0.2, 0.0), (1, 0, 0))
    assert f["mask"] == 0
    check_flux(f, euler_flux(rho, p, (u + u0, vt, 0.0), (1, 0, 0)), 1.0 + abs(u0) ** 3)


def vacuum_fan_state(rho, p, u, side, gamma=GAMMA):
    """Toro section 4.6: face state at x/t = 0 of a state next to vacuum on
    `side` ("right": the gas is on the left)."""
    c = math.sqrt(gamma * p / rho)
    s = 1.0 if side == "right" else -1.0  # direction of the expansion
    front = u + s * 2 * c / (gamma - 1)
    head = u - s * c
    if s * head >= 0:           # the fan has not reached the face: undisturbed gas
        return rho, u, p
    if s * front <= 0:          # vacuum at the face
        return 0.0, 0.0, 0.0
    cf = 2 / (gamma + 1) * (c + s * (gamma - 1) / 2 * u)
    uf = 2 / (gamma + 1) * (s * c + (gamma - 1) / 2 * u)
    return rho * (cf / c) ** (2 / (gamma - 1)), uf, p * (cf / c) ** (2 * gamma / (gamma - 1))


@pytest.mark.parametrize("u", [-3.5, -1.0, 0.0, 0.5, 1.5])
@pytest.mark.parametrize("side", ["right", "left"])
def test_exact_flux_of_a_rarefaction_into_vacuum(u, side):
    gas = (1.0, 1.0, u, 0.0, 0.0)
    vac = (0.0, 0.0, 0.0, 0.0, 0.0)
    L, R = (gas, vac) if side == "right" else (vac, gas)
    f = O.riemann_exact_mfv(L, R, (1, 0, 0))
    rho, uf, p = vacuum_fan_state(1.0, 1.0, u, side)
    check_flux(f, euler_flux(rho, p, (uf, 0.0, 0.0), (1, 0, 0)), 1.0 + abs(u) ** 3, tol=1e-12)


@pytest.mark.parametrize("seed", range(6))
def test_swapping_the_states_and_the_normal_negates_the_fluxes(seed):
    rng = np.random.default_rng(100 + seed)
    L = (rng.uniform(0.2, 3), rng.uniform(0.1, 4), *rng.normal(size=3))
    R = (rng.uniform(0.2, 3), rng.uniform(0.1, 4), *rng.normal(size=3))
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    for fn in (lambda a, b, m: O.riemann_mfv(a, b, m, 1e300), lambda a, b, m: O.riemann_mfv(a, b, m, 0.0),
               O.riemann_kt_mfv, O.riemann_exact_mfv):
        a, b = fn(L, R, n), fn(R, L, -n)
        scale = 10.0
        assert abs(a["Frho"] + b["Frho"]) <= 1e-13 * scale
        assert np.abs(a["Fv"] + b["Fv"]).max() <= 1e-13 * scale
        assert abs(a["Fp"] + b["Fp"]) <= 1e-13 * scale


def test_the_solver_falls_back_to_the_exact_flux_past_the_limiter():
    L, R = (1.0, 1.0, 0.2, 0.1, 0.0), (0.3, 0.4, -0.4, 0.0, 0.2)
    n = (1.0, 0.0, 0.0)
    full = O.riemann_mfv(L, R, n, 0.0)  # every estimate exceeds a zero limiter
    exact = O.riemann_exact_mfv(L, R, n)
    assert full["mask"] & 4
    for k in ("Frho", "Fp"):
        assert full[k] == exact[k]
    assert np.array_equal(full["Fv"], exact["Fv"])
    hllc = O.riemann_mfv(L, R, n, 1e300)
    assert hllc["mask"] == 1
    # HLLC approximates the same flux (within ~10 % across this jump)
    assert abs(hllc["Frho"] - exact["Frho"]) < 0.15 * abs(exact["Frho"])
    assert abs(hllc["Fp"] - exact["Fp"]) < 0.15 * abs(exact["Fp"])


# ------------------------------------------------------------ whole pass ---
def mfv_params(**kw):
    return O.make_params(method=O.METHOD_MFV, nthreads=4, **kw)


def with_steps(ic, ti):
    ic = dict(ic)
    ic["ti_step"] = np.asarray(ti, dtype=np.int64)
    return ic


def test_galilean_invariance_of_the_pass():
    """A uniform boost changes no acceleration, du/dt or mass flux.  The
    velocities keep one sign per axis in both frames: the reconstruction's
    sign-preserving limiter branches (reimann.h:199-203) depend on the sign
    of each velocity component and are not Galilean invariant."""
    ic = ics.random_state(10)
    N = ic["pos"].shape[0]
    ic = with_steps(ic, np.full(N, 4096))
    ic["vel"] = ic["vel"] + np.array([6.0, 5.0, 7.0])
    a, _, _ = O.run_pass(ic, mfv_params())
    ic2 = dict(ic)
    ic2["vel"] = ic["vel"] + np.array([0.7, 0.4, 1.1])
    b, _, _ = O.run_pass(ic2, mfv_params())
    for k, l1 in (("acc", "acc_l1"), ("dtu", "dtu_l1"), ("dt_mass", "dtmass_l1"), ("d_mass", "dmass_l1"),
                  ("grav_work", "gravwork_l1")):
        err = np.abs(b[k] - a[k]) / np.maximum(a[l1], 1e-300)
        assert err.max() < 1e-12, (k, err.max())
    assert np.abs(b["particle_vel"] - a["particle_vel"] - np.array([0.7, 0.4, 1.1])).max() < 1e-13


@pytest.mark.parametrize("steps", ["equal", "mixed"])
def test_mass_exchange_conserves_mass(steps):
    ic = ics.random_state(10)
    N = ic["pos"].shape[0]
    rng = np.random.default_rng(3)
    ti = np.full(N, 4096) if steps == "equal" else 2 ** rng.integers(10, 13, size=N)
    out, st, _ = O.run_pass(with_steps(ic, ti), mfv_params())
    assert st.force_interactions > 0
    total = np.abs(out["d_mass"]).sum()
    assert total > 0
    assert abs(out["d_mass"].sum()) < 1e-13 * total
    # particles with the smaller step give their neighbours the whole share,
    # equal steps split it; the mass fluxes themselves do not depend on steps
    ref, _, _ = O.run_pass(with_steps(ic, np.full(N, 4096)), mfv_params())
    assert np.array_equal(out["dt_mass"], ref["dt_mass"]) or np.abs(out["dt_mass"] - ref["dt_mass"]).max() < 1e-15


def test_a_lattice_at_rest_has_no_fluxes():
    ic = ics.perfect_lattice(8)
    out, _, _ = O.run_pass(with_steps(ic, np.full(ic["pos"].shape[0], 1024)), mfv_params())
    scale = out["acc_l1"].max()
    assert np.abs(out["acc"]).max() <= 1e-12 * scale
    assert np.abs(out["dt_mass"]).max() <= 1e-12 * out["dtmass_l1"].max() + 1e-300
    assert np.abs(out["particle_vel"]).max() == 0.0


def test_particle_velocity_of_a_uniform_flow_is_the_flow():
    ic = ics.random_state(8)
    V = np.array([0.25, -1.5, 0.75])
    ic["vel"] = np.tile(V, (ic["pos"].shape[0], 1))
    out, _, _ = O.run_pass(ic, mfv_params(), stages=("density",))
    assert np.abs(out["particle_vel"] - V).max() < 1e-14


def cubic_spline_w(q, h):
    """Cubic spline kernel W(r, h) in 3-d (Monaghan & Lattanzio 1985; support h)."""
    w = np.where(q < 0.5, 1 - 6 * q * q + 6 * q ** 3, np.where(q < 1, 2 * (1 - q) ** 3, 0.0))
    return 8 / math.pi * w / h ** 3


def test_particle_velocity_against_a_brute_force_sum():
    """0.7 v_i + 0.3 sum_j m_j W(r_ij, h_i) v_j / sum_j m_j W(r_ij, h_i) over
    the converged hsml, all pairs of a small periodic box."""
    ic = ics.random_state(6)
    out, _, _ = O.run_pass(ic, mfv_params(), stages=("density",))
    pos, vel, m, h = ic["pos"], ic["vel"], ic["mass"], out["hsml"]
    d = pos[:, None, :] - pos[None, :, :]
    d -= np.round(d)  # nearest image in the unit box
    r = np.sqrt((d * d).sum(axis=2))
    w = cubic_spline_w(r / h[:, None], h[:, None]) * m[None, :]
    ref = 0.7 * vel + 0.3 * (w @ vel) / w.sum(axis=1)[:, None]
    assert np.abs(out["particle_vel"] - ref).max() < 1e-12 * (1 + np.abs(vel).max())
