"""Kick, drift and timestep of the host mirror (gizmo_integrate.c) on CPU:
each update against its closed form, then a Sod shock tube integrated with
the host's kick / drift / timestep around the oracle's hydro pass, against
the exact solution.  Reference: kicks.c:255-402, predict.c:120-240,
timestep.c:71-77, 482-491, eos.c:58-61, 221-231."""
import math

import numpy as np
import pytest

import ics
import oracle as O
from host_mirror import fill, host_lib, read_state, views, write_state
from sod_exact import sod_profile, star_state

GAMMA = 5.0 / 3.0


@pytest.fixture()
def lib():
    lib = host_lib()
    yield lib
    lib.gizmo_host_release()
    lib.gizmo_host_free()


def small(lib, n=4):
    ic = ics.lattice_gaussian(n)
    fill(lib, ic)
    N = ic["pos"].shape[0]
    P, S, A, _ = views(lib, N)
    return ic, N, P, S, A


def test_kick_updates_velocity_and_energy(lib):
    ic, N, P, S, A = small(lib)
    rng = np.random.default_rng(1)
    acc = rng.normal(size=(N, 3))
    dtu = rng.normal(size=N)
    S["HydroAccel"] = acc
    S["DtInternalEnergy"] = dtu
    u0, v0 = S["InternalEnergy"].copy(), P["Vel"].copy()
    dt = 0.05
    lib.gizmo_kick_gas(dt, 0)
    np.testing.assert_array_equal(P["Vel"], v0 + acc * dt)
    np.testing.assert_array_equal(S["InternalEnergy"], u0 + dtu * dt)
    np.testing.assert_array_equal(S["VelPred"], ic["vel"])  # first half kick: predictions untouched
    lib.gizmo_kick_gas(dt, 1)
    np.testing.assert_array_equal(S["VelPred"], P["Vel"])
    np.testing.assert_array_equal(S["InternalEnergyPred"], S["InternalEnergy"])


def test_kick_energy_floors(lib):
    """A step below half the energy halves it (kicks.c:329); below MinEgySpec
    the energy is floored and DtInternalEnergy zeroed (eos.c:221-231)."""
    ic, N, P, S, A = small(lib)
    S["InternalEnergy"] = 1.0
    S["DtInternalEnergy"] = -30.0
    lib.gizmo_kick_gas(0.02, 0)                     # 1 - 0.6 = 0.4 < 0.5
    assert np.all(S["InternalEnergy"] == 0.5)
    A.MinEgySpec = 0.45
    S["DtInternalEnergy"] = -1.0
    lib.gizmo_kick_gas(0.1, 0)                      # 0.5 - 0.1 = 0.4 < 0.45
    assert np.all(S["InternalEnergy"] == 0.45) and np.all(S["DtInternalEnergy"] == 0)


def test_drift_moves_and_predicts(lib):
    ic, N, P, S, A = small(lib)
    rng = np.random.default_rng(2)
    P["Vel"] = rng.normal(size=(N, 3)) * 3.0
    S["HydroAccel"] = rng.normal(size=(N, 3))
    S["DtInternalEnergy"] = rng.normal(size=N) * 0.1
    divv = rng.normal(size=N) * 20.0  # some beyond the +-0.3 clamp at dt = 0.02
    P["Particle_DivVel"] = divv
    S["Density"] = 2.0
    x0, vp0, up0 = P["Pos"].copy(), S["VelPred"].copy(), S["InternalEnergyPred"].copy()
    dt = 0.02
    lib.gizmo_drift_gas(dt)
    np.testing.assert_allclose(P["Pos"], np.mod(x0 + P["Vel"] * dt, 1.0), rtol=0, atol=1e-15)
    assert np.all((P["Pos"] >= 0) & (P["Pos"] < 1))
    np.testing.assert_array_equal(S["VelPred"], vp0 + S["HydroAccel"] * dt)
    fac = np.clip(divv * dt, -0.3, 0.3)
    np.testing.assert_allclose(S["Density"], 2.0 * np.exp(-fac), rtol=1e-15)
    np.testing.assert_allclose(S["InternalEnergyPred"], up0 + S["DtInternalEnergy"] * dt, rtol=1e-15)
    np.testing.assert_allclose(S["Pressure"], (GAMMA - 1) * S["InternalEnergyPred"] * S["Density"], rtol=1e-15)


def test_courant_timestep(lib):
    """dt = CourantFac * 1.61199 h / NumNgb^(1/3) / (0.5 v_sig), at most
    MaxSizeTimestep; the global step is the smallest."""
    ic, N, P, S, A = small(lib)
    rng = np.random.default_rng(3)
    P["Hsml"] = rng.uniform(0.05, 0.2, N)
    P["NumNgb"] = 32.0 ** (1 / 3)
    S["MaxSignalVel"] = rng.uniform(1.0, 3.0, N)
    A.CourantFac, A.MaxSizeTimestep = 0.3, 1.0
    want = 0.3 * 1.61199 * P["Hsml"] / P["NumNgb"] / (0.5 * S["MaxSignalVel"])
    got = np.array([lib.gizmo_gas_timestep(i) for i in range(N)])
    np.testing.assert_allclose(got, want, rtol=1e-15)
    assert lib.gizmo_global_timestep() == pytest.approx(want.min(), rel=1e-15)
    A.MaxSizeTimestep = 0.5 * want.min()
    assert lib.gizmo_global_timestep() == A.MaxSizeTimestep


def test_exact_sod_profile_is_continuous():
    """The rarefaction fan meets the left state at its head and the star state
    at its tail; the star pressure and velocity are Toro's for these states."""
    ps, us = star_state(1.0, 1.0, 0.125, 0.1, 1.4)
    assert ps == pytest.approx(0.30313, rel=2e-5) and us == pytest.approx(0.92745, rel=2e-5)
    g = GAMMA
    ps, us = star_state(1.0, 1.0, 0.125, 0.1, g)
    cl = math.sqrt(g)
    tail = us - cl * (ps / 1.0) ** ((g - 1) / (2 * g))
    eps = 1e-9
    r0, u0, p0 = sod_profile(np.array([-cl - eps, -cl + eps, tail - eps, tail + eps]), 1.0, 1.0, 0.125, 0.1, g)
    np.testing.assert_allclose(r0[:2], 1.0, rtol=1e-7)
    np.testing.assert_allclose(r0[2], r0[3], rtol=1e-7)
    np.testing.assert_allclose(p0[2:], ps, rtol=1e-7)
    np.testing.assert_allclose(u0[2:], us, rtol=1e-7)


# ----------------------------------------------- Sod with the oracle pass ---
STATE_KEYS = ("hsml", "num_ngb", "dhsml", "divvel", "density", "pressure", "nvt", "cn", "fce",
              "grad_rho", "grad_p", "grad_v", "acc", "dtu", "vsig")


def oracle_forces(lib, N, box, gamma, first_step):
    """compute_hydro_densities_and_forces with the oracle as the pass: the
    inputs the host would upload (VelPred, InternalEnergyPred and the carried
    state), the outputs written back into P/SphP"""
    P, S, A, _ = views(lib, N)
    st = read_state(lib, N)
    ic = dict(pos=P["Pos"].copy(), vel=S["VelPred"].copy(), mass=P["Mass"].copy(),
              u=S["InternalEnergyPred"].copy(), **{k: st[k] for k in STATE_KEYS})
    params = O.make_params(box=tuple(box), gamma=gamma, first_step=int(first_step))
    out, _, _ = O.run_pass(ic, params=params)
    new = dict(ic)
    new.update({k: out[k] for k in STATE_KEYS})
    write_state(lib, new)


def run_sod_with_oracle(lib, ic, t_end, max_steps=400):
    """the loop of gizmo_run_until, with the oracle pass in place of the device"""
    N = ic["pos"].shape[0]
    box, g = ic["box"], ic["gamma"]
    P, S, A, _ = views(lib, N)
    oracle_forces(lib, N, box, g, True)
    steps = 0
    while A.Time < t_end and steps < max_steps:
        dt = lib.gizmo_global_timestep()
        last = A.Time + dt >= t_end
        if last:
            dt = t_end - A.Time
        lib.gizmo_kick_gas(0.5 * dt, 0)
        lib.gizmo_drift_gas(dt)
        A.Time = t_end if last else A.Time + dt
        oracle_forces(lib, N, box, g, False)
        lib.gizmo_kick_gas(0.5 * dt, 1)
        steps += 1
    return steps


def sod_setup(lib, n_left=32, n_cross=6):
    ic = ics.sod_tube(n_left=n_left, n_cross=n_cross)
    fill(lib, ic)
    N = ic["pos"].shape[0]
    P, S, A, _ = views(lib, N)
    A.GasGamma = ic["gamma"]
    A.CourantFac, A.MaxSizeTimestep = 0.2, 0.01
    return ic, N, P, S, A


def totals(P, S):
    m = P["Mass"]
    return (m.sum(), (m[:, None] * P["Vel"]).sum(0),
            (m * (S["InternalEnergy"] + 0.5 * (P["Vel"] ** 2).sum(1))).sum())


def sod_errors(P, S, t, gamma):
    """L1 errors of density, x-velocity and pressure against the exact
    solution, over particles between the two Riemann problems' fans
    (x in [0.3, 0.7], around the interface at x = 1/2), normalised by the
    mean exact value (velocity: by the star velocity)."""
    x = P["Pos"][:, 0]
    sel = (x > 0.3) & (x < 0.7)
    rho, u, p = sod_profile((x[sel] - 0.5) / t, 1.0, 1.0, 0.125, 0.1, gamma)
    _, us = star_state(1.0, 1.0, 0.125, 0.1, gamma)
    e_rho = np.mean(np.abs(S["Density"][sel] - rho)) / np.mean(rho)
    e_u = np.mean(np.abs(P["Vel"][sel, 0] - u)) / us
    e_p = np.mean(np.abs(S["Pressure"][sel] - p)) / np.mean(p)
    return e_rho, e_u, e_p


# L1 bounds of the Sod profiles at t = 0.1 with 64 particles across each half
# of the tube (measured: density 0.034, velocity 0.059, pressure 0.052; the
# pass smears the shock and contact over a few kernel lengths, and the
# unrelaxed lattice leaves a start-up bump where the fan begins)
SOD_L1 = {"density": 0.06, "velocity": 0.10, "pressure": 0.09}


def test_sod_shock_tube_with_oracle_pass(lib):
    """Sod (gamma 5/3) to t = 0.1: mass exact, momentum to rounding (the MFM
    fluxes are antisymmetric), total energy to the second-order error of the
    kick-drift-kick step; profiles within the L1 bounds; halving the particle
    spacing cuts the density error (the scheme converges to the exact
    solution)."""
    errs = {}
    for n_left, n_cross in ((32, 6), (64, 8)):
        ic, N, P, S, A = sod_setup(lib, n_left, n_cross)
        m0, p0, e0 = totals(P, S)
        steps = run_sod_with_oracle(lib, ic, 0.1)
        assert A.Time == 0.1 and 10 < steps < 400
        m1, p1, e1 = totals(P, S)
        assert m1 == m0
        assert np.abs(p1).max() <= 1e-12 * e0
        assert abs(e1 - e0) <= 1e-3 * e0
        errs[n_left] = sod_errors(P, S, 0.1, ic["gamma"])
        print(n_left, f"steps={steps} L1 rho, u, p = {errs[n_left]} dE/E={(e1 - e0) / e0:.2e}")
        lib.gizmo_host_free()
    e_rho, e_u, e_p = errs[64]
    assert e_rho < SOD_L1["density"] and e_u < SOD_L1["velocity"] and e_p < SOD_L1["pressure"]
    assert errs[64][0] < 0.7 * errs[32][0]
