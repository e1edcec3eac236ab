"""The C host mirror (libgizmo_host.so) as a drop-in for GIZMO's density(),
hydro_gradient_calc() and hydro_force() over P[]/SphP[] (proto.h:723,962,749),
driven through the reference's globals and active list, against the library
called directly and against the FP64 oracle."""
import numpy as np
import pytest

import gzm
import ics
import oracle as O
from host_mirror import codes, fill, host_lib, read_state, set_active, views
from parity import check_pass

pytestmark = pytest.mark.gpu

STATE_KEYS = ("hsml", "num_ngb", "dhsml", "divvel", "density", "pressure", "nvt", "cn", "fce",
              "grad_rho", "grad_p", "grad_v", "acc", "dtu", "vsig")


@pytest.fixture()
def lib():
    lib = host_lib()
    yield lib
    lib.gizmo_host_release()
    lib.gizmo_host_free()


def test_host_mirror_is_a_drop_in(lib):
    ic = ics.random_state(10)
    N = ic["pos"].shape[0]
    fill(lib, ic)
    before = len(codes)
    lib.compute_hydro_densities_and_forces()
    assert lib.gizmo_host_last_status() == 0 and len(codes) == before
    got = read_state(lib, N)
    out, _ = gzm.run_pass(ic)
    for k in STATE_KEYS:
        np.testing.assert_array_equal(got[k], np.asarray(out[k]).reshape(got[k].shape), err_msg=k)
    # the reference's failure path: endrun with the library's code; an
    # out-of-box position is rejected by the upload (invalid argument)
    P, _, _, _ = views(lib, N)
    P["Pos"][0, 0] = 2.0
    lib.density()
    assert codes[before:] == [1]


def test_host_mirror_partial_active_after_full_pass(lib):
    """Hierarchical timesteps: after a full pass, a pass over a random 10 % of
    a clustered set.  Inactive particles enter the active ones' gradient and
    force loops with the density, pressure, NV_T, gradients etc. that P/SphP
    hold (gradients.c:1477-2089, hydro_evaluate.h:146-498), so the active
    particles' results equal the oracle's on the same carried state, and the
    inactive particles' fields are left as they were."""
    ic = ics.plummer_clusters(32768, 50)
    N = ic["pos"].shape[0]
    fill(lib, ic)
    lib.compute_hydro_densities_and_forces()
    assert lib.gizmo_host_last_status() == 0
    s0 = read_state(lib, N)
    active = ics.random_subset(N, 0.1)
    set_active(lib, active)
    lib.compute_hydro_densities_and_forces()
    assert lib.gizmo_host_last_status() == 0
    s1 = read_state(lib, N)
    state = dict(ic)
    for k in STATE_KEYS:
        state[k] = s0[k]
    ref, st, _ = O.run_pass(state, active=active)
    errs, bad = check_pass(s1, ref, active=active)
    print({k: f"{v:.2e}" for k, v in errs.items()})
    assert not bad, bad
    np.testing.assert_array_equal(s1["wakeup"], ref["wakeup"])
    assert ref["wakeup"].sum() >= 0
    inactive = np.setdiff1d(np.arange(N), active)
    for k in STATE_KEYS:
        np.testing.assert_array_equal(s1[k][inactive], s0[k][inactive], err_msg=k)


def test_host_mirror_skips_converged_timebins(lib):
    """A particle whose TimeBin is negative on entry to density() has finished
    iterating (density_isactive, density.c:34): it is left out of the
    iteration, and its TimeBin is restored to -TimeBin - 1 on exit
    (density.c:869-872)."""
    ic = ics.random_state(10)
    N = ic["pos"].shape[0]
    fill(lib, ic)
    P, _, _, _ = views(lib, N)
    marked = np.arange(0, N, 9)
    P["TimeBin"] = 4
    P["TimeBin"][marked] = -5
    lib.density()
    assert lib.gizmo_host_last_status() == 0
    got = read_state(lib, N)
    assert np.all(got["timebin"] == 4)
    rest = np.setdiff1d(np.arange(N), marked).astype(np.int32)
    ref, _ = gzm.run_pass(ic, active=rest, stages=("density",))
    for k in ("hsml", "density", "pressure", "num_ngb", "cn", "nvt"):
        np.testing.assert_array_equal(got[k][rest], np.asarray(ref[k])[rest].reshape(got[k][rest].shape), err_msg=k)
    np.testing.assert_array_equal(got["hsml"][marked], ic["hsml"][marked])
    assert np.all(got["density"][marked] == 0)


def test_host_mirror_meshless_finite_volume(lib):
    """gizmo_host_set_method(1): the same entry points run the MFV pass, with
    the integer step 2^TimeBin of each particle and All.Timebase_interval,
    and fill ParticleVel, DtMass, dMass and GravWorkTerm as the library does
    when called directly."""
    ic = ics.random_state(10)
    N = ic["pos"].shape[0]
    bins = (10 + np.floor(3 * ics.uniform(3, 8, N))).astype(np.int32)
    fill(lib, ic)
    P, _, A, _ = views(lib, N)
    P["TimeBin"] = bins
    A.Timebase_interval = 2.0 ** -20
    assert lib.gizmo_host_set_method(1) == 0
    try:
        lib.compute_hydro_densities_and_forces()
        assert lib.gizmo_host_last_status() == 0
        got = read_state(lib, N)
    finally:
        lib.gizmo_host_set_method(0)
        lib.gizmo_host_release()
    ic2 = dict(ic, ti_step=(2 ** bins.astype(np.int64)))
    out, _ = gzm.run_pass(ic2, params=gzm.make_params(method=gzm.METHOD_MFV, timebase_interval=2.0 ** -20))
    for k in STATE_KEYS + ("particle_vel", "dt_mass", "grav_work"):
        np.testing.assert_array_equal(got[k], np.asarray(out[k]).reshape(got[k].shape), err_msg=k)
    # dMass: the neighbours' shares arrive by atomic adds, in a run-dependent order
    scale = np.abs(out["d_mass"]).max()
    assert scale > 0
    assert np.abs(got["d_mass"] - out["d_mass"]).max() <= 1e-13 * scale
