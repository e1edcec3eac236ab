"""The C host mirror's handling of the reference's active list before any
device work (runs on CPU-only hosts: the list is checked before the library
is called).  Reference: FirstActiveParticle / NextActiveParticle[]
(allvars.h:1652-1654), endrun (allvars.h:1635)."""
import numpy as np
import pytest

import ics
from host_mirror import codes, fill, host_lib, views


@pytest.fixture()
def lib():
    lib = host_lib()
    ic = ics.random_state(6)
    fill(lib, ic)
    yield lib
    lib.gizmo_host_release()
    lib.gizmo_host_free()


def test_numpy_views_alias_the_particle_structures(lib):
    N = 6 ** 3
    P, S, A, nxt = views(lib, N)
    assert P["Pos"].shape == (N, 3) and S["NV_T"].shape == (N, 3, 3)
    ic = ics.random_state(6)
    np.testing.assert_array_equal(P["Pos"], ic["pos"])
    np.testing.assert_array_equal(S["InternalEnergyPred"], ic["u"])
    assert A.DesNumNgb == 32.0
    np.testing.assert_array_equal(nxt[:-1], np.arange(1, N))  # gizmo_host_all_active
    assert nxt[-1] == -1


@pytest.mark.parametrize("kind", ["cycle", "repeat", "out_of_range"])
def test_malformed_active_list_ends_the_run(lib, kind):
    """A cyclic list, a particle listed twice or an index past NumPart is
    reported through endrun (code 1, invalid argument) instead of being
    followed past the staging buffers; nothing is sent to the device."""
    N = 6 ** 3
    _, _, _, nxt = views(lib, N)
    if kind == "cycle":
        nxt[N - 1] = 5           # ... -> N-1 -> 5 -> 6 -> ...
    elif kind == "repeat":
        nxt[10] = 3              # 0 .. 10 -> 3 -> 4 ...: 3 visited twice before the cycle closes
    else:
        nxt[20] = N + 7
    before = len(codes)
    lib.density()
    assert codes[before:] == [1]
    assert lib.gizmo_host_last_status() == 1
    assert b"malformed active list" in lib.gizmo_host_last_error()
