"""The HDF5 IC / snapshot layout (libgizmo_io.so) over the host mirror's
P[]/SphP[] (runs on CPU-only hosts).

Files "as scripts/make_IC.py writes them" (make_IC.py:96-165: h5py with numpy
defaults, i.e. float64 datasets, int64 IDs and integer attributes, a
MagneticField block and a PartType3 group beside the gas) are produced here
with the HDF5 C library called through ctypes, independently of the code
under test.  Reference behaviour: read_ic.c:105, 187-269, 698-833, 1379-1411;
init.c:503-507, 1067-1110; io.c:4067-4135, 4496-4566."""
import ctypes as C
import math
import os

import numpy as np
import pytest

import ics
from host_mirror import IOHeader, fill, host_lib, io_lib, read_state, views

HDF5 = "/opt/conda/lib/libhdf5.so"

pytestmark = pytest.mark.skipif(io_lib() is None or not os.path.exists(HDF5),
                                reason="libgizmo_io.so not built (no HDF5 C library)")


class H5:
    """The few HDF5 C calls an IC writer needs (h5py is not installed)."""

    def __init__(self):
        L = C.CDLL(HDF5)
        L.H5open()
        hid = C.c_int64
        for f in ("H5Fcreate", "H5Gcreate2", "H5Screate_simple", "H5Screate", "H5Acreate2", "H5Dcreate2"):
            getattr(L, f).restype = hid
        L.H5Fcreate.argtypes = [C.c_char_p, C.c_uint, hid, hid]
        L.H5Gcreate2.argtypes = [hid, C.c_char_p, hid, hid, hid]
        L.H5Screate_simple.argtypes = [C.c_int, C.POINTER(C.c_uint64), C.c_void_p]
        L.H5Screate.argtypes = [C.c_int]
        L.H5Acreate2.argtypes = [hid, C.c_char_p, hid, hid, hid, hid]
        L.H5Awrite.argtypes = [hid, hid, C.c_void_p]
        L.H5Dcreate2.argtypes = [hid, C.c_char_p, hid, hid, hid, hid, hid]
        L.H5Dwrite.argtypes = [hid, hid, hid, hid, hid, C.c_void_p]
        for f in ("H5Aclose", "H5Dclose", "H5Sclose", "H5Gclose", "H5Fclose"):
            getattr(L, f).argtypes = [hid]
        self.L = L
        self.types = {np.dtype(np.float64): hid.in_dll(L, "H5T_NATIVE_DOUBLE_g").value,
                      np.dtype(np.float32): hid.in_dll(L, "H5T_NATIVE_FLOAT_g").value,
                      np.dtype(np.int64): hid.in_dll(L, "H5T_NATIVE_LLONG_g").value,
                      np.dtype(np.int32): hid.in_dll(L, "H5T_NATIVE_INT_g").value}

    def space(self, a):
        if a.ndim == 0:
            return self.L.H5Screate(0)  # H5S_SCALAR
        dims = (C.c_uint64 * a.ndim)(*a.shape)
        return self.L.H5Screate_simple(a.ndim, dims, None)

    def attr(self, loc, name, value):
        a = np.ascontiguousarray(value)
        t = self.types[a.dtype]
        sp = self.space(a)
        h = self.L.H5Acreate2(loc, name.encode(), t, sp, 0, 0)
        assert h >= 0 and self.L.H5Awrite(h, t, a.ctypes.data) >= 0
        self.L.H5Aclose(h)
        self.L.H5Sclose(sp)

    def dataset(self, loc, name, value):
        a = np.ascontiguousarray(value)
        t = self.types[a.dtype]
        sp = self.space(a)
        d = self.L.H5Dcreate2(loc, name.encode(), t, sp, 0, 0, 0)
        assert d >= 0 and self.L.H5Dwrite(d, t, 0, 0, 0, a.ctypes.data) >= 0
        self.L.H5Dclose(d)
        self.L.H5Sclose(sp)

    def write(self, fname, header, groups):
        f = self.L.H5Fcreate(fname.encode(), 2, 0, 0)  # H5F_ACC_TRUNC
        assert f >= 0
        h = self.L.H5Gcreate2(f, b"Header", 0, 0, 0)
        for k, v in header.items():
            self.attr(h, k, v)
        self.L.H5Gclose(h)
        for gname, blocks in groups.items():
            g = self.L.H5Gcreate2(f, gname.encode(), 0, 0, 0)
            for k, v in blocks.items():
                self.dataset(g, k, v)
            self.L.H5Gclose(g)
        self.L.H5Fclose(f)


def make_ic_file(path, ngas=300, ngrains=50, seed=3, mass_table0=0.0, flag_double=0, pos=None):
    """Header and blocks laid out as make_IC.py:96-165 writes them."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1, (ngas, 3)) if pos is None else pos
    v = rng.normal(size=(ngas, 3))
    m = np.full(ngas, 1.0 / ngas)
    u = rng.uniform(0.5, 2.0, ngas)
    ids = np.arange(ngas, dtype=np.int64) + 1
    npart = np.array([ngas, 0, 0, ngrains, 0, 0], dtype=np.int64)
    mt = np.zeros(6)
    mt[0] = mass_table0
    header = {"NumPart_ThisFile": npart, "NumPart_Total": npart, "NumPart_Total_HighWord": 0 * npart,
              "MassTable": mt, "Time": np.float64(0.0), "NumFilesPerSnapshot": np.int64(1),
              "Flag_DoublePrecision": np.int64(flag_double)}
    gas = {"Coordinates": x, "Velocities": v, "ParticleIDs": ids, "InternalEnergy": u,
           "MagneticField": np.zeros((ngas, 3))}
    if mass_table0 == 0:
        gas["Masses"] = m
    grains = {"Coordinates": rng.uniform(0, 1, (ngrains, 3)), "Velocities": np.zeros((ngrains, 3)),
              "ParticleIDs": np.arange(ngrains, dtype=np.int64) + ngas + 1, "Masses": np.full(ngrains, 1e-3),
              "PICParticleType": np.full(ngrains, 3, dtype=np.int64)}
    groups = {"PartType0": gas}
    if ngrains:
        groups["PartType3"] = grains
    H5().write(str(path), header, groups)
    return dict(pos=x, vel=v, mass=m, u=u, ids=ids)


@pytest.fixture()
def libs():
    host = host_lib()
    io = io_lib()
    yield host, io
    host.gizmo_host_release()
    host.gizmo_host_free()


def set_box(host, L=1.0):
    A = views(host, 0)[2] if False else None  # (views needs allocated arrays; All is a plain global)
    from host_mirror import AllData
    A = AllData.in_dll(host, "All")
    A.BoxSize, A.DesNumNgb, A.MaxNumNgbDeviation, A.MaxHsml = L, 32.0, 0.05, 1e56
    A.Time = A.TimeBegin = 0.0
    return A


def read(io, path, input_double=0):
    h = IOHeader()
    rc = io.gizmo_read_ic(str(path).encode(), input_double, C.byref(h))
    return rc, h


def test_reads_the_make_ic_layout(libs, tmp_path):
    host, io = libs
    set_box(host)
    path = tmp_path / "ic.hdf5"
    src = make_ic_file(path)
    rc, h = read(io, path)
    assert rc == 0, io.gizmo_io_last_error()
    n = 300
    assert list(h.npart) == [n, 0, 0, 50, 0, 0] and h.num_files == 1 and h.flag_double == 0
    assert C.c_int.in_dll(host, "N_gas").value == n
    P, S, _, _ = views(host, n)
    f32 = lambda a: np.asarray(a, dtype=np.float32).astype(np.float64)  # MyInputFloat = float
    np.testing.assert_array_equal(P["Pos"], f32(src["pos"]))
    np.testing.assert_array_equal(P["Vel"], f32(src["vel"]))
    np.testing.assert_array_equal(S["VelPred"], P["Vel"])                 # init.c:507
    np.testing.assert_array_equal(P["Mass"], f32(src["mass"]))
    np.testing.assert_array_equal(S["InternalEnergy"], f32(src["u"]))
    np.testing.assert_array_equal(S["InternalEnergyPred"], S["InternalEnergy"])  # init.c:503
    np.testing.assert_array_equal(P["ID"], src["ids"])
    assert np.all(P["Type"] == 0) and np.all(P["Hsml"] == 0)


def test_double_precision_input(libs, tmp_path):
    host, io = libs
    set_box(host)
    path = tmp_path / "ic_double.hdf5"
    src = make_ic_file(path, flag_double=1)
    assert read(io, path, input_double=0)[0] == 10   # read_ic.c:827-831
    rc, _ = read(io, path, input_double=1)
    assert rc == 0
    P, S, _, _ = views(host, 300)
    np.testing.assert_array_equal(P["Pos"], src["pos"])
    np.testing.assert_array_equal(S["InternalEnergy"], src["u"])
    single = tmp_path / "ic_single.hdf5"
    make_ic_file(single, flag_double=0)
    assert read(io, single, input_double=1)[0] == 11  # read_ic.c:820-824


def test_mass_table_replaces_the_masses_block(libs, tmp_path):
    host, io = libs
    set_box(host)
    path = tmp_path / "ic_mt.hdf5"
    make_ic_file(path, mass_table0=2.5e-5, ngrains=0)
    rc, h = read(io, path)
    assert rc == 0, io.gizmo_io_last_error()
    P, _, _, _ = views(host, 300)
    assert np.all(P["Mass"] == 2.5e-5)  # read_ic.c:105


def test_positions_are_wrapped_into_the_periodic_box(libs, tmp_path):
    host, io = libs
    set_box(host, 2.0)
    pos = np.array([[-0.5, 2.5, 1.0], [4.25, -3.75, 0.0], [1.999, 0.001, -2.0]])
    path = tmp_path / "ic_wrap.hdf5"
    make_ic_file(path, ngas=3, ngrains=0, pos=pos)
    assert read(io, path, input_double=0)[0] == 0
    P, _, _, _ = views(host, 3)
    want = np.mod(np.asarray(pos, dtype=np.float32).astype(np.float64), 2.0)
    np.testing.assert_allclose(P["Pos"], want, rtol=0, atol=1e-12)
    assert np.all((P["Pos"] >= 0) & (P["Pos"] < 2.0))


def test_malformed_files_are_reported(libs, tmp_path):
    host, io = libs
    set_box(host)
    assert read(io, tmp_path / "missing.hdf5")[0] == 123
    path = tmp_path / "no_gas_group.hdf5"
    H5().write(str(path), {"NumPart_ThisFile": np.array([5, 0, 0, 0, 0, 0], dtype=np.int64),
                           "NumPart_Total": np.array([5, 0, 0, 0, 0, 0], dtype=np.int64),
                           "NumPart_Total_HighWord": np.zeros(6, dtype=np.int64), "MassTable": np.zeros(6),
                           "Time": np.float64(0), "NumFilesPerSnapshot": np.int64(1),
                           "Flag_DoublePrecision": np.int64(0)}, {})
    rc, _ = read(io, path)
    assert rc == 3001 and b"PartType0" in io.gizmo_io_last_error()
    multi = tmp_path / "multi.hdf5"
    H5().write(str(multi), {"NumPart_ThisFile": np.array([5, 0, 0, 0, 0, 0], dtype=np.int64),
                            "NumPart_Total": np.array([9, 0, 0, 0, 0, 0], dtype=np.int64),
                            "NumPart_Total_HighWord": np.zeros(6, dtype=np.int64), "MassTable": np.zeros(6),
                            "Time": np.float64(0), "NumFilesPerSnapshot": np.int64(2),
                            "Flag_DoublePrecision": np.int64(0)}, {})
    assert read(io, multi)[0] == 3003


@pytest.mark.parametrize("output_double", [1, 0])
def test_snapshot_round_trip(libs, tmp_path, output_double):
    """write_snapshot -> read_ic gives the particles back: bit for bit in
    double precision, to float rounding in single precision; the header
    records the time, box and precision (io.c:4496-4566)."""
    host, io = libs
    ic = ics.lattice_gaussian(8)
    n = ic["pos"].shape[0]
    fill(host, ic)
    A = set_box(host)
    A.Time = 0.125
    P, S, _, _ = views(host, n)
    P["ID"] = np.arange(n) * 7 + 3
    S["Density"] = 1.0 + 0.01 * np.arange(n)
    path = tmp_path / "snap.hdf5"
    assert io.gizmo_write_snapshot(str(path).encode(), output_double) == 0, io.gizmo_io_last_error()
    host.gizmo_host_free()
    rc, h = read(io, path, input_double=output_double)
    assert rc == 0, io.gizmo_io_last_error()
    assert h.time == 0.125 and h.box_size == 1.0 and h.flag_double == output_double
    assert list(h.npart) == [n, 0, 0, 0, 0, 0] and list(h.mass_table) == [0.0] * 6
    P, S, _, _ = views(host, n)
    conv = (lambda a: a) if output_double else (lambda a: np.asarray(a, np.float32).astype(np.float64))
    np.testing.assert_array_equal(P["Pos"], conv(ic["pos"]))
    np.testing.assert_array_equal(P["Vel"], conv(ic["vel"]))
    np.testing.assert_array_equal(P["Mass"], conv(ic["mass"]))
    np.testing.assert_array_equal(S["InternalEnergy"], conv(ic["u"]))
    np.testing.assert_array_equal(P["Hsml"], conv(ic["hsml"]))  # SmoothingLength block
    np.testing.assert_array_equal(P["ID"], np.arange(n) * 7 + 3)


@pytest.mark.parametrize("nside", [16, 32])
def test_initial_smoothing_lengths_on_a_lattice(libs, nside):
    """init.c:1067-1110 on an n^3 lattice (n a power of two): the smallest
    octree cell holding 2 DesNumNgb particle masses has side 4/n and holds 64
    particles, so every particle gets (3 * 32 / (4 pi * 64))^(1/3) * 4/n --
    within 1e-5 the radius that encloses 32 particles at the lattice's number
    density, (3 * 32 / (4 pi n^3))^(1/3)."""
    host, io = libs
    g = (np.arange(nside) + 0.5) / nside
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    n = pos.shape[0]
    ic = dict(pos=pos, vel=np.zeros((n, 3)), mass=np.full(n, 1.0 / n), u=np.ones(n), hsml=np.zeros(n))
    fill(host, ic)
    set_box(host)
    assert io.gizmo_setup_smoothinglengths() == 0
    h = views(host, n)[0]["Hsml"]
    expect = (3.0 * 32 / (4 * math.pi * 64)) ** 0.333333 * 4.0 / nside
    np.testing.assert_allclose(h, expect, rtol=1e-14)
    assert abs(expect / (3.0 * 32 / (4 * math.pi * n)) ** (1 / 3) - 1) < 1e-5
