"""ctypes view of libgizmo_host.so: the reference's particle structures
(include/gizmo_mfm_host.h) as numpy structured arrays over the library's own
P[] and SphP[], so tests fill and read them without per-particle loops."""
import ctypes as C
import os

import numpy as np

import gzm


class ParticleData(C.Structure):
    _fields_ = [("Pos", C.c_double * 3), ("Vel", C.c_double * 3), ("Mass", C.c_double), ("Hsml", C.c_double),
                ("NumNgb", C.c_double), ("DhsmlNgbFactor", C.c_double), ("Particle_DivVel", C.c_double),
                ("Type", C.c_int), ("TimeBin", C.c_int), ("ID", C.c_ulonglong)]


class GasCellData(C.Structure):
    _fields_ = [("Density", C.c_double), ("Pressure", C.c_double), ("InternalEnergy", C.c_double),
                ("InternalEnergyPred", C.c_double), ("VelPred", C.c_double * 3), ("HydroAccel", C.c_double * 3),
                ("DtInternalEnergy", C.c_double), ("MaxSignalVel", C.c_double), ("ConditionNumber", C.c_double),
                ("FaceClosureError", C.c_double), ("NV_T", (C.c_double * 3) * 3),
                ("GradDensity", C.c_double * 3), ("GradPressure", C.c_double * 3),
                ("GradVelocity", (C.c_double * 3) * 3), ("wakeup", C.c_int),
                ("ParticleVel", C.c_double * 3), ("DtMass", C.c_double), ("dMass", C.c_double),
                ("GravWorkTerm", C.c_double * 3)]


class AllData(C.Structure):
    _fields_ = [("BoxSize", C.c_double), ("DesNumNgb", C.c_double), ("MaxNumNgbDeviation", C.c_double),
                ("MinHsml", C.c_double), ("MaxHsml", C.c_double), ("Time", C.c_double), ("TimeBegin", C.c_double),
                ("MaxPart", C.c_int), ("CourantFac", C.c_double), ("MaxSizeTimestep", C.c_double),
                ("MinEgySpec", C.c_double), ("BoxLong", C.c_double * 3), ("GasGamma", C.c_double),
                ("Timebase_interval", C.c_double)]


_lib = None
_handler = None
codes = []  # endrun codes seen since the library was loaded


def host_lib():
    """libgizmo_host.so with an endrun handler that records the code instead
    of exiting (the reference's endrun aborts the run)."""
    global _lib, _handler
    if _lib is None:
        lib = C.CDLL(os.path.join(gzm.LIB_DIR, "libgizmo_host.so"))
        lib.gizmo_host_allocate.argtypes = [C.c_int]
        lib.gizmo_host_last_status.restype = C.c_int
        lib.gizmo_host_last_error.restype = C.c_char_p
        lib.gizmo_host_set_active.argtypes = [C.POINTER(C.c_int), C.c_int]
        lib.gizmo_host_set_method.argtypes = [C.c_int]
        lib.gizmo_kick_gas.argtypes = [C.c_double, C.c_int]
        lib.gizmo_drift_gas.argtypes = [C.c_double]
        lib.gizmo_gas_timestep.argtypes = [C.c_int]
        lib.gizmo_gas_timestep.restype = C.c_double
        lib.gizmo_global_timestep.restype = C.c_double
        lib.gizmo_run_until.argtypes = [C.c_double, C.c_int]
        lib.gizmo_box_side.argtypes = [C.c_int]
        lib.gizmo_box_side.restype = C.c_double
        lib.gizmo_gamma.restype = C.c_double
        handler_t = C.CFUNCTYPE(None, C.c_int, C.c_char_p)
        _handler = handler_t(lambda code, where: codes.append(code))
        lib.gizmo_host_set_endrun_handler(_handler)
        _lib = lib
    return _lib


def _array(lib, name, ctype, n):
    """numpy view of the n-element C array that global pointer `name` points to"""
    dt = np.dtype(ctype)
    assert dt.itemsize == C.sizeof(ctype), (name, dt.itemsize, C.sizeof(ctype))
    addr = C.c_void_p.in_dll(lib, name).value
    assert addr, f"{name} is not allocated"
    return np.frombuffer((C.c_char * (n * dt.itemsize)).from_address(addr), dtype=dt)


def views(lib, n):
    """(P, SphP, All, NextActiveParticle) as numpy views of the library's globals."""
    return (_array(lib, "P", ParticleData, n), _array(lib, "SphP", GasCellData, n), AllData.in_dll(lib, "All"),
            _array(lib, "NextActiveParticle", C.c_int, n))


def fill(lib, ic, state=False):
    """gizmo_host_allocate + the particle set `ic` (and, with state=True, the
    carried fields of a previous pass that `ic` holds)."""
    N = ic["pos"].shape[0]
    assert lib.gizmo_host_allocate(N) == 0
    P, S, A, _ = views(lib, N)
    A.BoxSize, A.DesNumNgb, A.MaxNumNgbDeviation = 1.0, 32.0, 0.05
    A.MinHsml, A.MaxHsml, A.Time, A.TimeBegin = 0.0, 1e56, 0.0, 0.0
    A.CourantFac, A.MaxSizeTimestep, A.MinEgySpec, A.GasGamma = 0.2, 0.0, 0.0, 0.0
    box = np.asarray(ic.get("box", (1.0, 1.0, 1.0)), dtype=np.float64)
    A.BoxSize = float(box[0])
    A.BoxLong[:] = list(box / box[0])
    P["Pos"] = ic["pos"]
    P["Vel"] = ic["vel"]
    S["VelPred"] = ic["vel"]
    P["Mass"] = ic["mass"]
    P["Hsml"] = ic["hsml"]
    P["Type"] = 0
    S["InternalEnergy"] = ic["u"]
    S["InternalEnergyPred"] = ic["u"]
    if state:
        write_state(lib, ic)
    return P, S


def write_state(lib, st):
    N = st["pos"].shape[0]
    P, S, _, _ = views(lib, N)
    P["Hsml"] = st["hsml"]
    P["NumNgb"] = st["num_ngb"]
    P["DhsmlNgbFactor"] = st["dhsml"]
    P["Particle_DivVel"] = st["divvel"]
    S["Density"] = st["density"]
    S["Pressure"] = st["pressure"]
    S["ConditionNumber"] = st["cn"]
    S["FaceClosureError"] = st["fce"]
    S["NV_T"] = np.asarray(st["nvt"]).reshape(N, 3, 3)
    S["GradDensity"] = st["grad_rho"]
    S["GradPressure"] = st["grad_p"]
    S["GradVelocity"] = np.asarray(st["grad_v"]).reshape(N, 3, 3)
    S["HydroAccel"] = st["acc"]
    S["DtInternalEnergy"] = st["dtu"]
    S["MaxSignalVel"] = st["vsig"]


def read_state(lib, n):
    """The pass outputs in the key names of the oracle and gzm.run_pass."""
    P, S, _, _ = views(lib, n)
    return {
        "hsml": P["Hsml"].copy(), "num_ngb": P["NumNgb"].copy(), "dhsml": P["DhsmlNgbFactor"].copy(),
        "divvel": P["Particle_DivVel"].copy(), "density": S["Density"].copy(), "pressure": S["Pressure"].copy(),
        "cn": S["ConditionNumber"].copy(), "fce": S["FaceClosureError"].copy(),
        "nvt": S["NV_T"].reshape(n, 9).copy(), "grad_rho": S["GradDensity"].copy(),
        "grad_p": S["GradPressure"].copy(), "grad_v": S["GradVelocity"].copy(), "acc": S["HydroAccel"].copy(),
        "dtu": S["DtInternalEnergy"].copy(), "vsig": S["MaxSignalVel"].copy(),
        "wakeup": S["wakeup"].astype(np.uint8).copy(), "timebin": P["TimeBin"].copy(),
        "particle_vel": S["ParticleVel"].copy(), "dt_mass": S["DtMass"].copy(), "d_mass": S["dMass"].copy(),
        "grav_work": S["GravWorkTerm"].copy(),
    }


def set_active(lib, idx):
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    assert lib.gizmo_host_set_active(idx.ctypes.data_as(C.POINTER(C.c_int)), int(idx.size)) == 0


_io = None


def io_lib():
    """libgizmo_io.so (HDF5 IC / snapshot layout over the same P[]/SphP[]);
    None when it was not built (no HDF5 library)."""
    global _io
    if _io is None:
        path = os.path.join(gzm.LIB_DIR, "libgizmo_io.so")
        if not os.path.exists(path):
            return None
        host_lib()  # the host library first: libgizmo_io.so resolves P, SphP, All there
        lib = C.CDLL(path)
        lib.gizmo_read_ic.argtypes = [C.c_char_p, C.c_int, C.c_void_p]
        lib.gizmo_write_snapshot.argtypes = [C.c_char_p, C.c_int]
        lib.gizmo_io_last_error.restype = C.c_char_p
        _io = lib
    return _io


class IOHeader(C.Structure):
    _fields_ = [("npart", C.c_int * 6), ("npart_total", C.c_uint * 6), ("npart_total_high", C.c_uint * 6),
                ("mass_table", C.c_double * 6), ("time", C.c_double), ("redshift", C.c_double),
                ("box_size", C.c_double), ("hubble_param", C.c_double), ("num_files", C.c_int),
                ("flag_double", C.c_int), ("comoving", C.c_int)]
