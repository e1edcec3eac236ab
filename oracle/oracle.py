"""ctypes wrapper around the CPU FP64 oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by bench.py's CPU-baseline leg
and by __graft_entry__.smoke() as the checker.  The product never uses it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_ub = C.POINTER(C.c_ubyte)


class Params(C.Structure):
    _fields_ = [
        ("box", C.c_double * 3),
        ("gamma", C.c_double),
        ("des_num_ngb", C.c_double),
        ("max_num_ngb_dev", C.c_double),
        ("min_hsml", C.c_double),
        ("max_hsml", C.c_double),
        ("first_step", C.c_int),
        ("maxiter", C.c_int),
        ("nthreads", C.c_int),
        ("brute_force", C.c_int),
        ("tree", C.c_int),
        ("method", C.c_int),
        ("timebase_interval", C.c_double),
    ]


class State(C.Structure):
    _fields_ = [
        ("n", C.c_int),
        ("pos", _dp), ("vel", _dp), ("mass", _dp), ("u", _dp),
        ("hsml", _dp), ("num_ngb", _dp), ("dhsml", _dp), ("divvel", _dp),
        ("density", _dp), ("pressure", _dp), ("nvt", _dp), ("cn", _dp), ("fce", _dp),
        ("grad_rho", _dp), ("grad_p", _dp), ("grad_v", _dp),
        ("acc", _dp), ("dtu", _dp), ("vsig", _dp),
        ("wakeup", _ub), ("is_active", _ub),
        ("divvel_l1", _dp), ("fce_l1", _dp), ("grad_l1", _dp), ("acc_l1", _dp), ("dtu_l1", _dp),
        ("tie_flag", _ub), ("grad_in", _dp),
        ("particle_vel", _dp), ("dt_mass", _dp), ("d_mass", _dp), ("grav_work", _dp),
        ("ti_step", C.POINTER(C.c_longlong)),
        ("pvel_l1", _dp), ("dtmass_l1", _dp), ("dmass_l1", _dp), ("gravwork_l1", _dp),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int),
        ("dens_interactions", C.c_longlong),
        ("grad_interactions", C.c_longlong),
        ("force_interactions", C.c_longlong),
        ("status", C.c_int),
        ("riemann_retries", C.c_longlong),
        ("exact_solver_calls", C.c_longlong),
    ]


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB_PATH)
        L.orc_kernel.argtypes = [C.c_double, C.c_double, _dp, _dp]
        L.orc_matrix_invert.argtypes = [_dp, _dp]
        L.orc_matrix_invert.restype = C.c_double
        L.orc_reconstruct.argtypes = [C.c_double, _dp, C.c_double, _dp, _dp, _dp, _dp, _dp, C.c_int]
        L.orc_slope_limiter.argtypes = [_dp, C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_int, C.c_double, C.c_double]
        L.orc_riemann.argtypes = [C.c_double, _dp, _dp, _dp, C.c_double, _dp]
        L.orc_riemann.restype = C.c_int
        L.orc_riemann_kt.argtypes = [C.c_double, _dp, _dp, _dp, _dp]
        L.orc_riemann_kt.restype = None
        L.orc_pstar_estimate.argtypes = [C.c_double, _dp, _dp, _dp, C.c_int, _dp]
        L.orc_pstar_estimate.restype = None
        L.orc_wavespeeds.argtypes = [C.c_double, _dp, _dp, _dp, C.c_double, _dp]
        L.orc_wavespeeds.restype = C.c_int
        L.orc_riemann_mfv.argtypes = [C.c_double, _dp, _dp, _dp, C.c_double, _dp]
        L.orc_riemann_mfv.restype = C.c_int
        L.orc_riemann_kt_mfv.argtypes = [C.c_double, _dp, _dp, _dp, _dp]
        L.orc_riemann_kt_mfv.restype = None
        L.orc_riemann_exact_mfv.argtypes = [C.c_double, _dp, _dp, _dp, _dp]
        L.orc_riemann_exact_mfv.restype = C.c_int
        L.orc_hsml_step.argtypes = [C.c_double, C.c_double, _dp, _dp, _dp, C.c_int, C.c_double, C.c_double,
                                    C.c_double, C.c_double]
        L.orc_hsml_step.restype = C.c_int
        for f in (L.orc_density, L.orc_gradients, L.orc_hydro_force):
            f.argtypes = [C.POINTER(Params), C.POINTER(State), _ip, C.c_int, C.POINTER(Stats)]
            f.restype = C.c_int
        L.orc_neighbor_counts.argtypes = [C.POINTER(Params), C.c_int, _dp, _dp, _ip, C.c_int, C.c_int, _ip]
        L.orc_neighbor_counts.restype = C.c_longlong
        L.orc_neighbor_lists.argtypes = [C.POINTER(Params), C.c_int, _dp, _dp, _ip, C.c_int, C.c_int,
                                         C.POINTER(C.c_longlong), _ip, C.c_longlong]
        L.orc_neighbor_lists.restype = C.c_longlong
        _lib = L
    return _lib


def ptr(a: np.ndarray, t=_dp):
    return a.ctypes.data_as(t)


METHOD_MFM, METHOD_MFV = 0, 1


def make_params(box=(1.0, 1.0, 1.0), gamma=5.0 / 3.0, des_num_ngb=32.0, max_num_ngb_dev=0.05,
                min_hsml=0.0, max_hsml=1e56, first_step=1, maxiter=150, nthreads=0,
                brute_force=0, tree=0, method=METHOD_MFM, timebase_interval=2.0 ** -20) -> Params:
    p = Params()
    p.box[:] = list(box)
    p.gamma = gamma
    p.des_num_ngb = des_num_ngb
    p.max_num_ngb_dev = max_num_ngb_dev
    p.min_hsml = min_hsml
    p.max_hsml = max_hsml
    p.first_step = first_step
    p.maxiter = maxiter
    p.nthreads = nthreads
    p.brute_force = brute_force
    p.tree = tree
    p.method = method
    p.timebase_interval = timebase_interval
    return p


# ---------------------------------------------------------------------------
# element functions
# ---------------------------------------------------------------------------
def kernel(u: float, h: float):
    wk, dwk = C.c_double(), C.c_double()
    lib().orc_kernel(u, h, C.byref(wk), C.byref(dwk))
    return wk.value, dwk.value


def matrix_invert(T):
    T = np.ascontiguousarray(np.asarray(T, dtype=np.float64).reshape(9))
    Ti = np.zeros(9)
    cn = lib().orc_matrix_invert(ptr(T), ptr(Ti))
    return Ti.reshape(3, 3), cn


def reconstruct(qi, gi, qj, gj, di, dj, mode=1):
    arr = [np.ascontiguousarray(np.asarray(x, dtype=np.float64)) for x in (gi, gj, di, dj)]
    ql, qr = C.c_double(), C.c_double()
    lib().orc_reconstruct(qi, ptr(arr[0]), qj, ptr(arr[1]), ptr(arr[2]), ptr(arr[3]),
                          C.byref(ql), C.byref(qr), mode)
    return ql.value, qr.value


def slope_limiter(grad, valmax, valmin, alim, h, shoot_tol, pos_preserve, d_max, val_cen):
    g = np.ascontiguousarray(np.asarray(grad, dtype=np.float64).copy())
    lib().orc_slope_limiter(ptr(g), valmax, valmin, alim, h, shoot_tol, pos_preserve, d_max, val_cen)
    return g


def riemann(L, R, n=(1.0, 0.0, 0.0), press_tot_limiter=1e300, gamma=5.0 / 3.0):
    """L, R = (rho, p, vx, vy, vz). Returns dict(P_M, S_M, Fv, Fp, mask)."""
    Lc = np.ascontiguousarray(np.asarray(L, dtype=np.float64))
    Rc = np.ascontiguousarray(np.asarray(R, dtype=np.float64))
    nc = np.ascontiguousarray(np.asarray(n, dtype=np.float64))
    out = np.zeros(6)
    mask = lib().orc_riemann(gamma, ptr(Lc), ptr(Rc), ptr(nc), press_tot_limiter, ptr(out))
    return dict(P_M=out[0], S_M=out[1], Fv=out[2:5].copy(), Fp=out[5], mask=mask)


def _states(L, R, n):
    return [np.ascontiguousarray(np.asarray(x, dtype=np.float64)) for x in (L, R, n)]


def riemann_kt(L, R, n=(1.0, 0.0, 0.0), gamma=5.0 / 3.0):
    """Kurganov-Tadmor/PWK flux alone. Returns dict(P_M, S_M, Fv, Fp)."""
    Lc, Rc, nc = _states(L, R, n)
    out = np.zeros(6)
    lib().orc_riemann_kt(gamma, ptr(Lc), ptr(Rc), ptr(nc), ptr(out))
    return dict(P_M=out[0], S_M=out[1], Fv=out[2:5].copy(), Fp=out[5])


def _mfv_out(out, mask=None):
    d = dict(P_M=out[0], S_M=out[1], Frho=out[2], Fv=out[3:6].copy(), Fp=out[6])
    if mask is not None:
        d["mask"] = mask
    return d


def riemann_mfv(L, R, n=(1.0, 0.0, 0.0), press_tot_limiter=1e300, gamma=5.0 / 3.0):
    """Riemann_solver() of an MFV run: dict(P_M, S_M, Frho, Fv, Fp, mask), fluxes
    through a face at rest in the frame of L and R."""
    Lc, Rc, nc = _states(L, R, n)
    out = np.zeros(7)
    mask = lib().orc_riemann_mfv(gamma, ptr(Lc), ptr(Rc), ptr(nc), press_tot_limiter, ptr(out))
    return _mfv_out(out, mask)


def riemann_kt_mfv(L, R, n=(1.0, 0.0, 0.0), gamma=5.0 / 3.0):
    """The KT flux alone, MFV branch: dict(P_M, S_M, Frho, Fv, Fp)."""
    Lc, Rc, nc = _states(L, R, n)
    out = np.zeros(7)
    lib().orc_riemann_kt_mfv(gamma, ptr(Lc), ptr(Rc), ptr(nc), ptr(out))
    return _mfv_out(out)


def riemann_exact_mfv(L, R, n=(1.0, 0.0, 0.0), gamma=5.0 / 3.0):
    """The exact solver with the face sampling (Godunov flux): dict(P_M, S_M,
    Frho, Fv, Fp, mask) with mask the status (0 or 1234545)."""
    Lc, Rc, nc = _states(L, R, n)
    out = np.zeros(7)
    e = lib().orc_riemann_exact_mfv(gamma, ptr(Lc), ptr(Rc), ptr(nc), ptr(out))
    return _mfv_out(out, e)


PSTAR_HLLC, PSTAR_ROE, PSTAR_PVRS = 0, 1, 2


def pstar_estimate(L, R, which, n=(1.0, 0.0, 0.0), gamma=5.0 / 3.0):
    """One star-state estimate (0 HLLC/Davis, 1 Roe, 2 PVRS): dict(P_M, S_M, S_L, S_R)."""
    Lc, Rc, nc = _states(L, R, n)
    out = np.zeros(4)
    lib().orc_pstar_estimate(gamma, ptr(Lc), ptr(Rc), ptr(nc), int(which), ptr(out))
    return dict(P_M=out[0], S_M=out[1], S_L=out[2], S_R=out[3])


def wavespeeds(L, R, n=(1.0, 0.0, 0.0), press_tot_limiter=1e300, gamma=5.0 / 3.0):
    """The estimate cascade: dict(P_M, S_M, S_L, S_R, which) with which 0 vacuum,
    1 HLLC, 2 Roe, 3 PVRS."""
    Lc, Rc, nc = _states(L, R, n)
    out = np.zeros(4)
    which = lib().orc_wavespeeds(gamma, ptr(Lc), ptr(Rc), ptr(nc), press_tot_limiter, ptr(out))
    return dict(P_M=out[0], S_M=out[1], S_L=out[2], S_R=out[3], which=which)


HSML_DONE, HSML_AGAIN, HSML_BRACKETED, HSML_COLLAPSED = 0, 1, 2, -1


def hsml_step(num_ngb, dhsml, h, left=0.0, right=0.0, iter=0, des=32.0, dev=0.05, min_hsml=0.0, max_hsml=1e56):
    """One hsml update of density()'s iteration: (result, h, left, right)."""
    hh, ll, rr = C.c_double(h), C.c_double(left), C.c_double(right)
    r = lib().orc_hsml_step(num_ngb, dhsml, C.byref(hh), C.byref(ll), C.byref(rr), iter, des, dev, min_hsml, max_hsml)
    return r, hh.value, ll.value, rr.value


# ---------------------------------------------------------------------------
# full pass
# ---------------------------------------------------------------------------
class OracleState:
    """Owns every array of one particle set (caller order)."""

    def __init__(self, ic: dict, active=None):
        N = ic["pos"].shape[0]
        self.n = N
        self.pos = np.ascontiguousarray(ic["pos"], dtype=np.float64)
        self.vel = np.ascontiguousarray(ic["vel"], dtype=np.float64)
        self.mass = np.ascontiguousarray(ic["mass"], dtype=np.float64)
        self.u = np.ascontiguousarray(ic["u"], dtype=np.float64)
        self.hsml = np.ascontiguousarray(ic["hsml"], dtype=np.float64).copy()
        self.cn = np.ascontiguousarray(ic.get("cn", np.ones(N)), dtype=np.float64).copy()
        z = lambda *shape: np.zeros(shape, dtype=np.float64)
        self.num_ngb = np.ascontiguousarray(ic.get("num_ngb", z(N)), dtype=np.float64).copy()
        self.dhsml = np.ascontiguousarray(ic.get("dhsml", z(N)), dtype=np.float64).copy()
        self.divvel = z(N)
        self.density = np.ascontiguousarray(ic.get("density", z(N)), dtype=np.float64).copy()
        self.pressure = np.ascontiguousarray(ic.get("pressure", z(N)), dtype=np.float64).copy()
        self.nvt = np.ascontiguousarray(ic.get("nvt", z(N, 9)), dtype=np.float64).reshape(N, 9).copy()
        self.fce = np.ascontiguousarray(ic.get("fce", z(N)), dtype=np.float64).copy()
        self.grad_rho = np.ascontiguousarray(ic.get("grad_rho", z(N, 3)), dtype=np.float64).copy()
        self.grad_p = np.ascontiguousarray(ic.get("grad_p", z(N, 3)), dtype=np.float64).copy()
        self.grad_v = np.ascontiguousarray(ic.get("grad_v", z(N, 3, 3)), dtype=np.float64).copy()
        self.acc = np.ascontiguousarray(ic.get("acc", z(N, 3)), dtype=np.float64).copy()
        self.dtu = np.ascontiguousarray(ic.get("dtu", z(N)), dtype=np.float64).copy()
        self.vsig = np.ascontiguousarray(ic.get("vsig", np.full(N, 1e-56)), dtype=np.float64).copy()
        self.wakeup = np.zeros(N, dtype=np.uint8)
        self.divvel_l1 = z(N)
        self.fce_l1 = z(N)
        self.grad_l1 = z(N, 15)
        self.grad_in = z(N, 15)
        self.acc_l1 = z(N, 3)
        self.dtu_l1 = z(N)
        self.tie_flag = np.zeros(N, dtype=np.uint8)
        self.particle_vel = np.ascontiguousarray(ic.get("particle_vel", self.vel), dtype=np.float64).copy()
        self.dt_mass = np.ascontiguousarray(ic.get("dt_mass", z(N)), dtype=np.float64).copy()
        self.d_mass = np.ascontiguousarray(ic.get("d_mass", z(N)), dtype=np.float64).copy()
        self.grav_work = np.ascontiguousarray(ic.get("grav_work", z(N, 3)), dtype=np.float64).copy()
        self.ti_step = (None if ic.get("ti_step") is None
                        else np.ascontiguousarray(ic["ti_step"], dtype=np.int64).copy())
        self.pvel_l1 = z(N, 3)
        self.dtmass_l1 = z(N)
        self.dmass_l1 = z(N)
        self.gravwork_l1 = z(N, 3)
        if active is None:
            active = np.arange(N, dtype=np.int32)
        self.active = np.ascontiguousarray(active, dtype=np.int32)
        self.is_active = np.zeros(N, dtype=np.uint8)
        self.is_active[self.active] = 1

    def cstruct(self) -> State:
        s = State()
        s.n = self.n
        for name in ("pos", "vel", "mass", "u", "hsml", "num_ngb", "dhsml", "divvel", "density",
                     "pressure", "nvt", "cn", "fce", "grad_rho", "grad_p", "grad_v", "acc", "dtu", "vsig"):
            setattr(s, name, ptr(getattr(self, name)))
        s.wakeup = ptr(self.wakeup, _ub)
        s.is_active = ptr(self.is_active, _ub)
        for name in ("divvel_l1", "fce_l1", "grad_l1", "acc_l1", "dtu_l1", "grad_in"):
            setattr(s, name, ptr(getattr(self, name)))
        s.tie_flag = ptr(self.tie_flag, _ub)
        for name in ("particle_vel", "dt_mass", "d_mass", "grav_work", "pvel_l1", "dtmass_l1", "dmass_l1",
                     "gravwork_l1"):
            setattr(s, name, ptr(getattr(self, name)))
        if self.ti_step is not None:
            s.ti_step = self.ti_step.ctypes.data_as(C.POINTER(C.c_longlong))
        return s

    def outputs(self) -> dict:
        names = ("hsml", "num_ngb", "dhsml", "divvel", "density", "pressure", "nvt", "cn", "fce",
                 "grad_rho", "grad_p", "grad_v", "acc", "dtu", "vsig", "wakeup",
                 "divvel_l1", "fce_l1", "grad_l1", "acc_l1", "dtu_l1", "tie_flag", "grad_in",
                 "particle_vel", "dt_mass", "d_mass", "grav_work", "pvel_l1", "dtmass_l1", "dmass_l1",
                 "gravwork_l1")
        return {k: getattr(self, k).copy() for k in names}


def run_pass(ic: dict, params: Params | None = None, active=None, stages=("density", "gradients", "force")):
    """Run the requested stages on a copy of `ic`; returns (outputs, stats, seconds per stage)."""
    params = params or make_params(box=tuple(ic["box"]))
    st = OracleState(ic, active)
    cs = st.cstruct()
    stats = Stats()
    L = lib()
    times = {}
    for stage, fn in (("density", L.orc_density), ("gradients", L.orc_gradients), ("force", L.orc_hydro_force)):
        if stage not in stages:
            continue
        t0 = time.perf_counter()
        rc = fn(C.byref(params), C.byref(cs), ptr(st.active, _ip), int(st.active.size), C.byref(stats))
        times[stage] = time.perf_counter() - t0
        if rc != 0:
            raise RuntimeError(f"oracle {stage} failed with status {rc}")
    out = st.outputs()
    return out, stats, times


def neighbor_counts(pos, hsml, mode, active=None, params=None):
    N = pos.shape[0]
    params = params or make_params()
    active = np.arange(N, dtype=np.int32) if active is None else np.ascontiguousarray(active, dtype=np.int32)
    counts = np.zeros(active.size, dtype=np.int32)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    hsml = np.ascontiguousarray(hsml, dtype=np.float64)
    total = lib().orc_neighbor_counts(C.byref(params), N, ptr(pos), ptr(hsml), ptr(active, _ip),
                                      int(active.size), mode, ptr(counts, _ip))
    return counts, total


def neighbor_lists(pos, hsml, mode, active=None, params=None):
    N = pos.shape[0]
    params = params or make_params()
    active = np.arange(N, dtype=np.int32) if active is None else np.ascontiguousarray(active, dtype=np.int32)
    counts, total = neighbor_counts(pos, hsml, mode, active, params)
    offsets = np.zeros(active.size + 1, dtype=np.int64)
    idx = np.zeros(max(int(total), 1), dtype=np.int32)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    hsml = np.ascontiguousarray(hsml, dtype=np.float64)
    got = lib().orc_neighbor_lists(C.byref(params), N, ptr(pos), ptr(hsml), ptr(active, _ip), int(active.size),
                                   mode, offsets.ctypes.data_as(C.POINTER(C.c_longlong)), ptr(idx, _ip),
                                   int(idx.size))
    assert got == total
    return offsets, idx[:total]


def force_pair(st: "OracleState", i: int, j: int, params=None):
    """F_v[3], F_p, v_sig, stage mask of one (i, j) force interaction."""
    params = params or make_params()
    L = lib()
    L.orc_force_pair.argtypes = [C.POINTER(Params), C.POINTER(State), C.c_int, C.c_int, _dp]
    L.orc_force_pair.restype = C.c_int
    out = np.zeros(6)
    cs = st.cstruct()
    rc = L.orc_force_pair(C.byref(params), C.byref(cs), int(i), int(j), ptr(out))
    return out, rc
