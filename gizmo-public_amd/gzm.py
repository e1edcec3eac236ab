"""ctypes binding of libgzm.so (include/gzm.h) plus a small driver that runs
the MFM pass on one particle set.

The library is the product: if lib/libgzm.so is missing or no GPU is present,
every entry point raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
# GZM_LIBRARY selects another build of the library (e.g. to compare two builds)
LIB_PATH = os.environ.get("GZM_LIBRARY") or os.path.join(LIB_DIR, "libgzm.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)

# every extern "C" symbol declared in include/gzm.h
EXPORTED_SYMBOLS = (
    "gzm_status_string", "gzm_last_error", "gzm_ctx_create", "gzm_ctx_destroy", "gzm_upload",
    "gzm_density", "gzm_gradients", "gzm_hydro_force", "gzm_download", "gzm_restore_initial_state",
    "gzm_neighbor_counts", "gzm_timing_get", "gzm_timing_reset", "gzm_upload_batch", "gzm_run_passes",
    "gzm_riemann_states", "gzm_riemann_states_noalloc", "gzm_riemann_states_mfv",
)


class Params(C.Structure):
    _fields_ = [
        ("box", C.c_double * 3),
        ("gamma", C.c_double),
        ("des_num_ngb", C.c_double),
        ("max_num_ngb_dev", C.c_double),
        ("min_hsml", C.c_double),
        ("max_hsml", C.c_double),
        ("first_step", C.c_int32),
        ("maxiter", C.c_int32),
        ("device", C.c_int32),
        ("precision", C.c_int32),
        ("cell_size", C.c_double),
        ("pair_list_cap", C.c_int32),
        ("search", C.c_int32),
        ("density_budget", C.c_int32),
        ("block_size", C.c_int32 * 3),
        ("wave_sweep_max", C.c_int32),
        ("method", C.c_int32),
        ("timebase_interval", C.c_double),
    ]


class HostView(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("pos", _dp), ("vel", _dp), ("mass", _dp), ("u", _dp),
        ("hsml", _dp), ("cn", _dp), ("vsig", _dp),
        ("num_ngb", _dp), ("dhsml", _dp), ("divvel", _dp), ("density", _dp), ("pressure", _dp),
        ("nvt", _dp), ("fce", _dp), ("grad_rho", _dp), ("grad_p", _dp), ("grad_v", _dp),
        ("acc", _dp), ("dtu", _dp), ("wakeup", _u8p),
        ("ti_step", C.POINTER(C.c_int64)),
        ("particle_vel", _dp), ("dt_mass", _dp), ("d_mass", _dp), ("grav_work", _dp),
    ]


class PassStats(C.Structure):
    _fields_ = [
        ("density_sweeps", C.c_int32),
        ("status", C.c_int32),
        ("density_interactions", C.c_int64),
        ("gradient_interactions", C.c_int64),
        ("force_interactions", C.c_int64),
        ("riemann_retries", C.c_int64),
        ("exact_solver_calls", C.c_int64),
        ("density_updates", C.c_int64),
        ("fp64_force_particles", C.c_int64),
        ("density_fp64_redo", C.c_int64),
    ]


class Timing(C.Structure):
    _fields_ = [
        ("density_ms", C.c_double), ("density_launches", C.c_int64),
        ("gradients_ms", C.c_double), ("gradients_launches", C.c_int64),
        ("force_ms", C.c_double), ("force_launches", C.c_int64),
        ("other_ms", C.c_double), ("other_launches", C.c_int64),
        ("graph_ms", C.c_double), ("graph_launches", C.c_int64),
        ("graph_passes", C.c_int64),
        ("force_deferred_ms", C.c_double), ("force_deferred_launches", C.c_int64),
        ("density_wave_ms", C.c_double), ("density_wave_launches", C.c_int64),
    ]


class GzmError(RuntimeError):
    def __init__(self, status: int, where: str, msg: str):
        super().__init__(f"{where}: status {status}: {msg}")
        self.status = status


_lib = None


def load_library():
    """Load libgzm.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} not found: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(LIB_PATH)
    L.gzm_status_string.argtypes = [C.c_int]
    L.gzm_status_string.restype = C.c_char_p
    L.gzm_last_error.argtypes = []
    L.gzm_last_error.restype = C.c_char_p
    L.gzm_ctx_create.argtypes = [C.POINTER(C.c_void_p), C.POINTER(Params)]
    L.gzm_ctx_destroy.argtypes = [C.c_void_p]
    L.gzm_upload.argtypes = [C.c_void_p, C.POINTER(HostView)]
    for f in (L.gzm_density, L.gzm_gradients, L.gzm_hydro_force):
        f.argtypes = [C.c_void_p, _ip, C.c_int64, C.POINTER(PassStats)]
    L.gzm_download.argtypes = [C.c_void_p, C.POINTER(HostView)]
    L.gzm_restore_initial_state.argtypes = [C.c_void_p]
    L.gzm_neighbor_counts.argtypes = [C.c_void_p, C.c_int32, _ip, C.c_int64, _ip, C.POINTER(C.c_int64)]
    L.gzm_timing_get.argtypes = [C.c_void_p, C.POINTER(Timing)]
    L.gzm_timing_reset.argtypes = [C.c_void_p]
    L.gzm_upload_batch.argtypes = [C.c_void_p, C.POINTER(HostView), C.c_int32, C.POINTER(C.c_double),
                                   C.POINTER(C.c_int64)]
    L.gzm_run_passes.argtypes = [C.c_void_p, C.c_int32, C.POINTER(PassStats)]
    for f in (L.gzm_riemann_states, L.gzm_riemann_states_noalloc):
        f.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_int64, _dp, _dp, _dp, _dp, _dp, _ip]
    L.gzm_riemann_states_mfv.argtypes = [C.c_int32, C.c_int32, C.c_double, C.c_int64, _dp, _dp, _dp, _dp, _dp, _ip]
    for name in EXPORTED_SYMBOLS:
        f = getattr(L, name)
        if name not in ("gzm_status_string", "gzm_last_error"):
            f.restype = C.c_int
    _lib = L
    return L


def _check(status: int, where: str):
    if status != 0:
        L = load_library()
        raise GzmError(status, where, L.gzm_last_error().decode(errors="replace"))


def _p(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


def make_params(box=(1.0, 1.0, 1.0), gamma=5.0 / 3.0, des_num_ngb=32.0, max_num_ngb_dev=0.05,
                min_hsml=0.0, max_hsml=1e56, first_step=1, maxiter=150, device=0, cell_size=0.0,
                pair_list_cap=0, precision=0, search=0, density_budget=0, block_size=(0, 0, 0),
                wave_sweep_max=0, method=0, timebase_interval=2.0 ** -20) -> Params:
    """precision 0: plain FP64 pass; 1: FP64 density stage, FP32 pair
    arithmetic in gradients and force with FP64 matrix inversion, limiter,
    exact Riemann solver and flux accumulation; 2: as 1 with FP32 density
    pair sums.  search 0: cell grid; 1: tree with per-node hsml maxima
    (clustered sets).  block_size: workgroup sizes of the density,
    gradient and force loops (0 = 256); wave_sweep_max: density sweeps over at
    most this many particles run a wavefront per particle (0 = 4096).
    method 0: meshless finite mass; 1: meshless finite volume (FP64 only),
    whose mass exchange uses timebase_interval as the physical time of one
    integer time unit."""
    p = Params()
    p.box[:] = list(box)
    p.gamma = gamma
    p.des_num_ngb = des_num_ngb
    p.max_num_ngb_dev = max_num_ngb_dev
    p.min_hsml = min_hsml
    p.max_hsml = max_hsml
    p.first_step = first_step
    p.maxiter = maxiter
    p.device = device
    p.precision = precision
    p.cell_size = cell_size
    p.pair_list_cap = pair_list_cap
    p.search = search
    p.density_budget = density_budget
    p.block_size[:] = [int(b) for b in block_size]
    p.wave_sweep_max = wave_sweep_max
    p.method = method
    p.timebase_interval = timebase_interval
    return p


METHOD_MFM, METHOD_MFV = 0, 1


class MFMContext:
    """One gzm_ctx: a resident particle set on one GPU."""

    STATE_FIELDS = ("hsml", "cn", "vsig", "num_ngb", "dhsml", "divvel", "density", "pressure",
                    "nvt", "fce", "grad_rho", "grad_p", "grad_v", "acc", "dtu")

    def __init__(self, params: Params | None = None, **kw):
        self.L = load_library()
        self.params = params or make_params(**kw)
        self.ctx = C.c_void_p()
        _check(self.L.gzm_ctx_create(C.byref(self.ctx), C.byref(self.params)), "gzm_ctx_create")
        self.n = 0

    def close(self):
        if self.ctx:
            self.L.gzm_ctx_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, ic: dict):
        """One periodic box (gzm_upload); with `box_offsets` and `box_sizes` in
        `ic`, many independent boxes (gzm_upload_batch)."""
        N = ic["pos"].shape[0]
        self.n = N
        keep = {}

        def arr(name, shape=None, required=False):
            a = ic.get(name)
            if a is None:
                if required:
                    raise ValueError(f"missing {name}")
                return None
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep[name] = a
            return a

        v = HostView()
        v.n = N
        v.pos = _p(arr("pos", required=True))
        v.vel = _p(arr("vel", required=True))
        v.mass = _p(arr("mass", required=True))
        v.u = _p(arr("u", required=True))
        v.hsml = _p(arr("hsml", required=True))
        for name in ("cn", "vsig", "num_ngb", "dhsml", "divvel", "density", "pressure", "nvt", "fce",
                     "grad_rho", "grad_p", "grad_v", "acc", "dtu", "particle_vel", "dt_mass", "d_mass",
                     "grav_work"):
            setattr(v, name, _p(arr(name)))
        if ic.get("ti_step") is not None:
            ti = np.ascontiguousarray(ic["ti_step"], dtype=np.int64)
            keep["ti_step"] = ti
            v.ti_step = ti.ctypes.data_as(C.POINTER(C.c_int64))
        if ic.get("box_offsets") is not None:
            offs = np.ascontiguousarray(ic["box_offsets"], dtype=np.int64)
            sizes = np.ascontiguousarray(ic["box_sizes"], dtype=np.float64).reshape(-1, 3)
            _check(self.L.gzm_upload_batch(self.ctx, C.byref(v), sizes.shape[0], _p(sizes),
                                           offs.ctypes.data_as(C.POINTER(C.c_int64))), "gzm_upload_batch")
        else:
            _check(self.L.gzm_upload(self.ctx, C.byref(v)), "gzm_upload")

    @staticmethod
    def _active(active):
        if active is None:
            return None, 0
        a = np.ascontiguousarray(active, dtype=np.int32)
        return a, a.size

    def density(self, active=None) -> PassStats:
        st = PassStats()
        a, na = self._active(active)
        _check(self.L.gzm_density(self.ctx, _p(a, _ip), na, C.byref(st)), "gzm_density")
        return st

    def gradients(self, active=None) -> PassStats:
        st = PassStats()
        a, na = self._active(active)
        _check(self.L.gzm_gradients(self.ctx, _p(a, _ip), na, C.byref(st)), "gzm_gradients")
        return st

    def hydro_force(self, active=None) -> PassStats:
        st = PassStats()
        a, na = self._active(active)
        _check(self.L.gzm_hydro_force(self.ctx, _p(a, _ip), na, C.byref(st)), "gzm_hydro_force")
        return st

    def restore_initial_state(self):
        _check(self.L.gzm_restore_initial_state(self.ctx), "gzm_restore_initial_state")

    def run_passes(self, n_passes: int) -> PassStats:
        """n_passes full passes captured in one HIP graph (stats summed over them)."""
        st = PassStats()
        _check(self.L.gzm_run_passes(self.ctx, int(n_passes), C.byref(st)), "gzm_run_passes")
        return st

    def download(self) -> dict:
        N = self.n
        out = {
            "hsml": np.zeros(N), "cn": np.zeros(N), "vsig": np.zeros(N), "num_ngb": np.zeros(N),
            "dhsml": np.zeros(N), "divvel": np.zeros(N), "density": np.zeros(N), "pressure": np.zeros(N),
            "nvt": np.zeros((N, 9)), "fce": np.zeros(N), "grad_rho": np.zeros((N, 3)),
            "grad_p": np.zeros((N, 3)), "grad_v": np.zeros((N, 3, 3)), "acc": np.zeros((N, 3)),
            "dtu": np.zeros(N), "wakeup": np.zeros(N, dtype=np.uint8),
        }
        if self.params.method == METHOD_MFV:
            out.update(particle_vel=np.zeros((N, 3)), dt_mass=np.zeros(N), d_mass=np.zeros(N),
                       grav_work=np.zeros((N, 3)))
        v = HostView()
        v.n = N
        for k, a in out.items():
            setattr(v, k, _p(a, _u8p if k == "wakeup" else _dp))
        _check(self.L.gzm_download(self.ctx, C.byref(v)), "gzm_download")
        return out

    def neighbor_counts(self, mode: int, active=None):
        a, na = self._active(active)
        n_out = na if a is not None else self.n
        counts = np.zeros(n_out, dtype=np.int32)
        total = C.c_int64()
        _check(self.L.gzm_neighbor_counts(self.ctx, mode, _p(a, _ip), na, _p(counts, _ip), C.byref(total)),
               "gzm_neighbor_counts")
        return counts, total.value

    def timing(self) -> dict:
        t = Timing()
        _check(self.L.gzm_timing_get(self.ctx, C.byref(t)), "gzm_timing_get")
        return {f[0]: getattr(t, f[0]) for f in Timing._fields_}

    def timing_reset(self):
        _check(self.L.gzm_timing_reset(self.ctx), "gzm_timing_reset")


def riemann_states(L, R, normal, limiter=None, mode=0, precision=0, gamma=5.0 / 3.0, device=0, noalloc=False):
    """The device Riemann solver (mode 0) or KT flux (mode 1) on arrays of face
    states L, R [n][5] = rho, p, v[3], unit normals [n][3] and pressure
    limiters [n]; returns (out [n][6] = P_M, S_M, F_v[3], F_p; stage masks).
    noalloc: through gzm_riemann_states_noalloc (n <= 65536, no device
    allocation)."""
    L = np.ascontiguousarray(L, dtype=np.float64).reshape(-1, 5)
    R = np.ascontiguousarray(R, dtype=np.float64).reshape(-1, 5)
    nrm = np.ascontiguousarray(normal, dtype=np.float64).reshape(-1, 3)
    n = L.shape[0]
    lim = None if limiter is None else np.ascontiguousarray(np.broadcast_to(limiter, (n,)), dtype=np.float64)
    out = np.zeros((n, 6))
    mask = np.zeros(n, dtype=np.int32)
    lib = load_library()
    name = "gzm_riemann_states_noalloc" if noalloc else "gzm_riemann_states"
    _check(getattr(lib, name)(device, mode, precision, gamma, n, _p(L), _p(R), _p(nrm), _p(lim), _p(out),
                              _p(mask, _ip)), name)
    return out, mask


def riemann_states_mfv(L, R, normal, limiter=None, mode=0, gamma=5.0 / 3.0, device=0):
    """The device MFV Riemann solver: mode 0 the whole solver, 1 the KT flux,
    2 the exact solver with its face sampling; returns (out [n][7] = P_M, S_M,
    F_rho, F_v[3], F_p; masks)."""
    L = np.ascontiguousarray(L, dtype=np.float64).reshape(-1, 5)
    R = np.ascontiguousarray(R, dtype=np.float64).reshape(-1, 5)
    nrm = np.ascontiguousarray(normal, dtype=np.float64).reshape(-1, 3)
    n = L.shape[0]
    lim = None if limiter is None else np.ascontiguousarray(np.broadcast_to(limiter, (n,)), dtype=np.float64)
    out = np.zeros((n, 7))
    mask = np.zeros(n, dtype=np.int32)
    lib = load_library()
    _check(lib.gzm_riemann_states_mfv(device, mode, gamma, n, _p(L), _p(R), _p(nrm), _p(lim), _p(out),
                                      _p(mask, _ip)), "gzm_riemann_states_mfv")
    return out, mask


def run_pass(ic: dict, active=None, params: Params | None = None, stages=("density", "gradients", "force")):
    """Upload `ic`, run the requested stages, return (outputs, stats dict)."""
    params = params or make_params(box=tuple(ic["box"]) if "box" in ic else (1.0, 1.0, 1.0))
    with MFMContext(params) as ctx:
        ctx.upload(ic)
        stats = {}
        if "density" in stages:
            s = ctx.density(active)
            stats.update(density_sweeps=s.density_sweeps, density_interactions=s.density_interactions,
                         density_updates=s.density_updates, density_fp64_redo=s.density_fp64_redo)
        if "gradients" in stages:
            s = ctx.gradients(active)
            stats.update(gradient_interactions=s.gradient_interactions)
        if "force" in stages:
            s = ctx.hydro_force(active)
            stats.update(force_interactions=s.force_interactions, riemann_retries=s.riemann_retries,
                         exact_solver_calls=s.exact_solver_calls, fp64_force_particles=s.fp64_force_particles)
        out = ctx.download()
        stats["timing"] = ctx.timing()
    return out, stats
