"""Error metrics shared by the parity tests and smoke().

Summed outputs (divergence, face-closure error, gradients, accelerations, du/dt)
are compared relative to the sum of the absolute values of the pair terms
behind them (`*_l1` from the oracle): a different summation order can change
such a sum by at most ~n*eps*sum|terms|, even where the sum itself cancels to
nearly zero (a uniform lattice, a cold uniform gas).  Positive sums (density,
neighbour number, hsml, signal speed) are compared with the plain relative error.
"""
import numpy as np

# tolerances of the plain FP64 pass against the FP64 oracle
TOL_DENSITY = 1e-12   # density, hsml, NumNgb, pressure (relative)
TOL_FLUX = 1e-10      # gradients, dv/dt, du/dt (relative to sum|terms|)

# Gradients: a difference Q_j - Q_i also carries the rounding of the inputs
# Q themselves (density and pressure agree with the oracle to ~1e-15, not
# bit for bit), which no sum over |pair terms| bounds where Q is uniform up to
# rounding.  The gradient error scale adds INPUT_ROUNDING * grad_in (grad_in
# = sum |w dp| (|Q_i| + |Q_j|) from the oracle): |error| <= TOL_FLUX *
# grad_l1 + INPUT_ROUNDING * grad_in, ten times the measured input error.
INPUT_ROUNDING = 1e-14

# ConditionNumber = |NV_T|_F |NV_T^-1|_F / 3 (system/system.c:355-391) is at
# least a third of the 2-norm condition number kappa of NV_T, and a relative
# perturbation d of NV_T (its sums in another order: d <= ~n eps ~ 1e-14 for
# n ~ 50 neighbours) moves the inverse by up to kappa d.  The condition number
# is therefore held to max(TOL_DENSITY, COND_ROUNDING (1 + 3 cn)): on a lattice
# (cn ~ 1) that is TOL_DENSITY, in a Plummer core (cn in the hundreds) a few
# 1e-12.
COND_ROUNDING = 1e-14

# tolerances of the FP32-bulk pass (precision 1) against the FP64 oracle
TOL32_DENSITY = 1e-5  # density, hsml, pressure (relative)
TOL32_FLUX = 1e-4     # dv/dt, du/dt (relative to sum|terms|)


def rel_errs(a, b):
    """element-wise |a - b| / |b| (a zero reference must match exactly)"""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.abs(b)
    bad = den == 0
    err = np.abs(a - b) / np.where(bad, 1.0, den)
    return np.where(bad, np.abs(a - b) * 1e300, err)


def l1_errs(a, b, l1):
    """element-wise |a - b| / sum|terms|; the scale is floored at 1e-3 of the
    field's median sum|terms| so that particles whose terms all vanish (a
    static cold gas) do not turn denormal-level differences into huge ratios."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    l1 = np.asarray(l1, dtype=np.float64).reshape(a.shape)
    pos = l1[l1 > 0]
    floor = 1e-3 * float(np.median(pos)) if pos.size else 0.0
    den = np.maximum(l1, floor)
    err = np.abs(a - b) / np.where(den > 0, den, 1.0)
    return np.where(den > 0, err, np.abs(a - b) * 1e300)  # all-zero field must match exactly


def rel_err(a, b):
    e = rel_errs(a, b)
    return float(e.max()) if e.size else 0.0


def l1_err(a, b, l1):
    e = l1_errs(a, b, l1)
    return float(e.max()) if e.size else 0.0


def grad_scale(ref, sel=lambda d, k: np.asarray(d[k])):
    """Error scale of the 15 gradient components (see INPUT_ROUNDING)."""
    l1 = sel(ref, "grad_l1")
    if "grad_in" in ref:
        l1 = l1 + (INPUT_ROUNDING / TOL_FLUX) * sel(ref, "grad_in")
    return l1


def compare_pass(out, ref, active=None):
    """Per-output errors of a GPU pass `out` against oracle outputs `ref`."""
    idx = slice(None) if active is None else np.asarray(active)
    sel = lambda d, k: np.asarray(d[k])[idx]
    e = {}
    for k in ("hsml", "density", "pressure", "num_ngb", "dhsml", "vsig"):
        e[k] = rel_err(sel(out, k), sel(ref, k))
    cn = np.abs(sel(ref, "cn"))
    cn_scale = np.maximum(1.0, COND_ROUNDING * (1.0 + 3.0 * cn) / TOL_DENSITY)
    ecn = rel_errs(sel(out, "cn"), sel(ref, "cn")) / cn_scale
    e["cn"] = float(ecn.max()) if ecn.size else 0.0
    e["nvt"] = l1_err(sel(out, "nvt"), sel(ref, "nvt"),
                      np.repeat(np.abs(sel(ref, "nvt")).max(axis=1, keepdims=True), 9, axis=1))
    e["divvel"] = l1_err(sel(out, "divvel"), sel(ref, "divvel"), sel(ref, "divvel_l1"))
    e["fce"] = l1_err(sel(out, "fce"), sel(ref, "fce"), sel(ref, "fce_l1"))
    g_out = np.concatenate([sel(out, "grad_rho"), sel(out, "grad_p"), sel(out, "grad_v").reshape(-1, 9)], axis=1)
    g_ref = np.concatenate([sel(ref, "grad_rho"), sel(ref, "grad_p"), sel(ref, "grad_v").reshape(-1, 9)], axis=1)
    e["gradients"] = l1_err(g_out, g_ref, grad_scale(ref, sel))
    e["acc"] = l1_err(sel(out, "acc"), sel(ref, "acc"), sel(ref, "acc_l1"))
    # du/dt: particles whose entropic-switch branch is decided by a P/rho tie
    # (oracle tie_flag) may legitimately take the other branch; they are held
    # to |error| <= sum|terms| and counted, every other particle to TOL_FLUX.
    tie = sel(ref, "tie_flag").astype(bool) if "tie_flag" in ref else np.zeros(len(sel(ref, "dtu")), bool)
    e["dtu"] = l1_err(sel(out, "dtu")[~tie], sel(ref, "dtu")[~tie], sel(ref, "dtu_l1")[~tie])
    e["dtu_tied"] = l1_err(sel(out, "dtu")[tie], sel(ref, "dtu")[tie], sel(ref, "dtu_l1")[tie]) if tie.any() else 0.0
    e["n_tied"] = int(tie.sum())
    return e


def check_pass(out, ref, active=None, tol_density=TOL_DENSITY, tol_flux=TOL_FLUX):
    e = compare_pass(out, ref, active)
    limits = {k: tol_density for k in ("hsml", "density", "pressure", "num_ngb", "dhsml", "cn", "vsig")}
    limits.update({k: tol_flux for k in ("nvt", "divvel", "fce", "gradients", "acc", "dtu")})
    limits["dtu_tied"] = 1.0 + 1e-12
    limits["n_tied"] = float("inf")
    bad = {k: (v, limits[k]) for k, v in e.items() if not (v <= limits[k])}
    return e, bad


def error_stats(out, ref, active=None):
    """Max, rms, 99.9th and 99.99th percentile of the per-particle error of
    each output against the oracle, with the normalisations of compare_pass,
    and the number of particles above the FP32-bulk tolerance (n_over); du/dt
    split into the particles whose entropic-switch branch is not / is a P/rho
    tie."""
    idx = slice(None) if active is None else np.asarray(active)
    sel = lambda d, k: np.asarray(d[k])[idx]
    errs = {}
    for k in ("hsml", "density", "pressure", "num_ngb", "vsig"):
        errs[k] = rel_errs(sel(out, k), sel(ref, k))
    g_out = np.concatenate([sel(out, "grad_rho"), sel(out, "grad_p"), sel(out, "grad_v").reshape(-1, 9)], axis=1)
    g_ref = np.concatenate([sel(ref, "grad_rho"), sel(ref, "grad_p"), sel(ref, "grad_v").reshape(-1, 9)], axis=1)
    gs = grad_scale(ref, sel)
    errs["grad_rho"] = l1_errs(g_out[:, 0:3], g_ref[:, 0:3], gs[:, 0:3])
    errs["grad_p"] = l1_errs(g_out[:, 3:6], g_ref[:, 3:6], gs[:, 3:6])
    errs["grad_v"] = l1_errs(g_out[:, 6:], g_ref[:, 6:], gs[:, 6:])
    errs["acc"] = l1_errs(sel(out, "acc"), sel(ref, "acc"), sel(ref, "acc_l1"))
    tie = sel(ref, "tie_flag").astype(bool) if "tie_flag" in ref else np.zeros(len(sel(ref, "dtu")), bool)
    errs["dtu"] = l1_errs(sel(out, "dtu")[~tie], sel(ref, "dtu")[~tie], sel(ref, "dtu_l1")[~tie])
    errs["dtu_tied"] = l1_errs(sel(out, "dtu")[tie], sel(ref, "dtu")[tie], sel(ref, "dtu_l1")[tie])
    tol = {"hsml": TOL32_DENSITY, "density": TOL32_DENSITY, "pressure": TOL32_DENSITY,
           "acc": TOL32_FLUX, "dtu": TOL32_FLUX}
    stats = {}
    for k, e in errs.items():
        e = np.asarray(e, dtype=np.float64)
        per = e.reshape(e.shape[0], -1).max(axis=1) if e.size else e  # worst component per particle
        stats[k] = {"max": float(per.max()) if per.size else 0.0,
                    "rms": float(np.sqrt(np.mean(per * per))) if per.size else 0.0,
                    "p99.9": float(np.quantile(per, 0.999)) if per.size else 0.0,
                    "p99.99": float(np.quantile(per, 0.9999)) if per.size else 0.0}
        if k in tol:
            stats[k]["n_over"] = int((per > tol[k]).sum())
    stats["n_tied"] = int(tie.sum())
    stats["n_particles"] = int(len(sel(ref, "hsml")))
    return stats


# MFV outputs (method 1): ParticleVel against the sum of |terms| of its
# weighted mean, DtMass / dMass / GravWorkTerm against the sums of |pair
# terms| (dMass: a particle's own shares and those its neighbours give it)
MFV_FIELDS = (("particle_vel", "pvel_l1"), ("dt_mass", "dtmass_l1"), ("d_mass", "dmass_l1"),
              ("grav_work", "gravwork_l1"))


def compare_mfv(out, ref, active=None):
    """Errors of the MFV outputs of a GPU pass against the oracle's (over the
    active particles, or every particle for d_mass, which inactive particles
    receive too); plus compare_pass's errors of the shared outputs."""
    e = compare_pass(out, ref, active)
    idx = slice(None) if active is None else np.asarray(active)
    for k, l1 in MFV_FIELDS:
        sel = slice(None) if k == "d_mass" else idx
        e[k] = l1_err(np.asarray(out[k])[sel], np.asarray(ref[k])[sel], np.asarray(ref[l1])[sel])
    return e


def check_mfv(out, ref, active=None, tol_density=TOL_DENSITY, tol_flux=TOL_FLUX):
    e = compare_mfv(out, ref, active)
    limits = {k: tol_density for k in ("hsml", "density", "pressure", "num_ngb", "dhsml", "cn", "vsig")}
    limits.update({k: tol_flux for k in ("nvt", "divvel", "fce", "gradients", "acc", "dtu")})
    limits.update({k: tol_flux for k, _ in MFV_FIELDS})
    limits["dtu_tied"] = 0.0  # MFV has no entropic switch, so no tied particles
    limits["n_tied"] = 0
    bad = {k: (v, limits[k]) for k, v in e.items() if not (v <= limits[k])}
    return e, bad
