"""Pins the oracle's Riemann-solver fallbacks, effective-face switches, MFM
entropic switch and the hsml iteration's neighbour-number corrections against
textbook formulas, closed forms derived by hand for special states, and
invariants -- never against a retyping of the oracle's own expressions.

Reference functions covered (restated in oracle/mfm_oracle.c):
  Kurganov-Tadmor/PWK flux            hydro/reimann.h:484-519
  HLLC / Roe / PVRS star estimates    hydro/reimann.h:527-593
  face volume-weight switch, fallback hydro/compute_finitevol_faces.h:21,96-104
  MFM entropic switch                 hydro/hydro_core_meshless.h:335-374
  ncorr from ConditionNumber and FCE  hydro/density.c:565-595
  hsml update of the iteration        hydro/density.c:655-846
  WAKEUP flags of inactive neighbours hydro/hydro_evaluate.h:485-495
"""
import math

import numpy as np
import pytest

import ics
import oracle as O

GAMMA = 5.0 / 3.0


def sound(rho, p, gamma=GAMMA):
    return math.sqrt(gamma * p / rho)


def random_state(rng, vscale=1.0):
    return (rng.uniform(0.1, 3.0), rng.uniform(0.05, 5.0), *(vscale * rng.normal(size=3)))


def unit(rng):
    n = rng.normal(size=3)
    return n / np.linalg.norm(n)


# ----------------------------------------------------------------- KT/PWK ---
@pytest.mark.parametrize("seed", range(5))
def test_kt_equal_states_give_the_physical_flux(seed):
    """One state on both sides: P_M = p, F_v = p n, F_p = p (v.n) and the
    returned signal speed c + |v.n| (a consistent flux reproduces the exact
    flux of a uniform state).  v.n must not vanish: the PWK speed is |v.n|
    when the momentum jump is zero, and a zero speed leaves 0/0."""
    rng = np.random.default_rng(seed)
    for _ in range(40):
        rho, p, *v = random_state(rng)
        n = unit(rng)
        vn = float(np.dot(v, n))
        if abs(vn) < 1e-3:
            continue
        out = O.riemann_kt((rho, p, *v), (rho, p, *v), n=n)
        assert out["P_M"] == pytest.approx(p, rel=1e-13)
        np.testing.assert_allclose(out["Fv"], p * n, rtol=1e-12, atol=1e-13 * p)
        assert out["Fp"] == pytest.approx(p * vn, rel=1e-12, abs=1e-14 * p)
        assert out["S_M"] == pytest.approx(sound(rho, p) + abs(vn), rel=1e-14)


def test_kt_mirror_symmetry():
    """Swapping the sides and reversing the normal is the same face seen from
    the other particle: P_M and the signal speed are unchanged, the momentum
    and energy fluxes change sign."""
    rng = np.random.default_rng(21)
    for _ in range(300):
        L, R, n = random_state(rng), random_state(rng), unit(rng)
        a = O.riemann_kt(L, R, n=n)
        b = O.riemann_kt(R, L, n=-n)
        assert b["P_M"] == pytest.approx(a["P_M"], rel=1e-12)
        assert b["S_M"] == pytest.approx(a["S_M"], rel=1e-14)
        np.testing.assert_allclose(b["Fv"], -a["Fv"], rtol=1e-11, atol=1e-12 * np.abs(a["Fv"]).max())
        assert b["Fp"] == pytest.approx(-a["Fp"], rel=1e-10, abs=1e-12 * abs(a["P_M"]))


@pytest.mark.parametrize("w", [0.7, -0.4])
def test_kt_uniform_flow_is_upwinded(w):
    """Equal density and velocity, different pressures: the momentum jump is
    zero, the PWK switch alpha vanishes and the wave speed is |v.n| on both
    sides.  The flux is then pure upwinding: for v.n = w > 0, f_rho_right = 0
    and P_M = p_R, F_v = p_R n, F_p = w p_R; for w < 0 the left pressure."""
    rho, pL, pR = 1.3, 0.8, 2.1
    n = np.array([2.0, -1.0, 2.0]) / 3.0
    t = np.array([1.0, 2.0, 0.0]) / math.sqrt(5.0)  # tangential part, n.t = 0
    v = w * n + 0.3 * t
    out = O.riemann_kt((rho, pL, *v), (rho, pR, *v), n=n)
    p_up = pR if w > 0 else pL
    assert out["P_M"] == pytest.approx(p_up, rel=1e-13)
    np.testing.assert_allclose(out["Fv"], p_up * n, rtol=1e-12, atol=1e-13)
    assert out["Fp"] == pytest.approx(w * p_up, rel=1e-12)


@pytest.mark.parametrize("w,pL,pR", [(0.5, 1.0, 3.0), (2.0, 0.6, 0.2), (1e-4, 1.0, 1.5)])
def test_kt_converging_flow_closed_form(w, pL, pR):
    """Equal densities approaching head-on with speed w along n (plus a common
    tangential velocity).  Worked by hand from the PWK flux: the momentum jump
    is 2 rho w, alpha = 2 rho w / sqrt(delta^2 + 4 rho^2 w^2) with
    delta = 1e-3 rho (c_L + c_R)/2, the common wave speed S = alpha max(c) + w,
    f_rho_right = -f_rho_left = -rho (w + S), hence
        P_M = (p_L + p_R)/2,
        F_v = (rho w (w + S) + P_M) n,
        F_p = (p_R - p_L)/2 [S - gamma (w + S)/(gamma - 1)],
    the last from h_R - h_L = gamma (p_R - p_L)/((gamma - 1) rho)."""
    rho = 0.9
    n = np.array([0.0, 0.6, 0.8])
    t = np.array([1.0, 0.0, 0.0])
    vL = w * n + 0.25 * t
    vR = -w * n + 0.25 * t
    cL, cR = sound(rho, pL), sound(rho, pR)
    delta = 1e-3 * rho * 0.5 * (cL + cR)
    alpha = 2 * rho * w / math.sqrt(delta**2 + 4 * rho**2 * w**2)
    S = alpha * max(cL, cR) + w
    out = O.riemann_kt((rho, pL, *vL), (rho, pR, *vR), n=n)
    P = 0.5 * (pL + pR)
    assert out["P_M"] == pytest.approx(P, rel=1e-12)
    np.testing.assert_allclose(out["Fv"], (rho * w * (w + S) + P) * n, rtol=1e-12, atol=1e-14)
    g = GAMMA
    assert out["Fp"] == pytest.approx(0.5 * (pR - pL) * (S - g * (w + S) / (g - 1)), rel=1e-11)
    assert out["S_M"] == pytest.approx(max(cL, cR) + w, rel=1e-14)


# ------------------------------------------------- star-state estimates -----
def _normal_speeds(L, R, n):
    return float(np.dot(L[2:], n)), float(np.dot(R[2:], n))


def _rankine_hugoniot(L, R, n, est):
    """Both outer waves of an HLLC-type star state satisfy the momentum jump
    condition (Toro, Riemann Solvers, eqs. 10.36 and 10.38):
        P_M = p_K + rho_K (S_K - u_K)(S_M - u_K),  K = L, R."""
    uL, uR = _normal_speeds(L, R, n)
    pl = L[1] + L[0] * (est["S_L"] - uL) * (est["S_M"] - uL)
    pr = R[1] + R[0] * (est["S_R"] - uR) * (est["S_M"] - uR)
    scale = max(L[1], R[1], L[0] * (est["S_L"] - uL) ** 2, R[0] * (est["S_R"] - uR) ** 2)
    assert est["P_M"] == pytest.approx(pl, abs=1e-12 * scale)
    assert est["P_M"] == pytest.approx(pr, abs=1e-12 * scale)


def test_hllc_estimate_obeys_jump_conditions_with_davis_speeds():
    rng = np.random.default_rng(5)
    done = 0
    for _ in range(400):
        L, R, n = random_state(rng, 0.5), random_state(rng, 0.5), unit(rng)
        uL, uR = _normal_speeds(L, R, n)
        cL, cR = sound(L[0], L[1]), sound(R[0], R[1])
        est = O.pstar_estimate(L, R, O.PSTAR_HLLC, n=n)
        if est["P_M"] <= 1e-56:
            continue  # floored: S_L = S_R = S_M = 0 by the reference's rule
        # wave speeds bracket every signal of both states (Davis-type bounds)
        assert est["S_L"] == pytest.approx(min(uL, uR) - max(cL, cR), rel=1e-14, abs=1e-15)
        assert est["S_R"] == pytest.approx(max(uL, uR) + max(cL, cR), rel=1e-14, abs=1e-15)
        _rankine_hugoniot(L, R, n, est)
        done += 1
    assert done > 300


def _roe_einfeldt_speeds(L, R, n, gamma=GAMMA):
    """Einfeldt's bounds from Roe averages (Toro section 10.5.1): square-root
    density weights of velocity and total specific enthalpy
    H = |v|^2/2 + gamma p / ((gamma - 1) rho), c~^2 = (gamma - 1)(H~ - |v~|^2/2)."""
    sl, sr = math.sqrt(L[0]), math.sqrt(R[0])
    vl, vr = np.array(L[2:]), np.array(R[2:])
    v = (sl * vl + sr * vr) / (sl + sr)
    H = lambda s: 0.5 * np.dot(s[2:], s[2:]) + gamma * s[1] / ((gamma - 1) * s[0])
    Ht = (sl * H(L) + sr * H(R)) / (sl + sr)
    ct = math.sqrt(max(1e-30, (gamma - 1) * (Ht - 0.5 * np.dot(v, v))))
    uL, uR = _normal_speeds(L, R, n)
    un = float(np.dot(v, n))
    cL, cR = sound(L[0], L[1]), sound(R[0], R[1])
    return min(uL - cL, un - ct), max(uR + cR, un + ct)


def test_roe_estimate_is_hllc_with_einfeldt_speeds():
    rng = np.random.default_rng(6)
    done = 0
    for _ in range(400):
        L, R, n = random_state(rng, 0.5), random_state(rng, 0.5), unit(rng)
        est = O.pstar_estimate(L, R, O.PSTAR_ROE, n=n)
        if est["P_M"] <= 1e-56:
            continue
        SL, SR = _roe_einfeldt_speeds(L, R, n)
        assert est["S_L"] == pytest.approx(SL, rel=1e-13, abs=1e-14)
        assert est["S_R"] == pytest.approx(SR, rel=1e-13, abs=1e-14)
        _rankine_hugoniot(L, R, n, est)
        done += 1
    assert done > 300


def test_roe_and_pvrs_return_an_equal_state():
    for rho, p, v in ((1.0, 1.0, 0.3), (2.5, 0.01, -1.0), (0.1, 10.0, 0.0)):
        s = (rho, p, v, 0.2, -0.1)
        for which in (O.PSTAR_ROE, O.PSTAR_PVRS):
            est = O.pstar_estimate(s, s, which)
            assert est["P_M"] == pytest.approx(p, rel=1e-13)
            assert est["S_M"] == pytest.approx(v, rel=1e-13, abs=1e-15)


def test_pvrs_matches_toro_primitive_variable_solver():
    """Toro eqs. 9.20-9.21 with rho_bar = (rho_L + rho_R)/2, a_bar = (c_L + c_R)/2:
        p* = (p_L + p_R)/2 - (u_R - u_L) rho_bar a_bar / 2,
        u* = (u_L + u_R)/2 - (p_R - p_L) / (2 rho_bar a_bar),
    and, where u* leaves [-S+, S+] (S+ the largest |u -+ c| of both states), the
    reference's clamp to that bound."""
    rng = np.random.default_rng(7)
    clamped = 0
    for _ in range(400):
        L, R, n = random_state(rng, 0.5), random_state(rng, 0.5), unit(rng)
        uL, uR = _normal_speeds(L, R, n)
        cL, cR = sound(L[0], L[1]), sound(R[0], R[1])
        rb, ab = 0.5 * (L[0] + R[0]), 0.5 * (cL + cR)
        p_star = 0.5 * (L[1] + R[1]) - 0.5 * (uR - uL) * rb * ab
        u_star = 0.5 * (uL + uR) - 0.5 * (R[1] - L[1]) / (rb * ab)
        s_plus = max(abs(uL - cL), abs(uR - cR), abs(uL + cL), abs(uR + cR))
        est = O.pstar_estimate(L, R, O.PSTAR_PVRS, n=n)
        if p_star <= 1e-56:
            assert est["P_M"] == 1e-56 and est["S_M"] == 0.0
            continue
        assert est["P_M"] == pytest.approx(p_star, rel=1e-12)
        if abs(u_star) > s_plus:
            clamped += 1
            assert est["S_M"] == pytest.approx(math.copysign(s_plus, u_star), rel=1e-14)
        else:
            assert est["S_M"] == pytest.approx(u_star, rel=1e-12, abs=1e-14)
        assert est["S_L"] == pytest.approx(-s_plus, rel=1e-14)
        assert est["S_R"] == pytest.approx(s_plus, rel=1e-14)
    # gas at rest against a near-vacuum: u* = (p_L - p_R)/(2 rho_bar a_bar)
    # = 1.41 exceeds the fastest signal c_L = 1.29 and is clamped to it
    L, R = (1.0, 1.0, 0.0, 0.0, 0.0), (1e-6, 1e-8, 0.0, 0.0, 0.0)
    est = O.pstar_estimate(L, R, O.PSTAR_PVRS)
    s_plus = sound(1.0, 1.0)
    rb, ab = 0.5 * (1.0 + 1e-6), 0.5 * (sound(1.0, 1.0) + sound(1e-6, 1e-8))
    assert 0.5 * (1.0 - 1e-8) / (rb * ab) > s_plus
    assert est["S_M"] == pytest.approx(s_plus, rel=1e-14)
    assert est["P_M"] == pytest.approx(0.5 * (1.0 + 1e-8), rel=1e-14)


def test_estimate_cascade_follows_the_limiter():
    """The HLLC estimate stands unless P_M is <= 0, NaN or above the pressure
    limiter; then the Roe estimate under the same rule; then PVRS whatever it
    gives.  A receding pair faster than the largest sound speed is vacuum."""
    rng = np.random.default_rng(8)
    seen = set()
    for _ in range(400):
        L, R, n = random_state(rng, 0.5), random_state(rng, 0.5), unit(rng)
        uL, uR = _normal_speeds(L, R, n)
        if uR - uL > max(sound(L[0], L[1]), sound(R[0], R[1])):
            continue
        e = [O.pstar_estimate(L, R, k, n=n) for k in (O.PSTAR_HLLC, O.PSTAR_ROE, O.PSTAR_PVRS)]
        ph, pr = e[0]["P_M"], e[1]["P_M"]
        for lim in (1.5 * ph, 0.5 * (ph + pr), 0.5 * min(ph, pr)):
            got = O.wavespeeds(L, R, n=n, press_tot_limiter=lim)
            if ph <= lim:
                want = 1
            elif pr <= lim:
                want = 2
            else:
                want = 3
            assert got["which"] == want
            assert got["P_M"] == e[want - 1]["P_M"] and got["S_M"] == e[want - 1]["S_M"]
            seen.add(want)
    assert seen == {1, 2, 3}
    vac = O.wavespeeds((1.0, 1.0, -3.0, 0, 0), (1.0, 1.0, 3.0, 0, 0))
    assert vac["which"] == 0 and vac["P_M"] == 1e-56 and vac["S_M"] == 0.0


# ----------------------------------------------------- two-particle pairs ---
def pair_state(rho=(1.0, 1.0), P=(1.0, 1.0), V=(1e-3, 1e-3), a=(1.0, 1.0), cn=(1.0, 1.0), fce=(0.0, 0.0),
               r=0.08, h=0.2):
    """Particles i (index 0) and j (1) on the x axis, i right of j by r, at
    rest, no gradients, NV_T^-1 = a_k I: dp = x_i - x_j = (r, 0, 0)."""
    pos = np.array([[0.5 + 0.5 * r, 0.5, 0.5], [0.5 - 0.5 * r, 0.5, 0.5]])
    rho = np.array(rho, dtype=float)
    ic = dict(pos=pos, vel=np.zeros((2, 3)), mass=rho * np.array(V), u=np.ones(2), hsml=np.full(2, h),
              density=rho, pressure=np.array(P, dtype=float),
              nvt=np.stack([ak * np.eye(3).reshape(9) for ak in a]), cn=np.array(cn, dtype=float),
              fce=np.array(fce, dtype=float), dhsml=np.ones(2), box=np.ones(3))
    return O.OracleState(ic)


def face_of(st):
    out, rc = O.force_pair(st, 0, 1)
    assert rc == 0
    return out[:3], out[3], int(out[5])


@pytest.mark.parametrize("ratio", [4.7, 4.8])
def test_face_volume_weight_switch(ratio):
    """Equal pressures at rest: the Riemann pressure is p and S_M = 0, so
    F_v = p A with A the effective face.  Equal hsml gives wk_i = wk_j = wk.
    V_j / V_i = 4.7 puts (|V_i - V_j| / min V)/3 = 1.233 below the 1.25 switch:
    the Lanson-Vila weights wt = V give |A| = wk (V_i a_i + V_j a_j) r.  At 4.8
    (1.267) both weights become V_i V_j (wk_i + wk_j)/(V_i wk_i + V_j wk_j),
    which for equal wk is the harmonic mean 2 V_i V_j/(V_i + V_j)."""
    r, h, p = 0.08, 0.2, 1.7
    Vi, Vj = 1e-3, ratio * 1e-3
    ai, aj = 3.0, 1.0
    st = pair_state(P=(p, p), V=(Vi, Vj), a=(ai, aj), r=r, h=h)
    Fv, Fp, mask = face_of(st)
    wk = O.kernel(r / h, h)[0]
    if ratio < 4.75:
        area = wk * (Vi * ai + Vj * aj) * r
    else:
        area = wk * 2 * Vi * Vj / (Vi + Vj) * (ai + aj) * r
    np.testing.assert_allclose(Fv, [p * area, 0.0, 0.0], rtol=1e-13, atol=1e-15 * p * area)
    assert abs(Fp) <= 1e-12 * p * area  # S_M = 0, no motion


@pytest.mark.parametrize("case", ["negative_dot", "cn_j"])
def test_face_fallback_to_the_kernel_gradient_face(case):
    """When A.dp < 0 (NV_T^-1 of j negative definite) or CN_j^2 exceeds
    1e12 + (1e6 - CN_i^2), the face becomes the SPH-like one along dp with
    |A| = -(wt_i V_i dwk_i + wt_j V_j dwk_j); equal volumes and hsml give
    |A| = -2 V^2 dwk."""
    r, h, p, V = 0.08, 0.2, 1.3, 2e-3
    if case == "negative_dot":
        st = pair_state(P=(p, p), V=(V, V), a=(1.0, -5.0), r=r, h=h)
    else:
        st = pair_state(P=(p, p), V=(V, V), a=(1.0, 1.0), cn=(1.0, 1.1e6), r=r, h=h)
    Fv, _, _ = face_of(st)
    dwk = O.kernel(r / h, h)[1]
    np.testing.assert_allclose(Fv, [p * (-2 * V * V * dwk), 0.0, 0.0], rtol=1e-13)


# The entropic switch.  With both particles at rest and no gradients the face
# states are the particles' own, v_frame = 0, PdV_i = PdV_j = 0 and the face
# velocities vanish, so du_new = 0 while the unswitched energy flux is
# du_old = |A| P_M S_M = S_M |F_v|.  The pair therefore returns F_p = S_M |F_v|
# where the switch keeps the Riemann energy flux and F_p = 0 where it replaces
# it.  R is particle i's state and L particle j's (the face normal points from
# j to i), so the contact moves towards j (S_M < 0) when P_i > P_j.
ENTROPIC_CASES = {
    # name: (rho_i, P_i, rho_j, P_j, cn_j, fce, switched, regime of |S_M|/c_min)
    # P_i/rho_i > P_j/rho_j, S_M < 0: dtoj = -du_old > 0 keeps the flux
    "ei_gt_ej_flux_to_j": (1.0, 1.2, 1.0, 1.0, 1.0, 0.0, False, "mid"),
    # P_i/rho_i > P_j/rho_j but P_i < P_j: S_M > 0, dtoj < 0 and not above
    # -du_new + ... = 0: switched
    "ei_gt_ej_flux_to_i": (0.5, 1.0, 1.0, 1.2, 1.0, 0.0, True, "mid"),
    # mirror: P_i/rho_i < P_j/rho_j, S_M > 0: dtoi = du_old > 0 keeps it
    "ei_lt_ej_flux_to_i": (1.0, 1.0, 1.0, 1.2, 1.0, 0.0, False, "mid"),
    "ei_lt_ej_flux_to_j": (1.0, 1.2, 0.5, 1.0, 1.0, 0.0, True, "mid"),
    # |S_M| below 1e-3 c: no comparison, the entropic form is used
    "tiny_contact_speed": (1.0, 1.0001, 1.0, 1.0, 1.0, 0.0, True, "tiny"),
    # CN_j^2 >= 1e6 - CN_i^2 forces the entropic form
    "ill_conditioned_j": (1.0, 1.2, 1.0, 1.0, 1001.0, 0.0, True, "mid"),
    # |S_M| >= c/2 and no face leak: the switch is not entered (CN_j would force it)
    "fast_contact": (1.0, 10.0, 1.0, 0.1, 1001.0, 0.0, False, "fast"),
    # face leak above 1 enters the switch even for a fast contact
    "fast_contact_leaky_face": (1.0, 10.0, 1.0, 0.1, 1001.0, 1.5, True, "fast"),
}


@pytest.mark.parametrize("name", list(ENTROPIC_CASES))
def test_entropic_switch_branches(name):
    rho_i, P_i, rho_j, P_j, cn_j, fce, switched, regime = ENTROPIC_CASES[name]
    st = pair_state(rho=(rho_i, rho_j), P=(P_i, P_j), cn=(1.0, cn_j), fce=(fce, fce))
    Fv, Fp, mask = face_of(st)
    assert mask == 1  # HLLC only: neither KT, the exact solver nor the retry
    # the pair's Riemann problem on its own (pinned above): j left, i right
    rs = O.riemann((rho_j, P_j, 0, 0, 0), (rho_i, P_i, 0, 0, 0), n=(1, 0, 0),
                   press_tot_limiter=1.1 * max(P_i, P_j))
    S_M = rs["S_M"]
    ratio = abs(S_M) / min(sound(rho_i, P_i), sound(rho_j, P_j))
    assert {"tiny": ratio < 1e-3, "mid": 1e-3 < ratio < 0.5, "fast": ratio > 0.5}[regime], ratio
    if name.endswith("flux_to_j"):
        assert S_M < 0
    if name.endswith("flux_to_i"):
        assert S_M > 0
    area_pm = np.linalg.norm(Fv)
    assert area_pm > 0 and abs(S_M) * area_pm > 0
    if switched:
        assert abs(Fp) <= 1e-14 * abs(S_M) * area_pm
    else:
        assert Fp == pytest.approx(S_M * area_pm, rel=1e-12)


# ------------------------------------------- neighbour-number corrections ---
def ncorr_bound(nc, iterations, des=32.0, dev0=0.05):
    """Half-width of the converged window: dev0 * nc, widened after the
    second sweep by exp(0.1 log(des/(16 dev)) iter) and capped at des/4
    (density.c:594).  The window only grows with iter, so the last sweep
    index bounds every particle's."""
    d, dv = des * nc, dev0 * nc
    it = max(iterations - 1, 0)
    if it > 1:
        dv = min(0.25 * d, dv * math.exp(0.1 * math.log(d / (16.0 * dv)) * it))
    return dv


def ncorr_cn(cn):
    """sqrt(1 + (cn - c0)/1e3) above c0 = 100, at most 2 (density.c:585)."""
    return 1.0 if cn <= 100.0 else min(math.sqrt(1.0 + (cn - 100.0) / 1000.0), 2.0)


@pytest.mark.parametrize("cn_in", [400.0, 1600.0, 5000.0])
def test_stored_condition_number_raises_the_neighbour_target(cn_in):
    """A large ConditionNumber from the previous step (above every value this
    lattice produces, so the first sweep does not replace it) raises the
    neighbour target to 32 ncorr with ncorr = sqrt(1 + (cn - 100)/1000) capped
    at 2: 36.5, 43.5 and 64 neighbours.  On convergence the particle stores the
    condition number of its own final NV_T for the next step (density.c:719)."""
    ic = ics.lattice_gaussian(12)
    ic = dict(ic, cn=np.full(ic["pos"].shape[0], cn_in))
    out, st, _ = O.run_pass(ic, stages=("density",))
    assert np.all(out["cn"] < 10.0)
    nc = ncorr_cn(cn_in)
    n_eff = out["num_ngb"] ** 3
    dev = ncorr_bound(nc, st.iterations)
    assert np.all(np.abs(n_eff - 32.0 * nc) <= dev * (1 + 1e-12)), (n_eff.min(), n_eff.max(), 32 * nc, dev)
    assert abs(n_eff.mean() - 32.0 * nc) < 0.5 * dev + 1e-9


def test_face_closure_error_raises_the_neighbour_target():
    """An isolated 6^3 block of particles in an empty periodic box: the corner
    particles see one-sided kernels with face closure errors above
    d00 = 0.35, and their target becomes 32 max(ncorr(cn), min(FCE/0.35, 2))
    (density.c:589-591) with the FCE of the last sweep.  A stored
    ConditionNumber of 0 is never replaced during the iteration
    (density.c:574 needs it positive), so the condition number's share of
    ncorr is exactly 1 here."""
    w, s = 6, 1.0 / 14
    g = 0.3 + s * np.arange(w)
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    N = pos.shape[0]
    ic = dict(pos=pos, vel=np.zeros((N, 3)), mass=np.full(N, 1.0 / N), u=np.ones(N),
              hsml=np.full(N, 0.15), cn=np.zeros(N), box=np.ones(3))
    out, st, _ = O.run_pass(ic, stages=("density",))
    fce = out["fce"]
    nc = np.where(fce > 0.35, np.maximum(1.0, np.minimum(fce / 0.35, 2.0)), 1.0)
    n_eff = out["num_ngb"] ** 3
    dev = np.array([ncorr_bound(k, st.iterations) for k in nc])
    assert np.all(np.abs(n_eff - 32.0 * nc) <= dev * (1 + 1e-12))
    corner = fce > 0.35
    edge = (pos == g[0]) | (pos == g[-1])
    np.testing.assert_array_equal(corner, edge.all(axis=1))  # exactly the 8 corners
    assert np.all(n_eff[corner] > 32.0 * 1.1)
    assert np.all(np.abs(n_eff[~corner] - 32.0) <= dev[~corner] * (1 + 1e-12))


# ------------------------------------------------------ the hsml update -----
def test_hsml_step_inside_the_window_keeps_hsml():
    r, h, l, rr = O.hsml_step(32.02, 1.0, 0.7, 0.6, 0.8)
    assert (r, h, l, rr) == (O.HSML_DONE, 0.7, 0.6, 0.8)


@pytest.mark.parametrize("n_ngb,dhsml,expect", [
    # without a bracket the step is the Newton step of a uniform medium,
    # N ~ h^3: h -> h (32/N)^(1/3) (slope DhsmlNgbFactor = 1) ...
    (16.0, 1.0, 2.0 ** (1 / 3)),
    (64.0, 1.0, 0.5 ** (1 / 3)),
    # ... with a slope the step scales, capped 0.231 beyond the uniform step
    (16.0, 0.5, 2.0 ** (1 / 6)),
    (4.0, 2.0, 2.0 * math.exp(0.231)),
    (16.0, 2.15, 2.0 ** (1 / 3) * math.exp(0.231)),
    (60.0, 3.0, (32 / 60) ** (1 / 3) * math.exp(-0.231)),
    # outside (0.1, 2) x 32 the uniform step alone, shrinking by at most e^-1.535
    (2.0, 5.0, 16.0 ** (1 / 3)),
    (100.0, 5.0, (32 / 100) ** (1 / 3)),
    (1e6, 1.0, math.exp(-1.535)),
    # fewer than one neighbour: e^1.4
    (0.5, 1.0, math.exp(1.4)),
])
def test_hsml_step_one_sided(n_ngb, dhsml, expect):
    r, h, l, rr = O.hsml_step(n_ngb, dhsml, 1.0)
    assert r == O.HSML_AGAIN
    assert h == pytest.approx(expect, rel=1e-14)
    # the current hsml becomes the bracket on its side
    assert (l, rr) == ((1.0, 0.0) if n_ngb < 32 else (0.0, 1.0))


def test_hsml_step_bracketed():
    # inside the bracket: the Newton step of a uniform medium, accepted
    r, h, l, rr = O.hsml_step(30.0, 1.0, 1.0, 0.9, 1.2)
    assert (r, l, rr) == (O.HSML_AGAIN, 1.0, 1.2) and h == pytest.approx((32 / 30) ** (1 / 3), rel=1e-14)
    # past the bracket: the step is clamped to it and replaced by the geometric
    # mean of (clamped h, Left, Right) = (1.1 * 1.0 * 1.1)^(1/3)
    r, h, l, rr = O.hsml_step(20.0, 1.0, 1.0, 0.9, 1.1)
    assert h == pytest.approx(1.21 ** (1 / 3), rel=1e-14)
    # from the third sweep a step smaller than a fifth of the bracket's log
    # width is widened to it: h = Left * (Right/Left)^0.2
    r, h, l, rr = O.hsml_step(31.5, 1.0, 1.0, 0.95, 1.5, iter=2)
    assert (l, rr) == (1.0, 1.5) and h == pytest.approx(1.5 ** 0.2, rel=1e-14)
    # a bracket narrower than 1e-3 of Left is accepted as it stands
    assert O.hsml_step(30.0, 1.0, 1.0002, 1.0, 1.0005) == (O.HSML_BRACKETED, 1.0002, 1.0, 1.0005)


def test_hsml_step_window_grows_with_the_sweep():
    """The window half-width grows from the third sweep on as
    dev (des / (16 dev))^(0.1 iter), at most des/4 (density.c:594): for
    des = 32, dev = 0.05 that is 0.05 * 40^(0.1 iter) -- 0.151 at sweep 3,
    0.455 at sweep 6 and the cap of 8 from sweep 15."""
    for it, width in ((1, 0.05), (3, 0.05 * 40 ** 0.3), (6, 0.05 * 40 ** 0.6), (15, 8.0)):
        inside, outside = 32.0 + 0.98 * width, 32.0 - 1.02 * width
        assert O.hsml_step(inside, 1.0, 1.0, iter=it)[0] == O.HSML_DONE, it
        assert O.hsml_step(outside, 1.0, 1.0, iter=it)[0] == O.HSML_AGAIN, it


def test_hsml_step_limits_and_collapse():
    # below the target at MaxHsml: set to MaxHsml and swept once more, then kept
    r, h, _, _ = O.hsml_step(10.0, 1.0, 0.9995, max_hsml=1.0)
    assert (r, h) == (O.HSML_AGAIN, 1.0)
    assert O.hsml_step(10.0, 1.0, 1.0, max_hsml=1.0)[0] == O.HSML_DONE
    # above the target at MinHsml: likewise
    r, h, _, _ = O.hsml_step(90.0, 1.0, 0.1001, min_hsml=0.1)
    assert (r, h) == (O.HSML_AGAIN, 0.1)
    assert O.hsml_step(90.0, 1.0, 0.1, min_hsml=0.1)[0] == O.HSML_DONE
    # no bracket on either side (h = 0): the iteration cannot go on
    assert O.hsml_step(10.0, 1.0, 0.0)[0] == O.HSML_COLLAPSED


# ------------------------------------------------------------- WAKEUP -------
def test_wakeup_flags_inactive_neighbours_of_a_fast_signal():
    """An inactive neighbour j of an active particle is flagged when the
    pair's signal speed exceeds WAKEUP = 4.1 times j's stored MaxSignalVel
    (hydro_evaluate.h:485-495, allvars.h:78).  A perfect 10^3 lattice at
    rest with uniform u: every pair's signal speed is c_i + c_j = 2c, the
    MaxSignalVel of the full pass.  The active set is the slab x < 0.2; the
    inactive particles store 2c/4.11 (flagged beside an active particle) or
    2c/4.09 (not flagged) by the parity of their y row, and so do the active
    ones, which are never flagged.  One active particle moves along -y at
    0.1c: its pairs with neighbours ahead of it approach, and the signal speed
    2c + 0.1c cos(theta) >= 2.05c passes 4.1 * 2c/4.09 = 2.005c, while the
    pairs beside and behind it (no approach) do not, whatever the particle's
    largest signal speed.  Inactive particles out of every active kernel are
    never flagged."""
    n = 10
    ic = ics.lattice_gaussian(n, sigma_v=0.0, jitter=0.0)
    full, _, _ = O.run_pass(ic)
    h = full["hsml"]
    assert np.ptp(h) == 0.0
    c = np.sqrt(GAMMA * full["pressure"] / full["density"])
    np.testing.assert_allclose(full["vsig"], 2 * c, rtol=1e-14)

    pos = ic["pos"]
    N = pos.shape[0]
    ix, iy = np.rint(pos[:, 0] * n - 0.5).astype(int), np.rint(pos[:, 1] * n - 0.5).astype(int)
    active = np.flatnonzero(ix < 2).astype(np.int32)
    is_active = ix < 2
    k = int(np.flatnonzero((ix == 1) & (iy == 4) & (np.rint(pos[:, 2] * n - 0.5) == 5))[0])
    state = dict(ic)
    state.update({key: full[key] for key in full if not key.endswith("_l1") and key != "wakeup"})
    slow = iy % 2 == 0
    state["vsig"] = np.where(slow, full["vsig"] / 4.11, full["vsig"] / 4.09)
    vel = np.zeros((N, 3))
    vel[k, 1] = -0.1 * c[k]
    state["vel"] = vel
    out, _, _ = O.run_pass(state, active=active)

    d = pos[None, :, :] - pos[active][:, None, :]
    d -= np.rint(d)                                   # periodic unit box
    r2 = (d ** 2).sum(axis=2)
    assert np.abs(r2 - h[0] ** 2).min() > 1e-9       # no pair on the kernel edge
    inside = (r2 < h[0] ** 2) & (r2 > 0)
    near_active = inside.any(axis=0)
    row_k = int(np.flatnonzero(active == k)[0])
    ahead_of_k = inside[row_k] & (d[row_k, :, 1] < 0)
    want = ~is_active & near_active & (slow | ahead_of_k)
    assert want.sum() > 50 and (want & ~slow).sum() == 3  # x row 2, y row 3
    assert (~is_active & near_active & ~slow & ~ahead_of_k).sum() > 50
    np.testing.assert_array_equal(out["wakeup"].astype(bool), want)
