"""Pins the oracle's element functions against closed forms, published
solutions and hand-computed cases (no formula of the oracle is retyped here)."""
import math

import numpy as np
import pytest

import oracle as O

NORM = 8.0 / math.pi  # cubic-spline 3D normalisation (kernel.h:63)


# ---------------------------------------------------------------- kernel ----
def test_kernel_values_and_continuity():
    h = 0.7
    w0, dw0 = O.kernel(0.0, h)
    assert w0 == pytest.approx(NORM / h**3, rel=1e-15)
    assert dw0 == 0.0
    # both branches meet at u = 1/2 with W = 1/4 and dW/du = -3/2 (unnormalised)
    wl, dwl = O.kernel(np.nextafter(0.5, 0), h)
    wr, dwr = O.kernel(0.5, h)
    assert wl == pytest.approx(wr, rel=1e-12)
    assert wr == pytest.approx(0.25 * NORM / h**3, rel=1e-14)
    assert dwr == pytest.approx(-1.5 * NORM / h**4, rel=1e-14)
    assert dwl == pytest.approx(dwr, rel=1e-12)
    for u in (1.0, 1.5, 3.0):
        assert O.kernel(u, h) == (0.0, 0.0)


def test_kernel_normalised_to_unit_volume_integral():
    # 4 pi int_0^h W(r/h) r^2 dr = 1
    h = 1.3
    u = np.linspace(0, 1, 20001)
    w = np.array([O.kernel(x, h)[0] for x in u])
    integral = 4 * math.pi * np.trapezoid(w * (u * h) ** 2, u * h)
    assert integral == pytest.approx(1.0, rel=1e-7)


def test_kernel_derivative_matches_finite_difference():
    h = 0.9
    for u in (0.1, 0.3, 0.45, 0.6, 0.8, 0.95):
        eps = 1e-6
        wp = O.kernel(u + eps, h)[0]
        wm = O.kernel(u - eps, h)[0]
        dw = O.kernel(u, h)[1]
        # dwk = dW/du / h (kernel.h: dwk carries hinv4 = hinv3 * hinv)
        assert dw == pytest.approx((wp - wm) / (2 * eps) / h, rel=1e-7)


# ---------------------------------------------------------------- inverse ---
def test_matrix_inverse_matches_linalg():
    rng = np.random.default_rng(3)
    for _ in range(50):
        A = rng.normal(size=(3, 3))
        T = A @ A.T + 0.1 * np.eye(3)
        Ti, cn = O.matrix_invert(T)
        np.testing.assert_allclose(Ti, np.linalg.inv(T), rtol=1e-10, atol=1e-12)
        fro = math.sqrt(np.sum(T * T) * np.sum(Ti * Ti)) / 3
        assert cn == pytest.approx(max(fro, 1.0), rel=1e-12)


def test_condition_number_closed_forms():
    Ti, cn = O.matrix_invert(np.eye(3))
    np.testing.assert_array_equal(Ti, np.eye(3))
    assert cn == 1.0
    a, b, c = 2.0, 0.5, 8.0
    Ti, cn = O.matrix_invert(np.diag([a, b, c]))
    np.testing.assert_allclose(np.diag(Ti), [1 / a, 1 / b, 1 / c], rtol=1e-15)
    assert cn == pytest.approx(math.sqrt((a * a + b * b + c * c) * (a**-2 + b**-2 + c**-2)) / 3, rel=1e-14)


def test_condition_number_sensitivity_within_parity_bound():
    """The parity bar of ConditionNumber, COND_ROUNDING (1 + 3 cn), bounds the
    change of cn when every entry of NV_T moves by up to 5e-15 relative (the
    rounding of a ~50-term sum in another order), from cn ~ 1 to ~1e4."""
    from parity import COND_ROUNDING

    rng = np.random.default_rng(11)
    worst = 0.0
    for scale in (1.0, 1e-1, 1e-2, 1e-3, 1e-4):
        for _ in range(200):
            Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
            T = Q @ np.diag([1.0, rng.uniform(scale, 1.0), scale]) @ Q.T
            T = 0.5 * (T + T.T)
            _, cn = O.matrix_invert(T)
            Tp = T * (1.0 + 5e-15 * rng.uniform(-1.0, 1.0, size=(3, 3)))
            _, cnp = O.matrix_invert(0.5 * (Tp + Tp.T))
            ratio = abs(cnp - cn) / cn / (COND_ROUNDING * (1.0 + 3.0 * cn))
            worst = max(worst, ratio)
    assert worst <= 1.0, worst


def test_singular_matrix_gives_zero_inverse():
    T = np.array([[1.0, 2.0, 3.0], [2.0, 4.0, 6.0], [1.0, 1.0, 1.0]])
    Ti, cn = O.matrix_invert(T)
    assert np.all(Ti == 0)
    assert cn == 1.0  # system.c:389: max(sqrt(|T|^2 * 0)/3, 1)


# ---------------------------------------------------------------- Riemann ---
# Toro, "Riemann Solvers and Numerical Methods for Fluid Dynamics", 3rd ed.,
# Table 4.3 (gamma = 1.4): star-region pressure and velocity of tests 1-5.
TORO = [
    ((1.0, 1.0, 0.0), (0.125, 0.1, 0.0), 0.30313, 0.92745),
    ((1.0, 0.4, -2.0), (1.0, 0.4, 2.0), 0.00189, 0.0),
    ((1.0, 1000.0, 0.0), (1.0, 0.01, 0.0), 460.894, 19.5975),
    ((1.0, 0.01, 0.0), (1.0, 100.0, 0.0), 46.0950, -6.19633),
    ((5.99924, 460.894, 19.5975), (5.99242, 46.0950, -6.19633), 1691.64, 8.68975),
]


@pytest.mark.parametrize("L,R,pstar,ustar", TORO)
def test_exact_solver_reproduces_toro_table(L, R, pstar, ustar):
    # a zero pressure limiter sends every state to the exact solver (reimann.h:364)
    out = O.riemann((L[0], L[1], L[2], 0, 0), (R[0], R[1], R[2], 0, 0), n=(1, 0, 0),
                    press_tot_limiter=0.0, gamma=1.4)
    assert out["mask"] & 4
    assert out["P_M"] == pytest.approx(pstar, rel=2e-5, abs=1e-5)
    assert out["S_M"] == pytest.approx(ustar, rel=2e-5, abs=1e-5)


def test_exact_solver_along_oblique_normal():
    n = np.array([1.0, 2.0, -2.0]) / 3.0
    L = (1.0, 1.0, *(0.0 * n))
    R = (0.125, 0.1, *(0.0 * n))
    out = O.riemann(L, R, n=n, press_tot_limiter=0.0, gamma=1.4)
    assert out["P_M"] == pytest.approx(0.30313, rel=2e-5)
    assert out["S_M"] == pytest.approx(0.92745, rel=2e-5)


def test_hllc_equal_states_return_the_state():
    for rho, p, v in ((1.0, 1.0, 0.3), (2.5, 0.01, -1.0), (0.1, 10.0, 0.0)):
        out = O.riemann((rho, p, v, 0.2, -0.1), (rho, p, v, 0.2, -0.1), n=(1, 0, 0), press_tot_limiter=1e300)
        assert out["mask"] == 1
        assert out["P_M"] == pytest.approx(p, rel=1e-14)
        assert out["S_M"] == pytest.approx(v, rel=1e-14, abs=1e-15)
        # MFM fluxes in the face frame: F_v = P_M n, F_p = P_M S_M (reimann.h:410-412)
        np.testing.assert_allclose(out["Fv"], [out["P_M"], 0, 0], rtol=1e-15)
        assert out["Fp"] == pytest.approx(out["P_M"] * out["S_M"], rel=1e-15)


def test_riemann_mirror_symmetry():
    rng = np.random.default_rng(11)
    for _ in range(200):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        L = (rng.uniform(0.1, 3), rng.uniform(0.01, 5), *rng.normal(size=3))
        R = (rng.uniform(0.1, 3), rng.uniform(0.01, 5), *rng.normal(size=3))
        lim = rng.choice([1e300, 1.1 * max(L[1], R[1])])
        a = O.riemann(L, R, n=n, press_tot_limiter=lim)
        b = O.riemann(R, L, n=-n, press_tot_limiter=lim)
        assert a["P_M"] == pytest.approx(b["P_M"], rel=1e-12)
        assert a["S_M"] == pytest.approx(-b["S_M"], rel=1e-10, abs=1e-12)


def test_vacuum_branch():
    # receding faster than the sound speed: P_M = MIN_REAL_NUMBER, S_M = 0 (reimann.h:535-538)
    out = O.riemann((1.0, 1.0, -5.0, 0, 0), (1.0, 1.0, 5.0, 0, 0), n=(1, 0, 0), press_tot_limiter=1e300)
    assert out["P_M"] == 1e-56 and out["S_M"] == 0.0


# ---------------------------------------------------------- reconstruction --
def test_reconstruct_zeroth_order_and_equal_values():
    z = (0.0, 0.0, 0.0)
    assert O.reconstruct(1.0, (5, 5, 5), 2.0, (5, 5, 5), (1, 1, 1), (1, 1, 1), mode=0) == (2.0, 1.0)
    assert O.reconstruct(3.0, (1, 0, 0), 3.0, (1, 0, 0), (0.5, 0, 0), z) == (3.0, 3.0)


def test_reconstruct_limits_hand_cases():
    z = (0.0, 0.0, 0.0)
    # no gradient: L = Q_j, R = Q_i
    assert O.reconstruct(1.0, z, 2.0, z, z, z) == (2.0, 1.0)
    # Q_i=1, Q_j=2: Qmax_eff=2.5, Qmin_eff=0.5, Qmed=1.5 +- 0.375 -> R <= 1.875, L >= 1.125
    ql, qr = O.reconstruct(1.0, (10.0, 0, 0), 2.0, (-10.0, 0, 0), (0.1, 0, 0), (0.1, 0, 0))
    assert qr == pytest.approx(1.875, rel=1e-15)
    assert ql == pytest.approx(1.125, rel=1e-15)
    # mirrored ordering Q_i > Q_j
    ql, qr = O.reconstruct(2.0, (-10.0, 0, 0), 1.0, (10.0, 0, 0), (0.1, 0, 0), (0.1, 0, 0))
    assert qr == pytest.approx(1.125, rel=1e-15)
    assert ql == pytest.approx(1.875, rel=1e-15)
    # one-signed values keep their sign: Qmin_eff = Qmin^2 / (Qmin + (Qmin - Qmin_eff))
    ql, qr = O.reconstruct(0.1, (-10.0, 0, 0), 1.0, z, (0.1, 0, 0), z)
    assert qr == pytest.approx(0.01 / (0.1 + 0.45), rel=1e-14)
    assert ql == 1.0


# -------------------------------------------------------------- limiter -----
def test_slope_limiter_hand_cases():
    # cfac = 1/(alim h |g|) * min(|min| + tol |max|, |max|) (gradients.c:537-556)
    g = O.slope_limiter((1.0, 0, 0), 0.1, -0.05, 0.25, 1.0, 0.0, 0, 0.0, 0.0)
    np.testing.assert_allclose(g, [0.2, 0, 0], rtol=1e-15)
    g = O.slope_limiter((1.0, 0, 0), 0.1, -0.05, 0.25, 1.0, 0.1, 0, 0.0, 0.0)
    np.testing.assert_allclose(g, [0.24, 0, 0], rtol=1e-14)
    # no limiting when cfac >= 1
    g = O.slope_limiter((0.1, 0.1, 0), 1.0, -1.0, 0.25, 1.0, 0.0, 0, 0.0, 0.0)
    np.testing.assert_allclose(g, [0.1, 0.1, 0], rtol=0)
    # positivity: value 1, min difference -0.9, gradient 10 over d_max = 2 -> (1 - fmin)/2/10
    g = O.slope_limiter((10.0, 0, 0), 5.0, -0.9, 0.25, 1.0, 0.0, 1, 2.0, 1.0)
    # by hand: f_corr = 0.9, overshoot cfac = 0.9 / 2.5 = 0.36; fmin = 0.05 (half the
    # value plus the min difference), positivity cfac = (1 - 0.05) / 2 / 10 = 0.0475
    np.testing.assert_allclose(g, [0.475, 0, 0], rtol=1e-14)
