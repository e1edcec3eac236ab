// mfm_device.h -- device-side physics of the MFM hydro pass, gfx950.
//
// Pure per-particle / per-pair functions used by the stage kernels.  Each one
// cites the reference lines whose behaviour it reproduces.  Arithmetic keeps
// the reference's operation order; the library is compiled with
// -ffp-contract=off so that products and sums round exactly as in the CPU
// pass.
//
// The per-pair functions are templates over the real type T: T = double is
// the plain FP64 pass, T = float the FP32-bulk pass (same steps, FP32
// arithmetic with the hardware reciprocal and square root).  Per-particle
// steps (hsml iteration, NV_T inversion, slope limiter, exact Riemann solver)
// are FP64 only.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace gzm {

// allvars.h:1443 (the literal, not 4*pi/3), kernel.h:63, allvars.h:1106-1112, allvars.h:78
constexpr double kNormCoeff = 4.188790204786;
constexpr double kKernelNorm = 8.0 / 3.14159265358979323846;
constexpr double kCondDanger = 1.0e3;
constexpr double kMinReal = 1e-56;
constexpr double kWakeup = 4.1;

// proto.h:123-124 (NaN behaviour matters: DMAX(NaN,b) = b)
__device__ __forceinline__ double dmax(double a, double b) { return (a > b) ? a : b; }
__device__ __forceinline__ double dmin(double a, double b) { return (a < b) ? a : b; }
template <typename T> __device__ __forceinline__ T tmax(T a, T b) { return (a > b) ? a : b; }
template <typename T> __device__ __forceinline__ T tmin(T a, T b) { return (a < b) ? a : b; }

// Arithmetic that differs between the two precisions.  FP64: IEEE division
// and square root (the reference's).  FP32: v_rcp_f32 / v_sqrt_f32 (1 ulp),
// and MIN_REAL_NUMBER (1e-56, allvars.h:78) replaced by the smallest power of
// ten that stays a normal float, since 1e-56 underflows to 0 in FP32.
template <typename T> struct Real;
template <> struct Real<double> {
    static constexpr double tiny = kMinReal;
    static __device__ __forceinline__ double div(double a, double b) { return a / b; }
    static __device__ __forceinline__ double rcp(double a) { return 1.0 / a; }
    static __device__ __forceinline__ double root(double a) { return sqrt(a); }
};
template <> struct Real<float> {
    static constexpr float tiny = 1e-37f;
    static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    static __device__ __forceinline__ float rcp(float a) { return __builtin_amdgcn_rcpf(a); }
    static __device__ __forceinline__ float root(float a) { return __builtin_amdgcn_sqrtf(a); }
};

// reimann.h:1-10 for the run's adiabatic index
struct GammaConsts {
    double G0, G1, G2, G3, G4, G5, G6, G7, G8, G9;
};

struct BoxInfo {
    double L[3];
    double half[3];
};

// ---------------------------------------------------------------------------
// cubic-spline kernel (kernel.h:182-197, 225-242, 373-374)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void kernel_hinv(T h, T &hinv, T &hinv3, T &hinv4)
{
    hinv = Real<T>::rcp(h);
    hinv3 = hinv * hinv * hinv;
    hinv4 = hinv3 * hinv;
}

template <typename T>
__device__ __forceinline__ void kernel_wdw(T u, T hinv3, T hinv4, T &wk, T &dwk)
{
    if (u >= 1) { wk = 0; dwk = 0; return; }
    T w, dw;
    if (u < T(0.5)) {
        dw = u * (T(18.0) * u - T(12.0));
        w = (T(1.0) + T(6.0) * (u - T(1.0)) * u * u);
    } else {
        T t1 = (T(1.0) - u);
        T t2 = t1 * t1;
        dw = T(-6.0) * t2;
        w = T(2.0) * t2 * t1;
    }
    dwk = dw * (T(kKernelNorm) * hinv4);
    wk = w * (T(kKernelNorm) * hinv3);
}

template <typename T>
__device__ __forceinline__ T kernel_w(T u, T hinv3)
{
    if (u >= 1) return 0;
    T w;
    if (u < T(0.5)) {
        w = (T(1.0) + T(6.0) * (u - T(1.0)) * u * u);
    } else {
        T t1 = (T(1.0) - u);
        T t2 = t1 * t1;
        w = T(2.0) * t2 * t1;
    }
    return w * (T(kKernelNorm) * hinv3);
}

// signed nearest-image wrap of one component (allvars.h:1519, NEAREST_XYZ)
__device__ __forceinline__ double wrap_axis(double x, double L, double half)
{
    return (x > half) ? (x - L) : ((x < -half) ? (x + L) : x);
}

// ---------------------------------------------------------------------------
// 3x3 symmetric inverse with the cofactor formula and Frobenius condition
// number (system/system.c:355-391).  The input is the symmetric matrix
// (a b c / b d e / c e f); element order inside every product follows the
// general 3x3 expression so the rounding is the reference's.
// Returns CN; inv = (i00 i01 i02 i11 i12 i22).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double invert_sym3(double a, double b, double c, double d, double e, double f,
                                              double inv[6])
{
    // T00=a T01=b T02=c T10=b T11=d T12=e T20=c T21=e T22=f
    double frob = 0;
    frob += a * a; frob += b * b; frob += c * c;
    frob += b * b; frob += d * d; frob += e * e;
    frob += c * c; frob += e * e; frob += f * f;
    double det = a * d * f + b * e * c + c * b * e - c * d * c - b * b * f - a * e * e;
    double i00 = 0, i01 = 0, i02 = 0, i11 = 0, i12 = 0, i22 = 0;
    if ((det != 0) && !isnan(det)) {
        i00 = (d * f - e * e) / det;
        i01 = (c * e - b * f) / det;
        i02 = (b * e - c * d) / det;
        i11 = (a * f - c * c) / det;
        i12 = (c * b - a * e) / det;
        i22 = (a * d - b * b) / det;
    }
    // Tinv row-major: i00 i01 i02 / i10=i01 i11 i12 / i20=i02 i21=i12 i22
    double frob_inv = 0;
    frob_inv += i00 * i00; frob_inv += i01 * i01; frob_inv += i02 * i02;
    frob_inv += i01 * i01; frob_inv += i11 * i11; frob_inv += i12 * i12;
    frob_inv += i02 * i02; frob_inv += i12 * i12; frob_inv += i22 * i22;
    inv[0] = i00; inv[1] = i01; inv[2] = i02; inv[3] = i11; inv[4] = i12; inv[5] = i22;
    return dmax(sqrt(frob * frob_inv) / 3, 1);
}

// ---------------------------------------------------------------------------
// Newton/bisection update of one particle after a density sweep
// (density.c:505-850).  Reads the raw sums of the sweep, writes the
// normalised per-particle fields; returns 1 if the particle needs another
// sweep, 0 if it is done, -1 on the Left==Right==0 failure.
// ---------------------------------------------------------------------------
struct HsmlParams {
    double des_num_ngb, max_num_ngb_dev, min_hsml, max_hsml;
    int first_step;
    int iter;
};

struct DensitySums {
    double ngb, rho, dhsml, divvel;
    double t00, t01, t02, t11, t12, t22;  // second moments
    double f0, f1, f2;                    // first moments (face terms, NV_T[1][0],[2][0],[2][1])
};

struct DensityResult {
    double num_ngb, dhsml, divvel, rho_sum, fce, cn_stored;
    double nvt[6];
};

__device__ __forceinline__ double ncorr_from_cn(double cn)
{
    double c0 = 0.1 * kCondDanger;
    double nc = 1;
    if (cn > c0) nc = sqrt(1.0 + (cn - c0) / kCondDanger);
    if (nc > 2) nc = 2;
    return nc;
}

// What the hsml step (density.c:594-850) decides on: the normalised
// neighbour number, its target and tolerance, and the condition numbers.
struct HsmlStep {
    double NumNgb, Dhsml, desnumngb, desnumngbdev, ConditionNumber, cn_stored;
};

// density.c:505-595 on one particle's sums: normalisation, NV_T inversion
// with preconditioning, face-closure error and the neighbour-number target;
// fills R but R.cn_stored (set by hsml_step).
__device__ inline void hsml_prepare(const HsmlParams &P, const DensitySums &S, double h, double cn_stored_in,
                                    DensityResult &R, HsmlStep &X)
{
    const int iter = P.iter;
    double desnumngb = P.des_num_ngb, desnumngbdev = P.max_num_ngb_dev;
    if (P.first_step) { if (P.max_num_ngb_dev > 0.05) desnumngbdev = 0.05; }
    const double desnumngbdev_0 = desnumngbdev;
    double NumNgb = S.ngb, Dhsml = S.dhsml, DivVel = S.divvel;
    if (NumNgb > 0) {
        Dhsml *= h / (3 * NumNgb);
        DivVel /= NumNgb;
        NumNgb *= kNormCoeff * pow(h, 3.0);
    } else {
        NumNgb = Dhsml = DivVel = 0;
    }
    if (Dhsml > -0.9) Dhsml = 1 / (1 + Dhsml); else Dhsml = 1;
    DivVel *= Dhsml;

    // NV_T inversion with diagonal preconditioning, face-closure error
    double ConditionNumber;
    {
        double V_i = kNormCoeff * pow(h, 3.0) / NumNgb;
        double dx_i = sqrt(V_i * (S.t00 + S.t11 + S.t22));
        double norm = pow(h, -1.0);
        double a = S.t00 / norm, b = S.t01 / norm, c = S.t02 / norm;
        double d = S.t11 / norm, e = S.t12 / norm, f = S.t22 / norm;
        const double cn_thresh = 10. * kCondDanger;
        double trace = a + d + f;
        double add = 1.05 * (trace / 3) / cn_thresh;
        double inv[6];
        // the reference loops until CN < 1e4 (density.c:543-552); the cap only
        // guarantees termination on a GPU (1.2^4096 overflows long before)
        for (int guard = 0; guard < 4096; guard++) {
            ConditionNumber = invert_sym3(a, b, c, d, e, f, inv);
            if (ConditionNumber < cn_thresh) break;
            a += add; d += add; f += add;
            add *= 1.2;
        }
        for (int k = 0; k < 6; k++) R.nvt[k] = inv[k] / norm;
        // Face_Area_OneSided_Estimator_out = 2 V_i NV_T^-1 . (f0 f1 f2), row by row
        const double fin[3] = {S.f0, S.f1, S.f2};
        const double row[3][3] = {{R.nvt[0], R.nvt[1], R.nvt[2]},
                                  {R.nvt[1], R.nvt[3], R.nvt[4]},
                                  {R.nvt[2], R.nvt[4], R.nvt[5]}};
        double leak = 0;
        for (int k1 = 0; k1 < 3; k1++) {
            double o = 0;
            for (int k2 = 0; k2 < 3; k2++) o += 2. * V_i * row[k1][k2] * fin[k2];
            leak += fabs(o) / 3;
        }
        R.fce = leak / (2. * 3 * pow(dx_i, 2.0));
    }

    // neighbour-number target (density.c:565-595)
    double cn_stored = cn_stored_in;
    if ((iter == 0) && (ConditionNumber > cn_stored) && (cn_stored > 0)) {
        double nc = ncorr_from_cn(cn_stored);
        double dn_ngb = fabs(NumNgb - P.des_num_ngb * nc) / (desnumngbdev_0 * nc);
        nc = ncorr_from_cn(ConditionNumber);
        double dn_ngb_alt = fabs(NumNgb - P.des_num_ngb * nc) / (desnumngbdev_0 * nc);
        dn_ngb = dmin(dn_ngb, dn_ngb_alt);
        if (dn_ngb < 10.0) cn_stored = ConditionNumber;
    }
    double ncorr_ngb = ncorr_from_cn(cn_stored);
    if (R.fce > 0.35) ncorr_ngb = dmax(ncorr_ngb, dmin(R.fce / 0.35, 2.));
    desnumngb = P.des_num_ngb * ncorr_ngb;
    desnumngbdev = desnumngbdev_0 * ncorr_ngb;
    if (iter > 1)
        desnumngbdev = dmin(0.25 * desnumngb, desnumngbdev * exp(0.1 * log(desnumngb / (16. * desnumngbdev)) * (double)iter));
    R.num_ngb = NumNgb;
    R.dhsml = Dhsml;
    R.divvel = DivVel;
    R.rho_sum = S.rho;
    X.NumNgb = NumNgb;
    X.Dhsml = Dhsml;
    X.desnumngb = desnumngb;
    X.desnumngbdev = desnumngbdev;
    X.ConditionNumber = ConditionNumber;
    X.cn_stored = cn_stored;
}

// density.c:659-850: the redo test, the bracket update and the next hsml.
// Returns 1 (iterate again), 0 (done) or -1 (no bracket); cn_out is the
// condition number to store.
__device__ inline int hsml_step(const HsmlParams &P, const HsmlStep &X, double &h, double &Left, double &Right,
                                double &cn_out)
{
    const int iter = P.iter;
    const double NumNgb = X.NumNgb, Dhsml = X.Dhsml, desnumngb = X.desnumngb, desnumngbdev = X.desnumngbdev;
    const double ConditionNumber = X.ConditionNumber;
    double cn_stored = X.cn_stored;
    const double maxsoft = P.max_hsml, minsoft = P.min_hsml;
    int redo = 0, maxflag = 0, minflag = 0;
    if ((NumNgb < (desnumngb - desnumngbdev) && h < 0.999 * maxsoft) ||
        (NumNgb > (desnumngb + desnumngbdev) && h > 1.001 * minsoft)) redo = 1;
    if ((h >= 0.999 * maxsoft) && (NumNgb < (desnumngb - desnumngbdev))) {
        redo = 0;
        if (h != maxsoft) { redo = 1; h = maxsoft; maxflag = 1; }
    }
    if ((h <= 1.001 * minsoft) && (NumNgb > (desnumngb + desnumngbdev))) {
        redo = 0;
        if (h != minsoft) { redo = 1; h = minsoft; minflag = 1; }
    }
    if (redo == 0) cn_stored = ConditionNumber;

    int result = redo;
    if (redo) {
        if (Left > 0 && Right > 0 && (Right - Left) < 1.0e-3 * Left) {
            cn_out = ConditionNumber;  // bracket narrow enough: accept (density.c:732-740)
            return 0;
        }
        if ((maxflag == 0) && (minflag == 0)) {
            if (NumNgb < (desnumngb - desnumngbdev)) Left = dmax(h, Left);
            else {
                if (Right != 0) { if (h < Right) Right = h; }
                else Right = h;
            }
            if (Right > 0 && Left > 0) {
                double maxjump = 0;
                if (iter > 1) maxjump = 0.2 * log(Right / Left);
                if (NumNgb > 1) {
                    double jumpvar = Dhsml * log(desnumngb / NumNgb) / 3;
                    if (iter > 1) { if (fabs(jumpvar) < maxjump) { jumpvar = (jumpvar < 0) ? -maxjump : maxjump; } }
                    h *= exp(jumpvar);
                } else {
                    h *= 2.0;
                }
                if ((h < Right) && (h > Left)) {
                    if (iter > 1) {
                        double hfac = exp(maxjump);
                        if (h > Right / hfac) h = Right / hfac;
                        if (h < Left * hfac) h = Left * hfac;
                    }
                } else {
                    if (h > Right) h = Right;
                    if (h < Left) h = Left;
                    h = pow(h * Left * Right, 1.0 / 3.0);
                }
            } else {
                if (Right == 0 && Left == 0) { cn_out = cn_stored; return -1; }
                if (Right == 0 && Left > 0) {
                    double fac_lim = (NumNgb > 1) ? log(desnumngb / NumNgb) / 3 : 1.4;
                    if ((NumNgb < 2 * desnumngb) && (NumNgb > 0.1 * desnumngb)) {
                        double slope = Dhsml;
                        if (iter > 2 && slope < 1) slope = 0.5 * (slope + 1);
                        double fac = fac_lim * slope;
                        if (iter >= 4) { if (Dhsml == 1) fac *= 10; }
                        if (fac < fac_lim + 0.231) h *= exp(fac);
                        else h *= exp(fac_lim + 0.231);
                    } else h *= exp(fac_lim);
                }
                if (Right > 0 && Left == 0) {
                    double fac_lim = (NumNgb > 1) ? log(desnumngb / NumNgb) / 3 : 1.4;
                    if (fac_lim < -1.535) fac_lim = -1.535;
                    if ((NumNgb < 2 * desnumngb) && (NumNgb > 0.1 * desnumngb)) {
                        double slope = Dhsml;
                        if (iter > 2 && slope < 1) slope = 0.5 * (slope + 1);
                        double fac = fac_lim * slope;
                        if (iter >= 4) { if (Dhsml == 1) fac *= 10; }
                        if (fac > fac_lim - 0.231) h *= exp(fac);
                        else h *= exp(fac_lim - 0.231);
                    } else h *= exp(fac_lim);
                }
            }
        }
        if (h < minsoft) h = minsoft;
        if (minflag == 1) h = minsoft;
        if (h > maxsoft) h = maxsoft;
        if (maxflag == 1) h = maxsoft;
    }
    cn_out = cn_stored;
    return result;
}

__device__ inline int hsml_update(const HsmlParams &P, const DensitySums &S, double &h, double &Left,
                                  double &Right, double cn_stored_in, DensityResult &R)
{
    HsmlStep X;
    hsml_prepare(P, S, h, cn_stored_in, R, X);
    return hsml_step(P, X, h, Left, Right, R.cn_stored);
}

// ---------------------------------------------------------------------------
// slope limiter (gradients.c:537-556)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void slope_limit(double g[3], double valmax, double valmin, double alim, double h,
                                            double shoot_tol, int pos_preserve, double d_max, double val_cen)
{
    double d_abs = 0.0;
    for (int k = 0; k < 3; k++) d_abs += g[k] * g[k];
    if (d_abs > 0) {
        d_abs = sqrt(d_abs);
        double cfac = 1 / (alim * h * d_abs);
        double fabs_max = fabs(valmax), fabs_min = fabs(valmin), abs_max = fabs_max, abs_min = fabs_min;
        if (abs_max < abs_min) { abs_max = fabs_min; abs_min = fabs_max; }
        double f_corr = dmin(abs_min + shoot_tol * abs_max, abs_max);
        cfac *= f_corr;
        if (pos_preserve == 1) {
            double fmin = dmin(val_cen, dmax(0, dmax(kMinReal * val_cen, dmin(0.5 * (val_cen + valmin), val_cen - f_corr))));
            cfac = dmin((((val_cen - fmin) / d_max) / d_abs), cfac);
        }
        if (cfac < 1) { for (int k = 0; k < 3; k++) g[k] *= cfac; }
    }
}

// NV_T . g (gradients.c:559-572, matrix branch; nvt = i00 i01 i02 i11 i12 i22)
__device__ __forceinline__ void apply_nvt(const double nvt[6], double g[3])
{
    double v0 = g[0], v1 = g[1], v2 = g[2];
    g[0] = nvt[0] * v0 + nvt[1] * v1 + nvt[2] * v2;
    g[1] = nvt[1] * v0 + nvt[3] * v1 + nvt[4] * v2;
    g[2] = nvt[2] * v0 + nvt[4] * v1 + nvt[5] * v2;
}

// FP32 only: a branch of the face construction, the limiter, the wavespeed
// estimate or the Riemann retry whose FP32 operand lies within this relative
// margin of its threshold may go the other way than in the FP64 pass (the
// FP32 operands carry rounding of ~1e-7 to 1e-6 relative); the pair is then
// evaluated in FP64 instead (kNearThreshold).
constexpr float kDecideMargin = 1e-5f;

template <typename T>
__device__ __forceinline__ bool near_threshold(T x, T threshold, T scale)
{
    return sizeof(T) == sizeof(float) && !(fabs(x - threshold) > T(kDecideMargin) * fabs(scale));
}

// ---------------------------------------------------------------------------
// face reconstruction and limiter (reimann.h:129-248), fac_minmax 0.5,
// fac_meddev 0.375.  R side is particle i, L side is particle j.
// The sign-preserving bounds jump where Qmax_eff (Qmax < 0) or Qmin_eff
// (Qmin > 0) crosses zero: from 0 to Qmax/2 (Qmin/2).  In FP32 the function
// returns true when that crossing lies within kDecideMargin of (Qmax - Qmin),
// where the rounded operands may take the other side than in FP64.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ bool reconstruct(T Qi, const T gi[3], T Qj, const T gj[3], const T di[3], const T dj[3],
                                            T &QL, T &QR, int mode)
{
    if (mode == 0) { QR = Qi; QL = Qj; return false; }
    if (Qi == Qj) { QL = QR = Qi; return false; }
    QR = Qi + gi[0] * di[0] + gi[1] * di[1] + gi[2] * di[2];
    QL = Qj + gj[0] * dj[0] + gj[1] * dj[1] + gj[2] * dj[2];
    T Qmax, Qmin;
    T Qmed = T(0.5) * (Qi + Qj);
    if (Qi < Qj) { Qmax = Qj; Qmin = Qi; } else { Qmax = Qi; Qmin = Qj; }
    T fac = T(0.5) * (Qmax - Qmin);
    T Qmax_eff = Qmax + fac;
    T Qmin_eff = Qmin - fac;
    // one test covers both bounds: when Qmin <= 0 <= Qmax, Qmin_eff is at
    // least (Qmax - Qmin) / 2 below zero and never near
    const bool near = near_threshold((Qmax < 0) ? Qmax_eff : Qmin_eff, T(0), Qmax - Qmin);
    if (Qmax < 0) { if (Qmax_eff > 0) Qmax_eff = Real<T>::div(Qmax * Qmax, (Qmax - (Qmax_eff - Qmax))); }
    if (Qmin > 0) { if (Qmin_eff < 0) Qmin_eff = Real<T>::div(Qmin * Qmin, (Qmin + (Qmin - Qmin_eff))); }
    fac = T(0.375) * (Qmax - Qmin);
    T Qmed_max = Qmed + fac;
    T Qmed_min = Qmed - fac;
    if (Qmed_max > Qmax_eff) Qmed_max = Qmax_eff;
    if (Qmed_min < Qmin_eff) Qmed_min = Qmin_eff;
    if (Qi < Qj) {
        if (QR < Qmin_eff) QR = Qmin_eff;
        if (QR > Qmed_max) QR = Qmed_max;
        if (QL > Qmax_eff) QL = Qmax_eff;
        if (QL < Qmed_min) QL = Qmed_min;
    } else {
        if (QR > Qmax_eff) QR = Qmax_eff;
        if (QR < Qmed_min) QR = Qmed_min;
        if (QL < Qmin_eff) QL = Qmin_eff;
        if (QL > Qmed_max) QL = Qmed_max;
    }
    return near;
}

// ---------------------------------------------------------------------------
// Riemann problem for MFM hydro (reimann.h:286-593, 685-1138).
// ---------------------------------------------------------------------------
template <typename T>
struct RState {
    T rho, p, v[3];
};

template <typename T>
struct ROut {
    T P_M, S_M, Fv[3], Fp;
};

// get_wavespeeds_and_pressure_star: Gaburov HLLC estimate, Roe, then PVRS
// (reimann.h:527-593).  Sets P_M and S_M.  Returns true (FP32 only, with
// P_M and S_M unset) when the vacuum test or a P_M > limiter test -- also the
// one after PVRS, which decides whether the exact solver runs -- is within
// kDecideMargin of its threshold.
//
// S_LR (may be null): the outer wave speeds of the estimate that stands (both
// 0 on the vacuum test and where P_M is floored), which the MFV fluxes read.
template <typename T>
__device__ inline bool wavespeeds_pstar(const RState<T> &L, const RState<T> &R, const GammaConsts &G, const T n[3],
                                        T vL, T vR, T cL, T cR, T hL, T hR, T limiter, ROut<T> &o,
                                        T *S_LR = nullptr)
{
    using M = Real<T>;
    const T cmax = tmax(cL, cR);
    if (near_threshold(vR - vL, cmax, cmax)) return true;
    if ((vR - vL) > cmax) {
        o.P_M = M::tiny; o.S_M = 0;
        if (S_LR) S_LR[0] = S_LR[1] = 0;
        return false;
    }
    T PT_L = L.p, PT_R = R.p;
    T S_L = tmin(vL, vR) - tmax(cL, cR);
    T S_R = tmax(vL, vR) + tmax(cL, cR);
    T rwL = L.rho * (S_L - vL);
    T rwR = R.rho * (S_R - vR);
    o.S_M = M::div(((PT_R - PT_L) + rwL * vL - rwR * vR), (rwL - rwR));
    o.P_M = M::div((PT_L * rwR - PT_R * rwL + rwL * rwR * (vR - vL)), (rwR - rwL));
    if (near_threshold(o.P_M, limiter, limiter)) return true;
    if (o.P_M <= M::tiny) { o.P_M = M::tiny; o.S_M = 0; S_L = S_R = 0; }
    if (S_LR) { S_LR[0] = S_L; S_LR[1] = S_R; }
    if ((o.P_M <= 0) || isnan(o.P_M) || (o.P_M > limiter)) {
        T sL = M::root(L.rho), sR = M::root(R.rho);
        T sinv = M::rcp(sL + sR);
        T vx = (sL * L.v[0] + sR * R.v[0]) * sinv;
        T vy = (sL * L.v[1] + sR * R.v[1]) * sinv;
        T vz = (sL * L.v[2] + sR * R.v[2]) * sinv;
        T v_roe = vx * n[0] + vy * n[1] + vz * n[2];
        T h_roe = (sL * hL + sR * hR) * sinv;
        T c_roe = M::root(tmax(T(1.e-30), T(G.G9) * (h_roe - T(0.5) * (vx * vx + vy * vy + vz * vz))));
        S_R = tmax(vR + cR, v_roe + c_roe);
        S_L = tmin(vL - cL, v_roe - c_roe);
        rwR = R.rho * (S_R - vR);
        rwL = -L.rho * (S_L - vL);
        o.S_M = M::div(((rwR * vR + rwL * vL) + (PT_L - PT_R)), (rwR + rwL));
        o.P_M = L.rho * (vL - S_L) * (vL - o.S_M) + PT_L;
        if (near_threshold(o.P_M, limiter, limiter)) return true;
        if (o.P_M <= M::tiny) { o.P_M = M::tiny; o.S_M = 0; S_L = S_R = 0; }
        if (S_LR) { S_LR[0] = S_L; S_LR[1] = S_R; }
        if ((o.P_M <= 0) || isnan(o.P_M) || (o.P_M > limiter)) {
            o.P_M = T(0.5) * ((PT_L + PT_R) + (vL - vR) * T(0.25) * (L.rho + R.rho) * (cL + cR));
            o.S_M = T(0.5) * (vR + vL) + M::div(T(2.0) * (PT_L - PT_R), ((L.rho + R.rho) * (cL + cR)));
            T S_plus = tmax(tmax(fabs(vL - cL), fabs(vR - cR)), tmax(fabs(vL + cL), fabs(vR + cR)));
            S_L = -S_plus;
            S_R = S_plus;
            if (o.S_M < S_L) o.S_M = S_L;
            if (o.S_M > S_R) o.S_M = S_R;
            if (o.P_M <= M::tiny) { o.P_M = M::tiny; o.S_M = 0; S_L = S_R = 0; }
            if (S_LR) { S_LR[0] = S_L; S_LR[1] = S_R; }
            if (near_threshold(o.P_M, limiter, limiter)) return true;
        }
    }
    return false;
}

// Kurganov-Tadmor/PWK flux, MFM branch (reimann.h:484-519).  Reached only on
// a NaN/negative HLLC pressure, so the force loop calls it out of line
// (riemann_kt); the deferred-pair kernels, whose every lane may take it,
// inline it (riemann_kt_body).
template <typename T>
__device__ __forceinline__ ROut<T> riemann_kt_body(T rL, T pL, T vLx, T vLy, T vLz, T rR, T pR, T vRx, T vRy, T vRz,
                                                   T n0, T n1, T n2, T vL, T vR, T cL, T cR, T hL, T hR)
{
    using M = Real<T>;
    ROut<T> o;
    T thr = T(0.001) * T(0.5) * (rL + rR) * T(0.5) * (cL + cR);
    T d0 = rR * vRx - rL * vLx, d1 = rR * vRy - rL * vLy, d2 = rR * vRz - rL * vLz;
    T dv2 = 0;
    dv2 += d0 * d0;
    dv2 += d1 * d1;
    dv2 += d2 * d2;
    T dvL = fabs(rR * vR - rL * vL);
    T alpha = M::div(dvL, M::root(thr * thr + dv2));
    T S_L = alpha * cL; if (vL > 0) S_L += vL; else S_L -= vL;
    T S_R = alpha * cR; if (vR > 0) S_R += vR; else S_R -= vR;
    T S_M = tmax(S_L, S_R);
    o.S_M = tmax(cL + fabs(vL), cR + fabs(vR));
    T frl = rL * (vL + S_M);
    T frr = rR * (vR - S_M);
    T denom = M::rcp(rL * vL - rR * vR + S_M * (rL + rR));
    o.P_M = (frl * pR - frr * pL) * denom;
    o.Fv[0] = (vRx - vLx) * frl * frr * denom + (o.P_M) * n0;
    o.Fv[1] = (vRy - vLy) * frl * frr * denom + (o.P_M) * n1;
    o.Fv[2] = (vRz - vLz) * frl * frr * denom + (o.P_M) * n2;
    o.Fp = (S_M * (frl * pR + frr * pL) + (hR - hL) * frl * frr) * denom;
    return o;
}

// exact iterative solver (reimann.h:685-724, 1034-1138); for MFM only P_M and
// S_M change (sample_reimann_standard returns at once, reimann.h:873-877).
// Rare (HLLC pressure above the limiter), so out of line with scalar
// arguments; returns (P_M, S_M, status) with status 0 or 1234545.
template <typename T>
__device__ __noinline__ ROut<T> riemann_kt(T rL, T pL, T vLx, T vLy, T vLz, T rR, T pR, T vRx, T vRy, T vRz, T n0,
                                           T n1, T n2, T vL, T vR, T cL, T cR, T hL, T hR)
{
    return riemann_kt_body<T>(rL, pL, vLx, vLy, vLz, rR, pR, vRx, vRy, vRz, n0, n1, n2, vL, vR, cL, cR, hL, hR);
}

struct ExactOut {
    double P_M, S_M;
    int status;
};

// iterative_Riemann_solver (reimann.h:1034-1100) with guess_for_pressure
// (1101-1138): star pressure and contact speed of two states with positive
// densities.  Returns false (P, S unset) when the states open a vacuum or
// the iteration does not converge.
__device__ __forceinline__ bool exact_star(double rL, double pL, double rR, double pR, double vL, double vR, double cL,
                                           double cR, const GammaConsts &G, double &P_out, double &S_out)
{
    double dvel = vR - vL;
    double check_vel = G.G4 * (cR + cL) - dvel;
    if (check_vel < 0) return false;
    // guess_for_pressure (reimann.h:1101-1138)
    double Pg;
    {
        double pv = 0.5 * (pL + pR) - 0.125 * (vR - vL) * (pL + pR) * (cL + cR);
        double pmin = dmin(pL, pR), pmax = dmax(pL, pR);
        if (pmin <= 0) Pg = 0.5 * (pmin + pmax);
        else {
            double qrat = pmax / pmin;
            if (qrat <= 2.0 && (pmin <= pv && pv <= pmax)) Pg = pv;
            else if (pv < pmin) {
                double pnu = (cL + cR) - G.G7 * (vR - vL);
                double pde = cL / pow(pL, G.G1) + cR / pow(pR, G.G1);
                Pg = pow(pnu / pde, G.G3);
            } else {
                double gel = sqrt((G.G5 / rL) / (G.G6 * pL + pv));
                double ger = sqrt((G.G5 / rR) / (G.G6 * pR + pv));
                double x = (gel * pL + ger * pR - (vR - vL)) / (gel + ger);
                if (x < pmin || x > pmax) x = pmin;
                Pg = x;
            }
        }
    }
    double tol = 100.0, W_L = 0, W_R = 0, Z_L, Z_R;
    int niter = 0;
    while ((tol > 1.e-6) && (niter < 1000)) {
        double Pg_prev = Pg;
        if (Pg > pL) {
            double a0 = G.G5 / rL, a1 = G.G6 * pL, a2 = sqrt(a0 / (Pg + a1));
            W_L = (Pg - pL) * a2;
            Z_L = a2 * (1.0 - 0.5 * (Pg - pL) / (a1 + Pg));
        } else {
            double pratio = Pg / pL;
            W_L = G.G4 * cL * (pow(pratio, G.G1) - 1);
            Z_L = 1 / (rL * cL) * pow(Pg / pL, -G.G2);
        }
        if (Pg > pR) {
            double a0 = G.G5 / rR, a1 = G.G6 * pR, a2 = sqrt(a0 / (Pg + a1));
            W_R = (Pg - pR) * a2;
            Z_R = a2 * (1.0 - 0.5 * (Pg - pR) / (a1 + Pg));
        } else {
            double pratio = Pg / pR;
            W_R = G.G4 * cR * (pow(pratio, G.G1) - 1);
            Z_R = 1 / (rR * cR) * pow(pratio, -G.G2);
        }
        if (niter < 500) Pg -= (W_L + W_R + dvel) / (Z_L + Z_R);
        else Pg -= 0.5 * (W_L + W_R + dvel) / (Z_L + Z_R);
        if (Pg < 0.1 * Pg_prev) Pg = 0.1 * Pg_prev;
        tol = 2.0 * fabs((Pg - Pg_prev) / (Pg + Pg_prev));
        niter++;
    }
    if (niter < 1000) {
        P_out = Pg;
        S_out = 0.5 * (vL + vR) + 0.5 * (W_R - W_L);
        return true;
    }
    return false;
}

__device__ __forceinline__ ExactOut riemann_exact_body(double rL, double pL, double rR, double pR, double vL,
                                                      double vR, double cL, double cR, double P_M_in, double S_M_in,
                                                      const GammaConsts &G)
{
    ExactOut o = {P_M_in, S_M_in, 0};
    if ((pL == 0 && pR == 0) || (rL == 0 && rR == 0)) { o.P_M = o.S_M = 0; return o; }
    if ((rL > 0) && (rR > 0)) {
        const bool converged = exact_star(rL, pL, rR, pR, vL, vR, cL, cR, G, o.P_M, o.S_M);
        if (!converged) {
            // vacuum inside the fan, sampled at S = 0 (reimann.h:818-846)
            double S_L = vL + G.G4 * cL, S_R = vR - G.G4 * cR;
            o.P_M = 0;
            if (0.0 <= S_L) o.S_M = S_L;
            else if (0.0 >= S_R) o.S_M = S_R;
            else o.S_M = 0.0;
        }
        return o;
    }
    if ((rL < 0) || (rR < 0)) { o.status = 1234545; return o; }
    if (rL > 0) { o.P_M = 0; o.S_M = vL + G.G4 * cL; }
    if (rR > 0) { o.P_M = 0; o.S_M = vR - G.G4 * cR; }
    return o;
}

// out of line for the force loop (rarely taken), inline for the deferred-pair kernels
__device__ __noinline__ ExactOut riemann_exact(double rL, double pL, double rR, double pR, double vL, double vR,
                                               double cL, double cR, double P_M_in, double S_M_in, GammaConsts G)
{
    return riemann_exact_body(rL, pL, rR, pR, vL, vR, cL, cR, P_M_in, S_M_in, G);
}

// Riemann_solver() for MFM hydro (reimann.h:286-388).  Returns a stage mask:
// bit0 HLLC, bit1 KT, bit2 exact; err gets 1234545 on the exact solver's
// negative-density exit.  The exact solver always runs in FP64.  DEFER: stop
// where the exact solver would be called and return kDeferBit instead.
// FP32: kNearBit when the wavespeed estimate stopped at a near-threshold test.
constexpr int kDeferBit = 32;
constexpr int kNearBit = 64;

template <typename T, bool DEFER = false, bool INLINE = false>
__device__ __forceinline__ int riemann_solve(const RState<T> &L, const RState<T> &R, const T n[3], T limiter,
                                             const GammaConsts &G, ROut<T> &o, int &err)
{
    using M = Real<T>;
    if ((L.p < 0 && R.p < 0) || (L.rho < 0) || (R.rho < 0)) { o.P_M = 0; return 0; }
    T cL = M::root(M::div(T(G.G0) * L.p, L.rho));
    T cR = M::root(M::div(T(G.G0) * R.p, R.rho));
    T uL = M::div(L.p, (T(G.G9) * L.rho));
    T uR = M::div(R.p, (T(G.G9) * R.rho));
    T hL = M::div(L.p, L.rho) + uL + T(0.5) * (L.v[0] * L.v[0] + L.v[1] * L.v[1] + L.v[2] * L.v[2]);
    T hR = M::div(R.p, R.rho) + uR + T(0.5) * (R.v[0] * R.v[0] + R.v[1] * R.v[1] + R.v[2] * R.v[2]);
    T vL = L.v[0] * n[0] + L.v[1] * n[1] + L.v[2] * n[2];
    T vR = R.v[0] * n[0] + R.v[1] * n[1] + R.v[2] * n[2];
    int mask = 1;
    if (wavespeeds_pstar(L, R, G, n, vL, vR, cL, cR, hL, hR, limiter, o)) return mask | kNearBit;
    for (int k = 0; k < 3; k++) o.Fv[k] = o.P_M * n[k];
    o.Fp = o.P_M * o.S_M;
    if ((o.P_M < 0) || isnan(o.P_M)) {
        if (INLINE)
            o = riemann_kt_body<T>(L.rho, L.p, L.v[0], L.v[1], L.v[2], R.rho, R.p, R.v[0], R.v[1], R.v[2], n[0], n[1],
                                   n[2], vL, vR, cL, cR, hL, hR);
        else
            o = riemann_kt<T>(L.rho, L.p, L.v[0], L.v[1], L.v[2], R.rho, R.p, R.v[0], R.v[1], R.v[2], n[0], n[1], n[2],
                              vL, vR, cL, cR, hL, hR);
        mask |= 2;
        // FP32: the exact solver's P_M and S_M replace the estimate's in the
        // entropic switch, so whether it runs matters; decided in FP64 near the limiter
        if (near_threshold(o.P_M, limiter, limiter)) return mask | kNearBit;
    }
    if ((o.P_M < 0) || isnan(o.P_M) || (o.P_M > limiter)) {
        if (DEFER) return mask | kDeferBit;
        const ExactOut e = INLINE ? riemann_exact_body(L.rho, L.p, R.rho, R.p, vL, vR, cL, cR, o.P_M, o.S_M, G)
                                  : riemann_exact(L.rho, L.p, R.rho, R.p, vL, vR, cL, cR, o.P_M, o.S_M, G);
        o.P_M = (T)e.P_M;
        o.S_M = (T)e.S_M;
        if (e.status) err = e.status;
        mask |= 4;
    }
    return mask;
}

// ---------------------------------------------------------------------------
// One (i, j) interaction of the MFM force loop (hydro_evaluate.h:180-516,
// compute_finitevol_faces.h:14-106, hydro_core_meshless.h:8-385).
// ---------------------------------------------------------------------------
// Quantities of one side of a pair.  `e` = P/rho and `cn` stay FP64 in both
// precisions: the entropic switch compares e_i with e_j (equal-u gas makes
// that a tie decided by rounding) and the face fallback compares CN^2 with
// 1e12, so both decisions are taken exactly as in the FP64 pass.
template <typename T>
struct PairSide {
    T v[3], h, hinv, rho, P, cs, V, dhsml, fce;
    T nvt[6], gr[3], gp[3], gv[9];
    double cn, e;
};

// i-side extras: kernel factors and the condition-number threshold
template <typename T>
struct PairI : PairSide<T> {
    T hinv3, hinv4;
    double cnumcrit2;  // CONDITION_NUMBER_DANGER^2 - CN_i^2 (hydro_evaluate.h:203-210)
};

// dp = x_i - x_j (nearest image), r2 = |dp|^2.  Returns the status code
// (0, 1234 or 1234545); Fv, Fp, vsig are the pair's flux terms for i, mask
// the stages used (bit0 HLLC, bit1 KT, bit2 exact, bit3 retry).  DEFER:
// returns kDeferred, with no flux terms, for a pair that would call the exact
// Riemann solver (the caller evaluates it later with DEFER = false).  FP32:
// a pair whose wavespeed estimate or first-order-retry test lies within
// kDecideMargin of its threshold returns kDeferred (DEFER) or kNearThreshold,
// with no flux terms; the caller evaluates it in FP64.
constexpr int kDeferred = -1;
constexpr int kNearThreshold = -2;

template <typename T, bool DEFER = false, bool INLINE = false>
__device__ __forceinline__ int force_pair(const PairI<T> &I, const PairSide<T> &J, const T dp[3], T r2,
                                          const GammaConsts &G, ROut<T> &ro, T Fv[3], T &Fp, T &vsig, int &mask)
{
    using M = Real<T>;
    int status = 0;
    const T r = M::root(r2);
    const T rinv = M::rcp(r);
    const T dv[3] = {I.v[0] - J.v[0], I.v[1] - J.v[1], I.v[2] - J.v[2]};
    vsig = I.cs + J.cs;
    const T vdotr2 = dp[0] * dv[0] + dp[1] * dv[1] + dp[2] * dv[2];
    if (vdotr2 < 0) vsig -= T(1.0) * vdotr2 * rinv;  // MFM: fac_mu = 1 (hydro_evaluate.h:235-245)
    T wk_i = 0, dwk_i = 0, wk_j = 0, dwk_j = 0;
    if (r < I.h) kernel_wdw(r * I.hinv, I.hinv3, I.hinv4, wk_i, dwk_i);
    if (r < J.h) {
        const T hinv3_j = J.hinv * J.hinv * J.hinv;
        kernel_wdw(r * J.hinv, hinv3_j, hinv3_j * J.hinv, wk_j, dwk_j);
    }
    Fv[0] = Fv[1] = Fv[2] = 0;
    Fp = 0;
    mask = 0;
    const double leak = 0.5 * ((double)I.fce + (double)J.fce);
    // FP32 only: a listed pair (r^2 < h^2 decided in FP64) whose FP32 r / h
    // rounds to 1 on both sides has no kernel weight; in FP64 its weights are
    // tiny but positive and the face vanishes, so give it no face here instead
    // of the 0/0 of the volume-weighted switch below.
    if (sizeof(T) == sizeof(float) && !(wk_i + wk_j > 0)) return 0;
    // FP32 only: a kernel weight near the edge of its support, w ~ (1 - u)^3,
    // carries the FP32 geometry's relative error ~3e-7 / (1 - u).  With
    // pressures of similar size such a pair weighs nothing in i's sums; across
    // a pressure contrast above 1e4 (a blast front) it can carry most of i's
    // energy flux with that error, so the pair is decided in FP64 by the
    // caller (at contrast C the FP32 error would reach ~(2C)^(1/3) 1.5e-7 of
    // the particle's sum).
    if (sizeof(T) == sizeof(float) && (tmax(I.P, J.P) > T(1e4) * tmin(I.P, J.P)))
        return DEFER ? kDeferred : kNearThreshold;
    // effective face (compute_finitevol_faces.h:14-106)
    T wt_i = I.V, wt_j = J.V;
    const T vratio = M::div(M::div(fabs(I.V - J.V), tmin(I.V, J.V)), T(3));
    if (near_threshold(vratio, T(1.25), T(1.25))) return DEFER ? kDeferred : kNearThreshold;
    if (vratio > T(1.25)) {
        wt_i = wt_j = M::div(I.V * J.V * (wk_i + wk_j), (I.V * wk_i + J.V * wk_j));
    }
    T A[3];
    T Anorm = 0, fdp = 0;
    {
        const T ti_d[3] = {I.nvt[0] * dp[0] + I.nvt[1] * dp[1] + I.nvt[2] * dp[2],
                           I.nvt[1] * dp[0] + I.nvt[3] * dp[1] + I.nvt[4] * dp[2],
                           I.nvt[2] * dp[0] + I.nvt[4] * dp[1] + I.nvt[5] * dp[2]};
        const T tj_d[3] = {J.nvt[0] * dp[0] + J.nvt[1] * dp[1] + J.nvt[2] * dp[2],
                           J.nvt[1] * dp[0] + J.nvt[3] * dp[1] + J.nvt[4] * dp[2],
                           J.nvt[2] * dp[0] + J.nvt[4] * dp[1] + J.nvt[5] * dp[2]};
#pragma unroll
        for (int k = 0; k < 3; k++) {
            A[k] = wk_i * wt_i * ti_d[k] + wk_j * wt_j * tj_d[k];
            Anorm += A[k] * A[k];
            fdp += A[k] * dp[k];
        }
    }
    const bool cn_switch = J.cn * J.cn > 1.0e12 + I.cnumcrit2;
    if (!cn_switch && near_threshold(fdp, T(0), (fabs(A[0]) + fabs(A[1]) + fabs(A[2])) * r))
        return DEFER ? kDeferred : kNearThreshold;
    if (cn_switch || (fdp < 0)) {
        T An = M::div(-(wt_i * I.V * dwk_i + wt_j * J.V * dwk_j), r);
        A[0] = An * dp[0];
        A[1] = An * dp[1];
        A[2] = An * dp[2];
        Anorm = An * An * r2;
    }
    Anorm = M::root(Anorm);
    if (Anorm == 0) return 0;

    const T n[3] = {M::div(A[0], Anorm), M::div(A[1], Anorm), M::div(A[2], Anorm)};
    // faces at the midpoint: s_i = -r/2, s_j = +r/2 (hydro_core_meshless.h:70-90)
    const T s_i = T(0) - T(0.5) * r, s_j = T(0) - (T(-0.5) * r);
    T di[3], dj[3], vf[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const T e = dp[k] * rinv;
        dj[k] = e * s_j;
        di[k] = e * s_i;
        vf[k] = rinv * (-s_i * J.v[k] + s_j * I.v[k]);
    }
    T fvi = 0, fvj = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) { fvi += I.v[k] * n[k]; fvj += J.v[k] * n[k]; }
    const T fadv = rinv * (-s_i * fvj + s_j * fvi);
    T v2a = 0;
    T vdotr2_phys = vdotr2;
    vdotr2_phys *= M::rcp(r * T(1.0));
    if (vdotr2_phys < 0) v2a = vdotr2_phys * vdotr2_phys;
    const T vdotf2 = fvi - fvj;
    if (vdotf2 < 0) v2a = tmax(v2a, vdotf2 * vdotf2);
    const int recon = (leak > 1) ? 0 : 1;
    RState<T> SL, SR;
    bool near = reconstruct(I.rho, I.gr, J.rho, J.gr, di, dj, SL.rho, SR.rho, recon);
    near |= reconstruct(I.P, I.gp, J.P, J.gp, di, dj, SL.p, SR.p, recon);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        near |= reconstruct(I.v[k], I.gv + 3 * k, J.v[k], J.gv + 3 * k, di, dj, SL.v[k], SR.v[k], recon);
        SL.v[k] -= vf[k];
        SR.v[k] -= vf[k];
    }
    if (sizeof(T) == sizeof(float) && near) return DEFER ? kDeferred : kNearThreshold;
    const T pti = I.P + I.rho * v2a;
    const T ptj = J.P + J.rho * v2a;
    T lim = T(1.1) * T(1.0) * tmax(pti, ptj);
    if (recon == 0) lim = tmax(lim, tmax(tmax(I.P, J.P), T(2.) * tmax(I.rho, J.rho) * v2a));
    // Riemann problem with the reference's two retries (hydro_core_meshless.h:209-265):
    // attempt 0 reconstructed states, 1 first-order states, 2 first-order with
    // zero velocities; one solver instance in the code
    int err = 0;
    T lim_a = lim;
    for (int attempt = 0; attempt < 3; attempt++) {
        mask |= riemann_solve<T, DEFER, INLINE>(SL, SR, n, lim_a, G, ro, err);
        if (sizeof(T) == sizeof(float) && (mask & kNearBit)) return DEFER ? kDeferred : kNearThreshold;
        if (DEFER && (mask & kDeferBit)) return kDeferred;
        const bool bad = (ro.P_M < 0) || isnan(ro.P_M);
        if (attempt == 0 && near_threshold(ro.P_M, T(1.4) * lim, lim)) return DEFER ? kDeferred : kNearThreshold;
        if (attempt == 0) {
            if (!(bad || (ro.P_M > T(1.4) * lim))) break;
            mask |= 8;
            lim_a = T(1.4) * lim;
        } else if (attempt == 1) {
            if (!bad) break;
            lim_a = T(2.0) * lim;
        } else {
            if (bad) status = 1234;
            break;
        }
        SR.p = I.P; SL.p = J.P;
        SR.rho = I.rho; SL.rho = J.rho;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            SR.v[k] = (attempt == 0) ? I.v[k] - vf[k] : T(0.0);
            SL.v[k] = (attempt == 0) ? J.v[k] - vf[k] : T(0.0);
        }
    }
    if (err) status = err;
    if ((ro.P_M > 0) && !isnan(ro.P_M)) {
#pragma unroll
        for (int k = 0; k < 3; k++) ro.Fp += vf[k] * ro.Fv[k];
        Fp = Anorm * ro.Fp;
#pragma unroll
        for (int k = 0; k < 3; k++) Fv[k] = Anorm * ro.Fv[k];
        // MFM adiabatic face correction (hydro_core_meshless.h:335-374)
        const T sm_c = M::div(fabs(ro.S_M), tmin(I.cs, J.cs));
        // FP32: the switch's decisions (|S_M|/c against 0.5 and 1e-3, and the
        // signs of the energy transfer to j or i) flip the pair's energy flux
        // by du_new - du_old; one within the FP32 margin of its threshold is
        // decided in FP64 by the caller.  S_M of nearly equal states comes
        // from a pressure difference: its FP32 error is ~eps c / gamma in
        // absolute terms whatever its size, so |S_M| / c carries ~1e-7
        // absolute error and the margins below hold that with room (0.03 +
        // threshold scales kDecideMargin to 3e-4 of 1e-3 and 1e-5 of 0.5).
        if (!(leak > 1) && near_threshold(sm_c, T(0.5), T(0.5 + 0.03))) return DEFER ? kDeferred : kNearThreshold;
        if ((sm_c < T(0.5)) || (leak > 1)) {
            int use = 1;
            const T fpm = ro.P_M * Anorm;
            const T pdv_fac = ro.P_M * vdotr2_phys / T(1.0);
            const T pdv_i = dwk_i * I.V * I.V * I.dhsml * pdv_fac;
            const T pdv_j = dwk_j * J.V * J.V * J.dhsml * pdv_fac;
            const T du_old = fpm * (ro.S_M + fadv);
            const T du_new = T(0.5) * (pdv_i - pdv_j + fpm * (fvi + fvj));
            const double cnum2 = J.cn * J.cn;
            if (cnum2 < I.cnumcrit2 && (I.e != J.e) && near_threshold(sm_c, T(1.e-3), T(1.e-3 + 0.03)))
                return DEFER ? kDeferred : kNearThreshold;
            if (sm_c > T(1.e-3) && cnum2 < I.cnumcrit2) {
                if (I.e != J.e) {
                    const T fpv = (I.e > J.e) ? fpm * fvj : fpm * fvi;
                    // dtoj = -du_old + fpm fvj, dtoi = du_old - fpm fvi: compared with 0 and with
                    // -du_new + fpm fvj (+du_new - fpm fvi); the S_M inside du_old carries
                    // ~eps c absolute error (above), hence the fpm c term
                    const T scale = fabs(du_old) + fabs(fpv) + fabs(du_new) + T(0.1) * fpm * tmin(I.cs, J.cs);
                    if (I.e > J.e) {
                        const T dtoj = -du_old + fpm * fvj;
                        if (near_threshold(dtoj, T(0), scale) || near_threshold(dtoj, -du_new + fpm * fvj, scale))
                            return DEFER ? kDeferred : kNearThreshold;
                        if (dtoj > 0) use = 0;
                        else if (dtoj < 0) { if (dtoj > -du_new + fpm * fvj) use = 0; }
                    } else {
                        const T dtoi = +du_old - fpm * fvi;
                        if (near_threshold(dtoi, T(0), scale) || near_threshold(dtoi, +du_new - fpm * fvi, scale))
                            return DEFER ? kDeferred : kNearThreshold;
                        if (dtoi > 0) use = 0;
                        else if (dtoi < 0) { if (dtoi > +du_new - fpm * fvi) use = 0; }
                    }
                }
            }
            if (cnum2 >= I.cnumcrit2) use = 1;
            if (use) Fp += du_new - du_old;
        }
    } else {
        Fp = 0;
        Fv[0] = Fv[1] = Fv[2] = 0;
    }
    return status;
}

}  // namespace gzm
