// mfv_device.h -- device physics of the meshless finite-volume (MFV) variant
// of the hydro pass, gfx950, FP64.
//
// GIZMO built with HYDRO_MESHLESS_FINITE_VOLUME instead of
// HYDRO_MESHLESS_FINITE_MASS moves its mesh-generating points with a smoothed
// velocity (HYDRO_FIX_MESH_MOTION = 5, allvars.h:94-98) rather than with the
// fluid, so mass crosses the faces: the Riemann solver returns the fluxes of
// mass, momentum and total energy through a face at rest in the states'
// frame (HLLC with its four-wave fluxes, the KT flux, or the exact solution
// sampled at the face), the force loop boosts them back to the simulation
// frame and exchanges mass between the pair.  Faces, reconstruction and the
// star-state estimates are those of the MFM pass (mfm_device.h).
#pragma once

#include "mfm_device.h"

namespace gzm {

struct ROutV {
    double P_M, S_M, Frho, Fv[3], Fp;
};

// HLLC_fluxes (reimann.h:603-675): the HLLC flux through the face (frame
// speed 0) from the estimate's P_M, S_M and outer wave speeds
__device__ inline void hllc_fluxes_mfv(const RState<double> &L, const RState<double> &R, ROutV &o, const double n[3],
                                       double vL, double vR, double hL, double hR, double S_L, double S_R)
{
    const double P_M = o.P_M, S_M = o.S_M, vf = 0;
    if ((P_M <= 0) || isnan(P_M)) return;
    if ((S_M == S_L) || (S_M == S_R) || (S_M == vf)) {
        o.Frho = 0;
        o.Fp = P_M * S_M;
        for (int k = 0; k < 3; k++) o.Fv[k] = P_M * n[k];
        return;
    }
    // one side's state (star or not) with its wave speed S and star factor
    auto side = [&](const RState<double> &Q, double vQ, double hQ, double S, bool star) {
        if (!star) {
            o.Frho = Q.rho * (vQ - vf);
            o.Fp = Q.rho * hQ * (vQ - vf) + Q.p * vf;
            for (int k = 0; k < 3; k++) o.Fv[k] = o.Frho * Q.v[k] + Q.p * n[k];
            return;
        }
        double nfac = Q.rho * (S - vQ) / (S - S_M);
        if (nfac < 0) nfac = 0;
        o.Frho = Q.rho * (vQ - S) + nfac * (S - vf);
        const double eK = Q.rho * hQ - Q.p;
        o.Fp = (Q.rho * hQ * vQ - eK * S) + (S - vf) * nfac * (eK / Q.rho + (S_M - vQ) * (S_M + Q.p / (Q.rho * (S - vQ))));
        const double pfac = nfac * (S - vf) * (S_M - vQ) + Q.p;
        for (int k = 0; k < 3; k++) o.Fv[k] = o.Frho * Q.v[k] + pfac * n[k];
    };
    if (vf < S_L) side(L, vL, hL, S_L, false);
    else if ((S_L <= vf) && (vf <= S_M)) side(L, vL, hL, S_L, true);
    else if ((S_M <= vf) && (vf <= S_R)) side(R, vR, hR, S_R, true);
    else side(R, vR, hR, S_R, false);
}

// Kurganov-Tadmor/PWK flux, MFV branch (reimann.h:484-507)
__device__ inline void riemann_kt_mfv(const RState<double> &L, const RState<double> &R, ROutV &o, const double n[3],
                                      double vL, double vR, double cL, double cR, double hL, double hR)
{
    const double thr = 0.001 * 0.5 * (L.rho + R.rho) * 0.5 * (cL + cR);
    double dv2 = 0;
    for (int k = 0; k < 3; k++) {
        const double d = R.rho * R.v[k] - L.rho * L.v[k];
        dv2 += d * d;
    }
    const double alpha = fabs(R.rho * vR - L.rho * vL) / sqrt(thr * thr + dv2);
    double S_L = alpha * cL; if (vL > 0) S_L += vL; else S_L -= vL;
    double S_R = alpha * cR; if (vR > 0) S_R += vR; else S_R -= vR;
    const double S_M = dmax(S_L, S_R);
    o.S_M = dmax(cL + fabs(vL), cR + fabs(vR));
    const double frl = L.rho * (vL + S_M);
    const double frr = R.rho * (vR - S_M);
    o.P_M = 0.5 * (R.p + L.p);
    o.Frho = 0.5 * (frl + frr);
    for (int k = 0; k < 3; k++) o.Fv[k] = 0.5 * (frl * L.v[k] + frr * R.v[k]) + (o.P_M) * n[k];
    o.Fp = 0.5 * (frl * hL + frr * hR + S_M * (R.p - L.p));
}

// The exact solution's face state at speed S = 0 (reimann.h:731-1030).  Frho,
// Fp, Fv hold density, pressure and velocity until face_to_flux_mfv.
// Vacuum on the left: the right state's fan (sample_reimann_vaccum_left)
__device__ inline void sample_vacuum_left(const RState<double> &R, ROutV &o, const GammaConsts &G, const double n[3],
                                          double vR, double cR)
{
    const double S = 0, S_R = vR - G.G4 * cR;
    if (S_R > S) {
        o.P_M = 0; o.S_M = S_R; o.Frho = 0;
    } else if (vR + cR > S) {
        const double C_eff = G.G5 * (cR - G.G7 * (vR - S));
        o.P_M = R.p * pow(C_eff / cR, G.G3);
        o.S_M = G.G5 * (-cR + G.G7 * vR + S);
        o.Frho = R.rho * pow(C_eff / cR, G.G4);
    } else {
        o.P_M = R.p; o.S_M = vR; o.Frho = R.rho;
    }
    o.Fp = o.P_M;
    for (int k = 0; k < 3; k++) o.Fv[k] = R.v[k] + (o.S_M - vR) * n[k];
}

// vacuum on the right: the left state's fan (sample_reimann_vaccum_right)
__device__ inline void sample_vacuum_right(const RState<double> &L, ROutV &o, const GammaConsts &G, const double n[3],
                                           double vL, double cL)
{
    const double S = 0, S_L = vL + G.G4 * cL;
    if (S_L < S) {
        o.P_M = 0; o.S_M = S_L; o.Frho = 0;
    } else if (vL - cL < S) {
        const double C_eff = G.G5 * (cL + G.G7 * (vL - S));
        o.P_M = L.p * pow(C_eff / cL, G.G3);
        o.S_M = G.G5 * (cL + G.G7 * vL + S);
        o.Frho = L.rho * pow(C_eff / cL, G.G4);
    } else {
        o.P_M = L.p; o.S_M = vL; o.Frho = L.rho;
    }
    o.Fp = o.P_M;
    for (int k = 0; k < 3; k++) o.Fv[k] = L.v[k] + (o.S_M - vL) * n[k];
}

// the standard fan (sample_reimann_standard) with star pressure P_M and
// contact speed S_M already in o
__device__ inline void sample_standard(const RState<double> &L, const RState<double> &R, ROutV &o,
                                       const GammaConsts &G, const double n[3], double vL, double vR, double cL,
                                       double cR)
{
    const double S = 0;
    // a data state, or a star state of density rs behind the contact
    auto data = [&](const RState<double> &Q) {
        o.Fp = Q.p; o.Frho = Q.rho;
        for (int k = 0; k < 3; k++) o.Fv[k] = Q.v[k];
    };
    auto star = [&](const RState<double> &Q, double vQ, double rs) {
        o.Frho = rs; o.Fp = o.P_M;
        for (int k = 0; k < 3; k++) o.Fv[k] = Q.v[k] + (o.S_M - vQ) * n[k];
    };
    if (S <= o.S_M) {
        if (o.P_M <= L.p) {  // left rarefaction
            if (S <= vL - cL) { data(L); return; }
            const double C_eff_L = cL * pow(o.P_M / L.p, G.G1);
            if (S > o.S_M - C_eff_L) { star(L, vL, L.rho * pow(o.P_M / L.p, G.G8)); return; }
            const double S_eff = G.G5 * (cL + G.G7 * vL + S);
            const double C_eff = G.G5 * (cL + G.G7 * (vL - S));
            o.Frho = L.rho * pow(C_eff / cL, G.G4);
            o.Fp = L.p * pow(C_eff / cL, G.G3);
            for (int k = 0; k < 3; k++) o.Fv[k] = L.v[k] + (S_eff - vL) * n[k];
            return;
        }
        if (L.p > 0) {  // left shock
            const double pml = o.P_M / L.p;
            const double S_L = vL - cL * sqrt(G.G2 * pml + G.G1);
            if (S <= S_L) { data(L); return; }
            star(L, vL, L.rho * (pml + G.G6) / (pml * G.G6 + 1.0));
        } else {
            star(L, vL, L.rho / G.G6);
        }
        return;
    }
    if (o.P_M > R.p) {  // right shock
        if (R.p > 0) {
            const double pmr = o.P_M / R.p;
            const double S_R = vR + cR * sqrt(G.G2 * pmr + G.G1);
            if (S >= S_R) { data(R); return; }
            star(R, vR, R.rho * (pmr + G.G6) / (pmr * G.G6 + 1.0));
        } else {
            star(R, vR, R.rho / G.G6);
        }
        return;
    }
    if (S >= vR + cR) { data(R); return; }  // right rarefaction
    const double C_eff_R = cR * pow(o.P_M / R.p, G.G1);
    if (S <= o.S_M + C_eff_R) { star(R, vR, R.rho * pow(o.P_M / R.p, G.G8)); return; }
    const double S_eff = G.G5 * (-cR + G.G7 * vR + S);
    const double C_eff = G.G5 * (cR - G.G7 * (vR - S));
    o.Frho = R.rho * pow(C_eff / cR, G.G4);
    o.Fp = R.p * pow(C_eff / cR, G.G3);
    for (int k = 0; k < 3; k++) o.Fv[k] = R.v[k] + (S_eff - vR) * n[k];
}

// convert_face_to_flux (reimann.h:1146-1166): face state -> Euler fluxes
// through a face at rest
__device__ inline void face_to_flux_mfv(ROutV &o, const double n[3], const GammaConsts &G)
{
    const double rho = o.Frho, P = o.Fp;
    double v[3], v_line = 0, h = 0;
    for (int k = 0; k < 3; k++) {
        v[k] = o.Fv[k];
        v_line += v[k] * n[k];
        h += v[k] * v[k];
    }
    h *= 0.5 * rho;
    h += (G.G0 / G.G9) * P;
    o.Fp = h * v_line;
    o.Frho = rho * v_line;
    for (int k = 0; k < 3; k++) o.Fv[k] = o.Frho * v[k] + P * n[k];
}

// Riemann_solver_exact with the MFV sampling (reimann.h:685-724); returns 0
// or 1234545
__device__ __forceinline__ int riemann_exact_mfv_body(const RState<double> &L, const RState<double> &R, ROutV &o,
                                                      const GammaConsts &G, const double *n, double vL, double vR,
                                                      double cL, double cR)
{
    if ((L.p == 0 && R.p == 0) || (L.rho == 0 && R.rho == 0)) {
        o.P_M = o.S_M = 0;
        o.Frho = o.Fp = o.Fv[0] = o.Fv[1] = o.Fv[2] = 0;
        return 0;
    }
    if ((L.rho > 0) && (R.rho > 0)) {
        double P, S;
        if (exact_star(L.rho, L.p, R.rho, R.p, vL, vR, cL, cR, G, P, S)) {
            o.P_M = P;
            o.S_M = S;
            sample_standard(L, R, o, G, n, vL, vR, cL, cR);
        } else {
            // vacuum inside the fan (sample_reimann_vaccum_internal); the
            // velocity between the fans keeps the reference's (1 - n_k) weight
            const double S_L = vL + G.G4 * cL, S_R = vR - G.G4 * cR, S0 = 0;
            if (S0 <= S_L) sample_vacuum_right(L, o, G, n, vL, cL);
            else if (S0 >= S_R) sample_vacuum_left(R, o, G, n, vR, cR);
            else {
                o.P_M = 0; o.S_M = S0; o.Frho = 0; o.Fp = o.P_M;
                for (int k = 0; k < 3; k++)
                    o.Fv[k] = (L.v[k] + (R.v[k] - L.v[k]) * (S0 - S_L) / (S_R - S_L)) * (1 - n[k]) + S0 * n[k];
            }
        }
    } else {
        if ((L.rho < 0) || (R.rho < 0)) return 1234545;
        if (L.rho > 0) sample_vacuum_right(L, o, G, n, vL, cL);
        if (R.rho > 0) sample_vacuum_left(R, o, G, n, vR, cR);
    }
    face_to_flux_mfv(o, n, G);
    return 0;
}

// out of line for the loops where it is rare (scalar arguments, result by value)
struct ExactOutV {
    ROutV o;
    int status;
};
__device__ __noinline__ ExactOutV riemann_exact_mfv(double rL, double pL, double vLx, double vLy, double vLz, double rR,
                                                    double pR, double vRx, double vRy, double vRz, double n0,
                                                    double n1, double n2, double vL, double vR, double cL, double cR,
                                                    GammaConsts G)
{
    const RState<double> L = {rL, pL, {vLx, vLy, vLz}}, R = {rR, pR, {vRx, vRy, vRz}};
    const double n[3] = {n0, n1, n2};
    ExactOutV e;
    e.o = ROutV{0, 0, 0, {0, 0, 0}, 0};
    e.status = riemann_exact_mfv_body(L, R, e.o, G, n, vL, vR, cL, cR);
    return e;
}

__device__ __noinline__ ROutV riemann_kt_mfv_ool(double rL, double pL, double vLx, double vLy, double vLz, double rR,
                                                 double pR, double vRx, double vRy, double vRz, double n0, double n1,
                                                 double n2, double vL, double vR, double cL, double cR, double hL,
                                                 double hR)
{
    const RState<double> L = {rL, pL, {vLx, vLy, vLz}}, R = {rR, pR, {vRx, vRy, vRz}};
    const double n[3] = {n0, n1, n2};
    ROutV o = {0, 0, 0, {0, 0, 0}, 0};
    riemann_kt_mfv(L, R, o, n, vL, vR, cL, cR, hL, hR);
    return o;
}

// Riemann_solver() for MFV hydro (reimann.h:286-388): HLLC with the HLLC
// fluxes, the KT flux on a failed estimate, the exact solver past the
// limiter.  Returns the stage mask (1 HLLC, 2 KT, 4 exact).  DEFER: stop
// where the exact solver would run and return kDeferBit.  INLINE: the KT flux
// and the exact solver inline (every lane may take them), else out of line.
template <bool DEFER = false, bool INLINE = false>
__device__ inline int riemann_solve_mfv(const RState<double> &L, const RState<double> &R, const double n[3],
                                        double limiter, const GammaConsts &G, ROutV &o, int &err)
{
    if ((L.p < 0 && R.p < 0) || (L.rho < 0) || (R.rho < 0)) { o.P_M = 0; return 0; }
    const double cL = sqrt(G.G0 * L.p / L.rho), cR = sqrt(G.G0 * R.p / R.rho);
    const double uL = L.p / (G.G9 * L.rho), uR = R.p / (G.G9 * R.rho);
    const double hL = L.p / L.rho + uL + 0.5 * (L.v[0] * L.v[0] + L.v[1] * L.v[1] + L.v[2] * L.v[2]);
    const double hR = R.p / R.rho + uR + 0.5 * (R.v[0] * R.v[0] + R.v[1] * R.v[1] + R.v[2] * R.v[2]);
    const double vL = L.v[0] * n[0] + L.v[1] * n[1] + L.v[2] * n[2];
    const double vR = R.v[0] * n[0] + R.v[1] * n[1] + R.v[2] * n[2];
    int mask = 1;
    ROut<double> e = {o.P_M, o.S_M, {0, 0, 0}, 0};
    double S_LR[2] = {0, 0};
    wavespeeds_pstar<double>(L, R, G, n, vL, vR, cL, cR, hL, hR, limiter, e, S_LR);
    o.P_M = e.P_M;
    o.S_M = e.S_M;
    if ((o.P_M >= 0) && !isnan(o.P_M)) hllc_fluxes_mfv(L, R, o, n, vL, vR, hL, hR, S_LR[0], S_LR[1]);
    if ((o.P_M < 0) || isnan(o.P_M)) {
        if (INLINE) riemann_kt_mfv(L, R, o, n, vL, vR, cL, cR, hL, hR);
        else
            o = riemann_kt_mfv_ool(L.rho, L.p, L.v[0], L.v[1], L.v[2], R.rho, R.p, R.v[0], R.v[1], R.v[2], n[0], n[1],
                                   n[2], vL, vR, cL, cR, hL, hR);
        mask |= 2;
    }
    if ((o.P_M < 0) || isnan(o.P_M) || (o.P_M > limiter)) {
        if (DEFER) return mask | kDeferBit;
        int ee;
        if (INLINE) {
            ee = riemann_exact_mfv_body(L, R, o, G, n, vL, vR, cL, cR);
        } else {
            const ExactOutV e = riemann_exact_mfv(L.rho, L.p, L.v[0], L.v[1], L.v[2], R.rho, R.p, R.v[0], R.v[1],
                                                  R.v[2], n[0], n[1], n[2], vL, vR, cL, cR, G);
            o = e.o;
            ee = e.status;
        }
        if (ee) err = ee;
        mask |= 4;
    }
    return mask;
}

// One (i, j) interaction of the MFV force loop (hydro_evaluate.h:180-298,
// compute_finitevol_faces.h:14-106, hydro_core_meshless.h:8-334 with the
// HYDRO_MESHLESS_FINITE_VOLUME branches).  The face moves with
// rinv (-s_i v_j + s_j w_i): w_i is i's ParticleVel, and the neighbour's
// element velocity is its VelPred (hydro_evaluate.h:195-196).  The signal
// speed takes 3 |v.r|/r (hydro_evaluate.h:240-242), the limiter is doubled
// (hydro_core_meshless.h:194-196), the fluxes are boosted back from the face
// frame (285-295) and there is no entropic face correction.  Writes the
// pair's Fv[3], Fp, Frho times the face area, vsig and the stage mask (bit3
// the first-order retry); returns 0, 1234 or 1234545.  DEFER: returns
// kDeferred, with no flux terms, for a pair that would call the exact solver
// (every first-order retry follows an exact solution, so retries happen only
// in the caller's later evaluation of such a pair with DEFER = false).
template <bool DEFER = false, bool INLINE = false>
__device__ inline int force_pair_mfv(const PairI<double> &I, const PairSide<double> &J, const double w_i[3],
                                     const double dp[3], double r2, const GammaConsts &G, double Fv[3], double &Fp,
                                     double &Frho, double &vsig, int &mask)
{
    int status = 0;
    const double r = sqrt(r2);
    const double rinv = 1 / r;
    const double dv[3] = {I.v[0] - J.v[0], I.v[1] - J.v[1], I.v[2] - J.v[2]};
    vsig = I.cs + J.cs;
    const double vdotr2 = dp[0] * dv[0] + dp[1] * dv[1] + dp[2] * dv[2];
    if (vdotr2 < 0) vsig -= 3 * 1.0 * vdotr2 * rinv;
    double wk_i = 0, dwk_i = 0, wk_j = 0, dwk_j = 0;
    if (r < I.h) kernel_wdw(r * I.hinv, I.hinv3, I.hinv4, wk_i, dwk_i);
    if (r < J.h) {
        const double hinv3_j = J.hinv * J.hinv * J.hinv;
        kernel_wdw(r * J.hinv, hinv3_j, hinv3_j * J.hinv, wk_j, dwk_j);
    }
    Fv[0] = Fv[1] = Fv[2] = 0;
    Fp = 0;
    Frho = 0;
    mask = 0;
    const double leak = 0.5 * (I.fce + J.fce);
    double wt_i = I.V, wt_j = J.V;
    if ((fabs(I.V - J.V) / dmin(I.V, J.V)) / 3 > 1.25) wt_i = wt_j = I.V * J.V * (wk_i + wk_j) / (I.V * wk_i + J.V * wk_j);
    double A[3], Anorm = 0, fdp = 0;
    {
        const double ti_d[3] = {I.nvt[0] * dp[0] + I.nvt[1] * dp[1] + I.nvt[2] * dp[2],
                                I.nvt[1] * dp[0] + I.nvt[3] * dp[1] + I.nvt[4] * dp[2],
                                I.nvt[2] * dp[0] + I.nvt[4] * dp[1] + I.nvt[5] * dp[2]};
        const double tj_d[3] = {J.nvt[0] * dp[0] + J.nvt[1] * dp[1] + J.nvt[2] * dp[2],
                                J.nvt[1] * dp[0] + J.nvt[3] * dp[1] + J.nvt[4] * dp[2],
                                J.nvt[2] * dp[0] + J.nvt[4] * dp[1] + J.nvt[5] * dp[2]};
        for (int k = 0; k < 3; k++) {
            A[k] = wk_i * wt_i * ti_d[k] + wk_j * wt_j * tj_d[k];
            Anorm += A[k] * A[k];
            fdp += A[k] * dp[k];
        }
    }
    if ((J.cn * J.cn > 1.0e12 + I.cnumcrit2) || (fdp < 0)) {
        const double An = -(wt_i * I.V * dwk_i + wt_j * J.V * dwk_j) / r;
        A[0] = An * dp[0];
        A[1] = An * dp[1];
        A[2] = An * dp[2];
        Anorm = An * An * r2;
    }
    Anorm = sqrt(Anorm);
    if (Anorm == 0) return 0;
    const double n[3] = {A[0] / Anorm, A[1] / Anorm, A[2] / Anorm};
    const double s_i = 0 - 0.5 * r, s_j = 0 - (-0.5 * r);
    double di[3], dj[3], vf[3];
    for (int k = 0; k < 3; k++) {
        const double e = dp[k] * rinv;
        dj[k] = e * s_j;
        di[k] = e * s_i;
        vf[k] = rinv * (-s_i * J.v[k] + s_j * w_i[k]);
    }
    double fvi = 0, fvj = 0;
    for (int k = 0; k < 3; k++) { fvi += I.v[k] * n[k]; fvj += J.v[k] * n[k]; }
    double v2a = 0;
    double vdotr2_phys = vdotr2;
    vdotr2_phys *= 1 / (r * 1.0);
    if (vdotr2_phys < 0) v2a = vdotr2_phys * vdotr2_phys;
    const double vdotf2 = fvi - fvj;
    if (vdotf2 < 0) v2a = dmax(v2a, vdotf2 * vdotf2);
    const int recon = (leak > 1) ? 0 : 1;
    RState<double> SL, SR;
    reconstruct(I.rho, I.gr, J.rho, J.gr, di, dj, SL.rho, SR.rho, recon);
    reconstruct(I.P, I.gp, J.P, J.gp, di, dj, SL.p, SR.p, recon);
    for (int k = 0; k < 3; k++) {
        reconstruct(I.v[k], I.gv + 3 * k, J.v[k], J.gv + 3 * k, di, dj, SL.v[k], SR.v[k], recon);
        SL.v[k] -= vf[k];
        SR.v[k] -= vf[k];
    }
    double lim = 1.1 * 1.0 * dmax(I.P + I.rho * v2a, J.P + J.rho * v2a);
    lim *= 2.0;
    if (recon == 0) lim = dmax(lim, dmax(dmax(I.P, J.P), 2. * dmax(I.rho, J.rho) * v2a));
    // the reconstructed states, then the two retries (hydro_core_meshless.h:209-265)
    ROutV ro = {0, 0, 0, {0, 0, 0}, 0};
    int err = 0;
    double lim_a = lim;
    for (int attempt = 0; attempt < 3; attempt++) {
        mask |= riemann_solve_mfv<DEFER, INLINE>(SL, SR, n, lim_a, G, ro, err);
        if (DEFER && (mask & kDeferBit)) return kDeferred;
        const bool bad = (ro.P_M < 0) || isnan(ro.P_M);
        if (attempt == 0) {
            if (!(bad || (ro.P_M > 1.4 * lim))) break;
            mask |= 8;
            lim_a = 1.4 * lim;
        } else if (attempt == 1) {
            if (!bad) break;
            lim_a = 2.0 * lim;
        } else {
            if (bad) status = 1234;
            break;
        }
        SR.p = I.P; SL.p = J.P;
        SR.rho = I.rho; SL.rho = J.rho;
        for (int k = 0; k < 3; k++) {
            SR.v[k] = (attempt == 0) ? I.v[k] - vf[k] : 0.0;
            SL.v[k] = (attempt == 0) ? J.v[k] - vf[k] : 0.0;
        }
    }
    if (err) status = err;
    if ((ro.P_M > 0) && !isnan(ro.P_M)) {
        for (int k = 0; k < 3; k++) {
            ro.Fp += vf[k] * ro.Fv[k];
            ro.Fp += (0.5 * vf[k] * vf[k]) * ro.Frho;
            ro.Fv[k] += vf[k] * ro.Frho;
        }
        Frho = Anorm * ro.Frho;
        Fp = Anorm * ro.Fp;
        for (int k = 0; k < 3; k++) Fv[k] = Anorm * ro.Fv[k];
    }
    return status;
}

// hydro_final_operations_and_cleanup with the MFV terms (hydro_toplevel.c:
// 661-665, 738-764): the signal speed includes the flow over the element,
// the conserved-variable sums become dv/dt and du/dt
__device__ __forceinline__ void mfv_final_ops(double acc[3], double &dte, double &vmax, double dtm, const double v[3],
                                              const double w[3], double u, double m_i, double h_i, double max_hsml)
{
    double v2_p = vmax * vmax;
    for (int k = 0; k < 3; k++) v2_p += (v[k] - w[k]) * (v[k] - w[k]);
    vmax = sqrt(v2_p);
    for (int k = 0; k < 3; k++) {
        dte -= (v[k] / 1.0) * acc[k];
        dte += 0.5 * (v[k] / 1.0) * (v[k] / 1.0) * dtm;
        acc[k] -= (v[k] / 1.0) * dtm;
        acc[k] /= m_i;
    }
    dte -= u * dtm;
    dte /= m_i;
    if (h_i >= 0.99 * max_hsml) dte = 0;
}

}  // namespace gzm
