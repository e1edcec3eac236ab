// gzm_mfv.hip -- HIP kernels of the meshless finite-volume (MFV) variant of
// the hydro pass for gfx950 (FP64): the force loop with mass fluxes.
//
// The density iteration (which also sums the smoothed mesh velocity,
// density_pair<.., MFV>), the gradients and the pair list are those of the
// MFM pass.  The force loop keeps its gather form (one thread per active
// particle, its flux sums in registers) with one scatter: the reference gives
// each neighbour its share of the exchanged mass (SphP[j].dMass -= ...,
// hydro_evaluate.h:384-392, under `omp atomic` there), an FP64 atomic add here.
// dMass is therefore summed in a run-dependent order; every other output is
// deterministic for a fixed particle order.
#include "grid_device.h"
#include "gzm_internal.h"
#include "mfv_device.h"

namespace gzm {

namespace {

// hydro_force_initial_operations_preloop, MFV fields (hydro_toplevel.c:
// 976-978): DtMass, dMass and GravWorkTerm of the active particles start at
// zero before any neighbour adds its share of dMass
__global__ void mfv_preloop_kernel(int n_active, const int *active, MfvArgs m)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_active) return;
    const int i = active[t];
    m.dt_mass[i] = 0;
    m.d_mass[i] = 0;
    for (int k = 0; k < 3; k++) m.grav_work[3 * (size_t)i + k] = 0;
}

// i's side of an MFV pair
struct MfvSide {
    PairI<double> I;
    double w[3];   // ParticleVel
    double ti, dt; // integer timestep (as a double) and physical timestep
    double m;
};

__device__ __forceinline__ void mfv_side(const ForceArgs &a, const MfvArgs &m, int i, const ForceRec &ri, MfvSide &S)
{
    unpack_side(ri, S.I);
    S.I.hinv3 = S.I.hinv * S.I.hinv * S.I.hinv;
    S.I.hinv4 = S.I.hinv3 * S.I.hinv;
    S.I.cnumcrit2 = kCondDanger * kCondDanger - S.I.cn * S.I.cn;
    for (int k = 0; k < 3; k++) S.w[k] = m.pvel[3 * (size_t)i + k];
    S.ti = m.ti ? m.ti[i] : 1.0;
    S.dt = S.ti * m.timebase;  // hydro_evaluate.h:82
    S.m = ri.mc.x;
}

// The mass exchange of one pair (hydro_evaluate.h:379-392) with the
// reference's 10 % limiter; j_is_active_for_fluxes is always 0 there
// (hydro_evaluate.h:158-168): a smaller step than j's gives j the whole
// share, an equal step half of it.  Returns i's share; j's goes to j now.
__device__ __forceinline__ double mfv_mass_exchange(const ForceArgs &a, const MfvArgs &m, const MfvSide &S, int j,
                                                    double Frho)
{
    const double ti_j = m.ti ? m.ti[j] : 1.0;
    double dmh = Frho * S.dt, lim;
    if (dmh > 0) lim = a.rec[j].mc.x; else lim = S.m;
    lim *= 0.1;
    if (fabs(dmh) > lim) dmh *= lim / fabs(dmh);
    double own = 0;
    if (S.ti < ti_j) {
        own += dmh;
        atomicAdd(&m.d_mass[j], -dmh);
    }
    if (S.ti == ti_j) {
        own += 0.5 * dmh;
        atomicAdd(&m.d_mass[j], -(0.5 * dmh));
    }
    return own;
}

// the final operations of one particle and its stores
__device__ __forceinline__ void mfv_store(const ForceArgs &a, const MfvArgs &m, int i, const double acc_in[3],
                                          double dte, double vmax, double dtm, const double gw[3], bool finish)
{
    double acc[3] = {acc_in[0], acc_in[1], acc_in[2]};
    const double m_i = a.posm[i].w;
    if (finish && m_i > 0) {
        const double4 v4 = a.velu[i];
        const double v[3] = {v4.x, v4.y, v4.z};
        const double w[3] = {m.pvel[3 * (size_t)i + 0], m.pvel[3 * (size_t)i + 1], m.pvel[3 * (size_t)i + 2]};
        mfv_final_ops(acc, dte, vmax, dtm, v, w, v4.w, m_i, a.hsml[i], a.max_hsml);
    }
    a.acc[3 * (size_t)i + 0] = acc[0];
    a.acc[3 * (size_t)i + 1] = acc[1];
    a.acc[3 * (size_t)i + 2] = acc[2];
    a.dtu[i] = dte;
    a.vsig[i] = vmax;
    m.dt_mass[i] = dtm;
    for (int k = 0; k < 3; k++) m.grav_work[3 * (size_t)i + k] = gw[k];
}

// One active particle's MFV force loop (hydro_evaluate.h:146-498).
// THREE_PHASE (pair lists of <= 64): walk the gradient kernel's pair list,
// leave the pairs that reach the exact Riemann solver to
// force_mfv_deferred_kernel (bit k of the particle's defer mask) and the
// particles whose pair set overflowed the list to this kernel's scanning
// form; write raw sums for force_mfv_finish_kernel.  Otherwise (the scanning
// form) every pair in place, in one pass, with the final operations: over
// every active particle (USE_LIST: the pair list where it holds the whole
// set), or over rows[0, *n_rows) (active entries), striding.
template <bool USE_LIST, bool THREE_PHASE>
__global__ __launch_bounds__(THREE_PHASE ? kMaxBlock : kBlock) void force_mfv_kernel(ForceArgs a, MfvArgs m,
                                                                                    const int *rows,
                                                                                    const int *n_rows)
{
    const int n = rows ? *n_rows : a.n_active;
    const int stride = (int)(gridDim.x * blockDim.x);
    unsigned long long cnt = 0, nretry = 0, nexact = 0;
    for (int base = xcd_block() * (int)blockDim.x; base < n; base += stride) {
        const int q = base + (int)threadIdx.x;
        bool valid = q < n;
        const int t = valid ? (rows ? rows[q] : q) : 0;
        const bool listed = USE_LIST && valid && a.list.count[t] <= a.list.cap;
        if (THREE_PHASE && valid && !listed) {  // evaluated by the scanning form
            valid = false;
            a.defer.count[t] = 0;
            a.defer.raw[t] = 0;
        }
        for_box(a.boxes, a.box_of, valid ? a.active[t] : 0, valid, [&](const BoxDesc &G) {
            const int i = a.active[t];
            const ForceRec ri = a.rec[i];
            MfvSide S;
            mfv_side(a, m, i, ri, S);
            double acc[3] = {0, 0, 0}, dte = 0, dtm = 0, dm = 0, gw[3] = {0, 0, 0}, vmax = kMinReal;
            int status = 0, n_deferred = 0;
            uint64_t dmask = 0;
            if ((S.m > 0) && (ri.th.x > 0)) {
                double out_vsig = S.I.cs;
                const double x_i[3] = {ri.xh.x, ri.xh.y, ri.xh.z};
                const double h2_i = S.I.h * S.I.h;
                auto visit = [&](int j, int k, bool check) {
                    const double4 xj = a.rec[j].xh;
                    double dp[3] = {x_i[0] - xj.x, x_i[1] - xj.y, x_i[2] - xj.z};
                    wrap_dp(G.box, dp);
                    const double r2 = dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2];
                    if (check) {
                        const double h_j = xj.w;
                        if ((r2 >= h2_i) && (r2 >= h_j * h_j)) return;
                        if (r2 <= 0) return;
                        if (!(a.rec[j].mc.x > 0) || !(a.rec[j].th.x > 0)) return;  // hydro_evaluate.h:149-150
                    }
                    PairSide<double> J;
                    unpack_side(a.rec[j], J);
                    double Fv[3], Fp, Frho, vsig;
                    int mask;
                    const int e = force_pair_mfv<THREE_PHASE>(S.I, J, S.w, dp, r2, a.G, Fv, Fp, Frho, vsig, mask);
                    cnt++;
                    if (THREE_PHASE && e == kDeferred) {
                        dmask |= 1ull << k;
                        n_deferred++;
                        return;
                    }
                    if (e > status) status = e;
                    nretry += (mask & 8) ? 1 : 0;
                    nexact += (mask & 4) ? 1 : 0;
                    dm += mfv_mass_exchange(a, m, S, j, Frho);
                    dtm += Frho;
                    for (int c = 0; c < 3; c++) gw[c] += Frho * dp[c];
                    for (int c = 0; c < 3; c++) acc[c] += Fv[c];
                    dte += Fp;
                    if (vsig > out_vsig) out_vsig = vsig;
                    if (a.is_active && !a.is_active[j]) {
                        if (vsig > kWakeup * a.vsig_old[j]) a.wakeup[j] = 1;  // WAKEUP (hydro_evaluate.h:485-495)
                    }
                };
                if (listed) {
                    const int nl = a.list.count[t];
                    const int *row = a.list.idx + (size_t)t * a.list.cap;
                    for (int k = 0; k < nl; k++) visit(row[k], k, false);
                } else {
                    for_each_neighbour_candidate(a.tree, G, x_i, S.I.h, true, [&](int j) { visit(j, 0, true); });
                }
                if (THREE_PHASE) {
                    // raw sums; force_mfv_finish_kernel adds the deferred pairs and finishes
                    atomicAdd(&m.d_mass[i], dm);
                    mfv_store(a, m, i, acc, dte, out_vsig, dtm, gw, false);
                    a.defer.mask[t] = dmask;
                    a.defer.count[t] = n_deferred;
                    a.defer.raw[t] = 1;
                    if (status) atomicMax(&a.ctr->status, status);
                    return;
                }
                if (vmax < out_vsig) vmax = out_vsig;
            }
            if (THREE_PHASE) {
                a.defer.count[t] = 0;
                a.defer.raw[t] = 0;
            }
            atomicAdd(&m.d_mass[i], dm);
            mfv_store(a, m, i, acc, dte, vmax, dtm, gw, true);
            if (status) atomicMax(&a.ctr->status, status);
        });
    }
    wave_add_u64(&a.ctr->interactions, cnt);
    wave_add_u64(&a.ctr->retries, nretry);
    wave_add_u64(&a.ctr->exact_calls, nexact);
}

// Deferred MFV pairs q in [q_begin, min(total, q_end)), one per lane, with
// the exact Riemann solver inline: both shares of the exchanged mass go out
// here; the flux terms go to the round's result slots (9 per pair: dv/dt (3),
// du/dt, DtMass and v_sig terms, GravWorkTerm (3)).
__global__ __launch_bounds__(kBlock) void force_mfv_deferred_kernel(ForceArgs a, MfvArgs m)
{
    const int q = a.defer.q_begin + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int q_end = min(*a.defer.total, a.defer.q_end);
    const bool valid = q < q_end;
    const int2 ij = valid ? a.defer.pair[q - a.defer.q_begin] : make_int2(0, 0);
    unsigned long long nretry = 0, nexact = 0;
    for_box(a.boxes, a.box_of, ij.x, valid, [&](const BoxDesc &G) {
        const int i = ij.x, j = ij.y;
        const ForceRec ri = a.rec[i];
        MfvSide S;
        mfv_side(a, m, i, ri, S);
        const double4 xj = a.rec[j].xh;
        double dp[3] = {ri.xh.x - xj.x, ri.xh.y - xj.y, ri.xh.z - xj.z};
        wrap_dp(G.box, dp);
        const double r2 = dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2];
        PairSide<double> J;
        unpack_side(a.rec[j], J);
        double Fv[3], Fp, Frho, vsig;
        int mask;
        const int e = force_pair_mfv<false, true>(S.I, J, S.w, dp, r2, a.G, Fv, Fp, Frho, vsig, mask);
        nretry += (mask & 8) ? 1 : 0;
        nexact += (mask & 4) ? 1 : 0;
        atomicAdd(&m.d_mass[i], mfv_mass_exchange(a, m, S, j, Frho));
        double *o = a.defer.result + 9 * (size_t)(q - a.defer.q_begin);
        o[0] = Fv[0]; o[1] = Fv[1]; o[2] = Fv[2]; o[3] = Fp; o[4] = Frho; o[5] = vsig;
        o[6] = Frho * dp[0]; o[7] = Frho * dp[1]; o[8] = Frho * dp[2];
        if (a.is_active && !a.is_active[j]) {
            if (vsig > kWakeup * a.vsig_old[j]) a.wakeup[j] = 1;
        }
        if (e) atomicMax(&a.ctr->status, e);
    });
    wave_add_u64(&a.ctr->retries, nretry);
    wave_add_u64(&a.ctr->exact_calls, nexact);
}

// Add the round's deferred pairs to the raw sums in list order and, for the
// particles whose deferred pairs are all in, apply the final operations.
__global__ __launch_bounds__(kBlock) void force_mfv_finish_kernel(ForceArgs a, MfvArgs m)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n_active || !a.defer.raw[t]) return;
    const int i = a.active[t];
    double acc[3] = {a.acc[3 * (size_t)i + 0], a.acc[3 * (size_t)i + 1], a.acc[3 * (size_t)i + 2]};
    double gw[3] = {m.grav_work[3 * (size_t)i + 0], m.grav_work[3 * (size_t)i + 1], m.grav_work[3 * (size_t)i + 2]};
    double dte = a.dtu[i], dtm = m.dt_mass[i], out_vsig = a.vsig[i];
    const int q0 = a.defer.offset[t], q1 = q0 + a.defer.count[t];
    for (int q = max(q0, a.defer.q_begin); q < min(q1, a.defer.q_end); q++) {
        const double *o = a.defer.result + 9 * (size_t)(q - a.defer.q_begin);
        acc[0] += o[0];
        acc[1] += o[1];
        acc[2] += o[2];
        dte += o[3];
        dtm += o[4];
        if (o[5] > out_vsig) out_vsig = o[5];
        gw[0] += o[6];
        gw[1] += o[7];
        gw[2] += o[8];
    }
    const bool done = q1 <= a.defer.q_end;
    double vmax = out_vsig;
    if (done) {
        vmax = kMinReal;
        if (vmax < out_vsig) vmax = out_vsig;
        a.defer.raw[t] = 0;
    }
    mfv_store(a, m, i, acc, dte, vmax, dtm, gw, done);
}

__global__ void collect_overflow_rows_kernel(int n_active, const int *count, int cap, int *rows, int *n_rows)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n_active && count[t] > cap) rows[atomicAdd(n_rows, 1)] = t;
}

// The MFV Riemann solver on independent face states, one per thread (the
// force loop's inline functions): mode 0 Riemann_solver with the HLLC
// fluxes, 1 the KT flux alone, 2 the exact solver with its face sampling
// alone; out [n][7] = P_M, S_M, F_rho, F_v[3], F_p; mask = stages (mode 0),
// or the exact solver's status negated.
__global__ __launch_bounds__(kBlock) void riemann_states_mfv_kernel(int mode, GammaConsts G, int64_t n,
                                                                    const double *L, const double *R,
                                                                    const double *normal, const double *limiter,
                                                                    double *out, int *mask)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    RState<double> SL, SR;
    SL.rho = L[5 * t + 0]; SL.p = L[5 * t + 1];
    SR.rho = R[5 * t + 0]; SR.p = R[5 * t + 1];
    double nv[3];
    for (int k = 0; k < 3; k++) {
        SL.v[k] = L[5 * t + 2 + k];
        SR.v[k] = R[5 * t + 2 + k];
        nv[k] = normal[3 * t + k];
    }
    ROutV o = {0, 0, 0, {0, 0, 0}, 0};
    int m = 0;
    if (mode == 0) {
        int err = 0;
        m = riemann_solve_mfv<false, true>(SL, SR, nv, limiter[t], G, o, err);
        if (err) m = -err;
    } else {
        // the face-frame quantities of riemann_solve_mfv
        const double cL = sqrt(G.G0 * SL.p / SL.rho), cR = sqrt(G.G0 * SR.p / SR.rho);
        const double uL = SL.p / (G.G9 * SL.rho), uR = SR.p / (G.G9 * SR.rho);
        const double hL = SL.p / SL.rho + uL + 0.5 * (SL.v[0] * SL.v[0] + SL.v[1] * SL.v[1] + SL.v[2] * SL.v[2]);
        const double hR = SR.p / SR.rho + uR + 0.5 * (SR.v[0] * SR.v[0] + SR.v[1] * SR.v[1] + SR.v[2] * SR.v[2]);
        const double vL = SL.v[0] * nv[0] + SL.v[1] * nv[1] + SL.v[2] * nv[2];
        const double vR = SR.v[0] * nv[0] + SR.v[1] * nv[1] + SR.v[2] * nv[2];
        if (mode == 1) {
            riemann_kt_mfv(SL, SR, o, nv, vL, vR, cL, cR, hL, hR);
            m = 2;
        } else {
            m = -riemann_exact_mfv_body(SL, SR, o, G, nv, vL, vR, cL, cR);
        }
    }
    double *y = out + 7 * t;
    y[0] = o.P_M; y[1] = o.S_M; y[2] = o.Frho;
    y[3] = o.Fv[0]; y[4] = o.Fv[1]; y[5] = o.Fv[2]; y[6] = o.Fp;
    mask[t] = m;
}

inline unsigned nblocks_mfv(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

void launch_mfv_preloop(hipStream_t s, const ForceArgs &a, const MfvArgs &m)
{
    if (a.n_active > 0)
        hipLaunchKernelGGL(mfv_preloop_kernel, dim3(nblocks_mfv(a.n_active)), dim3(kBlock), 0, s, a.n_active, a.active,
                           m);
}

void launch_force_mfv(hipStream_t s, const ForceArgs &a, const MfvArgs &m)
{
    if (a.n_active <= 0) return;
    if (a.list.idx && a.defer.mask) {
        const int b = a.block;
        hipLaunchKernelGGL((force_mfv_kernel<true, true>), dim3((unsigned)((a.n_active + b - 1) / b)), dim3(b), 0, s, a,
                           m, (const int *)nullptr, (const int *)nullptr);
    } else if (a.list.idx) {
        hipLaunchKernelGGL((force_mfv_kernel<true, false>), dim3(nblocks_mfv(a.n_active)), dim3(kBlock), 0, s, a, m,
                           (const int *)nullptr, (const int *)nullptr);
    } else {
        hipLaunchKernelGGL((force_mfv_kernel<false, false>), dim3(nblocks_mfv(a.n_active)), dim3(kBlock), 0, s, a, m,
                           (const int *)nullptr, (const int *)nullptr);
    }
}

void launch_force_mfv_overflow(hipStream_t s, const ForceArgs &a, const MfvArgs &m, int *rows, int *n_rows)
{
    if (a.n_active <= 0) return;
    hipLaunchKernelGGL(collect_overflow_rows_kernel, dim3(nblocks_mfv(a.n_active)), dim3(kBlock), 0, s, a.n_active,
                       a.list.count, a.list.cap, rows, n_rows);
    // a fixed grid strides over the rows, whose count is on the device
    hipLaunchKernelGGL((force_mfv_kernel<false, false>), dim3(1024), dim3(kBlock), 0, s, a, m, (const int *)rows,
                       (const int *)n_rows);
}

void launch_force_mfv_deferred(hipStream_t s, const ForceArgs &a, const MfvArgs &m, int n_pairs)
{
    if (n_pairs > 0)
        hipLaunchKernelGGL(force_mfv_deferred_kernel, dim3(nblocks_mfv(n_pairs)), dim3(kBlock), 0, s, a, m);
}

void launch_force_mfv_finish(hipStream_t s, const ForceArgs &a, const MfvArgs &m)
{
    if (a.n_active > 0)
        hipLaunchKernelGGL(force_mfv_finish_kernel, dim3(nblocks_mfv(a.n_active)), dim3(kBlock), 0, s, a, m);
}

void launch_riemann_states_mfv(hipStream_t s, int mode, const GammaConsts &G, int64_t n, const double *L,
                               const double *R, const double *normal, const double *limiter, double *out, int *mask)
{
    if (n > 0)
        hipLaunchKernelGGL(riemann_states_mfv_kernel, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                           mode, G, n, L, R, normal, limiter, out, mask);
}

}  // namespace gzm
