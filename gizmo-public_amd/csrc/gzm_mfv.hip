// gzm_mfv.hip -- HIP kernels of the meshless finite-volume (MFV) variant of
// the hydro pass for gfx950 (FP64): the smoothed mesh velocity after the
// density iteration, and the force loop with mass fluxes.
//
// The density iteration, the gradients and the pair list are those of the
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

// ParticleVel of the active particles with mass (HYDRO_FIX_MESH_MOTION = 5):
// the density loop's sum of m_j W_ij v_j over {j : r_ij < h_i} (density.c:
// 291-293 for r = 0, 331-334 otherwise, as v_i - (v_i - v_j)) mixed with the
// particle's own velocity, 0.7 v_i + 0.3 sum / (sum of m_j W_ij)
// (density.c:886-894).  Runs after the last sweep, whose hsml and density
// sum it reads, and before the finalize overwrites that sum.
__global__ __launch_bounds__(kBlock) void mfv_particle_vel_kernel(const BoxDesc *boxes, const int *box_of,
                                                                  TreeArgs tree, int n_active, const int *active,
                                                                  const double4 *posm, const double4 *velu,
                                                                  const double *hsml, const double *rho_sum,
                                                                  double *pvel)
{
    const int t = xcd_block() * (int)blockDim.x + threadIdx.x;
    bool valid = t < n_active;
    const int i = valid ? active[t] : 0;
    if (valid && !(posm[i].w > 0)) valid = false;
    for_box(boxes, box_of, i, valid, [&](const BoxDesc &G) {
        const double4 pi = posm[i];
        const double4 vi = velu[i];
        const double h = hsml[i];
        double hinv, hinv3, hinv4;
        kernel_hinv(h, hinv, hinv3, hinv4);
        const double h2 = h * h;
        const double xs[3] = {pi.x, pi.y, pi.z};
        const double v_i[3] = {vi.x, vi.y, vi.z};
        double s[3] = {0, 0, 0};
        for_each_neighbour_candidate(tree, G, xs, h, false, [&](int j) {
            const double4 pj = posm[j];
            double dp[3] = {pi.x - pj.x, pi.y - pj.y, pi.z - pj.z};
            wrap_dp(G.box, dp);
            const double r2 = dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2];
            if (!(r2 < h2 && pj.w > 0)) return;
            const double r = sqrt(r2);
            double wk, dwk;
            kernel_wdw(r * hinv, hinv3, hinv4, wk, dwk);
            const double mj_wk = pj.w * wk;
            const double4 vj = velu[j];
            const double v_j[3] = {vj.x, vj.y, vj.z};
            if (r > 0) {
                for (int k = 0; k < 3; k++) s[k] += mj_wk * (v_i[k] - (v_i[k] - v_j[k]));
            } else {
                for (int k = 0; k < 3; k++) s[k] += mj_wk * v_j[k];
            }
        });
        const double d = rho_sum[i];
        const double eps_pvel = 0.3;
        for (int k = 0; k < 3; k++)
            pvel[3 * (size_t)i + k] = (d > 0) ? (1. - eps_pvel) * v_i[k] + eps_pvel * s[k] / d : s[k];
    });
}

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

// One active particle's MFV force loop (hydro_evaluate.h:146-498, mass
// exchange 379-401) and final operations.  USE_LIST: the pair list of the
// gradient kernel (a particle whose pair set overflowed it scans the grid or
// tree instead); otherwise every particle scans.
template <bool USE_LIST>
__global__ __launch_bounds__(kBlock) void force_mfv_kernel(ForceArgs a, MfvArgs m)
{
    const int t = xcd_block() * (int)blockDim.x + threadIdx.x;
    const bool valid = t < a.n_active;
    unsigned long long cnt = 0, nretry = 0, nexact = 0;
    for_box(a.boxes, a.box_of, valid ? a.active[t] : 0, valid, [&](const BoxDesc &G) {
        const int i = a.active[t];
        const ForceRec ri = a.rec[i];
        const double m_i = ri.mc.x, rho_i = ri.th.x;
        PairI<double> I;
        unpack_side(ri, I);
        I.hinv3 = I.hinv * I.hinv * I.hinv;
        I.hinv4 = I.hinv3 * I.hinv;
        I.cnumcrit2 = kCondDanger * kCondDanger - I.cn * I.cn;
        const double w_i[3] = {m.pvel[3 * (size_t)i + 0], m.pvel[3 * (size_t)i + 1], m.pvel[3 * (size_t)i + 2]};
        const double ti_i = m.ti ? m.ti[i] : 1.0;
        const double dt_i = ti_i * m.timebase;  // hydro_evaluate.h:82
        double acc[3] = {0, 0, 0}, dte = 0, dtm = 0, dm = 0, gw[3] = {0, 0, 0}, vmax = kMinReal;
        int status = 0;
        if ((m_i > 0) && (rho_i > 0)) {
            double out_vsig = I.cs;
            const double x_i[3] = {ri.xh.x, ri.xh.y, ri.xh.z};
            const double h2_i = I.h * I.h;
            auto visit = [&](int j, bool check) {
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
                const int e = force_pair_mfv(I, J, w_i, dp, r2, a.G, Fv, Fp, Frho, vsig, mask);
                cnt++;
                if (e > status) status = e;
                nretry += (mask & 8) ? 1 : 0;
                nexact += (mask & 4) ? 1 : 0;
                // mass exchange with the reference's 10 % limiter; j_is_active_for_fluxes
                // is always 0 there (hydro_evaluate.h:158-168)
                const double ti_j = m.ti ? m.ti[j] : 1.0;
                double dmh = Frho * dt_i, lim;
                if (dmh > 0) lim = a.rec[j].mc.x; else lim = m_i;
                lim *= 0.1;
                if (fabs(dmh) > lim) dmh *= lim / fabs(dmh);
                if (ti_i < ti_j) {
                    dm += dmh;
                    atomicAdd(&m.d_mass[j], -dmh);
                }
                if (ti_i == ti_j) {
                    dm += 0.5 * dmh;
                    atomicAdd(&m.d_mass[j], -(0.5 * dmh));
                }
                dtm += Frho;
                for (int k = 0; k < 3; k++) gw[k] += Frho * dp[k];
                for (int k = 0; k < 3; k++) acc[k] += Fv[k];
                dte += Fp;
                if (vsig > out_vsig) out_vsig = vsig;
                if (a.is_active && !a.is_active[j]) {
                    if (vsig > kWakeup * a.vsig_old[j]) a.wakeup[j] = 1;  // WAKEUP (hydro_evaluate.h:485-495)
                }
            };
            if (USE_LIST && a.list.count[t] <= a.list.cap) {
                const int n = a.list.count[t];
                const int *row = a.list.idx + (size_t)t * a.list.cap;
                for (int k = 0; k < n; k++) visit(row[k], false);
            } else {
                for_each_neighbour_candidate(a.tree, G, x_i, I.h, true, [&](int j) { visit(j, true); });
            }
            if (vmax < out_vsig) vmax = out_vsig;
        }
        if (m_i > 0) {
            const double u = a.velu[i].w;
            mfv_final_ops(acc, dte, vmax, dtm, I.v, w_i, u, m_i, I.h, a.max_hsml);
        }
        a.acc[3 * (size_t)i + 0] = acc[0];
        a.acc[3 * (size_t)i + 1] = acc[1];
        a.acc[3 * (size_t)i + 2] = acc[2];
        a.dtu[i] = dte;
        a.vsig[i] = vmax;
        m.dt_mass[i] = dtm;
        for (int k = 0; k < 3; k++) m.grav_work[3 * (size_t)i + k] = gw[k];
        atomicAdd(&m.d_mass[i], dm);
        if (status) atomicMax(&a.ctr->status, status);
    });
    wave_add_u64(&a.ctr->interactions, cnt);
    wave_add_u64(&a.ctr->retries, nretry);
    wave_add_u64(&a.ctr->exact_calls, nexact);
}

inline unsigned nblocks_mfv(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

void launch_mfv_particle_vel(hipStream_t s, const BoxDesc *boxes, const int *box_of, const TreeArgs &tree,
                             int n_active, const int *active, const double4 *posm, const double4 *velu,
                             const double *hsml, const double *rho_sum, double *pvel)
{
    if (n_active > 0)
        hipLaunchKernelGGL(mfv_particle_vel_kernel, dim3(nblocks_mfv(n_active)), dim3(kBlock), 0, s, boxes, box_of,
                           tree, n_active, active, posm, velu, hsml, rho_sum, pvel);
}

void launch_force_mfv(hipStream_t s, const ForceArgs &a, const MfvArgs &m)
{
    if (a.n_active <= 0) return;
    const dim3 grid(nblocks_mfv(a.n_active)), block(kBlock);
    hipLaunchKernelGGL(mfv_preloop_kernel, grid, block, 0, s, a.n_active, a.active, m);
    if (a.list.idx) hipLaunchKernelGGL(force_mfv_kernel<true>, grid, block, 0, s, a, m);
    else hipLaunchKernelGGL(force_mfv_kernel<false>, grid, block, 0, s, a, m);
}

}  // namespace gzm
