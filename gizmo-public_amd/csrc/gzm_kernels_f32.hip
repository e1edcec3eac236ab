// gzm_kernels_f32.hip -- FP32-bulk kernels of the MFM hydro pass (gfx950).
//
// Same loops as gzm_kernels.hip (one thread per active particle, gather
// form), with the per-pair arithmetic in FP32:
//   * geometry: FP32 offsets from FP64 cell origins plus exact integer cell
//     differences (the F32Geom layout in gzm_internal.h);
//   * neighbour predicates: decided on the FP32 r^2 unless it lies within a
//     relative band of the threshold, where the pair is re-tested with the FP64
//     expression of the FP64 kernels, so the neighbour sets are those of the
//     FP64 pass for the same hsml, bit for bit;
//   * kernel weights, MLS sums, face areas, reconstruction, HLLC in FP32;
//   * FP64 kept for the hsml iteration (hsml_update), the NV_T inversion, the
//     slope limiter, the exact Riemann solver, the entropic-switch comparison
//     of P/rho, and the accumulation of the conserved fluxes (dv/dt, du/dt).
// Like the FP64 kernels, this file is compiled without floating-point
// contraction: the pair-list and scanning force loops (and the exact FP64
// neighbour predicate, exact_r2) must round identically wherever the same
// expression is inlined, and fused multiply-adds measured no faster here.
#include "grid_device.h"
#include "gzm_internal.h"

namespace gzm {

namespace {

// Bound on |r2_fp32 - r2_fp64| relative to the threshold h^2.  The offsets
// carry 2^-24 s rounding, a cell shift dc * s_f carries 2 * 2^-24 |dc| s, and
// the component sums, squares and their sum add 2^-24 each; for the
// neighbourhoods walked here (|dc| s ~ r ~ h, s <= kFineCells h) that is
// below 2e-6 h^2, so 2e-5 leaves a factor of ten.  Pairs inside the band are
// decided in FP64.
constexpr double kBand = 2e-5;
constexpr float kBandF = 2e-5f;

// A pair whose smaller hsml is below 1/kFineCells of the cell side (a dense
// clump in a grid sized for the mean hsml; every pair of the tree search)
// takes its displacement from the FP64 positions: the offsets' absolute
// rounding, 2^-24 s, would otherwise exceed ~1e-6 of the separation.
constexpr float kFineCells = 16.0f;

__device__ __forceinline__ float smax_of(const BoxDesc &G)
{
    return fmaxf(G.s_f[0], fmaxf(G.s_f[1], G.s_f[2]));
}

struct Band {
    float lo, hi;
};

__device__ __forceinline__ Band band_of(double t2)
{
    return {(float)(t2 * (1 - kBand)), (float)(t2 * (1 + kBand))};
}

// FP32 decision of r2 < t2: 1 inside, 0 outside, -1 undecided
__device__ __forceinline__ int decide(float r2, Band b) { return (r2 < b.lo) ? 1 : ((r2 >= b.hi) ? 0 : -1); }

// squared distance exactly as the FP64 kernels compute it (the neighbour
// predicate must round as in the FP64 kernels: no fused multiply-add, also
// should this file be built with contraction)
__device__ __forceinline__ double exact_r2(const BoxInfo &box, const double4 &pi, const double4 &pj, double dp[3])
{
#pragma clang fp contract(off)
    dp[0] = pi.x - pj.x;
    dp[1] = pi.y - pj.y;
    dp[2] = pi.z - pj.z;
    wrap_dp(box, dp);
    return dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2];
}

// FP32 displacement along one axis of cell-anchored positions: the cell shift
// dc * s rounded on its own, then added to the offset difference, so that the
// grid scan (shift hoisted per cell) and the pair list (shift per pair) give
// the same bits.
__device__ __forceinline__ float cell_shift(int dc, float s)
{
#pragma clang fp contract(off)
    return (float)dc * s;
}

__device__ __forceinline__ void unpack_cell(uint32_t p, int c[3])
{
    c[0] = (int)(p & 1023u);
    c[1] = (int)((p >> 10) & 1023u);
    c[2] = (int)((p >> 20) & 1023u);
}

// nearest image of a cell-index difference (valid while the pair is closer
// than half the box minus two cells, which the callers guarantee)
__device__ __forceinline__ int wrap_cell(int d, int nc)
{
    const int half = nc >> 1;
    return (d > half) ? d - nc : ((d < -half) ? d + nc : d);
}

__global__ void pack_geom_f32_kernel(int n, const double4 *posm, const double4 *velu, const BoxDesc *boxes,
                                     const int *box_of, float4 *offm, float4 *velf, uint32_t *cellc)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double4 p = posm[k], v = velu[k];
    const Grid &g = boxes[box_of[k]].grid;
    const int cx = cell_coord(p.x, g.inv_s[0], g.nc[0]);
    const int cy = cell_coord(p.y, g.inv_s[1], g.nc[1]);
    const int cz = cell_coord(p.z, g.inv_s[2], g.nc[2]);
    offm[k] = make_float4((float)(p.x - cx * g.s[0]), (float)(p.y - cy * g.s[1]), (float)(p.z - cz * g.s[2]),
                          (float)p.w);
    velf[k] = make_float4((float)v.x, (float)v.y, (float)v.z, 0.0f);
    cellc[k] = (uint32_t)cx | ((uint32_t)cy << 10) | ((uint32_t)cz << 20);
}

// ---------------------------------------------------------------------------
// density (density.c:252-347 pair sums in FP32 + 505-850 iteration in FP64)
// ---------------------------------------------------------------------------
// FP32 density sums carry rounding of up to ~1e-6 relative.  The hsml step's
// discrete decisions (the redo test against the neighbour-number tolerance,
// the bracket test, the Newton / bisection branches, the sign of a clamped
// jump) may then go the other way than with FP64 sums, moving hsml by a whole
// step.  A particle whose next hsml jumps when its neighbour number is moved
// by +-kSumMargin (relative) is close to such a decision: its sums are redone
// in FP64 (density_heavy_kernel) instead.
constexpr double kSumMargin = 2e-5;

__device__ inline bool hsml_step_near(const HsmlParams &P, const HsmlStep &X, double h, double left, double right)
{
    double hs[2];
    int rs[2];
#pragma unroll
    for (int k = 0; k < 2; k++) {
        HsmlStep Y = X;
        Y.NumNgb = X.NumNgb * (k ? 1 + kSumMargin : 1 - kSumMargin);
        double hh = h, l = left, r = right, cn;
        rs[k] = hsml_step(P, Y, hh, l, r, cn);
        hs[k] = hh;
    }
    // away from a decision the step moves hsml by ~Dhsml/3 of the relative
    // change, 7e-6 here; a decision moves it by ~1e-2 or more
    return rs[0] != rs[1] || !(fabs(hs[1] - hs[0]) <= 1e-4 * h);
}

template <bool TREE, int MAXB>
__global__ __launch_bounds__(MAXB) void density_sweep_f32_kernel(DensitySweepArgs a)
{
    const int t = xcd_block() * (int)blockDim.x + threadIdx.x;
    const int n_live = a.n_live_dev ? *a.n_live_dev : a.n_live;
    const bool valid = t < n_live;
    if (!valid && a.n_live_dev && t < a.n_live) a.redo[t] = 0;  // captured pass: clear up to the capacity
    unsigned long long cnt = 0;
    for_box(a.boxes, a.box_of, valid ? a.live[t] : 0, valid, [&](const BoxDesc &G) {
        const int i = a.live[t];
        const double4 pi = a.posm[i];
        const float4 oi = a.f32.offm[i];
        const float4 vi = a.f32.velf[i];
        int ci[3];
        unpack_cell(a.f32.cellc[i], ci);
        double h = a.hsml[i];
        const double h2 = h * h;
        const Band bh = band_of(h2);
        const float hinv = (float)(1.0 / h);
        const float hinv3 = hinv * hinv * hinv, hinv4 = hinv3 * hinv;
        float s_ngb = 0, s_rho = 0, s_dh = 0, s_div = 0;
        float t00 = 0, t01 = 0, t02 = 0, t11 = 0, t12 = 0, t22 = 0, f0 = 0, f1 = 0, f2 = 0;
        const double xs[3] = {pi.x, pi.y, pi.z};
        const int budget = a.budget;
        int visited = 0;
        bool heavy = false;
        unsigned long long lane_cnt = 0;
        // a candidate within kEdgeBand of the kernel edge (FP32 r^2 carries up
        // to 2e-6 h^2 of its own, inside the band)
        const float edge_lo = (float)(h2 * (1 - kEdgeBand)), edge_hi = (float)(h2 * (1 + kEdgeBand));
        bool edge = false;
        auto visit = [&](int k, float sh0, float sh1, float sh2, bool exact_geom) {
            const float4 oj = a.f32.offm[k];
            float dp[3];
            float r2;
            int in;
            if (!exact_geom) {
                dp[0] = sh0 + (oi.x - oj.x);
                dp[1] = sh1 + (oi.y - oj.y);
                dp[2] = sh2 + (oi.z - oj.z);
                r2 = dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2];
                in = decide(r2, bh);
                if (in < 0) {
                    double d64[3];
                    in = exact_r2(G.box, pi, a.posm[k], d64) < h2;
                }
            } else {
                double d64[3];
                const double r2e = exact_r2(G.box, pi, a.posm[k], d64);
                in = r2e < h2;
                dp[0] = (float)d64[0]; dp[1] = (float)d64[1]; dp[2] = (float)d64[2];
                r2 = (float)r2e;
            }
            if ((oj.w > 0) && (r2 >= edge_lo) && (r2 < edge_hi)) edge = true;
            if (!in || !(oj.w > 0)) return;
            lane_cnt++;
            const float r = Real<float>::root(r2);
            const float u = r * hinv;
            float wk, dwk;
            kernel_wdw(u, hinv3, hinv4, wk, dwk);
            s_ngb += wk;
            s_rho += oj.w * wk;
            s_dh += -(3 * hinv * wk + u * dwk);
            if (r > 0) {
                t00 += wk * dp[0] * dp[0];
                t01 += wk * dp[0] * dp[1];
                t02 += wk * dp[0] * dp[2];
                t11 += wk * dp[1] * dp[1];
                t12 += wk * dp[1] * dp[2];
                t22 += wk * dp[2] * dp[2];
                f0 += wk * dp[0];
                f1 += wk * dp[1];
                f2 += wk * dp[2];
                const float4 vj = a.f32.velf[k];
                const float dv0 = vi.x - vj.x, dv1 = vi.y - vj.y, dv2 = vi.z - vj.z;
                s_div -= Real<float>::div(dwk * (dp[0] * dv0 + dp[1] * dv1 + dp[2] * dv2), r);
            }
        };
        if (TREE) {
            tree_walk_grouped(a.tree, G, xs, h, [&](int k) -> bool {
                if (budget && ++visited > budget) {
                    heavy = true;
                    return false;
                }
                visit(k, 0.0f, 0.0f, 0.0f, true);
                return true;
            });
        } else {
            const bool fine_i = (float)h * kFineCells < smax_of(G);
            bool full[3];
            // the budget is charged a cell at a time, outside the candidate loop
            for_each_cell(G.grid, xs, h, full, [&](int cx, int cy, int cz, int beg, int end) -> bool {
                visited += end - beg;
                if (budget && visited > budget) {
                    heavy = true;
                    return false;
                }
                const bool exact_geom = full[0] || full[1] || full[2] || fine_i;
                const float sh0 = cell_shift(ci[0] - cx, G.s_f[0]);
                const float sh1 = cell_shift(ci[1] - cy, G.s_f[1]);
                const float sh2 = cell_shift(ci[2] - cz, G.s_f[2]);
                for (int k = beg; k < end; ++k) visit(k, sh0, sh1, sh2, exact_geom);
                return true;
            });
        }
        if (heavy) {  // left to the FP64 wavefront-per-particle kernel
            a.heavy_list[atomicAdd(a.n_heavy, 1)] = t;
            return;
        }
        DensitySums S = {s_ngb, s_rho, s_dh, s_div, t00, t01, t02, t11, t12, t22, f0, f1, f2};
        double left = a.left[i], right = a.right[i];
        DensityResult R;
        HsmlStep X;
        hsml_prepare(a.hp, S, h, a.cn[i], R, X);
        if (hsml_step_near(a.hp, X, h, left, right)) {  // decided in FP64 by density_heavy_kernel
            a.heavy_list[atomicAdd(a.n_heavy, 1)] = t;
            return;
        }
        cnt += lane_cnt;
        const int res = hsml_step(a.hp, X, h, left, right, R.cn_stored);
        if (res == 0 && edge && a.edge_list) a.edge_list[atomicAdd(a.n_edge, 1)] = i;
        a.hsml[i] = h;
        a.left[i] = left;
        a.right[i] = right;
        a.cn[i] = R.cn_stored;
        a.num_ngb[i] = R.num_ngb;
        a.dhsml[i] = R.dhsml;
        a.divvel[i] = R.divvel;
        a.rho[i] = R.rho_sum;
        a.fce[i] = R.fce;
#pragma unroll
        for (int k = 0; k < 6; k++) a.nvt[6 * (size_t)i + k] = R.nvt[k];
        a.redo[t] = (res == 1) ? 1 : 0;
        if (res < 0) atomicMax(&a.ctr->status, (int)6);
    });
    wave_add_u64(&a.ctr->interactions, cnt);
}

// ---------------------------------------------------------------------------
// gradients (gradients.c:1477-2089 pair loop in FP32, 1052-1253 finalize in FP64)
// ---------------------------------------------------------------------------
// x = hi + lo with hi = (float)x, lo = (float)(x - hi)
__device__ __forceinline__ float low_word(double x, float hi) { return (float)(x - (double)hi); }

__global__ void pack_grad_f32_kernel(int n, const float4 *offm, const float4 *velf, const double *hsml,
                                     const double *rho, const double *pres, GradRecF *rec)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float4 o = offm[k], v = velf[k];
    const double d = rho[k], P = pres[k];
    const float df = (float)d, Pf = (float)P;
    GradRecF r;
    r.xh = make_float4(o.x, o.y, o.z, (float)hsml[k]);
    r.vr = make_float4(v.x, v.y, v.z, df);
    r.pm = make_float4(Pf, o.w, low_word(d, df), low_word(P, Pf));
    rec[k] = r;
}

// One particle's FP32 force record (shared by pack_force_f32_kernel and the
// gradient kernel's epilogue, so that both write the same bits)
__device__ __forceinline__ ForceRecF make_force_rec_f(const float4 &o, const float4 &v, uint32_t cellc, double m,
                                                      double d, double P, double h, double dh, double cn, double fce,
                                                      const double *t, const double *g, double gamma)
{
    ForceRecF r;
    r.xh = make_float4(o.x, o.y, o.z, (float)h);
    r.vc = make_float4(v.x, v.y, v.z, (float)sqrt(gamma * P / d));  // eos.c:245-254
    r.th = make_float4((float)d, (float)P, (float)(m / d), (float)dh);
    r.t0 = make_float4((float)t[0], (float)t[1], (float)t[2], (float)t[3]);
    r.t1 = make_float4((float)t[4], (float)t[5], (float)(1.0 / h), (float)fce);
    r.g0 = make_float4((float)g[0], (float)g[1], (float)g[2], (float)g[3]);
    r.g1 = make_float4((float)g[4], (float)g[5], (float)g[6], (float)g[7]);
    r.g2 = make_float4((float)g[8], (float)g[9], (float)g[10], (float)g[11]);
    r.g3 = make_float4((float)g[12], (float)g[13], (float)g[14], __uint_as_float(cellc));
    r.cn = cn;
    r.e = P / d;
    return r;
}

constexpr int kNoSlotF = 0x7fffffff;  // above every sorted slot: the end of a merged row

template <bool TREE, int BS>
__global__ __launch_bounds__(BS, 4) void gradients_f32_kernel(GradientArgs a)
{
    __shared__ int stage[kListChunk * BS];  // as in gradients_kernel; BS = blockDim.x
    constexpr int bs = BS;
    const int t = xcd_block() * bs + threadIdx.x;
    const bool valid = t < a.n_active;
    unsigned long long cnt = 0;
    int listed = 0;
    for_box(a.boxes, a.box_of, valid ? a.active[t] : 0, valid, [&](const BoxDesc &G) {
        const int i = a.active[t];
        const double4 pi = a.posm[i];
        const GradRecF ri = a.recf[i];
        int ci[3];
        unpack_cell(a.f32.cellc[i], ci);
        const double h_i = a.hsml[i], m_i = pi.w;
        const double rho_i = a.rho[i], p_i = a.pres[i];
        float gr0 = 0, gr1 = 0, gr2 = 0, gp0 = 0, gp1 = 0, gp2 = 0;
        float gv00 = 0, gv01 = 0, gv02 = 0, gv10 = 0, gv11 = 0, gv12 = 0, gv20 = 0, gv21 = 0, gv22 = 0;
        float maxd = 0, mn_r = 0, mx_r = 0, mn_p = 0, mx_p = 0;
        float mn_v0 = 0, mx_v0 = 0, mn_v1 = 0, mx_v1 = 0, mn_v2 = 0, mx_v2 = 0;
        int *list_row = a.list.idx ? a.list.idx + (size_t)t * a.list.cap : nullptr;
        if ((m_i > 0) && (rho_i > 0) && (h_i > 0)) {  // GasGrad_isactive (gradients.c:66-75)
            const float hinv = (float)(1.0 / h_i);
            const float hinv3 = hinv * hinv * hinv;
            const double h2_i = h_i * h_i;
            const Band bi = band_of(h2_i);
            const float pos_min = (float)(h2_i * kBand);  // r2 below this: re-test r2 > 0 in FP64
            const float rho_f = ri.vr.w, p_f = ri.pm.x, rho_lo = ri.pm.z, p_lo = ri.pm.w;
            const double xs[3] = {pi.x, pi.y, pi.z};
            const double R = (h_i > G.hmax) ? h_i : G.hmax;
            const float h_fine = smax_of(G) / kFineCells;  // pairs with a smaller hsml: FP64 geometry
            const bool fine_i = ri.xh.w < h_fine;
            auto visit = [&](int k, float sh0, float sh1, float sh2, bool exact_geom) {
                if (k == i) return;  // r2 = 0 exactly
                const float4 xj = a.recf[k].xh;
                const float hj2 = xj.w * xj.w;
                float dp[3];
                float r2;
                int in_i, in_j, pos;
                if (!exact_geom && !(xj.w < h_fine)) {
                    dp[0] = sh0 + (ri.xh.x - xj.x);
                    dp[1] = sh1 + (ri.xh.y - xj.y);
                    dp[2] = sh2 + (ri.xh.z - xj.z);
                    r2 = dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2];
                    if ((r2 >= bi.hi) && (r2 >= hj2 * (1 + kBandF))) return;  // clearly outside both
                    in_i = decide(r2, bi);
                    in_j = decide(r2, Band{hj2 * (1 - kBandF), hj2 * (1 + kBandF)});
                    pos = (r2 >= pos_min) ? 1 : -1;
                    if ((in_i < 0) || (in_j < 0) || (pos < 0)) {
                        double d64[3];
                        const double r2e = exact_r2(G.box, pi, a.posm[k], d64);
                        const double h_j = a.hsml[k];
                        in_i = r2e < h2_i;
                        in_j = r2e < h_j * h_j;
                        pos = r2e > 0;
                    }
                } else {
                    double d64[3];
                    const double r2e = exact_r2(G.box, pi, a.posm[k], d64);
                    const double h_j = a.hsml[k];
                    in_i = r2e < h2_i;
                    in_j = r2e < h_j * h_j;
                    pos = r2e > 0;
                    dp[0] = (float)d64[0]; dp[1] = (float)d64[1]; dp[2] = (float)d64[2];
                    r2 = (float)r2e;
                }
                if (!pos || (!in_i && !in_j)) return;
                const float4 vr = a.recf[k].vr;
                const float4 pm = a.recf[k].pm;
                if (!((pm.y > 0) && (vr.w > 0))) return;
                // the list holds the force loop's pair set (m_j > 0, rho_j > 0,
                // hydro_evaluate.h:149-150); gradients also need h_j > 0
                if (list_row) {
                    if (listed < a.list.cap) {
                        const int c = listed & (kListChunk - 1);
                        stage[c * bs + threadIdx.x] = k;
                        if (c == kListChunk - 1) flush_list_chunk(stage, list_row + (listed - c), bs);
                    }
                    listed++;
                }
                if (!(xj.w > 0)) return;
                cnt++;
                const float r = Real<float>::root(r2);
                if (r > maxd) maxd = r;
                // differences of the split values (GradRecF)
                const float d_rho = (vr.w - rho_f) + (pm.z - rho_lo);
                minmax_check(d_rho, mn_r, mx_r);
                const float dpp = (pm.x - p_f) + (pm.w - p_lo);
                minmax_check(dpp, mn_p, mx_p);
                const float dv0 = vr.x - ri.vr.x, dv1 = vr.y - ri.vr.y, dv2 = vr.z - ri.vr.z;
                minmax_check(dv0, mn_v0, mx_v0);
                minmax_check(dv1, mn_v1, mx_v1);
                minmax_check(dv2, mn_v2, mx_v2);
                if (in_i) {  // r < h_i (the weight vanishes at the boundary)
                    const float wk_i = kernel_w(r * hinv, hinv3);
                    const float w0 = -wk_i * dp[0], w1 = -wk_i * dp[1], w2 = -wk_i * dp[2];
                    gr0 += w0 * d_rho; gr1 += w1 * d_rho; gr2 += w2 * d_rho;
                    gp0 += w0 * dpp; gp1 += w1 * dpp; gp2 += w2 * dpp;
                    gv00 += w0 * dv0; gv01 += w1 * dv0; gv02 += w2 * dv0;
                    gv10 += w0 * dv1; gv11 += w1 * dv1; gv12 += w2 * dv1;
                    gv20 += w0 * dv2; gv21 += w1 * dv2; gv22 += w2 * dv2;
                }
            };
            if (TREE) {
                // symmetric pair rows when the stage built them (as in gradients_kernel)
                const int n_a = a.brow ? a.dcache_n[i] : 0;
                const int n_b = a.brow ? a.bcount[i] : 0;
                if (a.brow && n_a <= kDensityCacheCap && h2_i <= a.dcache_r2[i] && n_b <= kPairRowCap) {
                    const int *ra = a.dcache_idx + (size_t)i * kDensityCacheCap;
                    const int *rb = a.brow + (size_t)i * kPairRowCap;
                    int ka = 0, kb = 0;
                    int ja = (n_a > 0) ? ra[0] : kNoSlotF, jb = (n_b > 0) ? rb[0] : kNoSlotF;
                    while (ka < n_a || kb < n_b) {
                        if (ja < jb) {
                            visit(ja, 0.0f, 0.0f, 0.0f, true);
                            ka++;
                            ja = (ka < n_a) ? ra[ka] : kNoSlotF;
                        } else {
                            visit(jb, 0.0f, 0.0f, 0.0f, true);
                            kb++;
                            jb = (kb < n_b) ? rb[kb] : kNoSlotF;
                        }
                    }
                } else {
                    tree_walk(a.tree, G, xs, h_i, true, [&](int k) { visit(k, 0.0f, 0.0f, 0.0f, true); });
                }
            } else {
                bool full[3];
                for_each_cell(G.grid, xs, R, full, [&](int cx, int cy, int cz, int beg, int end) {
                    const bool exact_geom = full[0] || full[1] || full[2] || fine_i;
                    const float sh0 = cell_shift(ci[0] - cx, G.s_f[0]);
                    const float sh1 = cell_shift(ci[1] - cy, G.s_f[1]);
                    const float sh2 = cell_shift(ci[2] - cz, G.s_f[2]);
                    for (int k = beg; k < end; ++k) visit(k, sh0, sh1, sh2, exact_geom);
                }, h_i);
            }
            if (list_row && listed < a.list.cap && (listed & (kListChunk - 1)))
                flush_list_chunk(stage, list_row + (listed & ~(kListChunk - 1)), bs);
        } else if ((m_i > 0) && (rho_i > 0)) {
            listed = a.list.cap + 1;  // the force loop still runs for i: send it to the scan
        }
        double nvt[6];
#pragma unroll
        for (int k = 0; k < 6; k++) nvt[k] = a.nvt[6 * (size_t)i + k];
        double gr[3] = {gr0, gr1, gr2}, gp[3] = {gp0, gp1, gp2};
        double gv0[3] = {gv00, gv01, gv02}, gv1[3] = {gv10, gv11, gv12}, gv2[3] = {gv20, gv21, gv22};
        apply_nvt(nvt, gr);
        apply_nvt(nvt, gp);
        apply_nvt(nvt, gv0);
        apply_nvt(nvt, gv1);
        apply_nvt(nvt, gv2);
        // SELFGRAVITY_OFF, no MAGNETIC/GALSF: h_lim = h, stol = 0.1 (gradients.c:1219-1232)
        const double cn = a.cn[i];
        const double d_max = (h_i > (double)maxd) ? h_i : (double)maxd;
        double a_lim = 0.25;
        if (cn > 100) a_lim = dmin(0.5, 0.25 + 0.25 * (cn - 100) / 100);
        slope_limit(gr, mx_r, mn_r, a_lim, h_i, 0, 1, d_max, rho_i);
        slope_limit(gp, mx_p, mn_p, a_lim, h_i, 0.1, 1, d_max, p_i);
        slope_limit(gv0, mx_v0, mn_v0, a_lim, h_i, 0.1, 0, 0, 0);
        slope_limit(gv1, mx_v1, mn_v1, a_lim, h_i, 0.1, 0, 0, 0);
        slope_limit(gv2, mx_v2, mn_v2, a_lim, h_i, 0.1, 0, 0, 0);
        double *o = a.grad + 15 * (size_t)i;
        o[0] = gr[0]; o[1] = gr[1]; o[2] = gr[2];
        o[3] = gp[0]; o[4] = gp[1]; o[5] = gp[2];
        o[6] = gv0[0]; o[7] = gv0[1]; o[8] = gv0[2];
        o[9] = gv1[0]; o[10] = gv1[1]; o[11] = gv1[2];
        o[12] = gv2[0]; o[13] = gv2[1]; o[14] = gv2[2];
        if (a.frecf_out) {  // every particle active: its force record, as pack_force_f32_kernel writes it
            const double g[15] = {gr[0], gr[1], gr[2], gp[0], gp[1], gp[2], gv0[0], gv0[1], gv0[2],
                                  gv1[0], gv1[1], gv1[2], gv2[0], gv2[1], gv2[2]};
            a.frecf_out[i] = make_force_rec_f(make_float4(ri.xh.x, ri.xh.y, ri.xh.z, 0.0f), ri.vr, a.f32.cellc[i], m_i,
                                              rho_i, p_i, h_i, a.dhsml[i], cn, a.fce[i], nvt, g, a.gamma);
        }
        if (a.list.idx) a.list.count[t] = listed;
    });
    if (a.list.idx) {
        const bool over = __any(listed > a.list.cap);
        if ((threadIdx.x & 63) == 0 && t < a.n_active) {
            a.list.overflow[t >> 6] = over ? 1 : 0;
            if (over) atomicAdd(&a.ctr->overflow_slices, 1);
        }
    }
    wave_add_u64(&a.ctr->interactions, cnt);
}

// ---------------------------------------------------------------------------
// force (hydro_evaluate.h:22-516 with the pair math in FP32, FP64 flux sums)
// ---------------------------------------------------------------------------
__global__ void pack_force_f32_kernel(int n, const float4 *offm, const float4 *velf, const uint32_t *cellc,
                                      const double4 *posm, const double *hsml, const double *rho,
                                      const double *pres, const double *dhsml, const double *cn, const double *fce,
                                      const double *nvt, const double *grad, double gamma, ForceRecF *rec)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    rec[k] = make_force_rec_f(offm[k], velf[k], cellc[k], posm[k].w, rho[k], pres[k], hsml[k], dhsml[k], cn[k], fce[k],
                              nvt + 6 * (size_t)k, grad + 15 * (size_t)k, gamma);
}

__device__ __forceinline__ void unpack_side_f(const ForceRecF &r, PairSide<float> &S)
{
    S.v[0] = r.vc.x; S.v[1] = r.vc.y; S.v[2] = r.vc.z; S.cs = r.vc.w;
    S.h = r.xh.w; S.hinv = r.t1.z;
    S.rho = r.th.x; S.P = r.th.y; S.V = r.th.z; S.dhsml = r.th.w;
    S.cn = r.cn; S.fce = r.t1.w; S.e = r.e;
    S.nvt[0] = r.t0.x; S.nvt[1] = r.t0.y; S.nvt[2] = r.t0.z; S.nvt[3] = r.t0.w;
    S.nvt[4] = r.t1.x; S.nvt[5] = r.t1.y;
    S.gr[0] = r.g0.x; S.gr[1] = r.g0.y; S.gr[2] = r.g0.z;
    S.gp[0] = r.g0.w; S.gp[1] = r.g1.x; S.gp[2] = r.g1.y;
    S.gv[0] = r.g1.z; S.gv[1] = r.g1.w; S.gv[2] = r.g2.x;
    S.gv[3] = r.g2.y; S.gv[4] = r.g2.z; S.gv[5] = r.g2.w;
    S.gv[6] = r.g3.x; S.gv[7] = r.g3.y; S.gv[8] = r.g3.z;
}

// the FP32 list kernel takes pair displacements from cell differences unless
// the search radius spans half the box on some axis
__device__ __forceinline__ bool list_needs_exact_geom(const BoxDesc &G)
{
    bool e = false;
#pragma unroll
    for (int ax = 0; ax < 3; ax++)
        if (!(G.hmax * G.grid.inv_s[ax] < 0.5 * G.grid.nc[ax] - 2.0)) e = true;
    return e;
}

// displacement x_i - x_j of a listed pair in FP32 (the scan's choice of
// geometry: cell offsets unless the walk spans half the box, the search is the
// tree, or the smaller hsml is below the cell side / kFineCells)
__device__ __forceinline__ void list_dp(const BoxDesc &G, bool exact_geom, const ForceRecF &ri, const int ci[3],
                                        const double4 &pi, const ForceRecF &Jr, const double4 *posm, int j,
                                        float dp[3])
{
    const float4 xj = Jr.xh;
    if (!exact_geom && !(fminf(ri.xh.w, xj.w) * kFineCells < smax_of(G))) {
        int cj[3];
        unpack_cell(__float_as_uint(Jr.g3.w), cj);
        dp[0] = cell_shift(wrap_cell(ci[0] - cj[0], G.grid.nc[0]), G.s_f[0]) + (ri.xh.x - xj.x);
        dp[1] = cell_shift(wrap_cell(ci[1] - cj[1], G.grid.nc[1]), G.s_f[1]) + (ri.xh.y - xj.y);
        dp[2] = cell_shift(wrap_cell(ci[2] - cj[2], G.grid.nc[2]), G.s_f[2]) + (ri.xh.z - xj.z);
    } else {
        double d64[3];
        exact_r2(G.box, pi, posm[j], d64);
        dp[0] = (float)d64[0]; dp[1] = (float)d64[1]; dp[2] = (float)d64[2];
    }
}

// FP64 pair side of particle k from the FP64 state, with the expressions of
// the FP64 force records (pack_force_kernel: cs, V, e, 1/h)
__device__ __forceinline__ void side_fp64(const ForceArgs &a, int k, PairSide<double> &S)
{
    const double4 p = a.posm[k], v = a.velu[k];
    const double d = a.rho[k], P = a.pres[k], h = a.hsml[k];
    S.v[0] = v.x; S.v[1] = v.y; S.v[2] = v.z;
    S.cs = sqrt(a.G.G0 * P / d);  // eos.c:245-254
    S.h = h; S.hinv = 1.0 / h;
    S.rho = d; S.P = P; S.V = p.w / d; S.dhsml = a.dhsml[k];
    S.cn = a.cn[k]; S.fce = a.fce[k]; S.e = P / d;
    const double *t = a.nvt + 6 * (size_t)k;
#pragma unroll
    for (int q = 0; q < 6; q++) S.nvt[q] = t[q];
    const double *g = a.grad + 15 * (size_t)k;
#pragma unroll
    for (int q = 0; q < 3; q++) { S.gr[q] = g[q]; S.gp[q] = g[3 + q]; }
#pragma unroll
    for (int q = 0; q < 9; q++) S.gv[q] = g[6 + q];
}

// The pair (i, j) of particle i's loop returned kNearThreshold: list i for the
// FP64 force redo and i and j for the FP64 gradient redo, each once.  An
// inactive j keeps its stored gradient (an output of an earlier pass, which
// the FP64 pass reads as it is), so only an active j is listed.
__device__ __forceinline__ void mark_redo(const ForceArgs &a, int i, int j)
{
    const int fi = atomicOr(&a.redo_flag[i], 3);
    if (!(fi & 1)) a.redo_list[atomicAdd(&a.redo_count[0], 1)] = i;
    if (!(fi & 2)) a.redo_list[a.n + atomicAdd(&a.redo_count[1], 1)] = i;
    if (a.is_active && !a.is_active[j]) return;
    const int fj = atomicOr(&a.redo_flag[j], 2);
    if (!(fj & 2)) a.redo_list[a.n + atomicAdd(&a.redo_count[1], 1)] = j;
}

__device__ __forceinline__ void side_i_f(const ForceRecF &ri, PairI<float> &I)
{
    unpack_side_f(ri, I);
    I.hinv3 = I.hinv * I.hinv * I.hinv;
    I.hinv4 = I.hinv3 * I.hinv;
    I.cnumcrit2 = kCondDanger * kCondDanger - I.cn * I.cn;
}

// USE_LIST: walk the pair list of the FP32 gradient kernel (particles that
// overflowed it are left to force_wave_fp64_kernel); otherwise (no list) scan
// the grid or tree with the same banded predicate as the gradient kernel.  Exact-solver pairs are summed
// after the others, by force_deferred_f32_kernel + force_finish_kernel
// (THREE_PHASE) or by a second walk over the pairs, as in the FP64 kernel.
template <bool USE_LIST, bool THREE_PHASE, bool TREE = false>
__global__ __launch_bounds__((USE_LIST && THREE_PHASE) ? kMaxBlock : kBlock) void force_f32_kernel(ForceArgs a)
{
    const int t = xcd_block() * (int)blockDim.x + threadIdx.x;
    bool valid = t < a.n_active;
    if (USE_LIST) {
        if (valid && a.list.count[t] > a.list.cap) {
            valid = false;
            if (THREE_PHASE) {  // evaluated by the wavefront loop for long pair sets
                a.defer.count[t] = 0;
                a.defer.raw[t] = 0;
            }
        }
    }
    unsigned long long cnt = 0, nretry = 0, nexact = 0;
    for_box(a.boxes, a.box_of, valid ? a.active[t] : 0, valid, [&](const BoxDesc &G) {
        const int i = a.active[t];
        const ForceRecF ri = a.recf[i];
        const double4 pi = a.posm[i];
        double acc[3] = {0, 0, 0}, dte = 0, vmax = kMinReal;
        const double m_i = pi.w;
        int status = 0;
        uint64_t dmask = 0;
        if ((m_i > 0) && (ri.th.x > 0)) {
            PairI<float> I;
            side_i_f(ri, I);
            int ci[3];
            unpack_cell(__float_as_uint(ri.g3.w), ci);
            double out_vsig = I.cs;
            ROut<float> ro = {0, 0, {0, 0, 0}, 0};
            int n_deferred = 0;
            auto add = [&](int j, const float Fv[3], float Fp, float vsig, int mask, int e) {
                if (e > status) status = e;
                nretry += (mask & 8) ? 1 : 0;
                nexact += (mask & 4) ? 1 : 0;
                acc[0] += (double)Fv[0];
                acc[1] += (double)Fv[1];
                acc[2] += (double)Fv[2];
                dte += (double)Fp;
                if ((double)vsig > out_vsig) out_vsig = vsig;
                if (a.is_active && !a.is_active[j]) {
                    if ((double)vsig > kWakeup * a.vsig_old[j]) a.wakeup[j] = 1;  // WAKEUP (hydro_evaluate.h:485-495)
                }
            };
            // pass 0: every pair but the deferred ones; pass 1: only those
            auto eval = [&](int j, int k, const PairSide<float> &J, const float dp[3], float r2, int pass) {
                float Fv[3], Fp, vsig;
                int mask;
                const int e = force_pair<float, true>(I, J, dp, r2, a.G, ro, Fv, Fp, vsig, mask);
                if (pass == 0) {
                    cnt++;
                    if (e == kDeferred) {
                        if (THREE_PHASE) dmask |= 1ull << k;
                        n_deferred++;
                        return;
                    }
                    add(j, Fv, Fp, vsig, mask, e);
                } else if (!THREE_PHASE && e == kDeferred) {
                    const int e2 = force_pair<float, false>(I, J, dp, r2, a.G, ro, Fv, Fp, vsig, mask);
                    if (e2 == kNearThreshold) mark_redo(a, i, j);
                    else add(j, Fv, Fp, vsig, mask, e2);
                }
            };
            const double h_i = a.hsml[i];
            const double h2_i = h_i * h_i;
            const Band bi = band_of(h2_i);
            const float pos_min = (float)(h2_i * kBand);
            const double xs[3] = {pi.x, pi.y, pi.z};
            const double R = (h_i > G.hmax) ? h_i : G.hmax;
            const bool exact_geom_list = a.tree.nodes != nullptr || list_needs_exact_geom(G);
            const float smax = smax_of(G);
            for (int pass = 0; pass < 2; pass++) {
                if (pass == 1 && (THREE_PHASE || n_deferred == 0)) break;
                if (USE_LIST) {
                    const int n = a.list.count[t];
                    const int *row = a.list.idx + (size_t)t * a.list.cap;
                    for (int k = 0; k < n; k++) {
                        const int j = row[k];
                        const ForceRecF &Jr = a.recf[j];
                        PairSide<float> J;
                        unpack_side_f(Jr, J);
                        float dp[3];
                        list_dp(G, exact_geom_list, ri, ci, pi, Jr, a.posm, j, dp);
                        const float r2 = dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2];
                        eval(j, k, J, dp, r2, pass);
                    }
                } else {
                    auto visit = [&](int j, float sh0, float sh1, float sh2, bool exact_geom) {
                        if (j == i) return;
                        const float4 xj = a.recf[j].xh;
                        const float hj2 = xj.w * xj.w;
                        float dp[3];
                        float r2;
                        int in_i, in_j, pos;
                        if (!exact_geom && !(fminf(ri.xh.w, xj.w) * kFineCells < smax)) {
                            dp[0] = sh0 + (ri.xh.x - xj.x);
                            dp[1] = sh1 + (ri.xh.y - xj.y);
                            dp[2] = sh2 + (ri.xh.z - xj.z);
                            r2 = dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2];
                            if ((r2 >= bi.hi) && (r2 >= hj2 * (1 + kBandF))) return;
                            in_i = decide(r2, bi);
                            in_j = decide(r2, Band{hj2 * (1 - kBandF), hj2 * (1 + kBandF)});
                            pos = (r2 >= pos_min) ? 1 : -1;
                            if ((in_i < 0) || (in_j < 0) || (pos < 0)) {
                                double d64[3];
                                const double r2e = exact_r2(G.box, pi, a.posm[j], d64);
                                const double h_j = a.hsml[j];
                                in_i = r2e < h2_i;
                                in_j = r2e < h_j * h_j;
                                pos = r2e > 0;
                            }
                        } else {
                            double d64[3];
                            const double r2e = exact_r2(G.box, pi, a.posm[j], d64);
                            const double h_j = a.hsml[j];
                            in_i = r2e < h2_i;
                            in_j = r2e < h_j * h_j;
                            pos = r2e > 0;
                            dp[0] = (float)d64[0]; dp[1] = (float)d64[1]; dp[2] = (float)d64[2];
                            r2 = (float)r2e;
                        }
                        if (!pos || (!in_i && !in_j)) return;
                        const ForceRecF &Jr = a.recf[j];
                        if (!(a.posm[j].w > 0) || !(Jr.th.x > 0)) return;  // hydro_evaluate.h:149-150
                        PairSide<float> J;
                        unpack_side_f(Jr, J);
                        eval(j, 0, J, dp, r2, pass);
                    };
                    if (TREE) {
                        tree_walk(a.tree, G, xs, h_i, true, [&](int j) { visit(j, 0.0f, 0.0f, 0.0f, true); });
                    } else {
                        bool full[3];
                        for_each_cell(G.grid, xs, R, full, [&](int cx, int cy, int cz, int beg, int end) {
                            const bool exact_geom = full[0] || full[1] || full[2];
                            const float sh0 = cell_shift(ci[0] - cx, G.s_f[0]);
                            const float sh1 = cell_shift(ci[1] - cy, G.s_f[1]);
                            const float sh2 = cell_shift(ci[2] - cz, G.s_f[2]);
                            for (int j = beg; j < end; ++j) visit(j, sh0, sh1, sh2, exact_geom);
                        }, h_i);
                    }
                }
            }
            if (THREE_PHASE) {
                // raw sums: finished by force_finish_kernel after the deferred pairs
                a.acc[3 * (size_t)i + 0] = acc[0];
                a.acc[3 * (size_t)i + 1] = acc[1];
                a.acc[3 * (size_t)i + 2] = acc[2];
                a.dtu[i] = dte;
                a.vsig[i] = out_vsig;
                a.defer.mask[t] = dmask;
                a.defer.count[t] = n_deferred;
                a.defer.raw[t] = 1;
                if (status) atomicMax(&a.ctr->status, status);
                return;
            }
            if (vmax < out_vsig) vmax = out_vsig;
            const double4 v = a.velu[i];
            const double vi[3] = {v.x, v.y, v.z};
            force_final_ops(acc, dte, vi, m_i, h_i, a.max_hsml);
        } else if (m_i > 0) {
            dte /= m_i;
        }
        if (THREE_PHASE) {
            a.defer.count[t] = 0;
            a.defer.raw[t] = 0;
        }
        a.acc[3 * (size_t)i + 0] = acc[0];
        a.acc[3 * (size_t)i + 1] = acc[1];
        a.acc[3 * (size_t)i + 2] = acc[2];
        a.dtu[i] = dte;
        a.vsig[i] = vmax;
        if (status) atomicMax(&a.ctr->status, status);
    });
    wave_add_u64(&a.ctr->interactions, cnt);
    wave_add_u64(&a.ctr->retries, nretry);
    wave_add_u64(&a.ctr->exact_calls, nexact);
}

// Deferred pairs of the FP32 list kernel, one per lane: exact-solver pairs
// (FP32 with the FP64 exact solver) and near-threshold pairs, whose particle
// this lists for the FP64 redo (its result slot stays zero).  Bounded to 4
// waves per SIMD: 153 VGPRs become 128 with 14 spilled, and the extra wave
// hides more of the divergent solver's latency (2 % faster, same bits).
__global__ __launch_bounds__(kBlock, 4) void force_deferred_f32_kernel(ForceArgs a)
{
    const int q = a.defer.q_begin + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int q_end = min(*a.defer.total, a.defer.q_end);
    const bool valid = q < q_end;
    const int2 ij = valid ? a.defer.pair[q - a.defer.q_begin] : make_int2(0, 0);
    unsigned long long nretry = 0, nexact = 0;
    for_box(a.boxes, a.box_of, ij.x, valid, [&](const BoxDesc &G) {
        const int i = ij.x, j = ij.y;
        const ForceRecF ri = a.recf[i];
        PairI<float> I;
        side_i_f(ri, I);
        int ci[3];
        unpack_cell(__float_as_uint(ri.g3.w), ci);
        const ForceRecF &Jr = a.recf[j];
        PairSide<float> J;
        unpack_side_f(Jr, J);
        float dp[3];
        list_dp(G, a.tree.nodes != nullptr || list_needs_exact_geom(G), ri, ci, a.posm[i], Jr, a.posm, j, dp);
        const float r2 = dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2];
        ROut<float> ro = {0, 0, {0, 0, 0}, 0};
        float Fv[3], Fp, vsig;
        int mask;
        int e = force_pair<float, false, true>(I, J, dp, r2, a.G, ro, Fv, Fp, vsig, mask);
        double *o = a.defer.result + 5 * (size_t)(q - a.defer.q_begin);
        if (e == kNearThreshold) {
            mark_redo(a, i, j);
            o[0] = 0; o[1] = 0; o[2] = 0; o[3] = 0; o[4] = 0;
            e = 0;
        } else {
            o[0] = Fv[0]; o[1] = Fv[1]; o[2] = Fv[2]; o[3] = Fp; o[4] = vsig;
            if (a.is_active && !a.is_active[j]) {
                if ((double)vsig > kWakeup * a.vsig_old[j]) a.wakeup[j] = 1;
            }
            nretry += (mask & 8) ? 1 : 0;
            nexact += (mask & 4) ? 1 : 0;
        }
        if (e) atomicMax(&a.ctr->status, e);
    });
    wave_add_u64(&a.ctr->retries, nretry);
    wave_add_u64(&a.ctr->exact_calls, nexact);
}

// The gradients of each particle listed by mark_redo, again, in FP64 from the
// FP64 state with the arithmetic of gradients_kernel (gradients.c:66-75,
// 1219-1232), written over the FP32 gradients.  One wavefront per listed
// particle: the lanes share its candidates (wave_pair_candidates) and reduce
// their partial sums, maxima and minima; lane 0 applies NV_T and the limiter.  Sums in another order than the FP64 pass: equal to ~1e-16.
__global__ __launch_bounds__(kBlock) void gradients_redo_fp64_kernel(ForceArgs a)
{
    const int n_redo = a.redo_count[1];
    const int lane = threadIdx.x & 63;
    const int n_waves = gridDim.x * (kBlock / 64);
    for (int w = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); w < n_redo; w += n_waves) {
        const int i = a.redo_list[a.n + w];
        const BoxDesc &G = a.boxes[a.box_of[i]];
        const double4 pi = a.posm[i], vi = a.velu[i];
        double g[15];
#pragma unroll
        for (int q = 0; q < 15; q++) g[q] = 0;
        double maxd = 0, mn_r = 0, mx_r = 0, mn_p = 0, mx_p = 0;
        double mn_v0 = 0, mx_v0 = 0, mn_v1 = 0, mx_v1 = 0, mn_v2 = 0, mx_v2 = 0;
        const double h_i = a.hsml[i], rho_i = a.rho[i], p_i = a.pres[i], m_i = pi.w;
        if ((m_i > 0) && (rho_i > 0) && (h_i > 0)) {
            double hinv, hinv3, hinv4;
            kernel_hinv(h_i, hinv, hinv3, hinv4);
            const double h2_i = h_i * h_i;
            const double xs[3] = {pi.x, pi.y, pi.z};
            const double R = (h_i > G.hmax) ? h_i : G.hmax;
            wave_pair_candidates(a.tree, G, xs, R, h_i, lane, [&](int j) {
                const double4 xj = a.posm[j];
                double dp[3] = {xs[0] - xj.x, xs[1] - xj.y, xs[2] - xj.z};
                wrap_dp(G.box, dp);
                const double r2 = dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2];
                const double h_j = a.hsml[j];
                if (r2 <= 0) return;
                if ((r2 >= h2_i) && (r2 >= h_j * h_j)) return;
                const double rho_j = a.rho[j];
                if (!((xj.w > 0) && (rho_j > 0))) return;
                if (!(h_j > 0)) return;
                const double4 vj = a.velu[j];
                const double r = sqrt(r2);
                const double wk_i = (r < h_i) ? kernel_w(r * hinv, hinv3) : 0.0;
                if (r > maxd) maxd = r;
                const double d_rho = rho_j - rho_i;
                minmax_check(d_rho, mn_r, mx_r);
                const double dpp = a.pres[j] - p_i;
                minmax_check(dpp, mn_p, mx_p);
                const double dv0 = vj.x - vi.x, dv1 = vj.y - vi.y, dv2 = vj.z - vi.z;
                minmax_check(dv0, mn_v0, mx_v0);
                minmax_check(dv1, mn_v1, mx_v1);
                minmax_check(dv2, mn_v2, mx_v2);
                if (r < h_i) {
                    const double w0 = -wk_i * dp[0], w1 = -wk_i * dp[1], w2 = -wk_i * dp[2];
                    g[0] += w0 * d_rho; g[1] += w1 * d_rho; g[2] += w2 * d_rho;
                    g[3] += w0 * dpp; g[4] += w1 * dpp; g[5] += w2 * dpp;
                    g[6] += w0 * dv0; g[7] += w1 * dv0; g[8] += w2 * dv0;
                    g[9] += w0 * dv1; g[10] += w1 * dv1; g[11] += w2 * dv1;
                    g[12] += w0 * dv2; g[13] += w1 * dv2; g[14] += w2 * dv2;
                }
            });
        }
#pragma unroll
        for (int q = 0; q < 15; q++) g[q] = wave_sum(g[q]);
        maxd = wave_max(maxd);
        mx_r = wave_max(mx_r); mn_r = wave_min(mn_r);
        mx_p = wave_max(mx_p); mn_p = wave_min(mn_p);
        mx_v0 = wave_max(mx_v0); mn_v0 = wave_min(mn_v0);
        mx_v1 = wave_max(mx_v1); mn_v1 = wave_min(mn_v1);
        mx_v2 = wave_max(mx_v2); mn_v2 = wave_min(mn_v2);
        if (lane == 0) {
            double nvt[6];
#pragma unroll
            for (int q = 0; q < 6; q++) nvt[q] = a.nvt[6 * (size_t)i + q];
            double gr[3] = {g[0], g[1], g[2]}, gp[3] = {g[3], g[4], g[5]};
            double gv0[3] = {g[6], g[7], g[8]}, gv1[3] = {g[9], g[10], g[11]}, gv2[3] = {g[12], g[13], g[14]};
            apply_nvt(nvt, gr);
            apply_nvt(nvt, gp);
            apply_nvt(nvt, gv0);
            apply_nvt(nvt, gv1);
            apply_nvt(nvt, gv2);
            const double cn = a.cn[i];
            const double d_max = (h_i > maxd) ? h_i : maxd;
            double a_lim = 0.25;
            if (cn > 100) a_lim = dmin(0.5, 0.25 + 0.25 * (cn - 100) / 100);
            slope_limit(gr, mx_r, mn_r, a_lim, h_i, 0, 1, d_max, rho_i);
            slope_limit(gp, mx_p, mn_p, a_lim, h_i, 0.1, 1, d_max, p_i);
            slope_limit(gv0, mx_v0, mn_v0, a_lim, h_i, 0.1, 0, 0, 0);
            slope_limit(gv1, mx_v1, mn_v1, a_lim, h_i, 0.1, 0, 0, 0);
            slope_limit(gv2, mx_v2, mn_v2, a_lim, h_i, 0.1, 0, 0, 0);
            double *o = a.grad + 15 * (size_t)i;
            o[0] = gr[0]; o[1] = gr[1]; o[2] = gr[2];
            o[3] = gp[0]; o[4] = gp[1]; o[5] = gp[2];
            o[6] = gv0[0]; o[7] = gv0[1]; o[8] = gv0[2];
            o[9] = gv1[0]; o[10] = gv1[1]; o[11] = gv1[2];
            o[12] = gv2[0]; o[13] = gv2[1]; o[14] = gv2[2];
            a.redo_flag[i] &= 1;
        }
    }
}

// The force loop of each listed particle in FP64 from the FP64 state: the
// FP64 pass's neighbour predicate, pair arithmetic and inline exact solver
// (hydro_evaluate.h:22-516).  One wavefront per listed particle: the lanes
// share its candidates (wave_pair_candidates) and reduce their flux sums;
// lane 0 applies the final operations.  REDO: the particles listed by
// mark_redo (with the redone gradients), overwriting their FP32 result and
// clearing their flags; their pairs were counted by the FP32 loop.  Otherwise:
// the particles whose pair set overflowed the list (one thread would walk
// their long pair sets alone), counted here.
template <bool REDO>
__global__ __launch_bounds__(kBlock) void force_wave_fp64_kernel(ForceArgs a, const int *list, const int *count)
{
    const int n_redo = *count;
    if (REDO && blockIdx.x == 0 && threadIdx.x == 0 && n_redo > 0) atomicAdd(&a.ctr->redo_n, n_redo);
    const int lane = threadIdx.x & 63;
    const int n_waves = gridDim.x * (kBlock / 64);
    unsigned long long cnt = 0, nretry = 0, nexact = 0;
    for (int w = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); w < n_redo; w += n_waves) {
        const int i = list[w];
        const BoxDesc &G = a.boxes[a.box_of[i]];
        const double4 pi = a.posm[i];
        const double m_i = pi.w;
        double acc[3] = {0, 0, 0}, dte = 0, out_vsig = kMinReal;
        int status = 0;
        const bool live = (m_i > 0) && (a.rho[i] > 0);
        const double h_i = a.hsml[i];
        if (live) {
            PairI<double> I;
            side_fp64(a, i, I);
            I.hinv3 = I.hinv * I.hinv * I.hinv;
            I.hinv4 = I.hinv3 * I.hinv;
            I.cnumcrit2 = kCondDanger * kCondDanger - I.cn * I.cn;
            out_vsig = I.cs;
            const double h2_i = h_i * h_i;
            const double xs[3] = {pi.x, pi.y, pi.z};
            const double R = (h_i > G.hmax) ? h_i : G.hmax;
            ROut<double> ro = {0, 0, {0, 0, 0}, 0};
            wave_pair_candidates(a.tree, G, xs, R, h_i, lane, [&](int j) {
                if (j == i) return;
                double dp[3];
                const double r2 = exact_r2(G.box, pi, a.posm[j], dp);
                const double h_j = a.hsml[j];
                if (!(r2 > 0) || (!(r2 < h2_i) && !(r2 < h_j * h_j))) return;
                if (!(a.posm[j].w > 0) || !(a.rho[j] > 0)) return;  // hydro_evaluate.h:149-150
                PairSide<double> J;
                side_fp64(a, j, J);
                double Fv[3], Fp, vsig;
                int mask;
                const int e = force_pair<double, false, true>(I, J, dp, r2, a.G, ro, Fv, Fp, vsig, mask);
                if (e > status) status = e;
                cnt++;
                nretry += (mask & 8) ? 1 : 0;
                nexact += (mask & 4) ? 1 : 0;
                acc[0] += Fv[0];
                acc[1] += Fv[1];
                acc[2] += Fv[2];
                dte += Fp;
                if (vsig > out_vsig) out_vsig = vsig;
                if (a.is_active && !a.is_active[j]) {
                    if (vsig > kWakeup * a.vsig_old[j]) a.wakeup[j] = 1;
                }
            });
        }
        acc[0] = wave_sum(acc[0]);
        acc[1] = wave_sum(acc[1]);
        acc[2] = wave_sum(acc[2]);
        dte = wave_sum(dte);
        out_vsig = wave_max(out_vsig);
        for (int off = 32; off > 0; off >>= 1) {
            const int o = __shfl_xor(status, off, 64);
            status = (o > status) ? o : status;
        }
        if (lane == 0) {
            double vmax = kMinReal;
            if (live) {
                if (vmax < out_vsig) vmax = out_vsig;
                const double4 v = a.velu[i];
                const double vi[3] = {v.x, v.y, v.z};
                force_final_ops(acc, dte, vi, m_i, h_i, a.max_hsml);
            }
            a.acc[3 * (size_t)i + 0] = acc[0];
            a.acc[3 * (size_t)i + 1] = acc[1];
            a.acc[3 * (size_t)i + 2] = acc[2];
            a.dtu[i] = dte;
            a.vsig[i] = vmax;
            if (REDO) a.redo_flag[i] = 0;
            if (status) atomicMax(&a.ctr->status, status);
        }
    }
    if (!REDO) {
        wave_add_u64(&a.ctr->interactions, cnt);
        wave_add_u64(&a.ctr->retries, nretry);
        wave_add_u64(&a.ctr->exact_calls, nexact);
    }
}

__global__ void collect_long_lists_kernel(int n_active, const int *active, const int *count, int cap, int *list,
                                          int *n_list)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n_active && count[t] > cap) list[atomicAdd(n_list, 1)] = active[t];
}

inline unsigned nblocks(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }
inline unsigned nblocks(long long n, int block) { return (unsigned)((n + block - 1) / block); }

}  // namespace

void launch_pack_geom_f32(hipStream_t s, int n, const double4 *posm, const double4 *velu, const BoxDesc *boxes,
                          const int *box_of, float4 *offm, float4 *velf, uint32_t *cellc)
{
    if (n > 0)
        hipLaunchKernelGGL(pack_geom_f32_kernel, dim3(nblocks(n)), dim3(kBlock), 0, s, n, posm, velu, boxes, box_of, offm,
                           velf, cellc);
}
void launch_density_sweep_f32(hipStream_t s, const DensitySweepArgs &a)
{
    if (a.n_live <= 0) return;
    // workgroups of 512 take their own instantiation: the bound shapes the register allocation
    const int b = a.block;
    const dim3 grid(nblocks(a.n_live, b)), block(b);
    if (b > kBlock) {
        if (a.tree.nodes) hipLaunchKernelGGL((density_sweep_f32_kernel<true, kMaxBlock>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((density_sweep_f32_kernel<false, kMaxBlock>), grid, block, 0, s, a);
    } else {
        if (a.tree.nodes) hipLaunchKernelGGL((density_sweep_f32_kernel<true, kBlock>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((density_sweep_f32_kernel<false, kBlock>), grid, block, 0, s, a);
    }
}
void launch_pack_grad_f32(hipStream_t s, int n, const float4 *offm, const float4 *velf, const double *hsml,
                          const double *rho, const double *pres, GradRecF *rec)
{
    if (n > 0) hipLaunchKernelGGL(pack_grad_f32_kernel, dim3(nblocks(n)), dim3(kBlock), 0, s, n, offm, velf, hsml, rho, pres, rec);
}
void launch_gradients_f32(hipStream_t s, const GradientArgs &a)
{
    if (a.n_active <= 0) return;
    const dim3 grid(nblocks(a.n_active, a.block)), block(a.block);
    if (a.tree.nodes) {
        if (a.block == 64) hipLaunchKernelGGL((gradients_f32_kernel<true, 64>), grid, block, 0, s, a);
        else if (a.block == 128) hipLaunchKernelGGL((gradients_f32_kernel<true, 128>), grid, block, 0, s, a);
        else if (a.block == 512) hipLaunchKernelGGL((gradients_f32_kernel<true, 512>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((gradients_f32_kernel<true, kBlock>), grid, block, 0, s, a);
    } else {
        if (a.block == 64) hipLaunchKernelGGL((gradients_f32_kernel<false, 64>), grid, block, 0, s, a);
        else if (a.block == 128) hipLaunchKernelGGL((gradients_f32_kernel<false, 128>), grid, block, 0, s, a);
        else if (a.block == 512) hipLaunchKernelGGL((gradients_f32_kernel<false, 512>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((gradients_f32_kernel<false, kBlock>), grid, block, 0, s, a);
    }
}
void launch_pack_force_f32(hipStream_t s, int n, const float4 *offm, const float4 *velf, const uint32_t *cellc,
                           const double4 *posm, const double *hsml, const double *rho, const double *pres,
                           const double *dhsml, const double *cn, const double *fce, const double *nvt,
                           const double *grad, double gamma, ForceRecF *rec)
{
    if (n > 0)
        hipLaunchKernelGGL(pack_force_f32_kernel, dim3(nblocks(n)), dim3(kBlock), 0, s, n, offm, velf, cellc, posm, hsml,
                           rho, pres, dhsml, cn, fce, nvt, grad, gamma, rec);
}
void launch_force_f32(hipStream_t s, const ForceArgs &a)
{
    if (a.n_active <= 0) return;
    // only the three-phase list kernel takes workgroups above kBlock
    const int b = (a.list.idx && a.defer.mask) ? a.block : min(a.block, kBlock);
    const dim3 grid(nblocks(a.n_active, b)), block(b);
    if (a.list.idx) {
        if (a.defer.mask) hipLaunchKernelGGL((force_f32_kernel<true, true>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((force_f32_kernel<true, false>), grid, block, 0, s, a);
    } else if (a.tree.nodes) {
        hipLaunchKernelGGL((force_f32_kernel<false, false, true>), grid, block, 0, s, a);
    } else {
        hipLaunchKernelGGL((force_f32_kernel<false, false>), grid, block, 0, s, a);
    }
}
void launch_force_redo_fp64(hipStream_t s, const ForceArgs &a)
{
    // fixed grids: the counts are known on the device only (graph capture)
    hipLaunchKernelGGL(gradients_redo_fp64_kernel, dim3(1024), dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(force_wave_fp64_kernel<true>, dim3(1024), dim3(kBlock), 0, s, a, (const int *)a.redo_list,
                       (const int *)a.redo_count);
}
void launch_force_long_lists(hipStream_t s, const ForceArgs &a, int *list, int *n_list)
{
    if (a.n_active <= 0) return;
    hipLaunchKernelGGL(collect_long_lists_kernel, dim3(nblocks(a.n_active)), dim3(kBlock), 0, s, a.n_active, a.active,
                       a.list.count, a.list.cap, list, n_list);
    hipLaunchKernelGGL(force_wave_fp64_kernel<false>, dim3(1024), dim3(kBlock), 0, s, a, (const int *)list,
                       (const int *)n_list);
}
void launch_force_deferred_f32(hipStream_t s, const ForceArgs &a, int n_pairs)
{
    if (n_pairs > 0) hipLaunchKernelGGL(force_deferred_f32_kernel, dim3(nblocks(n_pairs)), dim3(kBlock), 0, s, a);
}

}  // namespace gzm
