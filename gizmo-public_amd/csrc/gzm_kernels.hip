// gzm_kernels.hip -- HIP kernels of the MFM hydro pass for gfx950 (FP64).
//
// One thread per active particle i ("gather form"): the thread walks the
// cells of the grid that can hold neighbours, evaluates every pair (i, j) for
// its own particle only, and finishes the per-particle work (hsml update,
// gradient construction + limiter, flux sums -> dv/dt, du/dt) in registers.
// Consecutive threads hold consecutive particles of the Morton-sorted order,
// so the lanes of a wavefront walk almost the same cells and their neighbour
// loads hit the same cache lines.
//
// No pair writes to j (the reference disables its pair-symmetric writes too,
// hydro_evaluate.h:159-168, gradients.c:1554-1565), so every kernel is free of
// races and deterministic for a fixed particle order.
#include "grid_device.h"
#include "gzm_internal.h"

namespace gzm {

namespace {

// ---------------------------------------------------------------------------
// grid build
// ---------------------------------------------------------------------------
// global cell index of a position in box B (cell_base + linear cell id)
__device__ __forceinline__ int global_cell(const BoxDesc &B, double x, double y, double z)
{
    const Grid &g = B.grid;
    return B.cell_base + (cell_coord(x, g.inv_s[0], g.nc[0]) * g.nc[1] + cell_coord(y, g.inv_s[1], g.nc[1])) * g.nc[2] +
           cell_coord(z, g.inv_s[2], g.nc[2]);
}

// sort key: box index above the 30-bit Morton code of the cell in its box
__global__ void cell_keys_kernel(int n, const double *pos, const BoxDesc *boxes, const int *box_caller,
                                 uint64_t *keys, int *iota)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int b = box_caller ? box_caller[i] : 0;
    const Grid &g = boxes[b].grid;
    int cx = cell_coord(pos[3 * i + 0], g.inv_s[0], g.nc[0]);
    int cy = cell_coord(pos[3 * i + 1], g.inv_s[1], g.nc[1]);
    int cz = cell_coord(pos[3 * i + 2], g.inv_s[2], g.nc[2]);
    keys[i] = ((uint64_t)b << 30) | morton3(cx, cy, cz);
    iota[i] = i;
}

__global__ void gather_sorted_kernel(int n, const int *perm, const double *pos, const double *vel,
                                     const double *mass, const double *u, const int *box_caller, double4 *posm,
                                     double4 *velu, int *inv, int *box_of)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int i = perm[k];
    posm[k] = make_double4(pos[3 * i + 0], pos[3 * i + 1], pos[3 * i + 2], mass[i]);
    velu[k] = make_double4(vel[3 * i + 0], vel[3 * i + 1], vel[3 * i + 2], u[i]);
    inv[i] = k;
    box_of[k] = box_caller ? box_caller[i] : 0;
}

__global__ void gather_scalar_kernel(int n, const int *perm, const double *src, double *dst, int width)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * width) return;
    int k = (int)(t / width), c = (int)(t % width);
    dst[t] = src[(long long)perm[k] * width + c];
}

__global__ void scatter_scalar_kernel(int n, const int *perm, const double *src, double *dst, int width)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * width) return;
    int k = (int)(t / width), c = (int)(t % width);
    dst[(long long)perm[k] * width + c] = src[t];
}

__global__ void cell_ranges_kernel(int n, const double4 *posm, const BoxDesc *boxes, const int *box_of, int2 *range)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double4 p = posm[k];
    const int c = global_cell(boxes[box_of[k]], p.x, p.y, p.z);
    int cprev = -1, cnext = -1;
    if (k > 0) {
        double4 q = posm[k - 1];
        cprev = global_cell(boxes[box_of[k - 1]], q.x, q.y, q.z);
    }
    if (k + 1 < n) {
        double4 q = posm[k + 1];
        cnext = global_cell(boxes[box_of[k + 1]], q.x, q.y, q.z);
    }
    if (c != cprev) range[c].x = k;
    if (c != cnext) range[c].y = k + 1;
}

// Per-box max hsml (the search radius of the pair loops, force_update_hmax):
// block = one tile of up to kHmaxTile slots of one box; block maximum, then
// one atomic max on the bit pattern (valid for hsml >= 0).
__global__ __launch_bounds__(kBlock) void box_hmax_kernel(BoxDesc *boxes, const double *hsml, const int2 *tiles)
{
    __shared__ double red[kBlock / 64];
    const int2 tile = tiles[blockIdx.x];  // (box, first slot)
    const BoxDesc &B = boxes[tile.x];
    const int end = min(tile.y + kHmaxTile, B.first + B.count);
    double m = 0;
    for (int k = tile.y + threadIdx.x; k < end; k += kBlock) m = fmax(m, hsml[k]);
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / 64; w++) m = fmax(m, red[w]);
        atomicMax((unsigned long long *)&boxes[tile.x].hmax, (unsigned long long)__double_as_longlong(m));
    }
}

__global__ void clear_box_hmax_kernel(int n_boxes, BoxDesc *boxes)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < n_boxes) boxes[b].hmax = 0.0;
}

__global__ void flag_unconverged_kernel(const int *n_live_dev, Counters *ctr)
{
    if (threadIdx.x == 0 && *n_live_dev > 0) atomicMax(&ctr->status, 1155);
}

__global__ void clear_counters_kernel(Counters *ctr)
{
    if (threadIdx.x == 0) {
        ctr->interactions = 0;
        ctr->retries = 0;
        ctr->exact_calls = 0;
        ctr->redo_n = 0;
        ctr->status = 0;
        ctr->overflow_slices = 0;
    }
}

__global__ void map_active_kernel(int n_active, const int *active_caller, const int *inv, int *active_sorted)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_active) return;
    active_sorted[t] = inv[active_caller[t]];
}

__global__ void fill_int_kernel(int *p, int n, int v)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

__global__ void fill_u8_kernel(uint8_t *p, int n, uint8_t v)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

__global__ void fill_double_kernel(double *p, long long n, double v)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

__global__ void set_flags_kernel(int n_active, const int *active, uint8_t *flags)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n_active) flags[active[t]] = 1;
}

// ---------------------------------------------------------------------------
// density (density.c:252-347 pair sums + 505-850 iteration, fused)
// ---------------------------------------------------------------------------
__global__ void density_init_kernel(int n_active, const int *active, const double4 *posm, int *dcache_n,
                                    double *hsml, double *left, double *right, double max_hsml, int *live_flag)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_active) return;
    int i = active[t];
    if (dcache_n) dcache_n[i] = kDensityCacheCap + 1;  // no candidate row until this stage's first sweep
    int live = posm[i].w > 0;  // density_isactive: gas with mass > 0 (density.c:85)
    live_flag[t] = live;
    if (!live) return;
    left[i] = 0;
    right[i] = 0;
    double h = hsml[i];
    if ((h < 0) || !isfinite(h) || (h > 0.99 * max_hsml)) hsml[i] = 0.99 * max_hsml;  // density.c:484
}

// one density pair (density.c:252-347): j's terms into S when r_ij < h_i.
// EDGE: also set `edge` when r_ij^2 lies within kEdgeBand of h_i^2.  MFV:
// also m_j W_ij v_j into pv (density.c:291-293 for r = 0, 331-334 as
// v_i - (v_i - v_j) otherwise), the sum behind ParticleVel.
template <bool EDGE = false, bool MFV = false>
__device__ __forceinline__ bool density_pair(const BoxDesc &G, const double4 &pi, const double4 &vi,
                                             const DensitySweepArgs &a, int j, double h2, double hinv,
                                             double hinv3, double hinv4, DensitySums &S, bool *edge = nullptr,
                                             double *pv = nullptr, double *r2_out = nullptr)
{
    const double4 pj = a.posm[j];
    double dp[3] = {pi.x - pj.x, pi.y - pj.y, pi.z - pj.z};
    wrap_dp(G.box, dp);
    const double r2 = dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2];
    if (r2_out) *r2_out = r2;
    if (EDGE && pj.w > 0 && fabs(r2 - h2) <= kEdgeBand * h2) *edge = true;
    if (!(r2 < h2 && pj.w > 0)) return false;
    const double r = sqrt(r2);
    const double u = r * hinv;
    double wk, dwk;
    kernel_wdw(u, hinv3, hinv4, wk, dwk);
    S.ngb += wk;
    S.rho += pj.w * wk;
    S.dhsml += -(3 * hinv * wk + u * dwk);
    if (r > 0) {
        S.t00 += wk * dp[0] * dp[0];
        S.t01 += wk * dp[0] * dp[1];
        S.t02 += wk * dp[0] * dp[2];
        S.t11 += wk * dp[1] * dp[1];
        S.t12 += wk * dp[1] * dp[2];
        S.t22 += wk * dp[2] * dp[2];
        S.f0 += wk * dp[0];
        S.f1 += wk * dp[1];
        S.f2 += wk * dp[2];
        const double4 vj = a.velu[j];
        const double dv0 = vi.x - vj.x, dv1 = vi.y - vj.y, dv2 = vi.z - vj.z;
        if (MFV) {
            const double mj_wk = pj.w * wk;
            pv[0] += mj_wk * (vi.x - dv0);
            pv[1] += mj_wk * (vi.y - dv1);
            pv[2] += mj_wk * (vi.z - dv2);
        }
        S.divvel -= dwk * (dp[0] * dv0 + dp[1] * dv1 + dp[2] * dv2) / r;
    } else if (MFV) {
        const double4 vj = a.velu[j];
        const double mj_wk = pj.w * wk;
        pv[0] += mj_wk * vj.x;
        pv[1] += mj_wk * vj.y;
        pv[2] += mj_wk * vj.z;
    }
    return true;
}

// the hsml update of one particle from its sums, and its outputs; returns
// hsml_update's result (1 sweep again, 0 done, -1 bracket failure)
__device__ __forceinline__ int density_store(const DensitySweepArgs &a, int t, int i, double h, const DensitySums &S)
{
    double left = a.left[i], right = a.right[i];
    DensityResult R;
    const int res = hsml_update(a.hp, S, h, left, right, a.cn[i], R);
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
    return res;
}

// 4 waves per SIMD: fits 128 VGPRs without spills, and the extra wave hides
// the candidate loads (sweeps 10 % faster than at the compiler's 146 VGPRs)
template <bool TREE, int MAXB, bool MFV = false>
__global__ __launch_bounds__(MAXB, 4) void density_sweep_kernel(DensitySweepArgs a)
{
    const int t = xcd_block() * (int)blockDim.x + threadIdx.x;
    const int n_live = a.n_live_dev ? *a.n_live_dev : a.n_live;
    const bool valid = t < n_live;
    if (!valid && a.n_live_dev && t < a.n_live) a.redo[t] = 0;  // captured pass: clear up to the capacity
    unsigned long long cnt = 0;
    for_box(a.boxes, a.box_of, valid ? a.live[t] : 0, valid, [&](const BoxDesc &G) {
        const int i = a.live[t];
        const double4 pi = a.posm[i];
        const double4 vi = a.velu[i];
        const double h = a.hsml[i];
        double hinv, hinv3, hinv4;
        kernel_hinv(h, hinv, hinv3, hinv4);
        const double h2 = h * h;
        DensitySums S = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        double pv[3] = {0, 0, 0};
        const double xs[3] = {pi.x, pi.y, pi.z};
        const int budget = a.budget;
        int visited = 0;
        bool heavy = false;
        unsigned long long lane_cnt = 0;
        auto pair = [&](int j) {
            if (density_pair<false, MFV>(G, pi, vi, a, j, h2, hinv, hinv3, hinv4, S, nullptr, pv)) lane_cnt++;
        };
        if (TREE) {
            const int cap = kDensityCacheCap;
            int *row = a.dcache_idx ? a.dcache_idx + (size_t)i * cap : nullptr;
            const int n_row = (row && a.hp.iter > 0) ? a.dcache_n[i] : cap + 1;
            if (n_row <= cap && h2 <= a.dcache_r2[i]) {
                // every candidate within h is in the row, in walk order
                const int4 *row4 = reinterpret_cast<const int4 *>(row);
                for (int k = 0; k < n_row; k += 4) {
                    const int4 q = row4[k >> 2];
                    pair(q.x);
                    if (k + 1 < n_row) pair(q.y);
                    if (k + 2 < n_row) pair(q.z);
                    if (k + 3 < n_row) pair(q.w);
                }
            } else if (row && a.hp.iter == 0) {
                // first sweep: walk the cache radius, keep its candidates (the
                // candidate budget doubles with the wider radius)
                const double Rc = kDensityCacheFactor * h, Rc2 = Rc * Rc;
                int m = 0;
                tree_walk_grouped(a.tree, G, xs, Rc, [&](int j) -> bool {
                    if (budget && ++visited > 2 * budget) {
                        heavy = true;
                        return false;
                    }
                    double r2;
                    if (density_pair<false, MFV>(G, pi, vi, a, j, h2, hinv, hinv3, hinv4, S, nullptr, pv, &r2))
                        lane_cnt++;
                    if (r2 < Rc2) {
                        if (m < cap) row[m] = j;
                        m++;
                    }
                    return true;
                });
                a.dcache_n[i] = heavy ? cap + 1 : m;
                a.dcache_r2[i] = Rc2;
            } else if (row && budget) {
                // a later sweep without a usable row: one wavefront for this
                // particle's walk, so that it does not hold a wavefront of row scans
                heavy = true;
            } else {
                tree_walk_grouped(a.tree, G, xs, h, [&](int j) -> bool {
                    if (budget && ++visited > budget) {
                        heavy = true;
                        return false;
                    }
                    pair(j);
                    return true;
                });
            }
        } else {
            // the budget is charged a cell at a time, outside the candidate loop
            bool full[3];
            for_each_cell(G.grid, xs, h, full, [&](int, int, int, int beg, int end) -> bool {
                visited += end - beg;
                if (budget && visited > budget) {
                    heavy = true;
                    return false;
                }
                for (int j = beg; j < end; ++j) pair(j);
                return true;
            });
        }
        if (heavy) {  // left to density_heavy_kernel
            a.heavy_list[atomicAdd(a.n_heavy, 1)] = t;
            return;
        }
        cnt += lane_cnt;
        density_store(a, t, i, h, S);
        if (MFV) {
            a.pvel[3 * (size_t)i + 0] = pv[0];
            a.pvel[3 * (size_t)i + 1] = pv[1];
            a.pvel[3 * (size_t)i + 2] = pv[2];
        }
    });
    wave_add_u64(&a.ctr->interactions, cnt);
}

// The particles that passed the candidate budget: one wavefront each, its
// lanes sharing the walk (wave_candidates); the sums are reduced across the
// wavefront (another summation order than a single lane's, FP64 throughout).
// A fixed grid strides over the list, so the launch is sized independently of
// the (device-side) count.
template <bool MFV = false>
__global__ __launch_bounds__(kBlock) void density_heavy_kernel(DensitySweepArgs a)
{
    const int lane = threadIdx.x & 63;
    const int n_waves = (int)(gridDim.x * (kBlock / 64));
    const int n = a.wave_all ? (a.n_live_dev ? *a.n_live_dev : a.n_live) : *a.n_heavy;
    if (a.wave_all && a.n_live_dev) {  // captured pass: clear the flags up to the capacity
        for (int t = n + (int)(blockIdx.x * kBlock + threadIdx.x); t < a.n_live; t += (int)(gridDim.x * kBlock))
            a.redo[t] = 0;
    }
    for (int u = (int)((blockIdx.x * kBlock + threadIdx.x) >> 6); u < n; u += n_waves) {
        const int t = a.wave_all ? u : a.heavy_list[u];
        const int i = a.live[t];
        const BoxDesc &G = a.boxes[a.box_of[i]];
        const double4 pi = a.posm[i];
        const double4 vi = a.velu[i];
        const double h = a.hsml[i];
        double hinv, hinv3, hinv4;
        kernel_hinv(h, hinv, hinv3, hinv4);
        const double h2 = h * h;
        DensitySums S = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        double pv[3] = {0, 0, 0};
        const double xs[3] = {pi.x, pi.y, pi.z};
        unsigned long long c = 0;
        bool edge = false;
        wave_candidates(a.tree, G, xs, h, lane, [&](int j) {
            if (a.edge_list) {
                if (density_pair<true>(G, pi, vi, a, j, h2, hinv, hinv3, hinv4, S, &edge)) c++;
            } else {
                if (density_pair<false, MFV>(G, pi, vi, a, j, h2, hinv, hinv3, hinv4, S, nullptr, pv)) c++;
            }
        });
        const bool any_edge = __any(edge);
        S.ngb = wave_sum(S.ngb); S.rho = wave_sum(S.rho); S.dhsml = wave_sum(S.dhsml);
        S.divvel = wave_sum(S.divvel);
        S.t00 = wave_sum(S.t00); S.t01 = wave_sum(S.t01); S.t02 = wave_sum(S.t02);
        S.t11 = wave_sum(S.t11); S.t12 = wave_sum(S.t12); S.t22 = wave_sum(S.t22);
        S.f0 = wave_sum(S.f0); S.f1 = wave_sum(S.f1); S.f2 = wave_sum(S.f2);
        if (MFV) { pv[0] = wave_sum(pv[0]); pv[1] = wave_sum(pv[1]); pv[2] = wave_sum(pv[2]); }
        wave_add_u64(&a.ctr->interactions, c);
        if (lane == 0) {
            const int res = density_store(a, t, i, h, S);
            if (res == 0 && any_edge && a.edge_list) a.edge_list[atomicAdd(a.n_edge, 1)] = i;
            if (MFV) {
                a.pvel[3 * (size_t)i + 0] = pv[0];
                a.pvel[3 * (size_t)i + 1] = pv[1];
                a.pvel[3 * (size_t)i + 2] = pv[2];
            }
        }
    }
}

__global__ void edge_reset_kernel(int cap, const int *n_edge, const int *edge_list, const double *hsml_entry,
                                  double *hsml, double *left, double *right)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cap || t >= *n_edge) return;
    const int i = edge_list[t];
    hsml[i] = hsml_entry[i];
    left[i] = 0;
    right[i] = 0;
}

__global__ void density_finalize_kernel(int n_active, const int *active, const double4 *posm, const double4 *velu,
                                        const double *hsml, double *num_ngb, double *rho, double *pres, double gm1,
                                        double *pvel)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_active) return;
    int i = active[t];
    double m = posm[i].w;
    if (!(m > 0)) return;
    if (pvel && rho[i] > 0) {
        // MFV mesh motion 5 (density.c:886-894): 0.7 VelPred + 0.3 the kernel-weighted
        // mean velocity, over the density sum the last sweep left in rho
        const double4 v = velu[i];
        const double eps_pvel = 0.3, d_sum = rho[i];
        const double vv[3] = {v.x, v.y, v.z};
        for (int k = 0; k < 3; k++)
            pvel[3 * (size_t)i + k] = (1. - eps_pvel) * vv[k] + eps_pvel * pvel[3 * (size_t)i + k] / d_sum;
    }
    double h = hsml[i], nn = num_ngb[i], d;
    // density.c:967-978 (MFM: own mass times the kernel number density)
    if ((h > 0) && (nn > 0)) d = m * nn / (kNormCoeff * pow(h, 3.0));
    else if (h <= 0) d = 0;
    else d = m / (kNormCoeff * pow(h, 3.0));
    rho[i] = d;
    pres[i] = gm1 * velu[i].w * d;  // eos.c:58-61
    num_ngb[i] = (nn > 0) ? pow(nn, 1. / 3.) : 0;  // density.c:1020
}

// ---------------------------------------------------------------------------
// gradients (gradients.c:1477-2089 pair loop, 1052-1253 finalize)
// ---------------------------------------------------------------------------
__global__ void pack_grad_kernel(int n, const double4 *posm, const double4 *velu, const double *hsml,
                                 const double *rho, const double *pres, GradRec *rec)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double4 p = posm[k], v = velu[k];
    GradRec r;
    r.xh = make_double4(p.x, p.y, p.z, hsml[k]);
    r.vr = make_double4(v.x, v.y, v.z, rho[k]);
    r.pm = make_double2(pres[k], p.w);
    rec[k] = r;
}

// One particle's FP64 force record (shared by pack_force_kernel and the
// gradient kernel's epilogue, so that both write the same bits)
__device__ __forceinline__ ForceRec make_force_rec(const double4 &p, const double4 &v, double d, double P, double h,
                                                   double dh, double cn, double fce, const double *t, const double *g,
                                                   double gamma)
{
    ForceRec r;
    r.xh = make_double4(p.x, p.y, p.z, h);
    r.vc = make_double4(v.x, v.y, v.z, sqrt(gamma * P / d));  // eos.c:245-254
    r.th = make_double4(d, P, p.w / d, dh);
    r.mc = make_double4(p.w, cn, fce, P / d);
    r.t0 = make_double4(t[0], t[1], t[2], t[3]);
    r.t1 = make_double4(t[4], t[5], 1.0 / h, 0.0);
    r.g0 = make_double4(g[0], g[1], g[2], g[3]);
    r.g1 = make_double4(g[4], g[5], g[6], g[7]);
    r.g2 = make_double4(g[8], g[9], g[10], g[11]);
    r.g3 = make_double4(g[12], g[13], g[14], 0.0);
    return r;
}

constexpr int kNoSlot = 0x7fffffff;  // above every sorted slot: the end of a merged row

// Symmetric pair rows: particle j offers itself to particle i (0 < r_ij < h_j)
// unless j is in i's candidate row (r_ij below the row's radius) or i's row
// does not cover its own hsml (then i walks the tree anyway).
__device__ __forceinline__ void sym_offer(const SymRowArgs &a, const BoxDesc &G, const double4 &pj, double hj2, int j,
                                          int i)
{
    const double4 pi = a.posm[i];
    // r_ij as the gradient kernel computes it for i (x_i - x_j)
    double dp[3] = {pi.x - pj.x, pi.y - pj.y, pi.z - pj.z};
    wrap_dp(G.box, dp);
    const double r2 = dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2];
    if (!(r2 < hj2) || r2 <= 0) return;
    const double ri2 = a.dcache_r2[i];
    if (r2 < ri2) return;
    if (a.dcache_n[i] > kDensityCacheCap) return;
    const double h_i = a.hsml[i];
    if (!(h_i * h_i <= ri2)) return;
    const int k = atomicAdd(&a.bcount[i], 1);
    if (k < kPairRowCap) a.brow[(size_t)i * kPairRowCap + k] = j;
}

// Symmetric pair rows, scatter: each j takes the i from its own candidate row
// when that covers h_j; the other j (rows past their cap, hsml past the row's
// radius) are listed for sym_rows_wave_kernel.
__global__ __launch_bounds__(kBlock) void sym_rows_scatter_kernel(SymRowArgs a)
{
    const int j = blockIdx.x * kBlock + threadIdx.x;
    const bool valid = j < a.n;
    for_box(a.boxes, a.box_of, valid ? j : 0, valid, [&](const BoxDesc &G) {
        const double4 pj = a.posm[j];
        const double h_j = a.hsml[j];
        if (!(h_j > 0)) return;
        const double hj2 = h_j * h_j;
        const int n_j = a.dcache_n[j];
        if (n_j <= kDensityCacheCap && hj2 <= a.dcache_r2[j]) {
            const int *row = a.dcache_idx + (size_t)j * kDensityCacheCap;
            for (int k = 0; k < n_j; k++) sym_offer(a, G, pj, hj2, j, row[k]);
        } else {
            a.wlist[atomicAdd(a.wcount, 1)] = j;
        }
    });
}

// The listed j, one wavefront each: its lanes share the tree walk of radius h_j
// (wave_candidates).  A fixed grid strides over the device-side count.
__global__ __launch_bounds__(kBlock) void sym_rows_wave_kernel(SymRowArgs a)
{
    const int lane = threadIdx.x & 63;
    const int n_waves = (int)(gridDim.x * (kBlock / 64));
    const int n = *a.wcount;
    for (int u = (int)((blockIdx.x * kBlock + threadIdx.x) >> 6); u < n; u += n_waves) {
        const int j = a.wlist[u];
        const BoxDesc &G = a.boxes[a.box_of[j]];
        const double4 pj = a.posm[j];
        const double h_j = a.hsml[j];
        const double hj2 = h_j * h_j;
        const double xs[3] = {pj.x, pj.y, pj.z};
        wave_candidates(a.tree, G, xs, h_j, lane, [&](int i) { sym_offer(a, G, pj, hj2, j, i); });
    }
}

// Symmetric pair rows, sort: each row ascending (a handful of entries)
__global__ __launch_bounds__(kBlock) void sym_rows_sort_kernel(int n, int *brow, const int *bcount)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int c = bcount[i];
    if (c < 2 || c > kPairRowCap) return;
    int *r = brow + (size_t)i * kPairRowCap;
    for (int k = 1; k < c; k++) {
        const int v = r[k];
        int m = k - 1;
        while (m >= 0 && r[m] > v) {
            r[m + 1] = r[m];
            m--;
        }
        r[m + 1] = v;
    }
}

template <int BS, bool ROWS = false>
__global__ __launch_bounds__(BS) void gradients_kernel(GradientArgs a)
{
    // Pair-list entries are staged in LDS, column-major (entry c of thread x
    // at stage[c * BS + x], so same-c stores of a wave hit distinct banks),
    // and leave as whole 64-byte chunks: single 4-byte stores spread over the
    // loop would reach HBM as partial-line write-backs.  BS = blockDim.x.
    __shared__ int stage[kListChunk * BS];
    constexpr int bs = BS;
    const int t = xcd_block() * bs + threadIdx.x;
    const bool valid = t < a.n_active;
    unsigned long long cnt = 0;
    int listed = 0;
    for_box(a.boxes, a.box_of, valid ? a.active[t] : 0, valid, [&](const BoxDesc &G) {
        const int i = a.active[t];
        const GradRec ri = a.rec[i];
        double gr0 = 0, gr1 = 0, gr2 = 0, gp0 = 0, gp1 = 0, gp2 = 0;
        double gv00 = 0, gv01 = 0, gv02 = 0, gv10 = 0, gv11 = 0, gv12 = 0, gv20 = 0, gv21 = 0, gv22 = 0;
        double maxd = 0, mn_r = 0, mx_r = 0, mn_p = 0, mx_p = 0;
        double mn_v0 = 0, mx_v0 = 0, mn_v1 = 0, mx_v1 = 0, mn_v2 = 0, mx_v2 = 0;
        const double h_i = ri.xh.w, rho_i = ri.vr.w, p_i = ri.pm.x, m_i = ri.pm.y;
        int *list_row = a.list.idx ? a.list.idx + (size_t)t * a.list.cap : nullptr;
        if ((m_i > 0) && (rho_i > 0) && (h_i > 0)) {  // GasGrad_isactive (gradients.c:66-75)
            double hinv, hinv3, hinv4;
            kernel_hinv(h_i, hinv, hinv3, hinv4);
            const double h2_i = h_i * h_i;
            const double xs[3] = {ri.xh.x, ri.xh.y, ri.xh.z};
            auto visit = [&](int j) {
                const double4 xj = a.rec[j].xh;
                double dp[3] = {xs[0] - xj.x, xs[1] - xj.y, xs[2] - xj.z};
                wrap_dp(G.box, dp);
                const double r2 = dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2];
                const double h_j = xj.w;
                if (r2 <= 0) return;
                if ((r2 >= h2_i) && (r2 >= h_j * h_j)) return;
                const double4 vr = a.rec[j].vr;
                const double2 pm = a.rec[j].pm;
                if (!((pm.y > 0) && (vr.w > 0))) return;
                // the list holds the force loop's pair set (m_j > 0, rho_j > 0,
                // hydro_evaluate.h:149-150); gradients also need h_j > 0
                if (list_row) {
                    if (listed < a.list.cap) {
                        const int c = listed & (kListChunk - 1);
                        stage[c * bs + threadIdx.x] = j;
                        if (c == kListChunk - 1) flush_list_chunk(stage, list_row + (listed - c), bs);
                    }
                    listed++;
                }
                if (!(h_j > 0)) return;
                cnt++;
                const double r = sqrt(r2);
                const double wk_i = (r < h_i) ? kernel_w(r * hinv, hinv3) : 0.0;
                if (r > maxd) maxd = r;
                const double d_rho = vr.w - rho_i;
                minmax_check(d_rho, mn_r, mx_r);
                const double dpp = pm.x - p_i;
                minmax_check(dpp, mn_p, mx_p);
                const double dv0 = vr.x - ri.vr.x, dv1 = vr.y - ri.vr.y, dv2 = vr.z - ri.vr.z;
                minmax_check(dv0, mn_v0, mx_v0);
                minmax_check(dv1, mn_v1, mx_v1);
                minmax_check(dv2, mn_v2, mx_v2);
                if (r < h_i) {
                    const double w0 = -wk_i * dp[0], w1 = -wk_i * dp[1], w2 = -wk_i * dp[2];
                    gr0 += w0 * d_rho; gr1 += w1 * d_rho; gr2 += w2 * d_rho;
                    gp0 += w0 * dpp; gp1 += w1 * dpp; gp2 += w2 * dpp;
                    gv00 += w0 * dv0; gv01 += w1 * dv0; gv02 += w2 * dv0;
                    gv10 += w0 * dv1; gv11 += w1 * dv1; gv12 += w2 * dv1;
                    gv20 += w0 * dv2; gv21 += w1 * dv2; gv22 += w2 * dv2;
                }
            };
            const int n_a = ROWS ? a.dcache_n[i] : 0;
            const int n_b = ROWS ? a.bcount[i] : 0;
            if (ROWS && n_a <= kDensityCacheCap && h2_i <= a.dcache_r2[i] && n_b <= kPairRowCap) {
                // ascending merge of i's candidate row and its pair row
                const int *ra = a.dcache_idx + (size_t)i * kDensityCacheCap;
                const int *rb = a.brow + (size_t)i * kPairRowCap;
                int ka = 0, kb = 0;
                int ja = (n_a > 0) ? ra[0] : kNoSlot, jb = (n_b > 0) ? rb[0] : kNoSlot;
                while (ka < n_a || kb < n_b) {
                    if (ja < jb) {
                        visit(ja);
                        ka++;
                        ja = (ka < n_a) ? ra[ka] : kNoSlot;
                    } else {
                        visit(jb);
                        kb++;
                        jb = (kb < n_b) ? rb[kb] : kNoSlot;
                    }
                }
            } else {
                for_each_neighbour_candidate(a.tree, G, xs, h_i, true, visit);
            }
            // the partly filled last chunk (entries past `listed` are never read)
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
        if (a.frec_out) {  // every particle active: its force record, as pack_force_kernel writes it
            const double g[15] = {gr[0], gr[1], gr[2], gp[0], gp[1], gp[2], gv0[0], gv0[1], gv0[2],
                                  gv1[0], gv1[1], gv1[2], gv2[0], gv2[1], gv2[2]};
            a.frec_out[i] = make_force_rec(make_double4(ri.xh.x, ri.xh.y, ri.xh.z, m_i),
                                           make_double4(ri.vr.x, ri.vr.y, ri.vr.z, 0.0), rho_i, p_i, h_i,
                                           a.dhsml[i], cn, a.fce[i], nvt, g, a.gamma);
        }
        if (a.list.idx) a.list.count[t] = listed;
    });
    if (a.list.idx) {
        // a slice overflows when any of its lanes listed more than cap neighbours
        const bool over = __any(listed > a.list.cap);
        if ((threadIdx.x & 63) == 0 && t < a.n_active) {
            a.list.overflow[t >> 6] = over ? 1 : 0;
            if (over) atomicAdd(&a.ctr->overflow_slices, 1);
        }
    }
    wave_add_u64(&a.ctr->interactions, cnt);
}

// ---------------------------------------------------------------------------
// force (hydro_evaluate.h:22-516, hydro_core_meshless.h, compute_finitevol_faces.h,
// final operations hydro_toplevel.c:652-767)
// ---------------------------------------------------------------------------
__global__ void pack_force_kernel(int n, const double4 *posm, const double4 *velu, const double *hsml,
                                  const double *rho, const double *pres, const double *dhsml, const double *cn,
                                  const double *fce, const double *nvt, const double *grad, double gamma,
                                  ForceRec *rec)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    rec[k] = make_force_rec(posm[k], velu[k], rho[k], pres[k], hsml[k], dhsml[k], cn[k], fce[k], nvt + 6 * (size_t)k,
                            grad + 15 * (size_t)k, gamma);
}


// One particle's force loop.  USE_LIST: walk the pair list written by the
// gradient kernel (particles that overflowed it are skipped and left to
// force_wave_fp64_kernel); otherwise scan the grid or tree (no list: every
// particle).  Pairs that reach the exact Riemann
// solver are summed after all other pairs of the particle: THREE_PHASE (list
// kernel with defer buffers) leaves them to force_deferred_kernel and
// force_finish_kernel, otherwise the loop walks the pairs a second time and
// evaluates them in place -- the same order, so every variant gives the same
// bits.
template <bool USE_LIST, bool THREE_PHASE>
__global__ __launch_bounds__((USE_LIST && THREE_PHASE) ? kMaxBlock : kBlock) void force_kernel(ForceArgs a)
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
        const ForceRec ri = a.rec[i];
        double acc[3] = {0, 0, 0}, dte = 0, vmax = kMinReal;
        const double m_i = ri.mc.x, rho_i = ri.th.x;
        int status = 0;
        uint64_t dmask = 0;
        if ((m_i > 0) && (rho_i > 0)) {
            PairI<double> I;
            unpack_side(ri, I);
            const double x_i[3] = {ri.xh.x, ri.xh.y, ri.xh.z};
            I.hinv3 = I.hinv * I.hinv * I.hinv;
            I.hinv4 = I.hinv3 * I.hinv;
            I.cnumcrit2 = kCondDanger * kCondDanger - I.cn * I.cn;
            double out_vsig = I.cs;
            ROut<double> ro = {0, 0, {0, 0, 0}, 0};
            const double h2_i = I.h * I.h;
            int n_deferred = 0;
            auto add = [&](int j, const double Fv[3], double Fp, double vsig, int mask, int e) {
                if (e > status) status = e;
                nretry += (mask & 8) ? 1 : 0;
                nexact += (mask & 4) ? 1 : 0;
                acc[0] += Fv[0];
                acc[1] += Fv[1];
                acc[2] += Fv[2];
                dte += Fp;
                if (vsig > out_vsig) out_vsig = vsig;
                if (a.is_active && !a.is_active[j]) {
                    if (vsig > kWakeup * a.vsig_old[j]) a.wakeup[j] = 1;  // WAKEUP (hydro_evaluate.h:485-495)
                }
            };
            // pass 0: every pair but the deferred ones; pass 1: only those
            auto visit = [&](int j, int k, bool check, int pass) {
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
                double Fv[3], Fp, vsig;
                int mask;
                const int e = force_pair<double, true>(I, J, dp, r2, a.G, ro, Fv, Fp, vsig, mask);
                if (pass == 0) {
                    cnt++;
                    if (e == kDeferred) {
                        if (THREE_PHASE) dmask |= 1ull << k;
                        n_deferred++;
                        return;
                    }
                    add(j, Fv, Fp, vsig, mask, e);
                } else if (!THREE_PHASE && e == kDeferred) {
                    const int e2 = force_pair<double, false>(I, J, dp, r2, a.G, ro, Fv, Fp, vsig, mask);
                    add(j, Fv, Fp, vsig, mask, e2);
                }
            };
            for (int pass = 0; pass < 2; pass++) {
                if (pass == 1 && (THREE_PHASE || n_deferred == 0)) break;
                if (USE_LIST) {
                    // the list holds the pair set in the gradient kernel's scan order,
                    // which is this kernel's scan order as well (same predicate, same R)
                    const int n = a.list.count[t];
                    const int *row = a.list.idx + (size_t)t * a.list.cap;
                    for (int k = 0; k < n; k++) visit(row[k], k, false, pass);
                } else {
                    for_each_neighbour_candidate(a.tree, G, x_i, I.h, true, [&](int j) { visit(j, 0, true, pass); });
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
            force_final_ops(acc, dte, I.v, m_i, I.h, a.max_hsml);
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

// Deferred pairs q in [q_begin, min(total, q_end)): one pair per lane, with
// the exact Riemann solver.
__global__ __launch_bounds__(kBlock) void force_deferred_kernel(ForceArgs a)
{
    const int q = a.defer.q_begin + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int q_end = min(*a.defer.total, a.defer.q_end);
    const bool valid = q < q_end;
    const int2 ij = valid ? a.defer.pair[q - a.defer.q_begin] : make_int2(0, 0);
    unsigned long long nretry = 0, nexact = 0;
    for_box(a.boxes, a.box_of, ij.x, valid, [&](const BoxDesc &G) {
        const int i = ij.x, j = ij.y;
        const ForceRec ri = a.rec[i];
        PairI<double> I;
        unpack_side(ri, I);
        I.hinv3 = I.hinv * I.hinv * I.hinv;
        I.hinv4 = I.hinv3 * I.hinv;
        I.cnumcrit2 = kCondDanger * kCondDanger - I.cn * I.cn;
        const double4 xj = a.rec[j].xh;
        double dp[3] = {ri.xh.x - xj.x, ri.xh.y - xj.y, ri.xh.z - xj.z};
        wrap_dp(G.box, dp);
        const double r2 = dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2];
        PairSide<double> J;
        unpack_side(a.rec[j], J);
        ROut<double> ro = {0, 0, {0, 0, 0}, 0};
        double Fv[3], Fp, vsig;
        int mask;
        const int e = force_pair<double, false, true>(I, J, dp, r2, a.G, ro, Fv, Fp, vsig, mask);
        nretry += (mask & 8) ? 1 : 0;
        nexact += (mask & 4) ? 1 : 0;
        double *o = a.defer.result + 5 * (size_t)(q - a.defer.q_begin);
        o[0] = Fv[0]; o[1] = Fv[1]; o[2] = Fv[2]; o[3] = Fp; o[4] = vsig;
        if (a.is_active && !a.is_active[j]) {
            if (vsig > kWakeup * a.vsig_old[j]) a.wakeup[j] = 1;
        }
        if (e) atomicMax(&a.ctr->status, e);
    });
    wave_add_u64(&a.ctr->retries, nretry);
    wave_add_u64(&a.ctr->exact_calls, nexact);
}

// Add the deferred pairs of this round to the raw sums (in list order) and,
// for particles whose deferred pairs are all in, apply the final operations.
// `velu` and `hsml` give the FP64 velocity and hsml of i.
__global__ __launch_bounds__(kBlock) void force_finish_kernel(ForceArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n_active || !a.defer.raw[t]) return;
    const int i = a.active[t];
    double acc[3] = {a.acc[3 * (size_t)i + 0], a.acc[3 * (size_t)i + 1], a.acc[3 * (size_t)i + 2]};
    double dte = a.dtu[i], out_vsig = a.vsig[i];
    const int q0 = a.defer.offset[t], q1 = q0 + a.defer.count[t];
    for (int q = max(q0, a.defer.q_begin); q < min(q1, a.defer.q_end); q++) {
        const double *o = a.defer.result + 5 * (size_t)(q - a.defer.q_begin);
        acc[0] += o[0];
        acc[1] += o[1];
        acc[2] += o[2];
        dte += o[3];
        if (o[4] > out_vsig) out_vsig = o[4];
    }
    if (q1 <= a.defer.q_end) {
        double vmax = kMinReal;
        if (vmax < out_vsig) vmax = out_vsig;
        const double4 v = a.velu[i];
        const double vi[3] = {v.x, v.y, v.z};
        force_final_ops(acc, dte, vi, a.posm[i].w, a.hsml[i], a.max_hsml);
        out_vsig = vmax;
        a.defer.raw[t] = 0;
    }
    a.acc[3 * (size_t)i + 0] = acc[0];
    a.acc[3 * (size_t)i + 1] = acc[1];
    a.acc[3 * (size_t)i + 2] = acc[2];
    a.dtu[i] = dte;
    a.vsig[i] = out_vsig;
}

__global__ void defer_total_kernel(int n, const int *offset, const int *count, int *total)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *total = (n > 0) ? offset[n - 1] + count[n - 1] : 0;
}

// neighbour-set sizes for a fixed hsml (same predicates as the stage kernels)
__global__ __launch_bounds__(kBlock) void neighbor_count_kernel(const BoxDesc *boxes, const int *box_of,
                                                                TreeArgs tree, int mode, int n_active,
                                                                const int *active, const double4 *posm,
                                                                const double *hsml, int *counts, Counters *ctr)
{
    const int t = xcd_block() * kBlock + threadIdx.x;
    unsigned long long cnt = 0;
    const bool valid = t < n_active;
    for_box(boxes, box_of, valid ? active[t] : 0, valid, [&](const BoxDesc &G) {
        const int i = active[t];
        const double4 pi = posm[i];
        const double h = hsml[i], h2 = h * h;
        const double xs[3] = {pi.x, pi.y, pi.z};
        for_each_neighbour_candidate(tree, G, xs, h, mode != 0, [&](int j) {
            const double4 pj = posm[j];
            double dp[3] = {pi.x - pj.x, pi.y - pj.y, pi.z - pj.z};
            wrap_dp(G.box, dp);
            const double r2 = dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2];
            if (mode == 0) {
                if (r2 < h2) cnt++;
            } else {
                const double h_j = hsml[j];
                if ((r2 >= h2) && (r2 >= h_j * h_j)) return;
                if (r2 <= 0) return;
                cnt++;
            }
        });
        if (counts) counts[t] = (int)cnt;
    });
    wave_add_u64(&ctr->interactions, cnt);
}

inline unsigned nblocks(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }
inline unsigned nblocks(long long n, int block) { return (unsigned)((n + block - 1) / block); }


}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void launch_cell_keys(hipStream_t s, int n, const double *pos, const BoxDesc *boxes, const int *box_caller,
                      uint64_t *keys, int *iota)
{
    if (n > 0) hipLaunchKernelGGL(cell_keys_kernel, dim3(nblocks(n)), dim3(kBlock), 0, s, n, pos, boxes, box_caller, keys, iota);
}
void launch_gather_sorted(hipStream_t s, int n, const int *perm, const double *pos, const double *vel,
                          const double *mass, const double *u, const int *box_caller, double4 *posm,
                          double4 *velu, int *inv, int *box_of)
{
    if (n > 0)
        hipLaunchKernelGGL(gather_sorted_kernel, dim3(nblocks(n)), dim3(kBlock), 0, s, n, perm, pos, vel, mass, u, box_caller,
                           posm, velu, inv, box_of);
}
void launch_gather_scalar(hipStream_t s, int n, const int *perm, const double *src, double *dst, int width)
{
    if (((long long)n * width) > 0) hipLaunchKernelGGL(gather_scalar_kernel, dim3(nblocks((long long)n * width)), dim3(kBlock), 0, s, n, perm, src,
                       dst, width);
}
void launch_scatter_scalar(hipStream_t s, int n, const int *perm, const double *src, double *dst, int width)
{
    if (((long long)n * width) > 0) hipLaunchKernelGGL(scatter_scalar_kernel, dim3(nblocks((long long)n * width)), dim3(kBlock), 0, s, n, perm,
                       src, dst, width);
}
void launch_cell_ranges(hipStream_t s, int n, const double4 *posm, const BoxDesc *boxes, const int *box_of,
                        int2 *range, int ncell)
{
    if ((2LL * ncell) > 0) hipLaunchKernelGGL(fill_int_kernel, dim3(nblocks(2LL * ncell)), dim3(kBlock), 0, s, (int *)range, 2 * ncell, 0);
    if (n > 0) hipLaunchKernelGGL(cell_ranges_kernel, dim3(nblocks(n)), dim3(kBlock), 0, s, n, posm, boxes, box_of, range);
}
void launch_box_hmax(hipStream_t s, int n_boxes, BoxDesc *boxes, const double *hsml, int n_tiles, const int2 *tiles)
{
    if (n_boxes > 0) hipLaunchKernelGGL(clear_box_hmax_kernel, dim3(nblocks(n_boxes)), dim3(kBlock), 0, s, n_boxes, boxes);
    if (n_tiles > 0) hipLaunchKernelGGL(box_hmax_kernel, dim3(n_tiles), dim3(kBlock), 0, s, boxes, hsml, tiles);
}
void launch_flag_unconverged(hipStream_t s, const int *n_live_dev, Counters *ctr)
{
    hipLaunchKernelGGL(flag_unconverged_kernel, dim3(1), dim3(64), 0, s, n_live_dev, ctr);
}
void launch_clear_counters(hipStream_t s, Counters *ctr)
{
    hipLaunchKernelGGL(clear_counters_kernel, dim3(1), dim3(64), 0, s, ctr);
}
void launch_map_active(hipStream_t s, int n_active, const int *active_caller, const int *inv, int *active_sorted)
{
    if ((n_active) > 0) hipLaunchKernelGGL(map_active_kernel, dim3(nblocks(n_active)), dim3(kBlock), 0, s, n_active, active_caller, inv,
                       active_sorted);
}
void launch_density_init(hipStream_t s, int n_active, const int *active, const double4 *posm, int *dcache_n, double *hsml,
                         double *left, double *right, double max_hsml, int *live_flag)
{
    if ((n_active) > 0) hipLaunchKernelGGL(density_init_kernel, dim3(nblocks(n_active)), dim3(kBlock), 0, s, n_active, active, posm,
                       dcache_n, hsml, left, right, max_hsml, live_flag);
}
void launch_density_sweep(hipStream_t s, const DensitySweepArgs &a)
{
    if (a.n_live <= 0) return;
    // workgroups of 512 take their own instantiation: the bound shapes the register allocation
    const int b = a.block;
    const dim3 grid(nblocks(a.n_live, b)), block(b);
    if (a.pvel) {  // MFV: one instantiation per search, 256-thread workgroups
        if (a.tree.nodes) hipLaunchKernelGGL((density_sweep_kernel<true, kBlock, true>), dim3(nblocks(a.n_live)), dim3(kBlock), 0, s, a);
        else hipLaunchKernelGGL((density_sweep_kernel<false, kBlock, true>), dim3(nblocks(a.n_live)), dim3(kBlock), 0, s, a);
    } else if (b > kBlock) {
        if (a.tree.nodes) hipLaunchKernelGGL((density_sweep_kernel<true, kMaxBlock>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((density_sweep_kernel<false, kMaxBlock>), grid, block, 0, s, a);
    } else {
        if (a.tree.nodes) hipLaunchKernelGGL((density_sweep_kernel<true, kBlock>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((density_sweep_kernel<false, kBlock>), grid, block, 0, s, a);
    }
}
void launch_edge_reset(hipStream_t s, int cap, const int *n_edge, const int *edge_list, const double *hsml_entry,
                       double *hsml, double *left, double *right)
{
    if (cap > 0)
        hipLaunchKernelGGL(edge_reset_kernel, dim3(nblocks(cap)), dim3(kBlock), 0, s, cap, n_edge, edge_list,
                           hsml_entry, hsml, left, right);
}
void launch_density_finalize(hipStream_t s, int n_active, const int *active, const double4 *posm,
                             const double4 *velu, const double *hsml, double *num_ngb, double *rho, double *pres,
                             double gm1, double *pvel)
{
    if ((n_active) > 0) hipLaunchKernelGGL(density_finalize_kernel, dim3(nblocks(n_active)), dim3(kBlock), 0, s, n_active, active, posm,
                       velu, hsml, num_ngb, rho, pres, gm1, pvel);
}
void launch_pack_grad(hipStream_t s, int n, const double4 *posm, const double4 *velu, const double *hsml,
                      const double *rho, const double *pres, GradRec *rec)
{
    if ((n) > 0) hipLaunchKernelGGL(pack_grad_kernel, dim3(nblocks(n)), dim3(kBlock), 0, s, n, posm, velu, hsml, rho, pres, rec);
}
void launch_sym_rows(hipStream_t s, const SymRowArgs &a)
{
    if (a.n <= 0) return;
    launch_fill_int(s, a.bcount, a.n, 0);
    launch_fill_int(s, a.wcount, 1, 0);
    hipLaunchKernelGGL(sym_rows_scatter_kernel, dim3(nblocks(a.n)), dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(sym_rows_wave_kernel, dim3(2048), dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(sym_rows_sort_kernel, dim3(nblocks(a.n)), dim3(kBlock), 0, s, a.n, a.brow, a.bcount);
}
void launch_gradients(hipStream_t s, const GradientArgs &a)
{
    if (a.n_active <= 0) return;
    const dim3 grid(nblocks(a.n_active, a.block)), block(a.block);
    if (a.brow) {
        if (a.block == 64) hipLaunchKernelGGL((gradients_kernel<64, true>), grid, block, 0, s, a);
        else if (a.block == 128) hipLaunchKernelGGL((gradients_kernel<128, true>), grid, block, 0, s, a);
        else if (a.block == 512) hipLaunchKernelGGL((gradients_kernel<512, true>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((gradients_kernel<kBlock, true>), grid, block, 0, s, a);
        return;
    }
    if (a.block == 64) hipLaunchKernelGGL(gradients_kernel<64>, grid, block, 0, s, a);
    else if (a.block == 128) hipLaunchKernelGGL(gradients_kernel<128>, grid, block, 0, s, a);
    else if (a.block == 512) hipLaunchKernelGGL(gradients_kernel<512>, grid, block, 0, s, a);
    else hipLaunchKernelGGL(gradients_kernel<kBlock>, grid, block, 0, s, a);
}
void launch_pack_force(hipStream_t s, int n, const double4 *posm, const double4 *velu, const double *hsml,
                       const double *rho, const double *pres, const double *dhsml, const double *cn,
                       const double *fce, const double *nvt, const double *grad, double gamma, ForceRec *rec)
{
    if ((n) > 0) hipLaunchKernelGGL(pack_force_kernel, dim3(nblocks(n)), dim3(kBlock), 0, s, n, posm, velu, hsml, rho, pres,
                       dhsml, cn, fce, nvt, grad, gamma, rec);
}
void launch_force(hipStream_t s, const ForceArgs &a)
{
    if (a.n_active <= 0) return;
    // only the three-phase list kernel takes workgroups above kBlock
    const int b = (a.list.idx && a.defer.mask) ? a.block : min(a.block, kBlock);
    const dim3 grid(nblocks(a.n_active, b)), block(b);
    if (a.list.idx) {
        if (a.defer.mask) hipLaunchKernelGGL((force_kernel<true, true>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((force_kernel<true, false>), grid, block, 0, s, a);
    } else {
        hipLaunchKernelGGL((force_kernel<false, false>), grid, block, 0, s, a);
    }
}
namespace {

// one thread per active entry: its deferred pairs q = offset + r (r-th set
// bit k of its mask) that fall in the round get their particle and neighbour
// (i, list entry k), so that the pair kernel starts from both records' indices
__global__ __launch_bounds__(kBlock) void defer_owners_kernel(ForceArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n_active) return;
    const int c = a.defer.count[t];
    if (c == 0) return;
    const int q0 = a.defer.offset[t];
    if (q0 >= a.defer.q_end || q0 + c <= a.defer.q_begin) return;
    uint64_t m = a.defer.mask[t];
    for (int r = 0; r < c && m; r++) {
        const int k = __builtin_ctzll(m);
        m &= m - 1;
        const int q = q0 + r;
        if (q >= a.defer.q_begin && q < a.defer.q_end)
            a.defer.pair[q - a.defer.q_begin] = make_int2(a.active[t], a.list.idx[(size_t)t * a.list.cap + k]);
    }
}

}  // namespace

void launch_defer_owners(hipStream_t s, const ForceArgs &a)
{
    if (a.n_active > 0) hipLaunchKernelGGL(defer_owners_kernel, dim3(nblocks(a.n_active)), dim3(kBlock), 0, s, a);
}

void launch_force_deferred(hipStream_t s, const ForceArgs &a, int n_pairs)
{
    if (n_pairs > 0) hipLaunchKernelGGL(force_deferred_kernel, dim3(nblocks(n_pairs)), dim3(kBlock), 0, s, a);
}
void launch_force_finish(hipStream_t s, const ForceArgs &a)
{
    if (a.n_active > 0) hipLaunchKernelGGL(force_finish_kernel, dim3(nblocks(a.n_active)), dim3(kBlock), 0, s, a);
}
void launch_defer_total(hipStream_t s, int n_active, const int *offset, const int *count, int *total)
{
    hipLaunchKernelGGL(defer_total_kernel, dim3(1), dim3(64), 0, s, n_active, offset, count, total);
}
void launch_density_heavy(hipStream_t s, const DensitySweepArgs &a, int max_entries)
{
    if (max_entries <= 0) return;
    // up to 2048 workgroups of 4 wavefronts stride over the entries
    const long long blocks = ((long long)max_entries + 3) / 4;
    const dim3 grid((unsigned)(blocks < 2048 ? blocks : 2048));
    if (a.pvel) hipLaunchKernelGGL(density_heavy_kernel<true>, grid, dim3(kBlock), 0, s, a);
    else hipLaunchKernelGGL(density_heavy_kernel<false>, grid, dim3(kBlock), 0, s, a);
}

namespace {
__global__ void cell_hmax_kernel(int ncell, const int2 *range, const double *hsml, float *hmax)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int2 r = range[c];
    double m = 0.0;
    for (int k = r.x; k < r.y; ++k) m = (hsml[k] > m) ? hsml[k] : m;
    hmax[c] = __double2float_ru(m);
}
}  // namespace

void launch_cell_hmax(hipStream_t s, int ncell, const int2 *range, const double *hsml, float *hmax)
{
    if (ncell > 0)
        hipLaunchKernelGGL(cell_hmax_kernel, dim3(nblocks(ncell)), dim3(kBlock), 0, s, ncell, range, hsml, hmax);
}

void launch_neighbor_counts(hipStream_t s, const BoxDesc *boxes, const int *box_of, const TreeArgs &tree, int mode,
                            int n_active, const int *active, const double4 *posm, const double *hsml, int *counts,
                            Counters *ctr)
{
    if (n_active > 0)
        hipLaunchKernelGGL(neighbor_count_kernel, dim3(nblocks(n_active)), dim3(kBlock), 0, s, boxes, box_of, tree,
                           mode, n_active, active, posm, hsml, counts, ctr);
}
void launch_fill_int(hipStream_t s, int *p, int n, int value)
{
    if ((n) > 0) hipLaunchKernelGGL(fill_int_kernel, dim3(nblocks(n)), dim3(kBlock), 0, s, p, n, value);
}
void launch_fill_u8(hipStream_t s, uint8_t *p, int n, uint8_t value)
{
    if ((n) > 0) hipLaunchKernelGGL(fill_u8_kernel, dim3(nblocks(n)), dim3(kBlock), 0, s, p, n, value);
}
void launch_fill_double(hipStream_t s, double *p, long long n, double value)
{
    if ((n) > 0) hipLaunchKernelGGL(fill_double_kernel, dim3(nblocks(n)), dim3(kBlock), 0, s, p, n, value);
}
void launch_set_flags(hipStream_t s, int n_active, const int *active, uint8_t *flags)
{
    if ((n_active) > 0) hipLaunchKernelGGL(set_flags_kernel, dim3(nblocks(n_active)), dim3(kBlock), 0, s, n_active, active, flags);
}

}  // namespace gzm
