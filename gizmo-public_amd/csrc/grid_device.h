// grid_device.h -- device helpers shared by the stage kernels: launch
// geometry, the cell-grid walk, the periodic wrap and wavefront reductions.
#pragma once

#include <type_traits>

#include "gzm_internal.h"

namespace gzm {

// Call a walk's visitor; a visitor returning bool stops the walk with false.
template <typename F, typename... A>
__device__ __forceinline__ bool visit_more(F &f, A... args)
{
    if constexpr (std::is_void_v<decltype(f(args...))>) {
        f(args...);
        return true;
    } else {
        return f(args...);
    }
}

constexpr int kBlock = 256;
// largest workgroup of the per-particle loop kernels (density sweeps,
// gradients, the pair-list force loops), whose size gzm_params.block_size
// picks from 64, 128, 256 and 512 per stage.  The scanning force kernels stay
// at 256: a bound of 512 threads makes the compiler fit their 256 VGPRs to two
// waves per SIMD, with 3-5x the spills.
constexpr int kMaxBlock = 512;
constexpr int kXcds = 8;         // MI355X: 8 XCDs, each with its own L2
constexpr int kListChunk = 16;   // pair-list entries a lane stages before one 64-byte store

// Workgroups are dealt to the XCDs round-robin (block b runs on XCD b % 8).
// Renumber them so that each XCD receives one contiguous run of the
// Morton-ordered particles: its L2 then holds the cells its waves share,
// instead of every eighth block of the whole box.  A bijection on
// [0, gridDim.x) for any grid size.
__device__ __forceinline__ int xcd_block()
{
    const int nb = (int)gridDim.x, b = (int)blockIdx.x;
    const int q = nb / kXcds, r = nb % kXcds, x = b % kXcds;
    return x * q + (x < r ? x : r) + b / kXcds;
}

__device__ __forceinline__ int cell_coord(double x, double inv_s, int nc)
{
    int c = (int)floor(x * inv_s);
    if (c < 0) c = 0;
    if (c >= nc) c = nc - 1;
    return c;
}

__device__ __forceinline__ uint32_t spread10(uint32_t v)
{
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__device__ __forceinline__ uint32_t morton3(int cx, int cy, int cz)
{
    return (spread10((uint32_t)cx) << 2) | (spread10((uint32_t)cy) << 1) | spread10((uint32_t)cz);
}

__device__ __forceinline__ double axis_gap(double x, double c0, double s)
{
    double g = c0 - x;
    double g2 = x - (c0 + s);
    g = (g > g2) ? g : g2;
    return (g > 0) ? g : 0.0;
}

// Visit every cell that may hold a particle within distance R of (x, y, z),
// periodic; cells whose box lies farther than R are skipped.  f(cx, cy, cz,
// begin, end) receives the cell's unwrapped coordinates (outside [0, nc) when
// the walk crosses the periodic boundary; the image of a particle of that cell
// lies at x + (c - c_wrapped) * s) and its sorted slot range.  `full` reports
// the axes walked whole because R spans half the box.  An f returning bool
// ends the walk by returning false.  With h_pair >= 0 (pair searches: R is
// the box's max hsml) a cell farther than h_pair is visited only when its own
// max hsml reaches x: no particle of it can pair with x otherwise.
template <typename F>
__device__ __forceinline__ void for_each_cell(const Grid &g, const double xs[3], double R, bool full[3], F &&f,
                                              double h_pair = -1.0)
{
    int lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        // a radius that spans the axis (or is not finite) walks the whole axis once
        full[a] = !(R * g.inv_s[a] < 0.5 * g.nc[a] - 2.0);
        if (!full[a]) {
            lo[a] = (int)floor((xs[a] - R) * g.inv_s[a] - 1e-6);
            hi[a] = (int)floor((xs[a] + R) * g.inv_s[a] + 1e-6);
        } else {
            lo[a] = 0;
            hi[a] = g.nc[a] - 1;
        }
    }
    const double R2 = R * R * (1.0 + 1e-9);
    const double hp2 = h_pair * h_pair * (1.0 + 1e-9);
    for (int cx = lo[0]; cx <= hi[0]; ++cx) {
        const int wx = (cx < 0) ? cx + g.nc[0] : ((cx >= g.nc[0]) ? cx - g.nc[0] : cx);
        const double gx = full[0] ? 0.0 : axis_gap(xs[0], cx * g.s[0], g.s[0]);
        const double gx2 = gx * gx;
        if (gx2 > R2) continue;
        for (int cy = lo[1]; cy <= hi[1]; ++cy) {
            const int wy = (cy < 0) ? cy + g.nc[1] : ((cy >= g.nc[1]) ? cy - g.nc[1] : cy);
            const double gy = full[1] ? 0.0 : axis_gap(xs[1], cy * g.s[1], g.s[1]);
            const double gxy2 = gx2 + gy * gy;
            if (gxy2 > R2) continue;
            const int rowbase = (wx * g.nc[1] + wy) * g.nc[2];
            for (int cz = lo[2]; cz <= hi[2]; ++cz) {
                const int wz = (cz < 0) ? cz + g.nc[2] : ((cz >= g.nc[2]) ? cz - g.nc[2] : cz);
                const double gz = full[2] ? 0.0 : axis_gap(xs[2], cz * g.s[2], g.s[2]);
                const double d2 = gxy2 + gz * gz;
                if (d2 > R2) continue;
                if (h_pair >= 0 && d2 > hp2) {
                    const double hm = g.hmax[rowbase + wz];
                    if (d2 > hm * hm * (1.0 + 1e-9)) continue;
                }
                const int2 r = g.range[rowbase + wz];
                if (!visit_more(f, cx, cy, cz, r.x, r.y)) return;
            }
        }
    }
}

// Visit every sorted slot k whose cell may hold a particle within distance R
// of (x, y, z), periodic.
template <typename F>
__device__ __forceinline__ void for_each_candidate(const Grid &g, const double xs[3], double R, F &&f,
                                                   double h_pair = -1.0)
{
    bool full[3];
    for_each_cell(g, xs, R, full, [&](int, int, int, int begin, int end) {
        for (int k = begin; k < end; ++k)
            if (!visit_more(f, k)) return false;
        return true;
    }, h_pair);
}

__device__ __forceinline__ void wrap_dp(const BoxInfo &box, double dp[3])
{
    dp[1] = wrap_axis(dp[1], box.L[1], box.half[1]);
    dp[0] = wrap_axis(dp[0], box.L[0], box.half[0]);
    dp[2] = wrap_axis(dp[2], box.L[2], box.half[2]);
}

// periodic distance from x to the interval [lo, hi] (both inside [0, L))
__device__ __forceinline__ double periodic_gap(double x, double lo, double hi, double L)
{
    double g = lo - x;
    const double g2 = x - hi;
    g = (g > g2) ? g : g2;
    g = (g > 0) ? g : 0.0;
    const double a = x + L - hi, b = lo + L - x;  // the images x + L, x - L
    g = (a < g) ? a : g;
    g = (b < g) ? b : g;
    return g;
}

// Visit (f(slot)) every particle of box G whose position may lie within R of
// x -- and, for the pair search, within the hsml of a particle below the
// node -- in depth-first Morton order.  Conservative by the same 1e-9
// relative margin as the cell walk; the caller applies the exact predicate.
// An f returning bool ends the walk by returning false.  One step per node or
// leaf (the walk the pair searches of the gradient and force loops use: their
// tree path is the rare one there, and the form below costs them registers).
template <typename F>
__device__ __forceinline__ void tree_walk(const TreeArgs &T, const BoxDesc &G, const double x[3], double R,
                                          bool pairs, F &&f)
{
    int node = G.root;
    // a well-formed tree visits each of its 2 count - 1 nodes at most once
    for (int steps = 2 * G.count; node != kTreeEnd && steps > 0; steps--) {
        if (node < 0) {
            const int j = -node - 1;
            if (!visit_more(f, j)) return;
            node = T.leaf_esc[j];
            continue;
        }
        const TreeNode nd = T.nodes[node];
        const double hm = nd.lo.w;
        const double Rn = (pairs && hm > R) ? hm : R;
        const double gx = periodic_gap(x[0], nd.lo.x, nd.hi.x, G.box.L[0]);
        const double gy = periodic_gap(x[1], nd.lo.y, nd.hi.y, G.box.L[1]);
        const double gz = periodic_gap(x[2], nd.lo.z, nd.hi.z, G.box.L[2]);
        const int esc = __float_as_int(nd.hi.w);
        if (!(gx * gx + gy * gy + gz * gz <= Rn * Rn * (1.0 + 1e-9))) {
            node = esc;
            continue;
        }
        if (nd.link.z - nd.link.y < kTreeBucket) {
            for (int j = nd.link.y; j <= nd.link.z; j++)
                if (!visit_more(f, j)) return;
            node = esc;
        } else {
            node = nd.link.x;
        }
    }
}

// The same visits as tree_walk (radius R only: the density search), in two
// nested loops ("while-while"): each lane walks nodes until it reaches a slot
// range to scan (a bucket or a single leaf), and the lanes then scan their
// ranges together.  In tree_walk's single loop a wavefront's lanes stand at
// nodes and buckets in no common rhythm, so nearly every step pays for a
// bucket scan (the candidate loop with its pair body) with the few lanes that
// happen to be at a bucket; here every unfinished lane is in the scan.  Each
// lane visits its candidates in the same order either way.
template <typename F>
__device__ __forceinline__ void tree_walk_grouped(const TreeArgs &T, const BoxDesc &G, const double x[3], double R,
                                                  F &&f)
{
    int node = G.root;
    int steps = 2 * G.count;  // as in tree_walk
    for (;;) {
        int beg = 0, end = -1;
        while (node != kTreeEnd && steps > 0) {
            steps--;
            if (node < 0) {
                beg = end = -node - 1;
                node = T.leaf_esc[beg];
                break;
            }
            const TreeNode nd = T.nodes[node];
            const double gx = periodic_gap(x[0], nd.lo.x, nd.hi.x, G.box.L[0]);
            const double gy = periodic_gap(x[1], nd.lo.y, nd.hi.y, G.box.L[1]);
            const double gz = periodic_gap(x[2], nd.lo.z, nd.hi.z, G.box.L[2]);
            const int esc = __float_as_int(nd.hi.w);
            if (!(gx * gx + gy * gy + gz * gz <= R * R * (1.0 + 1e-9))) {
                node = esc;
                continue;
            }
            if (nd.link.z - nd.link.y < kTreeBucket) {
                beg = nd.link.y;
                end = nd.link.z;
                node = esc;
                break;
            }
            node = nd.link.x;
        }
        if (beg > end) return;
        for (int j = beg; j <= end; j++)
            if (!visit_more(f, j)) return;
    }
}

// Every sorted slot that may hold a neighbour of x: the cell grid within R
// (pair search: R = max(h_i, box hmax)), or the tree (R = h_i, node hmax).
template <typename F>
__device__ __forceinline__ void for_each_neighbour_candidate(const TreeArgs &T, const BoxDesc &G,
                                                             const double x[3], double h_i, bool pairs, F &&f)
{
    if (T.nodes) {
        tree_walk(T, G, x, h_i, pairs, f);
    } else {
        const double R = (pairs && G.hmax > h_i) ? G.hmax : h_i;
        for_each_candidate(G.grid, x, R, f, pairs ? h_i : -1.0);
    }
}

// Every candidate of one particle (search radius R around x), spread over the
// 64 lanes of a wavefront that walks the tree or the grid in step: each lane
// receives its share, f(k).  For the few particles whose search covers far
// more candidates than a lane can walk alone.  h_pair >= 0 (the pair search,
// R = max(h_i, box hsml max), h_pair = h_i): cells and nodes are also culled
// by the largest hsml they hold, as in the pair walks.
template <typename F>
__device__ __forceinline__ void wave_candidates(const TreeArgs &T, const BoxDesc &G, const double x[3], double R,
                                                int lane, F &&f, double h_pair = -1.0)
{
    if (T.nodes) {
        int node = G.root;
        for (int steps = 2 * G.count; node != kTreeEnd && steps > 0; steps--) {
            if (node < 0) {
                const int j = -node - 1;
                if (lane == 0) f(j);
                node = T.leaf_esc[j];
                continue;
            }
            const TreeNode nd = T.nodes[node];
            const double hm = nd.lo.w;
            const double Rn = (h_pair < 0) ? R : ((hm > h_pair) ? hm : h_pair);
            const double gx = periodic_gap(x[0], nd.lo.x, nd.hi.x, G.box.L[0]);
            const double gy = periodic_gap(x[1], nd.lo.y, nd.hi.y, G.box.L[1]);
            const double gz = periodic_gap(x[2], nd.lo.z, nd.hi.z, G.box.L[2]);
            const int esc = __float_as_int(nd.hi.w);
            if (!(gx * gx + gy * gy + gz * gz <= Rn * Rn * (1.0 + 1e-9))) {
                node = esc;
                continue;
            }
            if (nd.link.z - nd.link.y < 64) {
                if (nd.link.y + lane <= nd.link.z) f(nd.link.y + lane);
                node = esc;
            } else {
                node = nd.link.x;
            }
        }
    } else {
        bool full[3];
        for_each_cell(G.grid, x, R, full, [&](int, int, int, int beg, int end) {
            for (int k = beg + lane; k < end; k += 64) f(k);
        }, h_pair);
    }
}

// The pair-search candidates of one particle (R = max(h_i, box hsml max),
// cells and nodes culled by their hsml max against h_pair = h_i, as in the pair
// walks) spread over the 64 lanes of a wavefront: in the grid each lane takes
// whole cells of the search range in turn (at the usual cell sizes a cell holds
// a particle or two, so lanes sharing one cell would mostly idle), in the tree
// the lanes share each bucket.
template <typename F>
__device__ __forceinline__ void wave_pair_candidates(const TreeArgs &T, const BoxDesc &G, const double x[3],
                                                     double R, double h_pair, int lane, F &&f)
{
    if (T.nodes) {
        wave_candidates(T, G, x, R, lane, f, h_pair);
        return;
    }
    const Grid &g = G.grid;
    int lo[3], n[3];
    bool full[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        full[a] = !(R * g.inv_s[a] < 0.5 * g.nc[a] - 2.0);
        if (!full[a]) {
            lo[a] = (int)floor((x[a] - R) * g.inv_s[a] - 1e-6);
            n[a] = (int)floor((x[a] + R) * g.inv_s[a] + 1e-6) - lo[a] + 1;
        } else {
            lo[a] = 0;
            n[a] = g.nc[a];
        }
    }
    const double R2 = R * R * (1.0 + 1e-9);
    const double hp2 = h_pair * h_pair * (1.0 + 1e-9);
    const int total = n[0] * n[1] * n[2];
    for (int c = lane; c < total; c += 64) {
        const int cz = lo[2] + c % n[2];
        const int cy = lo[1] + (c / n[2]) % n[1];
        const int cx = lo[0] + c / (n[2] * n[1]);
        const int wx = (cx < 0) ? cx + g.nc[0] : ((cx >= g.nc[0]) ? cx - g.nc[0] : cx);
        const int wy = (cy < 0) ? cy + g.nc[1] : ((cy >= g.nc[1]) ? cy - g.nc[1] : cy);
        const int wz = (cz < 0) ? cz + g.nc[2] : ((cz >= g.nc[2]) ? cz - g.nc[2] : cz);
        const double gx = full[0] ? 0.0 : axis_gap(x[0], cx * g.s[0], g.s[0]);
        const double gy = full[1] ? 0.0 : axis_gap(x[1], cy * g.s[1], g.s[1]);
        const double gz = full[2] ? 0.0 : axis_gap(x[2], cz * g.s[2], g.s[2]);
        const double d2 = gx * gx + gy * gy + gz * gz;
        if (d2 > R2) continue;
        const int cell = (wx * g.nc[1] + wy) * g.nc[2] + wz;
        if (d2 > hp2) {
            const double hm = g.hmax[cell];
            if (d2 > hm * hm * (1.0 + 1e-9)) continue;
        }
        const int2 r = g.range[cell];
        for (int k = r.x; k < r.y; ++k) f(k);
    }
}

__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v)
{
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = (o > v) ? o : v;
    }
    return v;
}

__device__ __forceinline__ double wave_min(double v)
{
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = (o < v) ? o : v;
    }
    return v;
}

// Run f(desc) for this lane's particle with its box descriptor, one
// wave-uniform box at a time: the box index is taken from the first lane still
// to run (readfirstlane), so the descriptor's fields are uniform and a
// wavefront that straddles two boxes runs the body twice with the lanes of
// each box.
template <typename F>
__device__ __forceinline__ void for_box(const BoxDesc *__restrict__ boxes, const int *box_of, int i, bool valid,
                                        F &&f)
{
    bool todo = valid;
    const int my_b = valid ? box_of[i] : 0;
    while (todo) {
        const int b = __builtin_amdgcn_readfirstlane(my_b);
        if (my_b == b) {
            f(boxes[b]);
            todo = false;
        }
    }
}

// hydro_final_operations_and_cleanup (hydro_toplevel.c:738-764) on the raw
// sums of one particle: dE -= v . a before the mass division
__device__ __forceinline__ void force_final_ops(double acc[3], double &dte, const double v[3], double m_i,
                                                double h_i, double max_hsml)
{
    dte -= (v[0] / 1.0) * acc[0];
    acc[0] /= m_i;
    dte -= (v[1] / 1.0) * acc[1];
    acc[1] /= m_i;
    dte -= (v[2] / 1.0) * acc[2];
    acc[2] /= m_i;
    dte /= m_i;
    if (h_i >= 0.99 * max_hsml) dte = 0;
}

// wavefront-wide sum, then one atomic per wavefront
__device__ __forceinline__ void wave_add_u64(unsigned long long *dst, unsigned long long v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(dst, v);
}

// MINMAX_CHECK of the reference (else-if: a new minimum is not also tested
// as a maximum, gradients.c:29).  Written with selects: as two conditional
// stores the compiler may merge them into one store through a selected
// address, which moves the accumulators to scratch memory.
template <typename T>
__device__ __forceinline__ void minmax_check(T x, T &mn, T &mx)
{
    const bool lt = x < mn;
    const bool gt = !lt && (x > mx);
    mn = lt ? x : mn;
    mx = gt ? x : mx;
}

// One lane's staged list chunk (column threadIdx.x of `stage`, rows `stride`
// ints apart: the workgroup size) -> 64 bytes of its row.
__device__ __forceinline__ void flush_list_chunk(const int *stage, int *dst, int stride)
{
    int4 *d = reinterpret_cast<int4 *>(dst);
#pragma unroll
    for (int q = 0; q < kListChunk / 4; q++)
        d[q] = make_int4(stage[(4 * q + 0) * stride + threadIdx.x], stage[(4 * q + 1) * stride + threadIdx.x],
                         stage[(4 * q + 2) * stride + threadIdx.x], stage[(4 * q + 3) * stride + threadIdx.x]);
}

}  // namespace gzm
