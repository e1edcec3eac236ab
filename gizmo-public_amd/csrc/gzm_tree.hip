// gzm_tree.hip -- neighbour-search tree for strongly clustered particle sets.
//
// The cell grid sizes its cells for the mean smoothing length of a box; where
// hsml spans orders of magnitude (a Plummer-sphere cluster: dense cores in a
// sparse halo) the cells of the core hold thousands of particles each and
// every core particle tests them all.  The tree adapts to the density the way
// the reference's octree does (forcetree.c, walked by ngb.c:102-130), built
// for the GPU from the Morton-sorted particles in four data-parallel passes:
//
//   1. keys:   box index above a 63-bit Morton code of the position
//              (21 bits per axis for a single box);
//   2. build:  one thread per internal node of the binary radix tree over the
//              sorted keys, each finding its key range and split from the
//              common-prefix lengths of its neighbours (Karras, "Maximizing
//              Parallelism in the Construction of BVHs, Octrees, and k-d
//              Trees", HPG 2012, section 4); equal keys are split by slot;
//   3. links:  the stackless walk's escape link of every node: the node that
//              follows a subtree in depth-first order is the right child of
//              the node whose split lies at the subtree's last slot;
//   4. bounds: the boxes and hsml maxima of nodes of at most 32 slots from
//              their slot ranges, then those of the larger nodes bottom-up,
//              one launch per level: a node combines the children that
//              earlier launches completed.  No fences: on a chip whose XCDs
//              each have their own L2, a device-scope fence writes back and
//              invalidates L2, and a chain of them per node cost 63 ms for
//              8.4 M particles.
//
// The node hmax changes with every density stage; it is refreshed without
// the bottom-up pass:
// every node covers a contiguous slot range, whose maximum hsml is the larger
// of two overlapping power-of-two windows of a window-maximum table built in
// log2(n) data-parallel levels.  No pass needs a host synchronisation, so
// the refresh can sit in a captured graph.
#include "grid_device.h"
#include "gzm_internal.h"

namespace gzm {

namespace {

__device__ __forceinline__ uint64_t spread21(uint64_t v)
{
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

__device__ __forceinline__ uint64_t quantise(double x, double L, int bits)
{
    const double q = floor(x / L * (double)(1ull << bits));
    const double top = (double)((1ull << bits) - 1);
    return (uint64_t)(q < 0 ? 0.0 : (q > top ? top : q));
}

__global__ void tree_keys_kernel(int n, const double *pos, const BoxDesc *boxes, const int *box_caller, int bits,
                                 uint64_t *keys, int *iota)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int b = box_caller ? box_caller[i] : 0;
    const BoxInfo &B = boxes[b].box;
    const uint64_t m = (spread21(quantise(pos[3 * i + 0], B.L[0], bits)) << 2) |
                       (spread21(quantise(pos[3 * i + 1], B.L[1], bits)) << 1) |
                       spread21(quantise(pos[3 * i + 2], B.L[2], bits));
    keys[i] = ((uint64_t)b << (3 * bits)) | m;
    iota[i] = i;
}

// common-prefix length of the keys at local slots i and j of a box of m
// particles starting at slot `first`; -1 outside the box; equal keys compare
// their slot numbers
__device__ __forceinline__ int prefix(const uint64_t *keys, int first, int m, int i, int j)
{
    if (j < 0 || j >= m) return -1;
    const uint64_t a = keys[first + i], b = keys[first + j];
    if (a != b) return __clzll((long long)(a ^ b));
    return 64 + __clz(i ^ j);
}

__device__ __forceinline__ int leaf_id(int slot) { return -slot - 1; }

__global__ void tree_build_kernel(TreeBuild t)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= t.n) return;
    const BoxDesc &B = t.boxes[t.box_of[s]];
    const int first = B.first, m = B.count, i = s - first;
    if (m == 1) return;
    if (i >= m - 1) return;  // m - 1 internal nodes per box
    const int g0 = B.root;
    const uint64_t *k = t.keys;
    // direction of the node's range and the prefix length it must exceed
    const int d = (prefix(k, first, m, i, i + 1) > prefix(k, first, m, i, i - 1)) ? 1 : -1;
    const int dmin = prefix(k, first, m, i, i - d);
    int lmax = 2;
    while (prefix(k, first, m, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int step = lmax >> 1; step >= 1; step >>= 1)
        if (prefix(k, first, m, i, i + (l + step) * d) > dmin) l += step;
    const int j = i + l * d;
    // split: the last slot sharing more than the node's prefix with slot i
    const int dnode = prefix(k, first, m, i, j);
    int sp = 0, step = l;
    do {
        step = (step + 1) >> 1;
        if (prefix(k, first, m, i, i + (sp + step) * d) > dnode) sp += step;
    } while (step > 1);
    const int gamma = i + sp * d + (d < 0 ? -1 : 0);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    const int g = g0 + i;
    const int left = (lo == gamma) ? leaf_id(first + gamma) : g0 + gamma;
    const int right = (hi == gamma + 1) ? leaf_id(first + gamma + 1) : g0 + gamma + 1;
    t.child[g] = make_int2(left, right);
    t.range[g] = make_int2(first + lo, first + hi);
    t.split_right[first + gamma] = right;
}

__global__ void tree_links_kernel(TreeBuild t)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= t.n) return;
    const BoxDesc &B = t.boxes[t.box_of[s]];
    const int first = B.first, m = B.count, i = s - first;
    const int last_slot = first + m - 1;
    t.leaf_esc[s] = (s == last_slot) ? kTreeEnd : t.split_right[s];
    if (i < m - 1) {
        const int g = B.root + i;
        const int last = t.range[g].y;
        t.esc[g] = (last == last_slot) ? kTreeEnd : t.split_right[last];
        t.flag[g] = 0;
    }
}

// Node boxes and the max hsml below each node, bottom-up without fences.
// The flag of a node records the launch that completed it (0: not yet); a
// node combines only children completed by an earlier launch, whose boxes the
// launch boundary has made visible on every XCD.

// nodes of at most kScanSlots slots take their box straight from their
// contiguous slot range; the larger ones are combined from their children
constexpr int kScanSlots = 32;
// combine launches: more than the height of the tallest tree (a node's prefix
// length grows on every level down and is below 64 key bits + 32 slot bits)
constexpr int kCombineLaunches = 100;
constexpr int kScanLaunch = 1;

__global__ void tree_bounds_scan_kernel(TreeBuild t, const double *hsml)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    int g = -1;
    bool large = false;
    if (s < t.n) {
        const BoxDesc &B = t.boxes[t.box_of[s]];
        const int i = s - B.first;
        if (i < B.count - 1) {
            g = B.root + i;
            const int2 r = t.range[g];
            large = r.y - r.x + 1 > kScanSlots;
            if (!large) {
                double4 p = t.posm[r.x];
                double4 lo = make_double4(p.x, p.y, p.z, hsml[r.x]);
                double4 hi = make_double4(p.x, p.y, p.z, 0.0);
                for (int q = r.x + 1; q <= r.y; q++) {
                    p = t.posm[q];
                    const double h = hsml[q];
                    lo = make_double4(fmin(lo.x, p.x), fmin(lo.y, p.y), fmin(lo.z, p.z), h > lo.w ? h : lo.w);
                    hi = make_double4(fmax(hi.x, p.x), fmax(hi.y, p.y), fmax(hi.z, p.z), 0.0);
                }
                t.lo[g] = lo;
                t.hi[g] = hi;
                t.flag[g] = kScanLaunch;
            }
        }
    }
    // append the large nodes to the list, one counter update per wavefront
    const uint64_t mask = __ballot(large);
    if (mask == 0) return;
    const int lane = __lane_id();
    const int leader = __ffsll((unsigned long long)mask) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(t.large_count, __popcll(mask));
    base = __shfl(base, leader);
    if (large) t.large[base + __popcll(mask & ((1ull << lane) - 1))] = g;
}

__global__ void tree_bounds_combine_kernel(TreeBuild t, const double *hsml, int launch)
{
    const int count = *t.large_count;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < count; q += gridDim.x * blockDim.x) {
        const int g = t.large[q];
        if (__hip_atomic_load(&t.flag[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) continue;
        const int2 c = t.child[g];
        const int ch[2] = {c.x, c.y};
        bool ready = true;
#pragma unroll
        for (int k = 0; k < 2; k++) {
            if (ch[k] >= 0) {
                const int f = __hip_atomic_load(&t.flag[ch[k]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ready = ready && f != 0 && f != launch;
            }
        }
        if (!ready) continue;
        double4 lo[2], hi[2];
#pragma unroll
        for (int k = 0; k < 2; k++) {
            if (ch[k] < 0) {
                const int j = -ch[k] - 1;
                const double4 p = t.posm[j];
                lo[k] = make_double4(p.x, p.y, p.z, hsml[j]);
                hi[k] = make_double4(p.x, p.y, p.z, 0.0);
            } else {
                lo[k] = t.lo[ch[k]];
                hi[k] = t.hi[ch[k]];
            }
        }
        const double hmax = lo[0].w > lo[1].w ? lo[0].w : lo[1].w;
        t.lo[g] = make_double4(fmin(lo[0].x, lo[1].x), fmin(lo[0].y, lo[1].y), fmin(lo[0].z, lo[1].z), hmax);
        t.hi[g] = make_double4(fmax(hi[0].x, hi[1].x), fmax(hi[0].y, hi[1].y), fmax(hi[0].z, hi[1].z), 0.0);
        __hip_atomic_store(&t.flag[g], launch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// level 0 of the window-maximum table: hsml rounded up to FP32
__global__ void window_max_base_kernel(int n, const double *hsml, float *m0)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) m0[s] = __double2float_ru(hsml[s]);
}

// level k from level k - 1: windows of 2^k slots (cut at n)
__global__ void window_max_level_kernel(int n, int half, const float *prev, float *next)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const float a = prev[s];
    const float b = (s + half < n) ? prev[s + half] : a;
    next[s] = fmaxf(a, b);
}

// hmax of every internal node: two overlapping windows cover its slot range
__global__ void node_hmax_kernel(TreeBuild t)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= t.n) return;
    const BoxDesc &B = t.boxes[t.box_of[s]];
    const int i = s - B.first;
    if (i >= B.count - 1) return;
    const int g = B.root + i;
    const int2 r = t.range[g];
    const int len = r.y - r.x + 1;
    const int k = 31 - __clz(len);
    const float *m = t.window_max + (size_t)k * t.n;
    t.nodes[g].lo.w = fmaxf(m[r.x], m[r.y - (1 << k) + 1]);
}

// the walk's 48-byte records from the build's arrays (box rounded outwards)
__global__ void pack_nodes_kernel(TreeBuild t)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= t.n) return;
    const BoxDesc &B = t.boxes[t.box_of[s]];
    const int i = s - B.first;
    if (i >= B.count - 1) return;
    const int g = B.root + i;
    const double4 lo = t.lo[g], hi = t.hi[g];
    const int2 r = t.range[g];
    TreeNode nd;
    nd.lo = make_float4(__double2float_rd(lo.x), __double2float_rd(lo.y), __double2float_rd(lo.z),
                        __double2float_ru(lo.w));
    nd.hi = make_float4(__double2float_ru(hi.x), __double2float_ru(hi.y), __double2float_ru(hi.z),
                        __int_as_float(t.esc[g]));
    nd.link = make_int4(t.child[g].x, r.x, r.y, 0);
    t.nodes[g] = nd;
}

inline int blocks(int n) { return (n + kBlock - 1) / kBlock; }

}  // namespace

void launch_tree_keys(hipStream_t s, int n, const double *pos, const BoxDesc *boxes, const int *box_caller,
                      int bits_per_axis, uint64_t *keys, int *iota)
{
    if (n > 0)
        hipLaunchKernelGGL(tree_keys_kernel, dim3(blocks(n)), dim3(kBlock), 0, s, n, pos, boxes, box_caller,
                           bits_per_axis, keys, iota);
}

void launch_tree_build(hipStream_t s, const TreeBuild &b, const double *hsml)
{
    if (b.n <= 0) return;
    hipLaunchKernelGGL(tree_build_kernel, dim3(blocks(b.n)), dim3(kBlock), 0, s, b);
    hipLaunchKernelGGL(tree_links_kernel, dim3(blocks(b.n)), dim3(kBlock), 0, s, b);
    launch_fill_int(s, b.large_count, 1, 0);
    hipLaunchKernelGGL(tree_bounds_scan_kernel, dim3(blocks(b.n)), dim3(kBlock), 0, s, b, hsml);
    const int combine_blocks = std::min(blocks(b.n), 1024);
    for (int launch = kScanLaunch + 1; launch <= kScanLaunch + kCombineLaunches; launch++)
        hipLaunchKernelGGL(tree_bounds_combine_kernel, dim3(combine_blocks), dim3(kBlock), 0, s, b, hsml, launch);
    hipLaunchKernelGGL(pack_nodes_kernel, dim3(blocks(b.n)), dim3(kBlock), 0, s, b);
}

void launch_tree_hmax(hipStream_t s, const TreeBuild &b, const double *hsml)
{
    if (b.n <= 0) return;
    const dim3 grid(blocks(b.n)), block(kBlock);
    hipLaunchKernelGGL(window_max_base_kernel, grid, block, 0, s, b.n, hsml, b.window_max);
    for (int k = 1; k < b.levels; k++)
        hipLaunchKernelGGL(window_max_level_kernel, grid, block, 0, s, b.n, 1 << (k - 1),
                           b.window_max + (size_t)(k - 1) * b.n, b.window_max + (size_t)k * b.n);
    hipLaunchKernelGGL(node_hmax_kernel, grid, block, 0, s, b);
}

}  // namespace gzm
