// gzm_internal.h -- device data layout and kernel launch interface shared by
// gzm_kernels.hip (kernels) and gzm_api.hip (context, memory, C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mfm_device.h"

namespace gzm {

// Cell grid over the periodic box.  Particles are stored sorted by the Morton
// key of their cell, so each cell is one contiguous slot range; `range` is
// indexed by the linear cell id (cx*ny + cy)*nz + cz.
struct Grid {
    int nc[3];
    double inv_s[3];  // cells per unit length
    double s[3];      // cell side
    const int2 *range;
    const float *hmax;  // per cell: max hsml of its particles, rounded up (pair searches)
};

// One periodic box of a context.  A context holds one box, or many
// independent boxes (batched pass) whose particles are sorted box by box; the
// kernels read the descriptor of a particle's box (box_of[slot]) for its grid,
// periodic wrap and search radius.
struct BoxDesc {
    Grid grid;             // range points at the box's first cell
    BoxInfo box;           // side lengths and half side lengths
    double hmax;           // max hsml of the box (the pair loops' search radius)
    float s_f[3];          // cell side in FP32
    int cell_base;         // first cell of the box in the global range array
    int first, count;      // sorted slot range of the box
    int root;              // tree search: root node of the box (kTreeEnd if empty); with two or
                           // more particles an internal node, the first of the box's count - 1
};

// Tree search (gzm_params.search = 1): a bounding-volume hierarchy per box
// over the particles sorted by fine Morton key (binary radix tree from
// adjacent key prefixes), walked without a stack through "escape" links:
// from a node whose box is too far, continue at the node that follows its
// subtree in depth-first order.  Node ids: internal nodes >= 0, leaf (sorted
// slot) s as -(s + 1), end of walk kTreeEnd.  A node of at most kTreeBucket
// particles is not descended: its sorted slot range is scanned directly.
// The per-node hmax is the radius test of the pair search, as the octree
// node hmax of force_update_hmax (forcetree_update.c:348-476).
constexpr int kTreeEnd = 0x7fffffff;
constexpr int kTreeBucket = 8;

// What the walk reads of an internal node, in one 48-byte record (one
// dependent load per visited node): its box rounded outwards to FP32, its
// hsml maximum rounded up, the escape link, the left child and the slot
// range.  Rounding only enlarges the box, so the opening test stays
// conservative; the exact predicates decide every pair.
struct alignas(16) TreeNode {
    float4 lo;             // x, y, z rounded down; w = max hsml below the node, rounded up
    float4 hi;             // x, y, z rounded up; w = escape link (int bits)
    int4 link;             // left child, first slot, last slot, unused
};

struct TreeArgs {
    const TreeNode *nodes = nullptr;  // internal nodes; null: grid search
    const int *leaf_esc = nullptr;    // sorted slot: next node after the leaf
};

// Per-particle record read by the force loop for the neighbour j (and for i).
// 320 bytes: one aligned block so a gather touches 2.5 cache lines.
struct alignas(64) ForceRec {
    double4 xh;    // x, y, z, hsml
    double4 vc;    // vx, vy, vz, sound speed sqrt(gamma P / rho)
    double4 th;    // rho, P, V = m / rho, DhsmlNgbFactor
    double4 mc;    // m, ConditionNumber, FaceClosureError, P / rho
    double4 t0;    // NV_T: i00 i01 i02 i11
    double4 t1;    // NV_T: i12 i22, 1 / hsml, 0
    double4 g0;    // grad rho (3), grad P x
    double4 g1;    // grad P y, z, grad v row 0 (x, y)
    double4 g2;    // grad v row 0 z, row 1 (x, y, z)
    double4 g3;    // grad v row 2 (x, y, z), 0
};

// ForceRec -> the pair-side quantities (moves only)
__device__ __forceinline__ void unpack_side(const ForceRec &r, PairSide<double> &S)
{
    S.v[0] = r.vc.x; S.v[1] = r.vc.y; S.v[2] = r.vc.z; S.cs = r.vc.w;
    S.h = r.xh.w; S.hinv = r.t1.z;
    S.rho = r.th.x; S.P = r.th.y; S.V = r.th.z; S.dhsml = r.th.w;
    S.cn = r.mc.y; S.fce = r.mc.z; S.e = r.mc.w;
    S.nvt[0] = r.t0.x; S.nvt[1] = r.t0.y; S.nvt[2] = r.t0.z; S.nvt[3] = r.t0.w;
    S.nvt[4] = r.t1.x; S.nvt[5] = r.t1.y;
    S.gr[0] = r.g0.x; S.gr[1] = r.g0.y; S.gr[2] = r.g0.z;
    S.gp[0] = r.g0.w; S.gp[1] = r.g1.x; S.gp[2] = r.g1.y;
    S.gv[0] = r.g1.z; S.gv[1] = r.g1.w; S.gv[2] = r.g2.x;
    S.gv[3] = r.g2.y; S.gv[4] = r.g2.z; S.gv[5] = r.g2.w;
    S.gv[6] = r.g3.x; S.gv[7] = r.g3.y; S.gv[8] = r.g3.z;
}

// Per-particle record read by the gradient loop for the neighbour j.
struct alignas(32) GradRec {
    double4 xh;    // x, y, z, hsml
    double4 vr;    // vx, vy, vz, rho
    double2 pm;    // P, m
};

// ---------------------------------------------------------------------------
// FP32-bulk records.  A position is stored as the packed integer coordinates
// of its grid cell plus an FP32 offset from that cell's FP64 origin
// (x = c * s + off), so a pair displacement is (c_i - c_j) * s +
// (off_i - off_j): exact cell arithmetic plus an FP32 term of the size of a
// cell, whatever the box size.
// ---------------------------------------------------------------------------
struct F32Geom {
    const float4 *offm;        // off x, y, z, m (FP32), sorted order
    const float4 *velf;        // v x, y, z, 0 (FP32)
    const uint32_t *cellc;     // cx | cy << 10 | cz << 20 (cell within the particle's box)
};

// Per-particle record of the FP32 gradient loop (48 bytes).  rho and P are
// split into a high FP32 word and the FP32 rounding of the remainder, so that
// a pair difference (hi_j - hi_i) + (lo_j - lo_i) is the FP64 difference to
// FP32 rounding of the difference itself, not of the values: neighbour
// densities and pressures agree to many digits in smooth flow.
struct alignas(16) GradRecF {
    float4 xh;     // off x, y, z, hsml
    float4 vr;     // vx, vy, vz, rho (high word)
    float4 pm;     // P (high word), m, rho and P low words
};

// Per-particle record of the FP32 force loop (160 bytes); P / rho and the
// condition number stay FP64 because the loop compares them exactly.
struct alignas(16) ForceRecF {
    float4 xh;     // off x, y, z, hsml
    float4 vc;     // vx, vy, vz, sound speed
    float4 th;     // rho, P, V = m / rho, DhsmlNgbFactor
    float4 t0;     // NV_T: i00 i01 i02 i11
    float4 t1;     // NV_T: i12 i22, 1 / hsml, FaceClosureError
    float4 g0;     // grad rho (3), grad P x
    float4 g1;     // grad P y, z, grad v row 0 (x, y)
    float4 g2;     // grad v row 0 z, row 1 (x, y, z)
    float4 g3;     // grad v row 2 (x, y, z), packed cell coordinates (bits)
    double cn;     // ConditionNumber
    double e;      // P / rho
};

// Counters accumulated by the kernels (device memory, zeroed per stage).
struct Counters {
    unsigned long long interactions;
    unsigned long long retries;
    unsigned long long exact_calls;
    int status;
    int redo_n;  // FP32 force: particles redone in FP64 (summed over stages)
    int overflow_slices;
};

// candidates a density-sweep lane walks before it defers its particle to the
// second, load-balanced launch (gzm_params.density_budget = 0)
constexpr int kDensityBudget = 512;

// Density candidate cache of the FP64 tree sweeps: the first sweep of a stage
// walks the tree with radius kDensityCacheFactor * hsml and keeps, per
// particle, the candidates within that radius in walk order (at most
// kDensityCacheCap; more marks the particle uncached).  A later sweep whose
// hsml still lies inside the radius scans that row instead of the tree: the
// candidates within hsml come out of it in the order the walk would give.
constexpr double kDensityCacheFactor = 1.2;
constexpr int kDensityCacheCap = 96;  // a multiple of 4: rows are read as int4

// Symmetric pair rows of the FP64 gradient stage (tree search, every particle
// active).  The gradient pair set of i is {j : r_ij < h_i or r_ij < h_j}; with
// positions unchanged since the density stage, the part within i's candidate
// row radius is in i's row, and the rest (r_ij < h_j only) is in j's row or
// j's own walk, so each j that finds i there appends itself to i's pair row
// (at most kPairRowCap; more sends i back to the tree walk).  The tree walk
// visits slots in ascending order, and so does the gradient kernel's merge of
// the two sorted rows: same candidates, same order, same sums.
constexpr int kPairRowCap = 32;

struct SymRowArgs {
    const BoxDesc *boxes;
    const int *box_of;
    TreeArgs tree;
    int n;
    const double4 *posm;
    const double *hsml;
    const int *dcache_idx;
    const int *dcache_n;
    const double *dcache_r2;
    int *brow;             // kPairRowCap per particle
    int *bcount;           // zeroed before the scatter
    int *wlist;            // particles j whose own row does not cover h_j (walked by a wavefront each)
    int *wcount;
};
// a sweep over at most this many particles runs one wavefront per particle:
// they all fit on the GPU at once, and a walk shared by 64 lanes is shorter
constexpr int kWaveSweepMax = 4096;
// Relative band around h^2 in which a neighbour of an FP32-summed particle
// (precision 2) may lie on the other side of the FP64 pass's kernel edge: the
// converged hsml of FP32 sums differs from the FP64 pass's by up to ~3e-7
// (measured 2.3e-7 on 31.5 M particles), and FP32 r^2 carries up to 2e-6 h^2.
constexpr double kEdgeBand = 1e-5;

struct DensitySweepArgs {
    const BoxDesc *boxes;  // per box
    const int *box_of;     // box of each sorted slot
    TreeArgs tree;         // tree.nodes != NULL: walk the tree instead of the cell grid
    HsmlParams hp;
    const int *live;
    int n_live;            // live entries, or the capacity when n_live_dev is set
    const int *n_live_dev; // device count of live entries (graph-captured passes)
    const double4 *posm;   // x, y, z, m
    const double4 *velu;   // vx, vy, vz, u
    double *hsml, *left, *right, *cn;
    double *num_ngb, *dhsml, *divvel, *rho, *fce, *nvt;  // nvt: 6 per particle
    // MFV (FP64 sweeps only): the sweep's sum of m_j W_ij v_j per particle, 3
    // per particle; the finalize turns the last one into ParticleVel.  NULL: MFM
    double *pvel;
    int *redo;             // per live entry: 1 = sweep again
    // Load balance: a particle whose walk passes `budget` candidates (0: no
    // limit) stops and appends its live position to heavy_list (count
    // n_heavy); a second launch gives each such particle a whole wavefront,
    // so one expensive search neither holds a wavefront of cheap ones nor
    // runs on a single lane.
    int budget;
    int *heavy_list;
    int *n_heavy;
    int wave_all;          // wavefront-per-particle launch over all live entries (a small sweep)
    // Density candidate cache (tree search, FP64 sums; NULL otherwise): row
    // dcache_idx[i * kDensityCacheCap ...] of dcache_n[i] candidates within
    // sqrt(dcache_r2[i]) of particle i, in walk order; dcache_n[i] above the
    // cap: no valid row (set for every active particle by the stage's init).
    int *dcache_idx;
    int *dcache_n;
    double *dcache_r2;
    // FP32 density sums (precision 2): a particle that converges with a
    // candidate whose r^2 lies within kEdgeBand of h^2 is appended to
    // edge_list (count n_edge); the stage then redoes its whole iteration in
    // FP64, so that its hsml and neighbour sets are those of the FP64 pass.
    // NULL in the FP64 kernels' own sweeps.
    int *edge_list;
    int *n_edge;
    F32Geom f32;           // FP32-bulk kernels only
    Counters *ctr;
    int block;             // threads per workgroup of the sweep kernel (64 .. kMaxBlock)
};

// Pair list: row t (entry t of the active list) holds the force-loop pair
// set of that particle in scan order, idx[t * cap + k]; cap is a multiple of
// 16 so each row is a whole number of 64-byte chunks.  A particle with more
// than `cap` pairs (count > cap) is left to the scanning fallback; the 64
// consecutive entries of the active list that one wavefront handles form a
// slice, flagged when it holds such a particle (the fallback kernel runs
// only when some slice is flagged).
struct PairList {
    int *idx;              // (n_slices * 64 * cap)
    int *count;            // per active entry
    uint8_t *overflow;     // per slice
    int cap;
};

struct GradientArgs {
    const BoxDesc *boxes;
    const int *box_of;
    TreeArgs tree;
    const int *active;
    int n_active;
    const GradRec *rec;
    const double *nvt;     // 6 per particle
    const double *cn;
    double *grad;          // 15 per particle: rho(3), P(3), v(3x3 row-major)
    PairList list;         // written when list.idx != NULL
    const double4 *posm;   // FP64 positions and masses (FP32 kernels: exact re-tests)
    const double *hsml;    // FP64 hsml (FP32 kernels: exact re-tests)
    const double *rho;     // FP64 density and pressure of i (FP32 kernels: limiter)
    const double *pres;
    const GradRecF *recf;  // FP32 kernels
    F32Geom f32;
    Counters *ctr;
    // all particles active: the kernels also write each particle's force record
    // (what pack_force_kernel / pack_force_f32_kernel would write), else nullptr
    ForceRec *frec_out;
    ForceRecF *frecf_out;
    const double *dhsml, *fce;
    double gamma;
    int block;             // threads per workgroup of the gradient kernel (64 .. kMaxBlock)
    // symmetric pair rows (tree search; NULL: every particle walks the tree)
    const int *dcache_idx;
    const int *dcache_n;
    const double *dcache_r2;
    const int *brow;
    const int *bcount;
};

// Pairs whose Riemann problem reaches the exact solver are taken out of the
// force loop (a few percent of pairs in supersonic flow, but they would hold
// every lane of their wavefront through the iterative solver) and evaluated
// by a second kernel, one pair per lane; a third kernel adds them to each
// particle after its other pairs, in list order.
struct DeferArgs {
    uint64_t *mask;        // per active entry: list positions k deferred (bit k; lists of <= 64)
    int *count;            // per active entry: number of deferred pairs
    uint8_t *raw;          // per active entry: 1 = acc, dtu, vsig hold the raw sums of the loop
    const int *offset;     // exclusive scan of count
    const int *total;      // sum of count (device)
    double *result;        // per deferred pair of the round: dv/dt terms (3), du/dt term, v_sig
                           // (MFV: 9 per pair, see force_mfv_deferred_kernel)
    int2 *pair;            // per deferred pair of the round: particle i and neighbour j, written by
                           //   launch_defer_owners before the round's pair kernel
    int q_begin, q_end;    // deferred pairs handled by this round
};

struct ForceArgs {
    const BoxDesc *boxes;
    const int *box_of;
    TreeArgs tree;
    GammaConsts G;
    const int *active;
    int n_active;
    double max_hsml;
    const ForceRec *rec;
    const uint8_t *is_active;  // NULL when every particle is active (no wakeup test)
    const double *vsig_old;    // MaxSignalVel of inactive particles
    double *acc;               // 3 per particle
    double *dtu;
    double *vsig;
    uint8_t *wakeup;
    PairList list;         // list.idx == NULL: scan the grid for every particle
    DeferArgs defer;       // list kernel: defer.mask != NULL -> exact-solver pairs deferred
    const double4 *posm;   // FP64 positions and masses (FP32 kernels)
    const double4 *velu;   // FP64 velocities (FP32 kernels: final operations)
    const double *hsml;    // FP64 hsml (FP32 kernels: exact re-tests)
    const ForceRecF *recf; // FP32 kernels
    // FP32 kernels: a pair whose branch decision lies within the FP32 rounding
    // margin of its threshold sends both its particles to an FP64 gradient
    // redo (gradients of the FP64 pass, into grad) and the particle whose loop
    // met it to an FP64 force redo.  redo_flag per sorted slot (bit 0 force,
    // bit 1 gradients; zero between stages); redo_list[0, n) force and
    // redo_list[n, 2n) gradient entries, counted by redo_count[0] and [1]
    // (zeroed before each stage).
    const double *rho, *pres, *dhsml, *cn, *fce, *nvt;
    double *grad;
    int *redo_flag, *redo_list, *redo_count;
    int n;
    F32Geom f32;
    Counters *ctr;
    int block;             // threads per workgroup of the force loop kernel (64 .. kMaxBlock)
};

// MFV force loop: ParticleVel (3 per particle), the integer timestep of each
// particle as a double (null: equal for all), the physical time of one
// integer unit, and the MFV outputs DtMass, dMass (neighbours add their
// shares atomically), GravWorkTerm (3 per particle)
struct MfvArgs {
    const double *pvel;
    const double *ti;
    double timebase;
    double *dt_mass, *d_mass, *grav_work;
};

// launchers (gzm_kernels.hip); all asynchronous on `stream`
void launch_cell_keys(hipStream_t s, int n, const double *pos, const BoxDesc *boxes, const int *box_caller,
                      uint64_t *keys, int *iota);
void launch_gather_sorted(hipStream_t s, int n, const int *perm, const double *pos, const double *vel,
                          const double *mass, const double *u, const int *box_caller, double4 *posm,
                          double4 *velu, int *inv, int *box_of);
void launch_gather_scalar(hipStream_t s, int n, const int *perm, const double *src, double *dst, int width);
void launch_scatter_scalar(hipStream_t s, int n, const int *perm, const double *src, double *dst, int width);
void launch_cell_ranges(hipStream_t s, int n, const double4 *posm, const BoxDesc *boxes, const int *box_of,
                        int2 *range, int ncell);
// per-box max hsml into boxes[b].hmax; tiles = (box, first slot) of each
// block, each tile at most kHmaxTile slots of one box
constexpr int kHmaxTile = 4096;
void launch_box_hmax(hipStream_t s, int n_boxes, BoxDesc *boxes, const double *hsml, int n_tiles, const int2 *tiles);
// status 1155 when live particles remain after the captured sweeps
void launch_flag_unconverged(hipStream_t s, const int *n_live_dev, Counters *ctr);
void launch_clear_counters(hipStream_t s, Counters *ctr);
void launch_map_active(hipStream_t s, int n_active, const int *active_caller, const int *inv, int *active_sorted);
void launch_density_init(hipStream_t s, int n_active, const int *active, const double4 *posm, int *dcache_n, double *hsml,
                         double *left, double *right, double max_hsml, int *live_flag);
void launch_density_sweep(hipStream_t s, const DensitySweepArgs &a);
// precision 2: the listed particles (count on the device) restart their hsml
// iteration from the stage's entry hsml, with empty brackets
void launch_edge_reset(hipStream_t s, int cap, const int *n_edge, const int *edge_list, const double *hsml_entry,
                       double *hsml, double *left, double *right);
// pvel (MFV, else NULL): the last sweep's sums of m_j W_ij v_j become
// ParticleVel before the density sum they are divided by is replaced
void launch_density_finalize(hipStream_t s, int n_active, const int *active, const double4 *posm,
                             const double4 *velu, const double *hsml, double *num_ngb, double *rho,
                             double *pres, double gamma_minus_1, double *pvel);
void launch_pack_grad(hipStream_t s, int n, const double4 *posm, const double4 *velu, const double *hsml,
                      const double *rho, const double *pres, GradRec *rec);
void launch_sym_rows(hipStream_t s, const SymRowArgs &a);
void launch_gradients(hipStream_t s, const GradientArgs &a);
void launch_pack_force(hipStream_t s, int n, const double4 *posm, const double4 *velu, const double *hsml,
                       const double *rho, const double *pres, const double *dhsml, const double *cn,
                       const double *fce, const double *nvt, const double *grad, double gamma, ForceRec *rec);
void launch_force(hipStream_t s, const ForceArgs &a);  // list-driven if a.list.idx, else scanning
// deferred exact-solver pairs [a.defer.q_begin, a.defer.q_end), then the per-particle finish
void launch_force_deferred(hipStream_t s, const ForceArgs &a, int n_pairs);
// owner entry and list position of each deferred pair of the round [q_begin, q_end)
void launch_defer_owners(hipStream_t s, const ForceArgs &a);
void launch_force_finish(hipStream_t s, const ForceArgs &a);
void launch_defer_total(hipStream_t s, int n_active, const int *offset, const int *count, int *total);
// FP32-bulk kernels (gzm_kernels_f32.hip)
void launch_pack_geom_f32(hipStream_t s, int n, const double4 *posm, const double4 *velu, const BoxDesc *boxes,
                          const int *box_of, float4 *offm, float4 *velf, uint32_t *cellc);
void launch_density_sweep_f32(hipStream_t s, const DensitySweepArgs &a);
void launch_pack_grad_f32(hipStream_t s, int n, const float4 *offm, const float4 *velf, const double *hsml,
                          const double *rho, const double *pres, GradRecF *rec);
void launch_gradients_f32(hipStream_t s, const GradientArgs &a);
void launch_pack_force_f32(hipStream_t s, int n, const float4 *offm, const float4 *velf, const uint32_t *cellc,
                           const double4 *posm, const double *hsml, const double *rho, const double *pres,
                           const double *dhsml, const double *cn, const double *fce, const double *nvt,
                           const double *grad, double gamma, ForceRecF *rec);
void launch_force_f32(hipStream_t s, const ForceArgs &a);
void launch_force_redo_fp64(hipStream_t s, const ForceArgs &a);  // gradients, then force loops
// particles whose pair set overflowed the list: FP64, one wavefront each (n_list zeroed by the caller)
void launch_force_long_lists(hipStream_t s, const ForceArgs &a, int *list, int *n_list);
void launch_force_deferred_f32(hipStream_t s, const ForceArgs &a, int n_pairs);

// density sweep of the entries in a.heavy_list (*a.n_heavy of them), one
// wavefront each, FP64 sums (also for the FP32-bulk sweeps)
void launch_density_heavy(hipStream_t s, const DensitySweepArgs &a, int max_entries);

void launch_neighbor_counts(hipStream_t s, const BoxDesc *boxes, const int *box_of, const TreeArgs &tree, int mode,
                            int n_active, const int *active, const double4 *posm, const double *hsml, int *counts,
                            Counters *ctr);

// tree search (gzm_tree.hip)
struct TreeBuild {
    int n;
    const int *box_of;
    const BoxDesc *boxes;
    const uint64_t *keys;      // sorted fine keys
    const double4 *posm;
    int2 *child;               // internal nodes (n entries: at most n - 1 used)
    int2 *range;
    int *split_right;          // sorted slot s: right child of the node that splits after s
    int *esc, *leaf_esc;
    double4 *lo, *hi;          // FP64 node boxes (build only)
    int *flag;                 // internal nodes: launch of the bounds pass that completed the node
    int *large;                // internal nodes of more than 32 slots, in no particular order
    int *large_count;          // length of `large` (one int)
    TreeNode *nodes;           // the walk's records
    // node hmax refresh: max of hsml over a node's slot range from a table of
    // window maxima, level k holding max(hsml[s .. s + 2^k - 1]) rounded up to
    // FP32 (a bound, never below the FP64 value)
    float *window_max;         // levels * n
    int levels;                // 1 + floor(log2(largest box))
};
// fine keys: box index above a 3 * bits_per_axis-bit Morton code of the position
void launch_tree_keys(hipStream_t s, int n, const double *pos, const BoxDesc *boxes, const int *box_caller,
                      int bits_per_axis, uint64_t *keys, int *iota);
// radix tree, escape links, node boxes and hmax from the sorted keys and hsml
void launch_tree_build(hipStream_t s, const TreeBuild &b, const double *hsml);
// node hmax from the current hsml (window-maximum table; node boxes unchanged)
void launch_tree_hmax(hipStream_t s, const TreeBuild &b, const double *hsml);
// per-cell max hsml (rounded up to FP32) of the cells' slot ranges
void launch_cell_hmax(hipStream_t s, int ncell, const int2 *range, const double *hsml, float *hmax);
// MFV force loop (gzm_mfv.hip): the preloop that zeroes the active
// particles' DtMass, dMass and GravWorkTerm, then the loop
void launch_mfv_preloop(hipStream_t s, const ForceArgs &a, const MfvArgs &m);
// list kernel with deferral when a.defer.mask is set (then also
// launch_force_mfv_overflow for the pair sets the list could not hold, and
// the deferred pairs' rounds), else the one-pass list or scanning kernel
void launch_force_mfv(hipStream_t s, const ForceArgs &a, const MfvArgs &m);
void launch_force_mfv_overflow(hipStream_t s, const ForceArgs &a, const MfvArgs &m, int *rows, int *n_rows);
void launch_force_mfv_deferred(hipStream_t s, const ForceArgs &a, const MfvArgs &m, int n_pairs);
void launch_force_mfv_finish(hipStream_t s, const ForceArgs &a, const MfvArgs &m);
// the MFV Riemann solver on n independent face states (out [n][7])
void launch_riemann_states_mfv(hipStream_t s, int mode, const GammaConsts &G, int64_t n, const double *L,
                               const double *R, const double *normal, const double *limiter, double *out, int *mask);
void launch_fill_int(hipStream_t s, int *p, int n, int value);
void launch_fill_u8(hipStream_t s, uint8_t *p, int n, uint8_t value);
void launch_fill_double(hipStream_t s, double *p, long long n, double value);
void launch_set_flags(hipStream_t s, int n_active, const int *active, uint8_t *flags);
// the device Riemann solver on independent face states (gzm_elements.hip):
// mode 0 riemann_solve, mode 1 the KT flux alone; precision 0 FP64, else FP32
void launch_riemann_states(hipStream_t s, int mode, int precision, const GammaConsts &G, int64_t n, const double *L,
                           const double *R, const double *normal, const double *limiter, double *out, int *mask);

}  // namespace gzm
