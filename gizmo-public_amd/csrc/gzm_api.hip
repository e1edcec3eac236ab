// gzm_api.hip -- C ABI of libgzm.so: context, device memory, grid build and
// the three stage drivers.  See include/gzm.h for the contract.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <cstring>

#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_reduce.hpp>
#include <rocprim/device/device_select.hpp>

#include <algorithm>
#include <mutex>
#include <vector>

#include "../../include/gzm.h"
#include "grid_device.h"
#include "gzm_internal.h"

using namespace gzm;

namespace {

thread_local char g_last_error[512] = "";

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
}

#define GZM_HIP(call)                                                                           \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
            return GZM_ERR_HIP;                                                                 \
        }                                                                                       \
    } while (0)

enum TimerKind { kDensity = 0, kGradients = 1, kForce = 2, kOther = 3, kForceDeferred = 4, kDensityWave = 5 };

struct TimedLaunch {
    hipEvent_t a, b;
    int kind;
};

}  // namespace

// counter blocks, one per stage (a captured multi-pass graph accumulates each)
enum CounterSlot { kCtrDensity = 0, kCtrGradients = 1, kCtrForce = 2, kCtrMisc = 3, kCtrSlots = 4 };

struct gzm_ctx {
    gzm_params p;
    hipStream_t stream = nullptr;
    int n = 0;
    int ncell = 0;            // cells of all boxes
    GammaConsts G;
    // boxes: one for gzm_upload, many for gzm_upload_batch
    int n_boxes = 0;
    std::vector<BoxDesc> boxes_h;
    BoxDesc *d_boxes = nullptr;
    int *d_box_of = nullptr;  // box of each sorted slot
    int2 *d_hmax_tiles = nullptr;
    int n_hmax_tiles = 0;
    bool capturing = false;   // enqueueing into a graph: no host synchronisation, no timing events

    // caller-order staging
    double *d_stage = nullptr;
    size_t stage_doubles = 0;
    // sorting
    uint64_t *d_keys = nullptr, *d_keys_sorted = nullptr;
    int *d_iota = nullptr, *d_perm = nullptr, *d_inv = nullptr;
    void *d_temp = nullptr;
    size_t temp_bytes = 0;
    int2 *d_range = nullptr;
    float *d_cell_hmax = nullptr;  // per cell max hsml (pair-search culling)
    // sorted particle state
    double4 *d_posm = nullptr, *d_velu = nullptr;
    double *d_hsml = nullptr, *d_hsml0 = nullptr, *d_left = nullptr, *d_right = nullptr;
    double *d_cn = nullptr, *d_cn0 = nullptr, *d_vsig = nullptr, *d_vsig0 = nullptr;
    double *d_num_ngb = nullptr, *d_dhsml = nullptr, *d_divvel = nullptr, *d_rho = nullptr, *d_pres = nullptr;
    double *d_nvt = nullptr, *d_fce = nullptr, *d_grad = nullptr, *d_acc = nullptr, *d_dtu = nullptr;
    uint8_t *d_wakeup = nullptr, *d_is_active = nullptr;
    int *d_active_caller = nullptr, *d_active = nullptr, *d_live = nullptr, *d_live2 = nullptr, *d_flags = nullptr;
    int *d_count = nullptr, *d_count_heavy = nullptr;
    int *d_heavy_list = nullptr;  // density sweeps: live positions past the candidate budget
    // precision 2: particles whose iteration is redone in FP64 (kEdgeBand),
    // their count, and every particle's hsml at the start of the stage
    int *d_edge_list = nullptr, *d_n_edge = nullptr;
    double *d_hsml_entry = nullptr;
    // method 1 (MFV): ParticleVel [3n], integer timesteps as doubles, DtMass,
    // dMass (and its value at upload), GravWorkTerm [3n]
    double *d_pvel = nullptr, *d_ti = nullptr, *d_dtmass = nullptr, *d_dmass = nullptr, *d_dmass0 = nullptr,
           *d_gravwork = nullptr;
    double *d_hmax = nullptr;
    GradRec *d_grec = nullptr;
    ForceRec *d_frec = nullptr;
    // FP32-bulk records (precision 1 and 2)
    float4 *d_offm = nullptr, *d_velf = nullptr;
    uint32_t *d_cellc = nullptr;
    GradRecF *d_grecf = nullptr;
    ForceRecF *d_frecf = nullptr;
    int *d_redo_flag = nullptr, *d_redo_list = nullptr, *d_redo_count = nullptr;  // FP32 force: FP64 redo
    int *d_long_list = nullptr, *d_long_count = nullptr;  // force: particles whose pair list overflowed
    Counters *d_ctr = nullptr;
    int n_active = 0;
    bool all_active = true;
    std::vector<int32_t> active_h;  // caller active list now in d_active (when !all_active)
    bool have_density = false, have_gradients = false;
    bool hmax_valid = false;  // box (and node) hmax reflect the current hsml
    // exact-solver pairs deferred out of the force loop
    uint64_t *d_defer_mask = nullptr;
    int *d_defer_count = nullptr, *d_defer_off = nullptr, *d_defer_total = nullptr;
    uint8_t *d_defer_raw = nullptr;
    double *d_defer_result = nullptr;
    int2 *d_defer_pair = nullptr;      // per deferred pair of a round: particle and neighbour
    size_t defer_result_cap = 0;   // deferred pairs the result buffer holds
    int64_t last_deferred = 0;     // deferred pairs of the last force call
    // pair list written by the gradient kernel, read by the force kernel
    int *d_list = nullptr, *d_list_count = nullptr;
    uint8_t *d_list_over = nullptr;
    size_t list_entries = 0;
    int list_cap = 64;
    int block[3] = {kBlock, kBlock, kBlock};  // workgroup size of the density, gradient and force loops
    int wave_sweep_max = kWaveSweepMax;       // sweeps over at most this many particles: a wavefront each
    bool list_valid = false;
    bool frec_fresh = false;  // the gradient stage wrote every particle's force record
    int list_overflow_slices = 0;
    uint64_t active_version = 0, list_version = 0;
    std::vector<int> last_sweep_live;  // particles of each density sweep of the last eager call
    std::vector<int> last_edge_live;   // precision 2: particles of each FP64 redo sweep of the last eager call
    // neighbour-search tree (search = 1)
    int2 *d_t_child = nullptr, *d_t_range = nullptr;
    int *d_t_large = nullptr, *d_t_large_count = nullptr, *d_t_split = nullptr,
        *d_t_esc = nullptr, *d_t_leaf_esc = nullptr, *d_t_flag = nullptr;
    double4 *d_t_lo = nullptr, *d_t_hi = nullptr;
    float *d_t_wmax = nullptr;
    TreeNode *d_t_nodes = nullptr;
    int *d_dcache_idx = nullptr, *d_dcache_n = nullptr;  // density candidate cache (tree search)
    double *d_dcache_r2 = nullptr;
    int *d_brow = nullptr, *d_bcount = nullptr;          // symmetric pair rows of the gradient stage
    int *d_bwave = nullptr, *d_bwave_n = nullptr;        //   and the particles that walk for them
    bool dcache_all = false;  // every particle's candidate row built since the last upload
    TreeArgs tree;      // all null in grid mode
    TreeBuild tbuild;

    gzm_timing timing;
    std::vector<TimedLaunch> pending;
    std::vector<hipEvent_t> event_pool;
};

namespace {

template <class T>
gzm_status dev_alloc(T **p, size_t count)
{
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) {
        set_error("hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
        *p = nullptr;
        return GZM_ERR_NOMEM;
    }
    return GZM_OK;
}

template <class T>
void dev_free(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

hipEvent_t take_event(gzm_ctx *c)
{
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Bracket a launch with events on the context stream.
struct Timed {
    gzm_ctx *c;
    TimedLaunch t;
    Timed(gzm_ctx *ctx, int kind) : c(ctx)
    {
        t.kind = kind;
        t.a = t.b = nullptr;
        if (c->capturing) return;  // a captured graph is timed as a whole
        t.a = take_event(c);
        t.b = take_event(c);
        if (t.a) (void)hipEventRecord(t.a, c->stream);
    }
    ~Timed()
    {
        if (c->capturing) return;
        if (t.b) (void)hipEventRecord(t.b, c->stream);
        c->pending.push_back(t);
    }
};

// Called after a stream synchronize: fold pending event pairs into the totals.
void collect_timings(gzm_ctx *c)
{
    for (auto &t : c->pending) {
        float ms = 0;
        if (t.a && t.b && hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
            switch (t.kind) {
            case kDensity: c->timing.density_ms += ms; c->timing.density_launches++; break;
            case kGradients: c->timing.gradients_ms += ms; c->timing.gradients_launches++; break;
            case kForce: c->timing.force_ms += ms; c->timing.force_launches++; break;
            case kForceDeferred: c->timing.force_deferred_ms += ms; c->timing.force_deferred_launches++; break;
            case kDensityWave: c->timing.density_wave_ms += ms; c->timing.density_wave_launches++; break;
            default: c->timing.other_ms += ms; c->timing.other_launches++; break;
            }
        }
        if (t.a) c->event_pool.push_back(t.a);
        if (t.b) c->event_pool.push_back(t.b);
    }
    c->pending.clear();
}

gzm_status ensure_temp(gzm_ctx *c, size_t bytes)
{
    if (bytes <= c->temp_bytes) return GZM_OK;
    dev_free(c->d_temp);
    c->temp_bytes = 0;
    gzm_status s = dev_alloc((char **)&c->d_temp, bytes);
    if (s == GZM_OK) c->temp_bytes = bytes;
    return s;
}

gzm_status ensure_stage(gzm_ctx *c, size_t doubles)
{
    if (doubles <= c->stage_doubles) return GZM_OK;
    dev_free(c->d_stage);
    c->stage_doubles = 0;
    gzm_status s = dev_alloc(&c->d_stage, doubles);
    if (s == GZM_OK) c->stage_doubles = doubles;
    return s;
}

void free_particles(gzm_ctx *c)
{
    dev_free(c->d_keys); dev_free(c->d_keys_sorted); dev_free(c->d_iota); dev_free(c->d_perm); dev_free(c->d_inv);
    dev_free(c->d_posm); dev_free(c->d_velu);
    dev_free(c->d_hsml); dev_free(c->d_hsml0); dev_free(c->d_left); dev_free(c->d_right);
    dev_free(c->d_cn); dev_free(c->d_cn0); dev_free(c->d_vsig); dev_free(c->d_vsig0);
    dev_free(c->d_num_ngb); dev_free(c->d_dhsml); dev_free(c->d_divvel); dev_free(c->d_rho); dev_free(c->d_pres);
    dev_free(c->d_nvt); dev_free(c->d_fce); dev_free(c->d_grad); dev_free(c->d_acc); dev_free(c->d_dtu);
    dev_free(c->d_wakeup); dev_free(c->d_is_active);
    dev_free(c->d_active_caller); dev_free(c->d_active); dev_free(c->d_live); dev_free(c->d_live2);
    dev_free(c->d_flags); dev_free(c->d_grec); dev_free(c->d_frec);
    dev_free(c->d_heavy_list);
    dev_free(c->d_edge_list); dev_free(c->d_n_edge); dev_free(c->d_hsml_entry);
    dev_free(c->d_pvel); dev_free(c->d_ti); dev_free(c->d_dtmass); dev_free(c->d_dmass); dev_free(c->d_dmass0);
    dev_free(c->d_gravwork);
    dev_free(c->d_offm); dev_free(c->d_velf); dev_free(c->d_cellc); dev_free(c->d_grecf); dev_free(c->d_frecf);
    dev_free(c->d_redo_flag); dev_free(c->d_redo_list); dev_free(c->d_redo_count);
    dev_free(c->d_long_list); dev_free(c->d_long_count);
    dev_free(c->d_list); dev_free(c->d_list_count); dev_free(c->d_list_over);
    dev_free(c->d_box_of);
    dev_free(c->d_defer_mask); dev_free(c->d_defer_count); dev_free(c->d_defer_off); dev_free(c->d_defer_total);
    dev_free(c->d_defer_raw); dev_free(c->d_defer_result); dev_free(c->d_defer_pair);
    dev_free(c->d_t_child); dev_free(c->d_t_range); dev_free(c->d_t_large);
    dev_free(c->d_t_large_count);
    dev_free(c->d_t_split); dev_free(c->d_t_esc); dev_free(c->d_t_leaf_esc); dev_free(c->d_t_flag);
    dev_free(c->d_t_lo); dev_free(c->d_t_hi); dev_free(c->d_t_wmax); dev_free(c->d_t_nodes);
    dev_free(c->d_dcache_idx); dev_free(c->d_dcache_n); dev_free(c->d_dcache_r2);
    dev_free(c->d_brow); dev_free(c->d_bcount); dev_free(c->d_bwave); dev_free(c->d_bwave_n);
    c->dcache_all = false;
    c->tree = TreeArgs{};
    c->defer_result_cap = 0;
    c->list_entries = 0;
    c->list_valid = false;
    c->frec_fresh = false;
    c->n = 0;
}

void free_boxes(gzm_ctx *c)
{
    dev_free(c->d_boxes);
    dev_free(c->d_hmax_tiles);
    dev_free(c->d_range);
    dev_free(c->d_cell_hmax);
    c->n_hmax_tiles = 0;
    c->n_boxes = 0;
    c->ncell = 0;
    c->boxes_h.clear();
}

// host array (caller order, `width` doubles per particle) -> sorted device array
gzm_status upload_sorted(gzm_ctx *c, const double *host, double *dst, int width)
{
    size_t cnt = (size_t)c->n * width;
    gzm_status s = ensure_stage(c, cnt);
    if (s) return s;
    GZM_HIP(hipMemcpyAsync(c->d_stage, host, cnt * sizeof(double), hipMemcpyHostToDevice, c->stream));
    launch_gather_scalar(c->stream, c->n, c->d_perm, c->d_stage, dst, width);
    GZM_HIP(hipGetLastError());
    return GZM_OK;
}

// sorted device array -> host array in caller order
gzm_status download_sorted(gzm_ctx *c, const double *src, double *host, int width)
{
    size_t cnt = (size_t)c->n * width;
    gzm_status s = ensure_stage(c, cnt);
    if (s) return s;
    launch_scatter_scalar(c->stream, c->n, c->d_perm, src, c->d_stage, width);
    GZM_HIP(hipGetLastError());
    GZM_HIP(hipMemcpyAsync(host, c->d_stage, cnt * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    GZM_HIP(hipStreamSynchronize(c->stream));
    return GZM_OK;
}

gzm_status reset_counters(gzm_ctx *c, int slot)
{
    launch_clear_counters(c->stream, c->d_ctr + slot);
    GZM_HIP(hipGetLastError());
    return GZM_OK;
}

gzm_status read_counters(gzm_ctx *c, int slot, Counters *h)
{
    GZM_HIP(hipMemcpyAsync(h, c->d_ctr + slot, sizeof(Counters), hipMemcpyDeviceToHost, c->stream));
    GZM_HIP(hipStreamSynchronize(c->stream));
    return GZM_OK;
}

// Map a caller active list to sorted slots, sorted ascending so that
// consecutive threads hold spatially close particles.
gzm_status set_active(gzm_ctx *c, const int32_t *active, int64_t n_active)
{
    if (active == nullptr) {
        if (!c->all_active || c->n_active != c->n) {
            GZM_HIP(hipMemcpyAsync(c->d_active, c->d_iota, sizeof(int) * c->n, hipMemcpyDeviceToDevice, c->stream));
            c->active_version++;
        }
        c->n_active = c->n;
        c->all_active = true;
        return GZM_OK;
    }
    if (n_active < 0 || n_active > c->n) {
        set_error("n_active=%lld out of range [0, %d]", (long long)n_active, c->n);
        return GZM_ERR_ARG;
    }
    // the same list as the previous stage call: d_active already holds it, and
    // the pair list the gradient stage wrote for it stays valid for the force
    if (!c->all_active && n_active == c->n_active && (size_t)n_active == c->active_h.size() &&
        (n_active == 0 || memcmp(active, c->active_h.data(), sizeof(int32_t) * n_active) == 0))
        return GZM_OK;
    // validate indices on the host before any kernel uses them
    for (int64_t k = 0; k < n_active; k++) {
        if (active[k] < 0 || active[k] >= c->n) {
            set_error("active[%lld]=%d out of range", (long long)k, active[k]);
            return GZM_ERR_ARG;
        }
    }
    GZM_HIP(hipMemcpyAsync(c->d_active_caller, active, sizeof(int) * n_active, hipMemcpyHostToDevice, c->stream));
    {
        Timed tm(c, kOther);
        launch_map_active(c->stream, (int)n_active, c->d_active_caller, c->d_inv, c->d_live);
        size_t bytes = 0;
        GZM_HIP(rocprim::radix_sort_keys(nullptr, bytes, c->d_live, c->d_active, (int)n_active, 0, 32, c->stream));
        gzm_status s = ensure_temp(c, bytes);
        if (s) return s;
        GZM_HIP(rocprim::radix_sort_keys(c->d_temp, bytes, c->d_live, c->d_active, (int)n_active, 0, 32,
                                         c->stream));
    }
    c->n_active = (int)n_active;
    c->all_active = false;
    c->active_h.assign(active, active + n_active);
    c->active_version++;
    return GZM_OK;
}

F32Geom f32_geom(const gzm_ctx *c)
{
    F32Geom g;
    g.offm = c->d_offm;
    g.velf = c->d_velf;
    g.cellc = c->d_cellc;
    return g;
}

// per-particle arrays of the deferred exact-solver pairs (sized for all particles)
gzm_status ensure_defer(gzm_ctx *c)
{
    if (c->d_defer_mask) return GZM_OK;
    gzm_status s = dev_alloc(&c->d_defer_mask, (size_t)c->n);
    s = s ? s : dev_alloc(&c->d_defer_count, (size_t)c->n);
    s = s ? s : dev_alloc(&c->d_defer_off, (size_t)c->n);
    s = s ? s : dev_alloc(&c->d_defer_total, 1);
    s = s ? s : dev_alloc(&c->d_defer_raw, (size_t)c->n);
    return s;
}

// the gradient kernel writes a pair list sized for the active list
gzm_status ensure_list(gzm_ctx *c)
{
    size_t slices = ((size_t)c->n_active + 63) / 64;
    size_t entries = slices * (size_t)c->list_cap * 64;
    if (entries <= c->list_entries && c->d_list_count) return GZM_OK;
    dev_free(c->d_list); dev_free(c->d_list_count); dev_free(c->d_list_over);
    c->list_entries = 0;
    gzm_status s = dev_alloc(&c->d_list, entries);
    s = s ? s : dev_alloc(&c->d_list_count, (size_t)c->n + 64);
    s = s ? s : dev_alloc(&c->d_list_over, slices + 1);
    if (s) return s;
    c->list_entries = entries;
    return GZM_OK;
}

// per-box max hsml into the device box descriptors (no host round trip)
gzm_status compute_hmax(gzm_ctx *c)
{
    if (c->hmax_valid) return GZM_OK;
    c->hmax_valid = true;
    Timed tm(c, kOther);
    launch_box_hmax(c->stream, c->n_boxes, c->d_boxes, c->d_hsml, c->n_hmax_tiles, c->d_hmax_tiles);
    if (c->tree.nodes) launch_tree_hmax(c->stream, c->tbuild, c->d_hsml);
    else launch_cell_hmax(c->stream, c->ncell, c->d_range, c->d_hsml, c->d_cell_hmax);
    GZM_HIP(hipGetLastError());
    return GZM_OK;
}

// Ordered stream compaction: out = in[flags != 0], count into c->d_count and,
// when `count` is not NULL, to the host.  While capturing, the temporary
// storage must already be large enough (sized by the probe pass).
gzm_status compact(gzm_ctx *c, const int *in, const int *flags, int n, int *out, int *count, int *d_count = nullptr)
{
    if (!d_count) d_count = c->d_count;
    size_t bytes = 0;
    GZM_HIP(rocprim::select(nullptr, bytes, in, flags, out, d_count, (size_t)n, c->stream));
    if (c->capturing && bytes > c->temp_bytes) { set_error("temporary storage too small while capturing"); return GZM_ERR_STATE; }
    gzm_status s = ensure_temp(c, bytes);
    if (s) return s;
    GZM_HIP(rocprim::select(c->d_temp, bytes, in, flags, out, d_count, (size_t)n, c->stream));
    if (count) {
        GZM_HIP(hipMemcpyAsync(count, d_count, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        GZM_HIP(hipStreamSynchronize(c->stream));
    }
    return GZM_OK;
}

}  // namespace

extern "C" {

const char *gzm_status_string(int status)
{
    switch (status) {
    case GZM_OK: return "ok";
    case GZM_ERR_ARG: return "invalid argument";
    case GZM_ERR_HIP: return "HIP runtime error";
    case GZM_ERR_NOMEM: return "device allocation failed";
    case GZM_ERR_STATE: return "call out of order";
    case GZM_ERR_NO_DEVICE: return "no usable GPU";
    case GZM_ERR_BRACKET: return "hsml bracket collapsed to zero (density.c:780-783)";
    case GZM_ERR_HSML_NOT_CONVERGED: return "hsml iteration did not converge (endrun 1155)";
    case GZM_ERR_RIEMANN: return "no positive Riemann pressure (endrun 1234)";
    case GZM_ERR_RIEMANN_NEG_DENSITY: return "negative density into the exact Riemann solver (endrun 1234545)";
    default: return "unknown status";
    }
}

const char *gzm_last_error(void) { return g_last_error; }

gzm_status gzm_ctx_create(gzm_ctx **out, const gzm_params *params)
{
    if (!out || !params) { set_error("NULL argument"); return GZM_ERR_ARG; }
    *out = nullptr;
    for (int a = 0; a < 3; a++) {
        if (!(params->box[a] > 0)) { set_error("box[%d] must be > 0", a); return GZM_ERR_ARG; }
    }
    if (!(params->gamma > 1)) { set_error("gamma must be > 1"); return GZM_ERR_ARG; }
    if (params->precision < 0 || params->precision > 2) {
        set_error("precision must be 0 (plain FP64), 1 (FP32 bulk) or 2 (FP32 bulk with FP32 density sums)");
        return GZM_ERR_ARG;
    }
    if (params->search != 0 && params->search != 1) {
        set_error("search must be 0 (cell grid) or 1 (tree)");
        return GZM_ERR_ARG;
    }
    if (params->method != 0 && params->method != 1) {
        set_error("method must be 0 (meshless finite mass) or 1 (meshless finite volume)");
        return GZM_ERR_ARG;
    }
    if (params->method == 1 && params->precision != 0) {
        set_error("method 1 (meshless finite volume) runs in FP64 only: precision must be 0");
        return GZM_ERR_ARG;
    }
    if (params->method == 1 && !(params->timebase_interval >= 0 && isfinite(params->timebase_interval))) {
        set_error("timebase_interval must be finite and >= 0");
        return GZM_ERR_ARG;
    }
    if (params->pair_list_cap > 4096) { set_error("pair_list_cap %d exceeds 4096", params->pair_list_cap); return GZM_ERR_ARG; }
    for (int k = 0; k < 3; k++) {
        const int b = params->block_size[k];
        if (b != 0 && b != 64 && b != 128 && b != 256 && b != kMaxBlock) {
            set_error("block_size[%d] = %d: must be 0 (%d), 64, 128, 256 or %d", k, b, kBlock, kMaxBlock);
            return GZM_ERR_ARG;
        }
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) { set_error("no HIP device"); return GZM_ERR_NO_DEVICE; }
    if (params->device < 0 || params->device >= ndev) { set_error("device %d not present", params->device); return GZM_ERR_ARG; }
    GZM_HIP(hipSetDevice(params->device));
    gzm_ctx *c = new gzm_ctx();
    c->p = *params;
    // whole 64-byte chunks per row (the gradient kernel stores 16 entries at a time)
    c->list_cap = params->pair_list_cap == 0 ? 64 : params->pair_list_cap;
    if (c->list_cap > 0) c->list_cap = (c->list_cap + 15) / 16 * 16;
    for (int k = 0; k < 3; k++) c->block[k] = params->block_size[k] ? params->block_size[k] : kBlock;
    c->wave_sweep_max = params->wave_sweep_max == 0 ? kWaveSweepMax : (params->wave_sweep_max < 0 ? 0 : params->wave_sweep_max);
    memset(&c->timing, 0, sizeof(c->timing));
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("hipStreamCreate failed");
        delete c;
        return GZM_ERR_HIP;
    }
    double g = params->gamma;
    c->G.G0 = g;
    c->G.G1 = ((g - 1.0) / (2.0 * g));
    c->G.G2 = ((g + 1.0) / (2.0 * g));
    c->G.G3 = ((2.0 * g / (g - 1.0)));
    c->G.G4 = (2.0 / (g - 1.0));
    c->G.G5 = (2.0 / (g + 1.0));
    c->G.G6 = ((g - 1.0) / (g + 1.0));
    c->G.G7 = (0.5 * (g - 1.0));
    c->G.G8 = (1.0 / g);
    c->G.G9 = (g - 1.0);
    if (dev_alloc(&c->d_ctr, kCtrSlots) || dev_alloc(&c->d_count, 1) || dev_alloc(&c->d_count_heavy, 1) ||
        dev_alloc(&c->d_hmax, 1)) {
        gzm_ctx_destroy(c);
        return GZM_ERR_NOMEM;
    }
    *out = c;
    return GZM_OK;
}

gzm_status gzm_ctx_destroy(gzm_ctx *c)
{
    if (!c) return GZM_OK;
    (void)hipSetDevice(c->p.device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    collect_timings(c);
    free_particles(c);
    free_boxes(c);
    dev_free(c->d_stage);
    dev_free(c->d_temp);
    dev_free(c->d_ctr);
    dev_free(c->d_count);
    dev_free(c->d_count_heavy);
    dev_free(c->d_hmax);
    for (auto e : c->event_pool) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return GZM_OK;
}

}  // extern "C"

namespace {

// Per-box grids, the global cell-range array and the hmax tiles of an upload.
gzm_status setup_boxes(gzm_ctx *c, const gzm_host_view *v, int n_boxes, const double *box_sizes,
                       const int64_t *offsets)
{
    std::vector<BoxDesc> bx(n_boxes);
    std::vector<int2> tiles;
    int64_t ncell = 0;
    int inode = 0;
    for (int b = 0; b < n_boxes; b++) {
        BoxDesc &B = bx[b];
        memset(&B, 0, sizeof(B));
        const int64_t i0 = offsets[b], i1 = offsets[b + 1];
        const double *L = box_sizes + 3 * (size_t)b;
        // cell side from the parameter, or a factor (-cell_size, default 1) of
        // the mean initial hsml of the box
        double s_cell = c->p.cell_size;
        if (!(s_cell > 0)) {
            const double factor = (c->p.cell_size < 0) ? -c->p.cell_size : 1.0;
            double hsum = 0;
            int64_t cnt = 0;
            for (int64_t i = i0; i < i1; i++) {
                const double h = v->hsml[i];
                if (h > 0 && isfinite(h)) { hsum += h; cnt++; }
            }
            s_cell = cnt ? hsum / cnt : cbrt(L[0] * L[1] * L[2] * 8.0 / (double)std::max<int64_t>(1, i1 - i0));
            s_cell *= factor;
        }
        int64_t nb = 1;
        for (int ax = 0; ax < 3; ax++) {
            int nc = (int)(L[ax] / s_cell);
            nc = std::max(1, std::min(nc, 1024));
            B.grid.nc[ax] = nc;
            B.grid.inv_s[ax] = nc / L[ax];
            B.grid.s[ax] = L[ax] / nc;
            B.s_f[ax] = (float)B.grid.s[ax];
            B.box.L[ax] = L[ax];
            B.box.half[ax] = 0.5 * L[ax];
            nb *= nc;
        }
        B.cell_base = (int)ncell;
        ncell += nb;
        if (ncell > ((int64_t)1 << 30)) { set_error("grids of %lld cells too large", (long long)ncell); return GZM_ERR_ARG; }
        B.first = (int)i0;
        B.count = (int)(i1 - i0);
        // internal tree nodes: count - 1 per box, numbered box after box
        B.root = B.count >= 2 ? inode : (B.count == 1 ? -B.first - 1 : kTreeEnd);
        inode += std::max(0, B.count - 1);
        for (int64_t k = i0; k < i1; k += kHmaxTile) tiles.push_back(make_int2(b, (int)k));
    }
    free_boxes(c);
    gzm_status s = dev_alloc(&c->d_range, (size_t)ncell);
    s = s ? s : dev_alloc(&c->d_cell_hmax, (size_t)ncell);
    s = s ? s : dev_alloc(&c->d_boxes, (size_t)n_boxes);
    s = s ? s : dev_alloc(&c->d_hmax_tiles, tiles.size());
    if (s) { free_boxes(c); return s; }
    for (auto &B : bx) {
        B.grid.range = c->d_range + B.cell_base;
        B.grid.hmax = c->d_cell_hmax + B.cell_base;
    }
    GZM_HIP(hipMemcpyAsync(c->d_boxes, bx.data(), sizeof(BoxDesc) * n_boxes, hipMemcpyHostToDevice, c->stream));
    if (!tiles.empty())
        GZM_HIP(hipMemcpyAsync(c->d_hmax_tiles, tiles.data(), sizeof(int2) * tiles.size(), hipMemcpyHostToDevice,
                               c->stream));
    GZM_HIP(hipStreamSynchronize(c->stream));
    c->boxes_h = bx;
    c->n_boxes = n_boxes;
    c->ncell = (int)ncell;
    c->n_hmax_tiles = (int)tiles.size();
    return GZM_OK;
}

gzm_status upload_common(gzm_ctx *c, const gzm_host_view *v, int n_boxes, const double *box_sizes,
                         const int64_t *offsets)
{
    if (!c || !v) { set_error("NULL argument"); return GZM_ERR_ARG; }
    if (v->n <= 0 || v->n > (int64_t)0x7fffffff / 16) { set_error("n=%lld out of range", (long long)v->n); return GZM_ERR_ARG; }
    if (!v->pos || !v->vel || !v->mass || !v->u || !v->hsml) {
        set_error("pos, vel, mass, u and hsml are required");
        return GZM_ERR_ARG;
    }
    if (n_boxes < 1 || n_boxes > (1 << 20)) { set_error("n_boxes=%d out of range", n_boxes); return GZM_ERR_ARG; }
    if (offsets[0] != 0 || offsets[n_boxes] != v->n) {
        set_error("box offsets must start at 0 and end at n=%lld", (long long)v->n);
        return GZM_ERR_ARG;
    }
    for (int b = 0; b < n_boxes; b++) {
        if (offsets[b + 1] < offsets[b]) { set_error("box offsets must not decrease (box %d)", b); return GZM_ERR_ARG; }
        for (int ax = 0; ax < 3; ax++) {
            const double L = box_sizes[3 * (size_t)b + ax];
            if (!(L > 0) || !isfinite(L)) { set_error("box %d side %d must be > 0", b, ax); return GZM_ERR_ARG; }
        }
        // positions must lie inside their box: the grid maps [0, L) to cells
        for (int64_t i = offsets[b]; i < offsets[b + 1]; i++) {
            for (int ax = 0; ax < 3; ax++) {
                const double x = v->pos[3 * i + ax], L = box_sizes[3 * (size_t)b + ax];
                if (!(x >= 0.0 && x < L)) {
                    set_error("pos[%lld][%d]=%g outside [0, %g) of box %d", (long long)i, ax, x, L, b);
                    return GZM_ERR_ARG;
                }
            }
        }
    }
    GZM_HIP(hipSetDevice(c->p.device));
    const int n = (int)v->n;
    if (n != c->n) {
        free_particles(c);
        c->n = n;
        gzm_status s = GZM_OK;
        s = s ? s : dev_alloc(&c->d_keys, n);
        s = s ? s : dev_alloc(&c->d_keys_sorted, n);
        s = s ? s : dev_alloc(&c->d_iota, n);
        s = s ? s : dev_alloc(&c->d_perm, n);
        s = s ? s : dev_alloc(&c->d_inv, n);
        s = s ? s : dev_alloc(&c->d_box_of, n);
        s = s ? s : dev_alloc(&c->d_posm, n);
        s = s ? s : dev_alloc(&c->d_velu, n);
        double **scal[] = {&c->d_hsml, &c->d_hsml0, &c->d_left, &c->d_right, &c->d_cn, &c->d_cn0, &c->d_vsig,
                           &c->d_vsig0, &c->d_num_ngb, &c->d_dhsml, &c->d_divvel, &c->d_rho, &c->d_pres, &c->d_fce,
                           &c->d_dtu};
        for (auto pp : scal) s = s ? s : dev_alloc(pp, n);
        s = s ? s : dev_alloc(&c->d_nvt, (size_t)n * 6);
        s = s ? s : dev_alloc(&c->d_grad, (size_t)n * 15);
        s = s ? s : dev_alloc(&c->d_acc, (size_t)n * 3);
        s = s ? s : dev_alloc(&c->d_wakeup, n);
        s = s ? s : dev_alloc(&c->d_is_active, n);
        s = s ? s : dev_alloc(&c->d_active_caller, n);
        s = s ? s : dev_alloc(&c->d_active, n);
        s = s ? s : dev_alloc(&c->d_live, n);
        s = s ? s : dev_alloc(&c->d_live2, n);
        s = s ? s : dev_alloc(&c->d_flags, n);
        s = s ? s : dev_alloc(&c->d_heavy_list, n);
        if (c->p.precision == 0) {
            s = s ? s : dev_alloc(&c->d_grec, n);
            s = s ? s : dev_alloc(&c->d_frec, n);
        } else {
            s = s ? s : dev_alloc(&c->d_offm, n);
            s = s ? s : dev_alloc(&c->d_velf, n);
            s = s ? s : dev_alloc(&c->d_cellc, n);
            s = s ? s : dev_alloc(&c->d_grecf, n);
            s = s ? s : dev_alloc(&c->d_frecf, n);
            s = s ? s : dev_alloc(&c->d_redo_flag, n);
            s = s ? s : dev_alloc(&c->d_redo_list, 2 * (size_t)n);
            s = s ? s : dev_alloc(&c->d_redo_count, 2);
        }
        if (c->p.method == 1) {
            s = s ? s : dev_alloc(&c->d_pvel, (size_t)n * 3);
            s = s ? s : dev_alloc(&c->d_ti, n);
            s = s ? s : dev_alloc(&c->d_dtmass, n);
            s = s ? s : dev_alloc(&c->d_dmass, n);
            s = s ? s : dev_alloc(&c->d_dmass0, n);
            s = s ? s : dev_alloc(&c->d_gravwork, (size_t)n * 3);
        }
        if (c->p.precision == 2) {
            s = s ? s : dev_alloc(&c->d_edge_list, n);
            s = s ? s : dev_alloc(&c->d_n_edge, 1);
            s = s ? s : dev_alloc(&c->d_hsml_entry, n);
        }
        if (c->p.search == 1) {
            s = s ? s : dev_alloc(&c->d_t_child, n);
            s = s ? s : dev_alloc(&c->d_t_range, n);
            s = s ? s : dev_alloc(&c->d_t_large, n);
            s = s ? s : dev_alloc(&c->d_t_large_count, 1);
            s = s ? s : dev_alloc(&c->d_t_split, n);
            s = s ? s : dev_alloc(&c->d_t_esc, n);
            s = s ? s : dev_alloc(&c->d_t_leaf_esc, n);
            s = s ? s : dev_alloc(&c->d_t_flag, n);
            s = s ? s : dev_alloc(&c->d_t_lo, n);
            s = s ? s : dev_alloc(&c->d_t_hi, n);
            s = s ? s : dev_alloc(&c->d_t_nodes, n);
            s = s ? s : dev_alloc(&c->d_dcache_idx, (size_t)n * kDensityCacheCap);
            s = s ? s : dev_alloc(&c->d_dcache_n, n);
            s = s ? s : dev_alloc(&c->d_dcache_r2, n);
            s = s ? s : dev_alloc(&c->d_brow, (size_t)n * kPairRowCap);
            s = s ? s : dev_alloc(&c->d_bcount, n);
            s = s ? s : dev_alloc(&c->d_bwave, n);
            s = s ? s : dev_alloc(&c->d_bwave_n, 1);
        }
        if (s) { free_particles(c); return s; }
    }
    if (c->d_redo_flag) launch_fill_int(c->stream, c->d_redo_flag, n, 0);
    gzm_status s = setup_boxes(c, v, n_boxes, box_sizes, offsets);
    if (s) return s;

    // inputs to staging (caller order), keys, stable sort by (box, cell key)
    s = ensure_stage(c, (size_t)n * 8);
    if (s) return s;
    double *st_pos = c->d_stage, *st_vel = c->d_stage + 3 * (size_t)n, *st_mass = c->d_stage + 6 * (size_t)n,
           *st_u = c->d_stage + 7 * (size_t)n;
    GZM_HIP(hipMemcpyAsync(st_pos, v->pos, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
    GZM_HIP(hipMemcpyAsync(st_vel, v->vel, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
    GZM_HIP(hipMemcpyAsync(st_mass, v->mass, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    GZM_HIP(hipMemcpyAsync(st_u, v->u, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    std::vector<int> box_caller(n);
    for (int b = 0; b < n_boxes; b++)
        for (int64_t i = offsets[b]; i < offsets[b + 1]; i++) box_caller[i] = b;
    int *d_box_caller = c->d_flags;  // free until the first stage call
    GZM_HIP(hipMemcpyAsync(d_box_caller, box_caller.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
    int box_bits = 0;
    while ((1 << box_bits) < n_boxes) box_bits++;
    const bool tree = c->p.search == 1;
    // tree: Morton bits per axis that fit beside the box index in 64 bits
    const int tree_bits = std::min(21, (64 - box_bits) / 3);
    const int key_bits = box_bits + (tree ? 3 * tree_bits : 30);
    {
        Timed tm(c, kOther);
        if (tree)
            launch_tree_keys(c->stream, n, st_pos, c->d_boxes, d_box_caller, tree_bits, c->d_keys, c->d_iota);
        else
            launch_cell_keys(c->stream, n, st_pos, c->d_boxes, d_box_caller, c->d_keys, c->d_iota);
        size_t bytes = 0;
        GZM_HIP(rocprim::radix_sort_pairs(nullptr, bytes, c->d_keys, c->d_keys_sorted, c->d_iota, c->d_perm, n, 0,
                                          key_bits, c->stream));
        s = ensure_temp(c, bytes);
        if (s) return s;
        GZM_HIP(rocprim::radix_sort_pairs(c->d_temp, bytes, c->d_keys, c->d_keys_sorted, c->d_iota, c->d_perm, n, 0,
                                          key_bits, c->stream));
        launch_gather_sorted(c->stream, n, c->d_perm, st_pos, st_vel, st_mass, st_u, d_box_caller, c->d_posm,
                             c->d_velu, c->d_inv, c->d_box_of);
        if (!tree) launch_cell_ranges(c->stream, n, c->d_posm, c->d_boxes, c->d_box_of, c->d_range, c->ncell);
        if (c->p.precision != 0)
            launch_pack_geom_f32(c->stream, n, c->d_posm, c->d_velu, c->d_boxes, c->d_box_of, c->d_offm, c->d_velf,
                                 c->d_cellc);
        GZM_HIP(hipGetLastError());
    }
    // d_iota was consumed as sort values; rebuild it as the identity for "all active"
    std::vector<int> iota(n);
    for (int i = 0; i < n; i++) iota[i] = i;
    GZM_HIP(hipMemcpyAsync(c->d_iota, iota.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->stream));

    // carried state
    s = upload_sorted(c, v->hsml, c->d_hsml, 1);
    if (s) return s;
    if (v->cn) { s = upload_sorted(c, v->cn, c->d_cn, 1); if (s) return s; }
    else launch_fill_double(c->stream, c->d_cn, n, 1.0);  // init.c:514
    if (v->vsig) { s = upload_sorted(c, v->vsig, c->d_vsig, 1); if (s) return s; }
    else launch_fill_double(c->stream, c->d_vsig, n, kMinReal);
    GZM_HIP(hipStreamSynchronize(c->stream));
    GZM_HIP(hipMemcpyAsync(c->d_hsml0, c->d_hsml, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    GZM_HIP(hipMemcpyAsync(c->d_cn0, c->d_cn, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    GZM_HIP(hipMemcpyAsync(c->d_vsig0, c->d_vsig, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    if (tree) {
        TreeBuild &t = c->tbuild;
        t.n = n;
        t.box_of = c->d_box_of;
        t.boxes = c->d_boxes;
        t.keys = c->d_keys_sorted;
        t.posm = c->d_posm;
        t.child = c->d_t_child; t.range = c->d_t_range;
        t.large = c->d_t_large; t.large_count = c->d_t_large_count; t.split_right = c->d_t_split;
        t.esc = c->d_t_esc; t.leaf_esc = c->d_t_leaf_esc; t.lo = c->d_t_lo; t.hi = c->d_t_hi; t.flag = c->d_t_flag;
        t.nodes = c->d_t_nodes;
        int largest = 1;
        for (const BoxDesc &B : c->boxes_h) largest = std::max(largest, B.count);
        int levels = 1;
        while ((2 << (levels - 1)) <= largest) levels++;  // 1 + floor(log2(largest))
        if (levels != t.levels || !c->d_t_wmax) {
            dev_free(c->d_t_wmax);
            s = dev_alloc(&c->d_t_wmax, (size_t)levels * n);
            if (s) return s;
        }
        t.levels = levels;
        t.window_max = c->d_t_wmax;
        {
            Timed tm(c, kOther);
            launch_tree_build(c->stream, t, c->d_hsml);
        }
        GZM_HIP(hipGetLastError());
        c->tree.nodes = c->d_t_nodes;
        c->tree.leaf_esc = c->d_t_leaf_esc;
        // new positions: no candidate row is valid
        launch_fill_int(c->stream, c->d_dcache_n, n, kDensityCacheCap + 1);
        c->dcache_all = false;
    }

    // optional state of particles that may stay inactive
    struct Opt { const double *h; double *d; int w; } opts[] = {
        {v->num_ngb, c->d_num_ngb, 1}, {v->dhsml, c->d_dhsml, 1}, {v->divvel, c->d_divvel, 1},
        {v->density, c->d_rho, 1}, {v->pressure, c->d_pres, 1}, {v->fce, c->d_fce, 1}};
    for (auto &o : opts) {
        if (o.h) { s = upload_sorted(c, o.h, o.d, o.w); if (s) return s; }
        else launch_fill_double(c->stream, o.d, (long long)n * o.w, 0.0);
    }
    if (v->nvt) {
        std::vector<double> t6((size_t)n * 6);
        for (int i = 0; i < n; i++) {
            const double *t = v->nvt + 9 * (size_t)i;
            double *o = t6.data() + 6 * (size_t)i;
            o[0] = t[0]; o[1] = t[1]; o[2] = t[2]; o[3] = t[4]; o[4] = t[5]; o[5] = t[8];
        }
        s = upload_sorted(c, t6.data(), c->d_nvt, 6);
        if (s) return s;
    } else launch_fill_double(c->stream, c->d_nvt, (long long)n * 6, 0.0);
    if (v->grad_rho && v->grad_p && v->grad_v) {
        std::vector<double> g((size_t)n * 15);
        for (int i = 0; i < n; i++) {
            double *o = g.data() + 15 * (size_t)i;
            for (int k = 0; k < 3; k++) o[k] = v->grad_rho[3 * (size_t)i + k];
            for (int k = 0; k < 3; k++) o[3 + k] = v->grad_p[3 * (size_t)i + k];
            for (int k = 0; k < 9; k++) o[6 + k] = v->grad_v[9 * (size_t)i + k];
        }
        s = upload_sorted(c, g.data(), c->d_grad, 15);
        if (s) return s;
    } else launch_fill_double(c->stream, c->d_grad, (long long)n * 15, 0.0);
    // force outputs of particles that stay inactive round-trip unchanged
    if (v->acc) { s = upload_sorted(c, v->acc, c->d_acc, 3); if (s) return s; }
    else launch_fill_double(c->stream, c->d_acc, (long long)n * 3, 0.0);
    if (v->dtu) { s = upload_sorted(c, v->dtu, c->d_dtu, 1); if (s) return s; }
    else launch_fill_double(c->stream, c->d_dtu, n, 0.0);
    launch_fill_u8(c->stream, c->d_wakeup, n, 0);
    if (c->p.method == 1) {
        // MFV state: the integer timesteps as doubles (exact below 2^53, as the
        // reference converts them, hydro_evaluate.h:82), ParticleVel, and the
        // mass fluxes of particles that stay inactive
        if (v->ti_step) {
            std::vector<double> ti(n);
            for (int i = 0; i < n; i++) ti[i] = (double)v->ti_step[i];
            s = upload_sorted(c, ti.data(), c->d_ti, 1);
            if (s) return s;
        } else launch_fill_double(c->stream, c->d_ti, n, 1.0);
        struct MOpt { const double *h; double *d; int w; } mopts[] = {
            {v->particle_vel, c->d_pvel, 3}, {v->dt_mass, c->d_dtmass, 1}, {v->d_mass, c->d_dmass, 1},
            {v->grav_work, c->d_gravwork, 3}};
        for (auto &o : mopts) {
            if (o.h) { s = upload_sorted(c, o.h, o.d, o.w); if (s) return s; }
            else launch_fill_double(c->stream, o.d, (long long)n * o.w, 0.0);
        }
        GZM_HIP(hipMemcpyAsync(c->d_dmass0, c->d_dmass, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    }
    GZM_HIP(hipGetLastError());
    GZM_HIP(hipStreamSynchronize(c->stream));
    collect_timings(c);
    c->have_density = (v->density != nullptr && v->nvt != nullptr);
    c->have_gradients = (v->grad_rho != nullptr);
    c->all_active = false;
    c->n_active = 0;
    c->active_h.clear();
    c->list_valid = false;
    c->frec_fresh = false;
    c->hmax_valid = false;
    return GZM_OK;
}


gzm_status restore_state(gzm_ctx *c)
{
    c->hmax_valid = false;
    c->frec_fresh = false;
    GZM_HIP(hipMemcpyAsync(c->d_hsml, c->d_hsml0, sizeof(double) * c->n, hipMemcpyDeviceToDevice, c->stream));
    GZM_HIP(hipMemcpyAsync(c->d_cn, c->d_cn0, sizeof(double) * c->n, hipMemcpyDeviceToDevice, c->stream));
    GZM_HIP(hipMemcpyAsync(c->d_vsig, c->d_vsig0, sizeof(double) * c->n, hipMemcpyDeviceToDevice, c->stream));
    if (c->d_dmass0)
        GZM_HIP(hipMemcpyAsync(c->d_dmass, c->d_dmass0, sizeof(double) * c->n, hipMemcpyDeviceToDevice, c->stream));
    return GZM_OK;
}

// What a captured pass cannot learn on the host: the number of density
// sweeps and whether the pair list overflowed, both taken from a probe pass
// on the same initial state.
struct Plan {
    int sweeps;
    bool overflow;
    int64_t deferred;             // exact-solver pairs deferred out of the force loop
    std::vector<int> sweep_live;  // particles of each density sweep
    std::vector<int> edge_live;   // precision 2: particles of each FP64 redo sweep
};

gzm_status density_impl(gzm_ctx *c, const Plan *gp, gzm_pass_stats *stats)
{
    gzm_status s;
    Counters *ctr = c->d_ctr + kCtrDensity;
    c->hmax_valid = false;
    if (!gp) {
        s = reset_counters(c, kCtrDensity);
        if (s) return s;
    }
    const int na = c->n_active;
    int n_live = 0;
    // candidate cache of the FP64 tree sweeps (not with FP32 sums: their FP64
    // redo sweeps restart the iteration of a few particles)
    const bool dcache = c->tree.nodes && c->d_dcache_idx && c->p.precision != 2;
    if (na > 0) {
        {
            Timed tm(c, kOther);
            launch_density_init(c->stream, na, c->d_active, c->d_posm, dcache ? c->d_dcache_n : nullptr, c->d_hsml,
                                c->d_left, c->d_right, c->p.max_hsml, c->d_flags);
        }
        s = compact(c, c->d_active, c->d_flags, na, c->d_live, gp ? nullptr : &n_live);
        if (s) return s;
    }
    const bool fp32_sums = c->p.precision == 2;
    if (fp32_sums && na > 0) {
        // the stage's entry hsml, from which an FP64 redo restarts, and an empty redo list
        GZM_HIP(hipMemcpyAsync(c->d_hsml_entry, c->d_hsml, sizeof(double) * (size_t)c->n, hipMemcpyDeviceToDevice,
                               c->stream));
        launch_fill_int(c->stream, c->d_n_edge, 1, 0);
    }
    DensitySweepArgs a;
    memset(&a, 0, sizeof(a));
    a.boxes = c->d_boxes;
    a.box_of = c->d_box_of;
    a.tree = c->tree;
    a.hp.des_num_ngb = c->p.des_num_ngb;
    a.hp.max_num_ngb_dev = c->p.max_num_ngb_dev;
    a.hp.min_hsml = c->p.min_hsml;
    a.hp.max_hsml = c->p.max_hsml;
    a.hp.first_step = c->p.first_step;
    a.posm = c->d_posm;
    a.velu = c->d_velu;
    a.hsml = c->d_hsml; a.left = c->d_left; a.right = c->d_right; a.cn = c->d_cn;
    a.num_ngb = c->d_num_ngb; a.dhsml = c->d_dhsml; a.divvel = c->d_divvel; a.rho = c->d_rho;
    a.fce = c->d_fce; a.nvt = c->d_nvt;
    a.pvel = (c->p.method == 1) ? c->d_pvel : nullptr;
    if (dcache) {
        a.dcache_idx = c->d_dcache_idx;
        a.dcache_n = c->d_dcache_n;
        a.dcache_r2 = c->d_dcache_r2;
    }
    a.redo = c->d_flags;
    a.budget = c->p.density_budget == 0 ? kDensityBudget : (c->p.density_budget < 0 ? 0 : c->p.density_budget);
    a.heavy_list = c->d_heavy_list;
    a.n_heavy = c->d_count_heavy;
    a.f32 = f32_geom(c);
    a.ctr = ctr;
    a.block = c->block[0];
    if (fp32_sums) {
        a.edge_list = c->d_edge_list;
        a.n_edge = c->d_n_edge;
    }
    bool sums_fp32 = fp32_sums;  // false for the FP64 redo sweeps
    // one sweep over up to `cap` live entries: the budgeted launch (entries
    // past the budget are listed on the device), then one wavefront per listed
    // entry
    // (a sweep over at most wave_sweep_max entries gives every entry a wavefront)
    auto sweep = [&](int cap, int expected) -> gzm_status {
        if (expected <= c->wave_sweep_max) {
            Timed tm(c, kDensityWave);
            DensitySweepArgs w = a;
            w.wave_all = 1;
            launch_density_heavy(c->stream, w, expected);
            return GZM_OK;
        }
        if (a.budget) launch_fill_int(c->stream, c->d_count_heavy, 1, 0);
        {
            Timed tm(c, kDensity);
            if (sums_fp32) launch_density_sweep_f32(c->stream, a);
            else launch_density_sweep(c->stream, a);
        }
        if (a.budget) {
            Timed tm(c, kDensityWave);
            launch_density_heavy(c->stream, a, cap);
        }
        return GZM_OK;
    };
    int iter = 0, sweeps = 0;
    if (!gp) {
        c->last_sweep_live.clear();
        c->last_edge_live.clear();
    }
    int64_t updates = 0;
    int status = GZM_OK;
    int n_edge = 0;
    if (gp) {
        // captured: every sweep is sized for all active particles and reads the
        // live count from the device; the probe pass fixed the number of sweeps
        if (na > 0) {
            for (int sw = 0; sw < gp->sweeps; sw++) {
                a.hp.iter = sw;
                a.live = c->d_live;
                a.n_live = na;
                a.n_live_dev = c->d_count;
                s = sweep(na, sw < (int)gp->sweep_live.size() ? gp->sweep_live[sw] : na);
                if (s) return s;
                s = compact(c, c->d_live, c->d_flags, na, c->d_live2, nullptr);
                if (s) return s;
                std::swap(c->d_live, c->d_live2);
            }
            launch_flag_unconverged(c->stream, c->d_count, ctr);
        }
    } else {
        while (n_live > 0) {
            updates += n_live;
            a.hp.iter = iter;
            a.live = c->d_live;
            a.n_live = n_live;
            c->last_sweep_live.push_back(n_live);
            s = sweep(n_live, n_live);
            if (s) return s;
            GZM_HIP(hipGetLastError());
            sweeps++;
            int n_next = 0;
            s = compact(c, c->d_live, c->d_flags, n_live, c->d_live2, &n_next);
            if (s) return s;
            std::swap(c->d_live, c->d_live2);
            n_live = n_next;
            if (n_live > 0) {
                iter++;
                if (iter > c->p.maxiter) { status = GZM_ERR_HSML_NOT_CONVERGED; break; }
            }
        }
    }
    // Precision 2: the particles that converged next to a kernel edge
    // (kEdgeBand) repeat their whole iteration with FP64 sums, from the entry
    // hsml, so that their hsml and neighbour sets are the FP64 pass's.  Their
    // sweeps count into the misc counter slot, not the density interactions.
    if (fp32_sums && na > 0 && status == GZM_OK) {
        sums_fp32 = false;
        a.edge_list = nullptr;
        a.n_edge = nullptr;
        a.ctr = c->d_ctr + kCtrMisc;
        if (gp) {
            const int ne = gp->edge_live.empty() ? 0 : gp->edge_live[0];
            if (ne > 0) {
                launch_edge_reset(c->stream, ne, c->d_n_edge, c->d_edge_list, c->d_hsml_entry, c->d_hsml, c->d_left,
                                  c->d_right);
                GZM_HIP(hipMemcpyAsync(c->d_live, c->d_edge_list, sizeof(int) * (size_t)ne, hipMemcpyDeviceToDevice,
                                       c->stream));
                GZM_HIP(hipMemcpyAsync(c->d_count, c->d_n_edge, sizeof(int), hipMemcpyDeviceToDevice, c->stream));
                for (int sw = 0; sw < (int)gp->edge_live.size(); sw++) {
                    a.hp.iter = sw;
                    a.live = c->d_live;
                    a.n_live = ne;
                    a.n_live_dev = c->d_count;
                    s = sweep(ne, gp->edge_live[sw]);
                    if (s) return s;
                    s = compact(c, c->d_live, c->d_flags, ne, c->d_live2, nullptr);
                    if (s) return s;
                    std::swap(c->d_live, c->d_live2);
                }
                launch_flag_unconverged(c->stream, c->d_count, ctr);
            }
        } else {
            GZM_HIP(hipMemcpyAsync(&n_edge, c->d_n_edge, sizeof(int), hipMemcpyDeviceToHost, c->stream));
            GZM_HIP(hipStreamSynchronize(c->stream));
            if (n_edge > 0) {
                s = reset_counters(c, kCtrMisc);
                if (s) return s;
                {
                    Timed tm(c, kOther);
                    launch_edge_reset(c->stream, n_edge, c->d_n_edge, c->d_edge_list, c->d_hsml_entry, c->d_hsml,
                                      c->d_left, c->d_right);
                }
                GZM_HIP(hipMemcpyAsync(c->d_live, c->d_edge_list, sizeof(int) * (size_t)n_edge,
                                       hipMemcpyDeviceToDevice, c->stream));
                n_live = n_edge;
                int it = 0;
                while (n_live > 0) {
                    a.hp.iter = it;
                    a.live = c->d_live;
                    a.n_live = n_live;
                    a.n_live_dev = nullptr;
                    c->last_edge_live.push_back(n_live);
                    s = sweep(n_live, n_live);
                    if (s) return s;
                    GZM_HIP(hipGetLastError());
                    int n_next = 0;
                    s = compact(c, c->d_live, c->d_flags, n_live, c->d_live2, &n_next);
                    if (s) return s;
                    std::swap(c->d_live, c->d_live2);
                    n_live = n_next;
                    if (n_live > 0 && ++it > c->p.maxiter) { status = GZM_ERR_HSML_NOT_CONVERGED; break; }
                }
                Counters hm;
                s = read_counters(c, kCtrMisc, &hm);
                if (s) return s;
                if (hm.status == 6 && status == GZM_OK) status = GZM_ERR_BRACKET;
            }
        }
        a.ctr = ctr;
    }
    if (na > 0) {
        Timed tm(c, kOther);
        launch_density_finalize(c->stream, na, c->d_active, c->d_posm, c->d_velu, c->d_hsml, c->d_num_ngb,
                                c->d_rho, c->d_pres, c->G.G9, (c->p.method == 1) ? c->d_pvel : nullptr);
    }
    GZM_HIP(hipGetLastError());
    c->have_density = true;
    c->have_gradients = false;
    c->list_valid = false;
    c->frec_fresh = false;
    if (dcache && na == c->n) c->dcache_all = true;  // every particle went through the first sweep
    if (gp) return GZM_OK;
    Counters h;
    s = read_counters(c, kCtrDensity, &h);
    if (s) return s;
    collect_timings(c);
    if (h.status == 6 && status == GZM_OK) status = GZM_ERR_BRACKET;
    if (stats) {
        stats->density_sweeps = sweeps;
        stats->density_interactions = (int64_t)h.interactions;
        stats->density_updates = updates;
        stats->density_fp64_redo = n_edge;
        stats->status = status;
    }
    if (status != GZM_OK) set_error("%s", gzm_status_string(status));
    return (gzm_status)status;
}

gzm_status gradients_impl(gzm_ctx *c, const Plan *gp, gzm_pass_stats *stats)
{
    gzm_status s;
    if (!gp) {
        s = reset_counters(c, kCtrGradients);
        if (s) return s;
    }
    s = compute_hmax(c);
    if (s) return s;
    {
        Timed tm(c, kOther);
        if (c->p.precision != 0)
            launch_pack_grad_f32(c->stream, c->n, c->d_offm, c->d_velf, c->d_hsml, c->d_rho, c->d_pres, c->d_grecf);
        else
            launch_pack_grad(c->stream, c->n, c->d_posm, c->d_velu, c->d_hsml, c->d_rho, c->d_pres, c->d_grec);
    }
    GradientArgs a;
    memset(&a, 0, sizeof(a));
    a.block = c->block[1];
    a.boxes = c->d_boxes;
    a.box_of = c->d_box_of;
    a.tree = c->tree;
    a.active = c->d_active;
    a.n_active = c->n_active;
    a.rec = c->d_grec;
    a.nvt = c->d_nvt;
    a.cn = c->d_cn;
    a.grad = c->d_grad;
    a.posm = c->d_posm;
    a.hsml = c->d_hsml;
    a.rho = c->d_rho;
    a.pres = c->d_pres;
    a.recf = c->d_grecf;
    a.f32 = f32_geom(c);
    a.ctr = c->d_ctr + kCtrGradients;
    // every particle active: the gradient kernel also writes the force records
    // (the force stage then skips its pack kernel)
    const bool fuse_records = c->all_active && c->n_active == c->n;
    if (fuse_records) {
        if (c->p.precision != 0) a.frecf_out = c->d_frecf;
        else a.frec_out = c->d_frec;
        a.dhsml = c->d_dhsml;
        a.fce = c->d_fce;
        a.gamma = c->p.gamma;
    }
    if (c->n_active > 0 && c->list_cap > 0) {
        if (c->capturing && ((size_t)c->n_active + 63) / 64 * (size_t)c->list_cap * 64 > c->list_entries) {
            set_error("pair list too small while capturing");
            return GZM_ERR_STATE;
        }
        s = ensure_list(c);
        if (s) return s;
        a.list.idx = c->d_list;
        a.list.count = c->d_list_count;
        a.list.overflow = c->d_list_over;
        a.list.cap = c->list_cap;
    }
    // symmetric pair rows instead of the pair walk (tree search, FP64 density
    // sums, every particle active, candidate rows of every particle built since
    // the upload)
    if (c->p.precision != 2 && c->tree.nodes && c->d_brow && c->dcache_all && fuse_records && c->n_active > 0) {
        SymRowArgs r;
        r.boxes = c->d_boxes;
        r.box_of = c->d_box_of;
        r.tree = c->tree;
        r.n = c->n;
        r.posm = c->d_posm;
        r.hsml = c->d_hsml;
        r.dcache_idx = c->d_dcache_idx;
        r.dcache_n = c->d_dcache_n;
        r.dcache_r2 = c->d_dcache_r2;
        r.brow = c->d_brow;
        r.bcount = c->d_bcount;
        r.wlist = c->d_bwave;
        r.wcount = c->d_bwave_n;
        {
            Timed tm(c, kGradients);
            launch_sym_rows(c->stream, r);
        }
        a.dcache_idx = c->d_dcache_idx;
        a.dcache_n = c->d_dcache_n;
        a.dcache_r2 = c->d_dcache_r2;
        a.brow = c->d_brow;
        a.bcount = c->d_bcount;
    }
    if (c->n_active > 0) {
        Timed tm(c, kGradients);
        if (c->p.precision != 0) launch_gradients_f32(c->stream, a);
        else launch_gradients(c->stream, a);
    }
    GZM_HIP(hipGetLastError());
    c->have_gradients = true;
    c->frec_fresh = fuse_records && c->n_active > 0;
    c->list_valid = (c->n_active > 0 && c->list_cap > 0);
    c->list_version = c->active_version;
    if (gp) {
        c->list_overflow_slices = gp->overflow ? 1 : 0;
        return GZM_OK;
    }
    Counters h;
    s = read_counters(c, kCtrGradients, &h);
    if (s) return s;
    collect_timings(c);
    if (stats) stats->gradient_interactions = (int64_t)h.interactions;
    c->list_overflow_slices = h.overflow_slices;
    return GZM_OK;
}

gzm_status force_impl(gzm_ctx *c, const Plan *gp, gzm_pass_stats *stats)
{
    gzm_status s;
    if (!gp) {
        s = reset_counters(c, kCtrForce);
        if (s) return s;
    }
    s = compute_hmax(c);
    if (s) return s;
    {
        Timed tm(c, kOther);
        if (c->frec_fresh) {
            // written by the gradient stage of this pass
        } else if (c->p.precision != 0)
            launch_pack_force_f32(c->stream, c->n, c->d_offm, c->d_velf, c->d_cellc, c->d_posm, c->d_hsml, c->d_rho,
                                  c->d_pres, c->d_dhsml, c->d_cn, c->d_fce, c->d_nvt, c->d_grad, c->p.gamma,
                                  c->d_frecf);
        else
            launch_pack_force(c->stream, c->n, c->d_posm, c->d_velu, c->d_hsml, c->d_rho, c->d_pres, c->d_dhsml,
                              c->d_cn, c->d_fce, c->d_nvt, c->d_grad, c->p.gamma, c->d_frec);
        if (!c->all_active) {
            launch_fill_u8(c->stream, c->d_is_active, c->n, 0);
            launch_set_flags(c->stream, c->n_active, c->d_active, c->d_is_active);
            launch_fill_u8(c->stream, c->d_wakeup, c->n, 0);
        }
    }
    ForceArgs a;
    memset(&a, 0, sizeof(a));
    a.block = c->block[2];
    a.boxes = c->d_boxes;
    a.box_of = c->d_box_of;
    a.tree = c->tree;
    a.G = c->G;
    a.active = c->d_active;
    a.n_active = c->n_active;
    a.max_hsml = c->p.max_hsml;
    a.rec = c->d_frec;
    a.is_active = c->all_active ? nullptr : c->d_is_active;
    a.vsig_old = c->d_vsig;
    a.acc = c->d_acc;
    a.dtu = c->d_dtu;
    a.vsig = c->d_vsig;
    a.wakeup = c->d_wakeup;
    a.posm = c->d_posm;
    a.velu = c->d_velu;
    a.hsml = c->d_hsml;
    a.recf = c->d_frecf;
    a.rho = c->d_rho;
    a.pres = c->d_pres;
    a.dhsml = c->d_dhsml;
    a.cn = c->d_cn;
    a.fce = c->d_fce;
    a.nvt = c->d_nvt;
    a.grad = c->d_grad;
    a.redo_flag = c->d_redo_flag;
    a.redo_list = c->d_redo_list;
    a.redo_count = c->d_redo_count;
    a.n = c->n;
    a.f32 = f32_geom(c);
    a.ctr = c->d_ctr + kCtrForce;
    const bool use_list = c->list_valid && c->list_version == c->active_version;
    if (use_list) {
        a.list.idx = c->d_list;
        a.list.count = c->d_list_count;
        a.list.overflow = c->d_list_over;
        a.list.cap = c->list_cap;
    }
    const bool mfv = c->p.method == 1;
    MfvArgs m;
    memset(&m, 0, sizeof(m));
    if (mfv) {
        m.pvel = c->d_pvel;
        m.ti = c->d_ti;
        m.timebase = c->p.timebase_interval;
        m.dt_mass = c->d_dtmass;
        m.d_mass = c->d_dmass;
        m.grav_work = c->d_gravwork;
    }
    // exact-solver pairs leave the list loop for a kernel of their own (lists of <= 64)
    const bool three_phase = use_list && c->list_cap <= 64 && c->n_active > 0;
    if (three_phase) {
        s = ensure_defer(c);
        if (s) return s;
        a.defer.mask = c->d_defer_mask;
        a.defer.count = c->d_defer_count;
        a.defer.raw = c->d_defer_raw;
        a.defer.offset = c->d_defer_off;
        a.defer.total = c->d_defer_total;
    }
    if (c->n_active > 0) {
        Timed tm(c, kForce);
        // a kernel, not a memset: this stage may be captured into a graph
        if (c->p.precision != 0) launch_fill_int(c->stream, c->d_redo_count, 2, 0);
        if (mfv) {
            launch_mfv_preloop(c->stream, a, m);
            launch_force_mfv(c->stream, a, m);
        } else {
            auto launch = (c->p.precision != 0) ? launch_force_f32 : launch_force;
            launch(c->stream, a);
        }
        if ((!mfv || three_phase) && use_list && c->list_overflow_slices > 0) {
            // particles whose pair set overflowed the list: one wavefront each, FP64
            if (!c->d_long_list) {
                s = dev_alloc(&c->d_long_list, (size_t)c->n);
                s = s ? s : dev_alloc(&c->d_long_count, 1);
                if (s) return s;
            }
            launch_fill_int(c->stream, c->d_long_count, 1, 0);
            if (mfv) launch_force_mfv_overflow(c->stream, a, m, c->d_long_list, c->d_long_count);
            else launch_force_long_lists(c->stream, a, c->d_long_list, c->d_long_count);
        }
    }
    GZM_HIP(hipGetLastError());
    if (three_phase) {
        size_t bytes = 0;
        GZM_HIP(rocprim::exclusive_scan(nullptr, bytes, c->d_defer_count, c->d_defer_off, 0, (size_t)c->n_active,
                                        rocprim::plus<int>(), c->stream));
        if (c->capturing && bytes > c->temp_bytes) { set_error("temporary storage too small while capturing"); return GZM_ERR_STATE; }
        s = ensure_temp(c, bytes);
        if (s) return s;
        {
            Timed tm(c, kForceDeferred);
            GZM_HIP(rocprim::exclusive_scan(c->d_temp, bytes, c->d_defer_count, c->d_defer_off, 0, (size_t)c->n_active,
                                            rocprim::plus<int>(), c->stream));
            launch_defer_total(c->stream, c->n_active, c->d_defer_off, c->d_defer_count, c->d_defer_total);
        }
        int64_t D = 0;
        if (gp) {
            D = gp->deferred;
        } else {
            int h_total = 0;
            GZM_HIP(hipMemcpyAsync(&h_total, c->d_defer_total, sizeof(int), hipMemcpyDeviceToHost, c->stream));
            GZM_HIP(hipStreamSynchronize(c->stream));
            D = h_total;
            // results of one round: all deferred pairs, or 2 per particle if more
            const size_t want = (size_t)std::min<int64_t>(D, std::max<int64_t>(1 << 20, 2 * (int64_t)c->n));
            if (want > c->defer_result_cap) {
                dev_free(c->d_defer_result);
                dev_free(c->d_defer_pair);
                c->defer_result_cap = 0;
                s = dev_alloc(&c->d_defer_result, want * (mfv ? 9 : 5));
                s = s ? s : dev_alloc(&c->d_defer_pair, want);
                if (s) return s;
                c->defer_result_cap = want;
            }
        }
        c->last_deferred = D;
        if (D > 0 && c->defer_result_cap == 0) { set_error("deferred-pair buffer missing while capturing"); return GZM_ERR_STATE; }
        const int64_t chunk = D > 0 ? (int64_t)c->defer_result_cap : 1;
        a.defer.result = c->d_defer_result;
        a.defer.pair = c->d_defer_pair;
        for (int64_t q0 = 0; q0 == 0 || q0 < D; q0 += chunk) {
            const int64_t q1 = std::min<int64_t>(D, q0 + chunk);
            a.defer.q_begin = (int)q0;
            a.defer.q_end = (int)q1;
            Timed tm(c, kForceDeferred);
            if (q1 > q0) launch_defer_owners(c->stream, a);
            if (mfv) {
                launch_force_mfv_deferred(c->stream, a, m, (int)(q1 - q0));
                launch_force_mfv_finish(c->stream, a, m);
            } else {
                if (c->p.precision != 0) launch_force_deferred_f32(c->stream, a, (int)(q1 - q0));
                else launch_force_deferred(c->stream, a, (int)(q1 - q0));
                launch_force_finish(c->stream, a);
            }
            if (D == 0) break;
        }
        GZM_HIP(hipGetLastError());
    }
    if (c->p.precision != 0 && c->n_active > 0) {
        Timed tm(c, kForceDeferred);
        launch_force_redo_fp64(c->stream, a);
        GZM_HIP(hipGetLastError());
    }
    if (gp) return GZM_OK;
    Counters h;
    s = read_counters(c, kCtrForce, &h);
    if (s) return s;
    collect_timings(c);
    if (stats) {
        stats->force_interactions = (int64_t)h.interactions;
        stats->riemann_retries = (int64_t)h.retries;
        stats->exact_solver_calls = (int64_t)h.exact_calls;
        stats->fp64_force_particles = (int64_t)h.redo_n;
        stats->status = h.status;
    }
    if (h.status != 0) {
        set_error("%s", gzm_status_string(h.status));
        return (gzm_status)h.status;
    }
    return GZM_OK;
}

// Arguments of the two Riemann-state entry points, checked before any device
// work; G receives the gamma constants.
gzm_status riemann_args(int32_t device, int32_t mode, int32_t precision, double gamma, int64_t n, int64_t n_max,
                        const double *L, const double *R, const double *normal, const double *limiter,
                        const double *out, const int32_t *mask, GammaConsts &G)
{
    if (!L || !R || !normal || !out || !mask || (mode == 0 && !limiter)) { set_error("NULL argument"); return GZM_ERR_ARG; }
    if (mode != 0 && mode != 1) { set_error("mode must be 0 (Riemann_solver) or 1 (KT flux)"); return GZM_ERR_ARG; }
    if (precision < 0 || precision > 2) { set_error("precision must be 0, 1 or 2"); return GZM_ERR_ARG; }
    if (!(gamma > 1)) { set_error("gamma must be > 1"); return GZM_ERR_ARG; }
    if (n < 0 || n > n_max) {
        set_error("n=%lld out of range [0, %lld]", (long long)n, (long long)n_max);
        return GZM_ERR_ARG;
    }
    if (n == 0) return GZM_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) { set_error("no HIP device"); return GZM_ERR_NO_DEVICE; }
    if (device < 0 || device >= ndev) { set_error("device %d not present", device); return GZM_ERR_ARG; }
    GZM_HIP(hipSetDevice(device));
    G.G0 = gamma;
    G.G1 = ((gamma - 1.0) / (2.0 * gamma));
    G.G2 = ((gamma + 1.0) / (2.0 * gamma));
    G.G3 = ((2.0 * gamma / (gamma - 1.0)));
    G.G4 = (2.0 / (gamma - 1.0));
    G.G5 = (2.0 / (gamma + 1.0));
    G.G6 = ((gamma - 1.0) / (gamma + 1.0));
    G.G7 = (0.5 * (gamma - 1.0));
    G.G8 = (1.0 / gamma);
    G.G9 = (gamma - 1.0);
    return GZM_OK;
}

// doubles of device memory per face state: L[5], R[5], normal[3], limiter, out[6]
constexpr size_t kRiemannDoubles = 5 + 5 + 3 + 1 + 6;

// Copy n face states into the device buffers d (n * kRiemannDoubles doubles)
// and dm (n ints), solve them on the null stream, copy the results back.
gzm_status riemann_run(int32_t mode, int32_t precision, const GammaConsts &G, int64_t n, const double *L,
                       const double *R, const double *normal, const double *limiter, double *out, int32_t *mask,
                       double *d, int *dm)
{
    const size_t nn = (size_t)n;
    double *dL = d, *dR = d + 5 * nn, *dn = d + 10 * nn, *dlim = d + 13 * nn, *dout = d + 14 * nn;
    hipError_t e = hipMemcpy(dL, L, sizeof(double) * 5 * nn, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dR, R, sizeof(double) * 5 * nn, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dn, normal, sizeof(double) * 3 * nn, hipMemcpyHostToDevice);
    if (e == hipSuccess && limiter) e = hipMemcpy(dlim, limiter, sizeof(double) * nn, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        launch_riemann_states(nullptr, mode, precision, G, n, dL, dR, dn, dlim, dout, dm);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, dout, sizeof(double) * 6 * nn, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(mask, dm, sizeof(int) * nn, hipMemcpyDeviceToHost);
    if (e != hipSuccess) { set_error("Riemann states: %s", hipGetErrorString(e)); return GZM_ERR_HIP; }
    return GZM_OK;
}

// Largest n of gzm_riemann_states_noalloc, and the size of its staging area:
// the face states, results and masks of one call (kNoAllocStates * (20
// doubles + 1 int) = 10 MiB + 256 KiB), which every device the library
// runs on holds until the process ends.
constexpr int64_t kNoAllocStates = GZM_RIEMANN_NOALLOC_MAX;
constexpr size_t kNoAllocDoubles = (size_t)kNoAllocStates * kRiemannDoubles + (size_t)kNoAllocStates / 2;

}  // namespace

// Device array behind gzm_riemann_states_noalloc.  In host code this name is a
// host-side stand-in, not a device address: host code uses only the address
// that hipGetSymbolAddress returns for the current device (scratch_of_device).
static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxScratchDevices] = {};

// The device address of `scratch` on `device`, which must be current: looked
// up on the first call on that device, then kept per device.
gzm_status scratch_of_device(int device, double **ptr)
{
    if (device >= kMaxScratchDevices) { set_error("device %d beyond %d", device, kMaxScratchDevices); return GZM_ERR_ARG; }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
            return GZM_ERR_HIP;
        }
        g_scratch_ptr[device] = static_cast<double *>(p);
    }
    *ptr = g_scratch_ptr[device];
    return GZM_OK;
}

}  // namespace

extern "C" {


gzm_status gzm_upload(gzm_ctx *c, const gzm_host_view *v)
{
    if (!c || !v) { set_error("NULL argument"); return GZM_ERR_ARG; }
    const int64_t offsets[2] = {0, v->n};
    return upload_common(c, v, 1, c->p.box, offsets);
}

gzm_status gzm_upload_batch(gzm_ctx *c, const gzm_host_view *v, int32_t n_boxes, const double *box_sizes,
                            const int64_t *box_offsets)
{
    if (!c || !v || !box_sizes || !box_offsets) { set_error("NULL argument"); return GZM_ERR_ARG; }
    return upload_common(c, v, n_boxes, box_sizes, box_offsets);
}

gzm_status gzm_restore_initial_state(gzm_ctx *c)
{
    if (!c || c->n == 0) { set_error("no particles uploaded"); return GZM_ERR_STATE; }
    return restore_state(c);
}

gzm_status gzm_density(gzm_ctx *c, const int32_t *active, int64_t n_active, gzm_pass_stats *stats)
{
    if (!c || c->n == 0) { set_error("no particles uploaded"); return GZM_ERR_STATE; }
    GZM_HIP(hipSetDevice(c->p.device));
    gzm_status s = set_active(c, active, n_active);
    if (s) return s;
    return density_impl(c, nullptr, stats);
}

gzm_status gzm_gradients(gzm_ctx *c, const int32_t *active, int64_t n_active, gzm_pass_stats *stats)
{
    if (!c || c->n == 0) { set_error("no particles uploaded"); return GZM_ERR_STATE; }
    if (!c->have_density) { set_error("gradients need density results (run gzm_density or upload them)"); return GZM_ERR_STATE; }
    GZM_HIP(hipSetDevice(c->p.device));
    gzm_status s = set_active(c, active, n_active);
    if (s) return s;
    return gradients_impl(c, nullptr, stats);
}

gzm_status gzm_hydro_force(gzm_ctx *c, const int32_t *active, int64_t n_active, gzm_pass_stats *stats)
{
    if (!c || c->n == 0) { set_error("no particles uploaded"); return GZM_ERR_STATE; }
    if (!c->have_density) { set_error("force needs density results"); return GZM_ERR_STATE; }
    GZM_HIP(hipSetDevice(c->p.device));
    gzm_status s = set_active(c, active, n_active);
    if (s) return s;
    return force_impl(c, nullptr, stats);
}

gzm_status gzm_run_passes(gzm_ctx *c, int32_t n_passes, gzm_pass_stats *stats)
{
    if (!c || c->n == 0) { set_error("no particles uploaded"); return GZM_ERR_STATE; }
    if (n_passes < 1 || n_passes > 100000) { set_error("n_passes=%d out of range [1, 100000]", n_passes); return GZM_ERR_ARG; }
    GZM_HIP(hipSetDevice(c->p.device));
    // probe pass (eager): sizes every buffer, gives the sweep count and the list overflow
    gzm_pass_stats probe;
    memset(&probe, 0, sizeof(probe));
    gzm_status s = restore_state(c);
    s = s ? s : set_active(c, nullptr, 0);
    s = s ? s : density_impl(c, nullptr, &probe);
    s = s ? s : gradients_impl(c, nullptr, &probe);
    s = s ? s : force_impl(c, nullptr, &probe);
    if (s) return s;
    const Plan gp = {probe.density_sweeps, c->list_overflow_slices > 0, c->last_deferred, c->last_sweep_live,
                     c->last_edge_live};
    {
        // the captured compactions run over the whole active list
        size_t bytes = 0;
        GZM_HIP(rocprim::select(nullptr, bytes, c->d_active, c->d_flags, c->d_live2, c->d_count, (size_t)c->n_active,
                                c->stream));
        s = ensure_temp(c, bytes);
        if (s) return s;
    }
    GZM_HIP(hipStreamSynchronize(c->stream));
    collect_timings(c);

    hipGraph_t graph = nullptr;
    GZM_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    for (int k = 0; k < kCtrSlots; k++) launch_clear_counters(c->stream, c->d_ctr + k);
    for (int p = 0; p < n_passes && s == GZM_OK; p++) {
        s = restore_state(c);
        s = s ? s : density_impl(c, &gp, nullptr);
        s = s ? s : gradients_impl(c, &gp, nullptr);
        s = s ? s : force_impl(c, &gp, nullptr);
    }
    c->capturing = false;
    const hipError_t ec = hipStreamEndCapture(c->stream, &graph);
    if (s != GZM_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return s;
    }
    if (ec != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        set_error("hipStreamEndCapture failed: %s", hipGetErrorString(ec));
        return GZM_ERR_HIP;
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
        set_error("hipGraphInstantiate failed: %s", hipGetErrorString(ei));
        return GZM_ERR_HIP;
    }
    hipEvent_t e0 = take_event(c), e1 = take_event(c);
    if (e0) (void)hipEventRecord(e0, c->stream);
    const hipError_t el = hipGraphLaunch(exec, c->stream);
    if (e1) (void)hipEventRecord(e1, c->stream);
    const hipError_t es = hipStreamSynchronize(c->stream);
    (void)hipGraphExecDestroy(exec);
    float ms = 0;
    if (e0 && e1 && el == hipSuccess && es == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) {
        c->timing.graph_ms += ms;
        c->timing.graph_launches++;
        c->timing.graph_passes += n_passes;
    }
    if (e0) c->event_pool.push_back(e0);
    if (e1) c->event_pool.push_back(e1);
    if (el != hipSuccess || es != hipSuccess) {
        set_error("graph launch failed: %s", hipGetErrorString(el != hipSuccess ? el : es));
        return GZM_ERR_HIP;
    }
    Counters h[kCtrSlots];
    GZM_HIP(hipMemcpy(h, c->d_ctr, sizeof(h), hipMemcpyDeviceToHost));
    int status = 0;
    for (int k = 0; k < kCtrSlots; k++) status = std::max(status, h[k].status);
    if (status == 6) status = GZM_ERR_BRACKET;
    if (stats) {
        stats->density_sweeps = gp.sweeps;
        stats->density_interactions = (int64_t)h[kCtrDensity].interactions;
        stats->gradient_interactions = (int64_t)h[kCtrGradients].interactions;
        stats->force_interactions = (int64_t)h[kCtrForce].interactions;
        stats->riemann_retries = (int64_t)h[kCtrForce].retries;
        stats->exact_solver_calls = (int64_t)h[kCtrForce].exact_calls;
        stats->fp64_force_particles = (int64_t)h[kCtrForce].redo_n;
        stats->density_updates = 0;
        stats->status = status;
    }
    if (status != 0) {
        set_error("%s", gzm_status_string(status));
        return (gzm_status)status;
    }
    return GZM_OK;
}

gzm_status gzm_download(gzm_ctx *c, gzm_host_view *v)
{
    if (!c || !v) { set_error("NULL argument"); return GZM_ERR_ARG; }
    if (c->n == 0 || v->n != c->n) { set_error("view size %lld does not match uploaded %d", (long long)v->n, c->n); return GZM_ERR_ARG; }
    GZM_HIP(hipSetDevice(c->p.device));
    GZM_HIP(hipStreamSynchronize(c->stream));
    const int n = c->n;
    struct Out { double *h; const double *d; int w; } outs[] = {
        {v->hsml, c->d_hsml, 1}, {v->cn, c->d_cn, 1}, {v->vsig, c->d_vsig, 1}, {v->num_ngb, c->d_num_ngb, 1},
        {v->dhsml, c->d_dhsml, 1}, {v->divvel, c->d_divvel, 1}, {v->density, c->d_rho, 1},
        {v->pressure, c->d_pres, 1}, {v->fce, c->d_fce, 1}, {v->acc, c->d_acc, 3}, {v->dtu, c->d_dtu, 1},
        {v->particle_vel, c->d_pvel, 3}, {v->dt_mass, c->d_dtmass, 1}, {v->d_mass, c->d_dmass, 1},
        {v->grav_work, c->d_gravwork, 3}};
    for (auto &o : outs) {
        if (!o.h || !o.d) continue;
        gzm_status s = download_sorted(c, o.d, o.h, o.w);
        if (s) return s;
    }
    if (v->nvt) {
        std::vector<double> t6((size_t)n * 6);
        gzm_status s = download_sorted(c, c->d_nvt, t6.data(), 6);
        if (s) return s;
        for (int i = 0; i < n; i++) {
            const double *t = t6.data() + 6 * (size_t)i;
            double *o = v->nvt + 9 * (size_t)i;
            o[0] = t[0]; o[1] = t[1]; o[2] = t[2];
            o[3] = t[1]; o[4] = t[3]; o[5] = t[4];
            o[6] = t[2]; o[7] = t[4]; o[8] = t[5];
        }
    }
    if (v->grad_rho || v->grad_p || v->grad_v) {
        std::vector<double> g((size_t)n * 15);
        gzm_status s = download_sorted(c, c->d_grad, g.data(), 15);
        if (s) return s;
        for (int i = 0; i < n; i++) {
            const double *o = g.data() + 15 * (size_t)i;
            if (v->grad_rho) for (int k = 0; k < 3; k++) v->grad_rho[3 * (size_t)i + k] = o[k];
            if (v->grad_p) for (int k = 0; k < 3; k++) v->grad_p[3 * (size_t)i + k] = o[3 + k];
            if (v->grad_v) for (int k = 0; k < 9; k++) v->grad_v[9 * (size_t)i + k] = o[6 + k];
        }
    }
    if (v->wakeup) {
        std::vector<uint8_t> w(n);
        GZM_HIP(hipMemcpyAsync(w.data(), c->d_wakeup, n, hipMemcpyDeviceToHost, c->stream));
        std::vector<int> perm(n);
        GZM_HIP(hipMemcpyAsync(perm.data(), c->d_perm, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
        GZM_HIP(hipStreamSynchronize(c->stream));
        for (int k = 0; k < n; k++) v->wakeup[perm[k]] = w[k];
    }
    return GZM_OK;
}

gzm_status gzm_neighbor_counts(gzm_ctx *c, int32_t mode, const int32_t *active, int64_t n_active, int32_t *counts,
                               int64_t *total)
{
    if (!c || c->n == 0) { set_error("no particles uploaded"); return GZM_ERR_STATE; }
    if (mode != 0 && mode != 1) { set_error("mode must be 0 or 1"); return GZM_ERR_ARG; }
    GZM_HIP(hipSetDevice(c->p.device));
    // counts come back in the order of `active`, so map without re-sorting
    int na;
    if (active == nullptr) {
        na = c->n;
        GZM_HIP(hipMemcpyAsync(c->d_live, c->d_inv, sizeof(int) * c->n, hipMemcpyDeviceToDevice, c->stream));
    } else {
        if (n_active < 0 || n_active > c->n) { set_error("n_active out of range"); return GZM_ERR_ARG; }
        for (int64_t k = 0; k < n_active; k++)
            if (active[k] < 0 || active[k] >= c->n) { set_error("active[%lld] out of range", (long long)k); return GZM_ERR_ARG; }
        na = (int)n_active;
        GZM_HIP(hipMemcpyAsync(c->d_active_caller, active, sizeof(int) * na, hipMemcpyHostToDevice, c->stream));
        launch_map_active(c->stream, na, c->d_active_caller, c->d_inv, c->d_live);
    }
    gzm_status s = reset_counters(c, kCtrMisc);
    if (s) return s;
    s = compute_hmax(c);
    if (s) return s;
    launch_neighbor_counts(c->stream, c->d_boxes, c->d_box_of, c->tree, mode, na, c->d_live, c->d_posm, c->d_hsml,
                           counts ? c->d_flags : nullptr, c->d_ctr + kCtrMisc);
    GZM_HIP(hipGetLastError());
    if (counts) GZM_HIP(hipMemcpyAsync(counts, c->d_flags, sizeof(int) * na, hipMemcpyDeviceToHost, c->stream));
    Counters h;
    s = read_counters(c, kCtrMisc, &h);
    if (s) return s;
    if (total) *total = (int64_t)h.interactions;
    c->all_active = false;  // d_live was used as scratch; the active list is rebuilt on the next stage call
    c->active_h.clear();
    c->n_active = 0;
    return GZM_OK;
}

gzm_status gzm_riemann_states(int32_t device, int32_t mode, int32_t precision, double gamma, int64_t n,
                              const double *L, const double *R, const double *normal, const double *limiter,
                              double *out, int32_t *mask)
{
    GammaConsts G;
    gzm_status s = riemann_args(device, mode, precision, gamma, n, (int64_t)1 << 28, L, R, normal, limiter, out, mask, G);
    if (s || n == 0) return s;
    const size_t nn = (size_t)n;
    double *d = nullptr;
    int *dm = nullptr;
    s = dev_alloc(&d, nn * kRiemannDoubles);
    if (s) return s;
    s = dev_alloc(&dm, nn);
    if (s) { dev_free(d); return s; }
    s = riemann_run(mode, precision, G, n, L, R, normal, limiter, out, mask, d, dm);
    dev_free(d);
    dev_free(dm);
    return s;
}

gzm_status gzm_riemann_states_noalloc(int32_t device, int32_t mode, int32_t precision, double gamma, int64_t n,
                                      const double *L, const double *R, const double *normal,
                                      const double *limiter, double *out, int32_t *mask)
{
    GammaConsts G;
    gzm_status s = riemann_args(device, mode, precision, gamma, n, kNoAllocStates, L, R, normal, limiter, out, mask, G);
    if (s || n == 0) return s;
    double *d = nullptr;
    s = scratch_of_device(device, &d);
    if (s) return s;
    // riemann_run's layout for n states inside the area of kNoAllocStates, the masks after all the doubles
    return riemann_run(mode, precision, G, n, L, R, normal, limiter, out, mask, d,
                       reinterpret_cast<int *>(d + (size_t)kNoAllocStates * kRiemannDoubles));
}

gzm_status gzm_riemann_states_mfv(int32_t device, int32_t mode, double gamma, int64_t n, const double *L,
                                  const double *R, const double *normal, const double *limiter, double *out,
                                  int32_t *mask)
{
    if (mode < 0 || mode > 2) { set_error("mode must be 0 (Riemann_solver), 1 (KT flux) or 2 (exact solver)"); return GZM_ERR_ARG; }
    GammaConsts G;
    // riemann_args checks pointers, n, gamma and the device (mode 0 needs the limiter)
    gzm_status s = riemann_args(device, mode == 0 ? 0 : 1, 0, gamma, n, (int64_t)1 << 27, L, R, normal, limiter, out,
                                mask, G);
    if (s || n == 0) return s;
    const size_t nn = (size_t)n;
    double *d = nullptr;
    int *dm = nullptr;
    s = dev_alloc(&d, nn * (5 + 5 + 3 + 1 + 7));
    if (s) return s;
    s = dev_alloc(&dm, nn);
    if (s) { dev_free(d); return s; }
    double *dL = d, *dR = d + 5 * nn, *dn = d + 10 * nn, *dlim = d + 13 * nn, *dout = d + 14 * nn;
    hipError_t e = hipMemcpy(dL, L, sizeof(double) * 5 * nn, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dR, R, sizeof(double) * 5 * nn, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dn, normal, sizeof(double) * 3 * nn, hipMemcpyHostToDevice);
    if (e == hipSuccess && limiter) e = hipMemcpy(dlim, limiter, sizeof(double) * nn, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        launch_riemann_states_mfv(nullptr, mode, G, n, dL, dR, dn, dlim, dout, dm);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, dout, sizeof(double) * 7 * nn, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(mask, dm, sizeof(int) * nn, hipMemcpyDeviceToHost);
    dev_free(d);
    dev_free(dm);
    if (e != hipSuccess) { set_error("MFV Riemann states: %s", hipGetErrorString(e)); return GZM_ERR_HIP; }
    return GZM_OK;
}

gzm_status gzm_timing_get(gzm_ctx *c, gzm_timing *t)
{
    if (!c || !t) { set_error("NULL argument"); return GZM_ERR_ARG; }
    (void)hipStreamSynchronize(c->stream);
    collect_timings(c);
    *t = c->timing;
    return GZM_OK;
}

gzm_status gzm_timing_reset(gzm_ctx *c)
{
    if (!c) { set_error("NULL argument"); return GZM_ERR_ARG; }
    (void)hipStreamSynchronize(c->stream);
    collect_timings(c);
    memset(&c->timing, 0, sizeof(c->timing));
    return GZM_OK;
}

}  // extern "C"
