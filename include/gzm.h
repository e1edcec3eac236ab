/*
 * gzm.h -- C ABI of the MI355X-native MFM hydro pass (libgzm.so).
 *
 * The library replaces the arithmetic and neighbour search behind GIZMO's
 * three hydro entry points for one rank's particle set on one GPU:
 *
 *   gzm_density()      replaces density()              (hydro/density.c:451-1044,
 *                                                        proto.h:723)
 *   gzm_gradients()    replaces hydro_gradient_calc()  (hydro/gradients.c:577-1469,
 *                                                        proto.h:962)
 *   gzm_hydro_force()  replaces hydro_force()          (hydro/hydro_toplevel.c:1045-1057,
 *                                                        proto.h:749)
 *
 * The reference runs each as an instance of its neighbour-operator template
 * (system/code_block_xchange_perform_ops.h) over the active list
 * FirstActiveParticle/NextActiveParticle[] (allvars.h:1652-1654) with an
 * octree search (ngb.c:102-130).  Here the search is a cell grid over
 * particles sorted by cell key, or (gzm_params.search = 1) a radix tree over
 * particles sorted by Morton key, and every stage is a HIP kernel for gfx950.
 *
 * Conventions
 *  - Plain pointers and sizes only.  Host arrays are in the caller's particle
 *    order, FP64, contiguous: pos/vel/acc [n][3], nvt [n][9] row-major,
 *    grad_rho/grad_p [n][3], grad_v [n][3][3] with grad_v[i][k][d] = dv_k/dx_d
 *    (the reference's SphP.Gradients.Velocity[k][d]).
 *  - Active lists hold caller indices; NULL means "all particles".
 *  - Calls are synchronous with respect to the host; device state stays
 *    resident between calls.  A context is not re-entrant; use one per host
 *    thread.  The library never aborts: failures the reference turns into
 *    endrun(code) come back as that code (1155, 1234, 1234545).
 */
#ifndef GZM_H
#define GZM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gzm_status {
    GZM_OK = 0,
    GZM_ERR_ARG = 1,            /* bad argument (NULL pointer, n <= 0, size mismatch) */
    GZM_ERR_HIP = 2,            /* HIP runtime error (message via gzm_last_error) */
    GZM_ERR_NOMEM = 3,          /* device allocation failed */
    GZM_ERR_STATE = 4,          /* call out of order (e.g. force before density) */
    GZM_ERR_NO_DEVICE = 5,      /* no usable GPU */
    GZM_ERR_BRACKET = 6,        /* Left == Right == 0 in the hsml iteration (density.c:780-783) */
    GZM_ERR_HSML_NOT_CONVERGED = 1155, /* density.c:859 */
    GZM_ERR_RIEMANN = 1234,            /* hydro_core_meshless.h:262 */
    GZM_ERR_RIEMANN_NEG_DENSITY = 1234545 /* reimann.h:714 */
} gzm_status;

/* Run-time parameters: the All.* fields the three stages read
 * (begrun.c:1227-1242 tags, allvars.h:1091-1113 compile-time defaults). */
typedef struct gzm_params {
    double box[3];          /* boxSize_X/Y/Z; the box is periodic in all axes */
    double gamma;           /* EOS_GAMMA, 5/3 by default */
    double des_num_ngb;     /* All.DesNumNgb */
    double max_num_ngb_dev; /* All.MaxNumNgbDeviation */
    double min_hsml;        /* All.MinHsml */
    double max_hsml;        /* All.MaxHsml */
    int32_t first_step;     /* 1 when All.Time == All.TimeBegin (density.c:509) */
    int32_t maxiter;        /* MAXITER, 150 (allvars.h:1294) */
    int32_t device;         /* HIP device ordinal */
    int32_t precision;      /* 0: plain FP64 pass; 1: FP32 bulk: density stage as in
                               the FP64 pass (same hsml, densities and neighbour sets,
                               bit for bit), gradient and force pairs in FP32 against
                               FP64 anchors with FP64 limiter, exact Riemann solver and
                               flux sums; 2: as 1 with the density pair sums in FP32 too
                               (hsml within ~1e-7 of the FP64 pass, so a few pairs per
                               million sit on the other side of the kernel edge) */
    double cell_size;       /* grid cell side; 0: the mean initial hsml of each box;
                               < 0: -cell_size times that mean */
    int32_t pair_list_cap;  /* neighbours kept per particle in the pair list that the
                               gradient loop hands to the force loop: 0 = default (64),
                               rounded up to a multiple of 16, at most 4096; < 0 = no
                               list (the force loop scans the grid); the 64-particle
                               slices holding a particle with more pairs fall back to
                               the scan */
    int32_t search;         /* neighbour search: 0 = cell grid (default); 1 = tree over
                               Morton-sorted particles with per-node hsml maxima, for
                               strongly clustered sets */
    int32_t density_budget; /* candidates a density-sweep thread tests before its particle
                               is deferred to a second launch over such particles only
                               (load balance where a few hsml jump far in one iteration):
                               0 = default (512), < 0 = no limit; results do not depend
                               on it */
    int32_t block_size[3];  /* threads per workgroup of the per-particle loop kernels of
                               [0] the density sweeps, [1] the gradient loop, [2] the force
                               loop: 64, 128, 256 or 512 (the force loop's grid-scanning
                               fallbacks stay at 256); 0 = 256.  Results do not depend on
                               it */
    int32_t wave_sweep_max; /* a density sweep over at most this many particles gives each
                               one a wavefront whose lanes share its walk: 0 = default
                               (4096), < 0 = never.  Sums then run in another order (lane
                               partial sums), within the parity bar */
    int32_t method;         /* 0: meshless finite mass (HYDRO_MESHLESS_FINITE_MASS, default);
                               1: meshless finite volume (HYDRO_MESHLESS_FINITE_VOLUME) with
                               the smoothed mesh motion HYDRO_FIX_MESH_MOTION = 5
                               (allvars.h:94-98): density() also sets ParticleVel, and
                               hydro_force() exchanges mass (DtMass, dMass, GravWorkTerm).
                               Method 1 runs in FP64 (precision 0) only */
    double timebase_interval; /* method 1: physical time of one integer time unit,
                               UNIT_INTEGERTIME_IN_PHYSICAL (allvars.h:908), which turns
                               the mass flux into the exchanged mass dMass */
} gzm_params;

/* Host arrays of one particle set, caller order.  Inputs are read by
 * gzm_upload; every non-NULL output is written by gzm_download. */
typedef struct gzm_host_view {
    int64_t n;
    /* inputs: P[].Pos, SphP[].VelPred, P[].Mass, SphP[].InternalEnergyPred */
    const double *pos;
    const double *vel;
    const double *mass;
    const double *u;
    /* state carried between passes (in at upload, out at download):
     * P[].Hsml, SphP[].ConditionNumber, SphP[].MaxSignalVel */
    double *hsml;
    double *cn;
    double *vsig;
    /* state of inactive particles that active ones read (in at upload when
     * non-NULL; recomputed for active particles by the stages) */
    double *num_ngb;   /* P[].NumNgb (NumNgb^(1/3) after density) */
    double *dhsml;     /* P[].DhsmlNgbFactor */
    double *divvel;    /* P[].Particle_DivVel */
    double *density;   /* SphP[].Density */
    double *pressure;  /* SphP[].Pressure */
    double *nvt;       /* SphP[].NV_T (inverse, after density) [n][9] */
    double *fce;       /* SphP[].FaceClosureError */
    double *grad_rho;  /* SphP[].Gradients.Density [n][3] */
    double *grad_p;    /* SphP[].Gradients.Pressure [n][3] */
    double *grad_v;    /* SphP[].Gradients.Velocity [n][3][3] */
    /* force outputs: SphP[].HydroAccel, SphP[].DtInternalEnergy, wakeup flags
     * (acc and dtu are also read at upload when non-NULL, so that inactive
     * particles keep their values) */
    double *acc;
    double *dtu;
    uint8_t *wakeup;
    /* method 1 (MFV) only; each may be NULL.  ti_step: P[].dt_step, the
     * integer timestep (GET_PARTICLE_INTEGERTIME with WAKEUP, allvars.h:914),
     * NULL = equal for every particle; particle_vel: SphP[].ParticleVel [n][3]
     * (out after density; in at upload for particles that stay inactive);
     * dt_mass, grav_work [n][3]: SphP[].DtMass, GravWorkTerm (out); d_mass:
     * SphP[].dMass (in at upload and out: inactive particles keep their value
     * plus the shares their active neighbours give them) */
    const int64_t *ti_step;
    double *particle_vel;
    double *dt_mass;
    double *d_mass;
    double *grav_work;
} gzm_host_view;

typedef struct gzm_pass_stats {
    int32_t density_sweeps;          /* iterations of the hsml loop */
    int32_t status;
    int64_t density_interactions;    /* sum over sweeps of |{j : r_ij^2 < h_i^2}| */
    int64_t gradient_interactions;   /* sum over active i of the pair-set size */
    int64_t force_interactions;      /* idem for the force loop */
    int64_t riemann_retries;         /* pairs that needed the first-order retry */
    int64_t exact_solver_calls;      /* pairs that reached the exact solver */
    int64_t density_updates;         /* sum over sweeps of the particles swept */
    int64_t fp64_force_particles;    /* FP32-bulk force loop: particles whose loop was redone
                                        in FP64 because one of its pairs' branch decisions lay
                                        within FP32 rounding of its threshold */
    int64_t density_fp64_redo;       /* precision 2: particles whose hsml iteration was redone
                                        with FP64 sums because they converged with a neighbour
                                        within 1e-5 of the kernel edge (their sweeps are not in
                                        density_interactions) */
} gzm_pass_stats;

/* Per-kernel device time (HIP events on the context stream), accumulated
 * since the last gzm_timing_reset. */
typedef struct gzm_timing {
    double density_ms;   int64_t density_launches;  /* density sweeps, one thread per particle */
    double gradients_ms; int64_t gradients_launches;
    double force_ms;     int64_t force_launches;
    double other_ms;     int64_t other_launches;   /* grid build, packing, finalize */
    double graph_ms;     int64_t graph_launches;   /* gzm_run_passes graphs, whole-graph time */
    int64_t graph_passes;                          /* passes inside those graphs */
    double force_deferred_ms; int64_t force_deferred_launches;  /* exact-solver pairs deferred
                                                   out of the force loop, their scan and finish */
    double density_wave_ms; int64_t density_wave_launches;  /* density sweeps run one wavefront
                                                   per particle: the particles past the candidate
                                                   budget, and sweeps of few particles */
} gzm_timing;

typedef struct gzm_ctx gzm_ctx;

const char *gzm_status_string(int status);
const char *gzm_last_error(void);

/* Context: owns the stream, the cell grid and every device array. */
gzm_status gzm_ctx_create(gzm_ctx **ctx, const gzm_params *params);
gzm_status gzm_ctx_destroy(gzm_ctx *ctx);

/* Copy one particle set to the device, sort it by cell and build the grid
 * (the role of domain_Decomposition + force_treebuild, domain.c:106,
 * gravity/forcetree.c:109, for this rank). */
gzm_status gzm_upload(gzm_ctx *ctx, const gzm_host_view *view);

/* Many independent periodic boxes in one context (batched pass; no
 * counterpart in the reference, which runs one rank's particle set): box b
 * holds the caller rows [box_offsets[b], box_offsets[b+1]) of `view` and has
 * side lengths box_sizes[3b..3b+2]; gzm_params.box is not used.  Every stage
 * call then runs over all boxes at once, each particle seeing only its own
 * box, with the results of running each box alone. */
gzm_status gzm_upload_batch(gzm_ctx *ctx, const gzm_host_view *view, int32_t n_boxes, const double *box_sizes,
                            const int64_t *box_offsets);

/* density(): hsml iteration to convergence over the active list, then the
 * final density, pressure and NumNgb^(1/3) (density.c:451-1044). */
gzm_status gzm_density(gzm_ctx *ctx, const int32_t *active, int64_t n_active, gzm_pass_stats *stats);

/* hydro_gradient_calc(): least-squares gradients of rho, P, v with the slope
 * limiter (gradients.c:577-1469). */
gzm_status gzm_gradients(gzm_ctx *ctx, const int32_t *active, int64_t n_active, gzm_pass_stats *stats);

/* hydro_force(): MFM faces, HLLC fluxes, dv/dt, du/dt, v_sig, wakeup flags
 * (hydro_toplevel.c:961-1057, hydro_evaluate.h:22-516); with method 1 the MFV
 * fluxes of mass, momentum and energy, DtMass, dMass and GravWorkTerm
 * (hydro_core_meshless.h with HYDRO_MESHLESS_FINITE_VOLUME, hydro_evaluate.h:
 * 379-401).  dMass receives each neighbour's share by atomic adds, so its
 * last bits depend on their order; every other output is deterministic. */
gzm_status gzm_hydro_force(gzm_ctx *ctx, const int32_t *active, int64_t n_active, gzm_pass_stats *stats);

/* Copy results back into the caller-order host arrays of `view`. */
gzm_status gzm_download(gzm_ctx *ctx, gzm_host_view *view);

/* Restore hsml, cn and vsig to the values given at upload (repeated passes
 * over the same input, e.g. for timing). Device-side copy only. */
gzm_status gzm_restore_initial_state(gzm_ctx *ctx);

/* n_passes full passes over the uploaded state (restore initial state,
 * density, gradients, force; all particles active) enqueued into one HIP
 * graph and launched once.  One eager probe pass first fixes what the graph
 * cannot decide on the host (number of density sweeps, pair-list overflow);
 * the graph replays the same input, so every pass repeats the probe's
 * decisions.  stats hold the sums over the n_passes passes (density_sweeps:
 * per pass); gzm_timing's graph fields hold the graph's device time.  The
 * outputs left on the device are those of the last pass.  No counterpart in
 * the reference: for callers that repeat passes without host round trips. */
gzm_status gzm_run_passes(gzm_ctx *ctx, int32_t n_passes, gzm_pass_stats *stats);

/* Neighbour-set sizes for the current hsml, caller order of `active`:
 * mode 0 density set {j : r2 < h_i^2}, mode 1 pair set
 * {j : r2 > 0 and (r2 < h_i^2 or r2 < h_j^2)}.  counts may be NULL. */
gzm_status gzm_neighbor_counts(gzm_ctx *ctx, int32_t mode, const int32_t *active, int64_t n_active,
                               int32_t *counts, int64_t *total);

/* Riemann_solver() for MFM hydro (hydro/reimann.h:286-388) on n independent
 * face states, one per GPU thread, with the same device code as the force
 * loop: element-level checks of branches that particle sets rarely reach.
 * Host arrays: L and R [n][5] = rho, p, v[3]; normal [n][3] (unit); limiter
 * [n] = press_tot_limiter (hydro_core_meshless.h:182-200); out [n][6] = P_M,
 * S_M, F_v[3], F_p; mask [n] = stages used (1 HLLC, 2 KT, 4 exact; 64 in FP32
 * when an estimate lay within the FP32 margin of its threshold, where the
 * force loop would redo the pair in FP64), or -1234545 on the exact
 * solver's negative-density exit.  mode 0: the whole solver; mode 1: the
 * Kurganov-Tadmor/PWK flux alone (reimann.h:484-519; limiter unused).
 * precision 0: FP64; 1 or 2: FP32 with the FP64 exact solver. */
gzm_status gzm_riemann_states(int32_t device, int32_t mode, int32_t precision, double gamma, int64_t n,
                              const double *L, const double *R, const double *normal, const double *limiter,
                              double *out, int32_t *mask);

/* gzm_riemann_states without device allocation: the same arguments, kernel
 * and results, with the face states staged through a fixed device array of
 * the library (one per GPU, 10 MiB + 256 KiB, held until the process ends)
 * instead of buffers allocated per call.  n is at most 65536
 * (GZM_RIEMANN_NOALLOC_MAX); a larger n returns GZM_ERR_ARG.  No counterpart
 * in the reference: for callers that cannot call hipMalloc (harnesses that
 * forbid it, tight loops).  Every call on one GPU uses that GPU's array, so
 * calls that overlap on one GPU (from several host threads) race; calls are
 * synchronous, so sequential calls are safe. */
#define GZM_RIEMANN_NOALLOC_MAX 65536
gzm_status gzm_riemann_states_noalloc(int32_t device, int32_t mode, int32_t precision, double gamma, int64_t n,
                                      const double *L, const double *R, const double *normal,
                                      const double *limiter, double *out, int32_t *mask);

/* Riemann_solver() of a meshless finite-volume run (reimann.h:286-388 with
 * the HYDRO_MESHLESS_FINITE_VOLUME branches) on n independent face states
 * in FP64, the device code of the MFV force loop: mode 0 the whole solver
 * (HLLC with the four-wave fluxes, the KT flux, the exact solver past the
 * limiter), 1 the KT flux alone (reimann.h:484-507), 2 the exact solver with
 * its face sampling alone (reimann.h:685-1030, 1146-1166).  L, R [n][5] =
 * rho, p, v[3]; normal [n][3]; limiter [n] (mode 0); out [n][7] = P_M, S_M,
 * F_rho, F_v[3], F_p through a face at rest in the states' frame; mask [n] =
 * stages used (mode 0: 1 HLLC, 2 KT, 4 exact), or -1234545 on the exact
 * solver's negative-density exit.  Element-level checks of branches particle
 * sets rarely reach (vacuum fans, vacuum inside the fan). */
gzm_status gzm_riemann_states_mfv(int32_t device, int32_t mode, double gamma, int64_t n, const double *L,
                                  const double *R, const double *normal, const double *limiter, double *out,
                                  int32_t *mask);

gzm_status gzm_timing_get(gzm_ctx *ctx, gzm_timing *timing);
gzm_status gzm_timing_reset(gzm_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
