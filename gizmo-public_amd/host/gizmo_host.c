/*
 * gizmo_host.c -- GIZMO's hydro entry points over P[]/SphP[] for one rank,
 * running on the MI355X library (include/gzm.h).  See
 * include/gizmo_mfm_host.h for the contract.
 *
 * Each entry point gathers the fields the stage reads from the particle
 * structures into contiguous arrays, calls the library, and scatters the
 * results back for the active particles -- the role the reference's
 * particle2in / out2particle functions play around its neighbour loops
 * (density.c:116-244, gradients.c:254-532, hydro_toplevel.c:397-638).
 */
#include "gizmo_mfm_host.h"

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gzm.h"

struct particle_data *P = NULL;
struct gas_cell_data *SphP = NULL;
struct global_data_all_processes All;
int NumPart = 0, N_gas = 0;
int FirstActiveParticle = -1;
int *NextActiveParticle = NULL;

static gzm_ctx *g_ctx = NULL;
static gzm_params g_params;      /* parameters g_ctx was created with */
static int g_device = 0;
static int g_method = 0;         /* 0 MFM, 1 MFV */
static int g_status = 0;
static int g_resident = 0;       /* the device holds this pass's particle set (uploaded by density()) */
static gizmo_endrun_handler g_endrun = NULL;
static char g_err[512];

/* host staging arrays (caller = P[] index order) */
static int g_cap = 0;
static double *h_pos, *h_vel, *h_mass, *h_u, *h_hsml, *h_cn, *h_vsig, *h_num_ngb, *h_dhsml, *h_divvel;
static double *h_rho, *h_pres, *h_nvt, *h_fce, *h_grho, *h_gp, *h_gv, *h_acc, *h_dtu;
static double *h_pvel, *h_dtm, *h_dm, *h_gw;  /* MFV: ParticleVel, DtMass, dMass, GravWorkTerm */
static int64_t *h_ti;                          /* MFV: integer timesteps */
static unsigned char *h_wakeup;
static int *h_active;
static unsigned char *h_seen;    /* active-list walk: particles already listed */

void gizmo_host_set_endrun_handler(gizmo_endrun_handler handler) { g_endrun = handler; }

void endrun(int code)
{
    if (g_endrun) {
        g_endrun(code, g_err);
        return;
    }
    fprintf(stderr, "endrun called with an error level of %d: %s\n", code, g_err);
    fflush(stderr);
    exit(code);
}

int gizmo_host_last_status(void) { return g_status; }
const char *gizmo_host_last_error(void) { return g_err; }
int gizmo_host_set_device(int device)
{
    g_device = device;
    return 0;
}

int gizmo_host_set_method(int method)
{
    if (method != 0 && method != 1) return -1;
    g_method = method;
    return 0;
}

static void free_staging(void)
{
    double **arrs[] = {&h_pos, &h_vel, &h_mass, &h_u, &h_hsml, &h_cn, &h_vsig, &h_num_ngb, &h_dhsml, &h_divvel,
                       &h_rho, &h_pres, &h_nvt, &h_fce, &h_grho, &h_gp, &h_gv, &h_acc, &h_dtu,
                       &h_pvel, &h_dtm, &h_dm, &h_gw};
    for (size_t k = 0; k < sizeof(arrs) / sizeof(arrs[0]); k++) { free(*arrs[k]); *arrs[k] = NULL; }
    free(h_ti); h_ti = NULL;
    free(h_wakeup); h_wakeup = NULL;
    free(h_active); h_active = NULL;
    free(h_seen); h_seen = NULL;
    g_cap = 0;
}

static int ensure_staging(int n)
{
    if (n <= g_cap) return 0;
    free_staging();
    size_t N = (size_t)n;
    h_pos = malloc(3 * N * sizeof(double)); h_vel = malloc(3 * N * sizeof(double));
    h_mass = malloc(N * sizeof(double)); h_u = malloc(N * sizeof(double)); h_hsml = malloc(N * sizeof(double));
    h_cn = malloc(N * sizeof(double)); h_vsig = malloc(N * sizeof(double)); h_num_ngb = malloc(N * sizeof(double));
    h_dhsml = malloc(N * sizeof(double)); h_divvel = malloc(N * sizeof(double)); h_rho = malloc(N * sizeof(double));
    h_pres = malloc(N * sizeof(double)); h_nvt = malloc(9 * N * sizeof(double)); h_fce = malloc(N * sizeof(double));
    h_grho = malloc(3 * N * sizeof(double)); h_gp = malloc(3 * N * sizeof(double)); h_gv = malloc(9 * N * sizeof(double));
    h_acc = malloc(3 * N * sizeof(double)); h_dtu = malloc(N * sizeof(double));
    h_pvel = malloc(3 * N * sizeof(double)); h_dtm = malloc(N * sizeof(double)); h_dm = malloc(N * sizeof(double));
    h_gw = malloc(3 * N * sizeof(double)); h_ti = malloc(N * sizeof(int64_t));
    h_wakeup = malloc(N); h_active = malloc(N * sizeof(int)); h_seen = calloc(N, 1);
    if (!h_pos || !h_vel || !h_mass || !h_u || !h_hsml || !h_cn || !h_vsig || !h_num_ngb || !h_dhsml || !h_divvel ||
        !h_rho || !h_pres || !h_nvt || !h_fce || !h_grho || !h_gp || !h_gv || !h_acc || !h_dtu || !h_wakeup || !h_active ||
        !h_seen || !h_pvel || !h_dtm || !h_dm || !h_gw || !h_ti) {
        free_staging();
        return -1;
    }
    g_cap = n;
    return 0;
}

int gizmo_host_allocate(int n_gas)
{
    gizmo_host_free();
    if (n_gas <= 0) return -1;
    P = calloc((size_t)n_gas, sizeof(struct particle_data));
    SphP = calloc((size_t)n_gas, sizeof(struct gas_cell_data));
    NextActiveParticle = malloc((size_t)n_gas * sizeof(int));
    if (!P || !SphP || !NextActiveParticle) { gizmo_host_free(); return -1; }
    NumPart = N_gas = n_gas;
    All.MaxPart = n_gas;
    for (int i = 0; i < n_gas; i++) {
        SphP[i].ConditionNumber = 1;   /* init.c:514 */
        SphP[i].FaceClosureError = 0;  /* init.c:516 */
        SphP[i].MaxSignalVel = 1e-56;
    }
    gizmo_host_all_active();
    return 0;
}

void gizmo_host_free(void)
{
    free(P); P = NULL;
    free(SphP); SphP = NULL;
    free(NextActiveParticle); NextActiveParticle = NULL;
    NumPart = N_gas = 0;
    FirstActiveParticle = -1;
    g_resident = 0;
}

void gizmo_host_all_active(void)
{
    for (int i = 0; i < NumPart; i++) NextActiveParticle[i] = (i + 1 < NumPart) ? i + 1 : -1;
    FirstActiveParticle = (NumPart > 0) ? 0 : -1;
}

int gizmo_host_set_active(const int *idx, int n)
{
    for (int k = 0; k < n; k++)
        if (idx[k] < 0 || idx[k] >= NumPart) return -1;
    FirstActiveParticle = (n > 0) ? idx[0] : -1;
    for (int k = 0; k < n; k++) NextActiveParticle[idx[k]] = (k + 1 < n) ? idx[k + 1] : -1;
    return 0;
}

void gizmo_host_release(void)
{
    if (g_ctx) gzm_ctx_destroy(g_ctx);
    g_ctx = NULL;
    g_resident = 0;
}

double gizmo_box_side(int k) { return All.BoxSize * ((All.BoxLong[k] > 0) ? All.BoxLong[k] : 1.0); }
double gizmo_gamma(void) { return (All.GasGamma > 1) ? All.GasGamma : 5.0 / 3.0; /* EOS_GAMMA (allvars.h:1092) */ }

static int fail(int status, const char *where)
{
    g_status = status;
    snprintf(g_err, sizeof(g_err), "%s: %s (%s)", where, gzm_status_string(status), gzm_last_error());
    return status;
}

/* FirstActiveParticle/NextActiveParticle -> gas indices in h_active (the
 * loops' condition P[i].Type == 0; mass and density tests happen inside the
 * library).  skip_converged: leave out particles whose TimeBin is negative,
 * the mark density() puts on particles that have finished iterating
 * (density_isactive, density.c:34).  A list that leaves [0, NumPart), visits a
 * particle twice or runs longer than NumPart steps is reported through
 * endrun instead of being followed. */
static int collect_active(int skip_converged, int *n_out)
{
    int n = 0, steps = 0, bad = 0;
    for (int i = FirstActiveParticle; i >= 0; i = NextActiveParticle[i]) {
        if (i >= NumPart || ++steps > NumPart) { bad = 1; break; }
        if (i >= N_gas || P[i].Type != 0) continue;
        if (h_seen[i]) { bad = 1; break; }
        h_seen[i] = 1;
        if (skip_converged && P[i].TimeBin < 0) continue;
        h_active[n++] = i;
    }
    /* clear the marks of the walk (listed or skipped) */
    steps = 0;
    for (int i = FirstActiveParticle; i >= 0 && i < NumPart && steps <= NumPart; i = NextActiveParticle[i], steps++)
        if (i < N_gas) h_seen[i] = 0;
    if (bad) {
        g_status = GZM_ERR_ARG;
        snprintf(g_err, sizeof(g_err), "malformed active list (index outside [0, %d), repeated or cyclic)", NumPart);
        return GZM_ERR_ARG;
    }
    *n_out = n;
    return 0;
}

static int ensure_ctx(void)
{
    gzm_params p;
    memset(&p, 0, sizeof(p));
    for (int k = 0; k < 3; k++) p.box[k] = gizmo_box_side(k);
    p.gamma = gizmo_gamma();
    p.des_num_ngb = All.DesNumNgb;
    p.max_num_ngb_dev = All.MaxNumNgbDeviation;
    p.min_hsml = All.MinHsml;
    p.max_hsml = All.MaxHsml;
    p.first_step = (All.Time == All.TimeBegin);
    p.maxiter = 150; /* MAXITER (allvars.h:1294) */
    p.device = g_device;
    p.precision = 0;
    p.cell_size = 0;
    p.pair_list_cap = 0;
    p.method = g_method;
    p.timebase_interval = g_method ? All.Timebase_interval : 0;
    /* the context (stream, device buffers) lives across passes; it is made
     * again only when a parameter it was created with changes (the first
     * step's tolerance, a new box size or device) */
    if (g_ctx && memcmp(&p, &g_params, sizeof(p)) == 0) return 0;
    if (g_ctx) gzm_ctx_destroy(g_ctx);
    g_ctx = NULL;
    g_resident = 0;
    int s = gzm_ctx_create(&g_ctx, &p);
    if (s) return fail(s, "gzm_ctx_create");
    g_params = p;
    return 0;
}

/* P[]/SphP[] -> staging -> device: positions, velocities, masses, energies and
 * the whole carried state (hsml, NumNgb, DhsmlNgbFactor, DivVel, density,
 * pressure, NV_T, ConditionNumber, FaceClosureError, gradients,
 * MaxSignalVel), so that particles outside the active list enter the
 * gradient and force loops of active neighbours with their stored values, as
 * the reference's loops read P[j]/SphP[j] (gradients.c:1477-2089,
 * hydro_evaluate.h:146-498) */
static int upload(void)
{
    int n = N_gas;
    if (ensure_staging(n)) return fail(GZM_ERR_NOMEM, "host staging");
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) {
            h_pos[3 * i + k] = P[i].Pos[k];
            h_vel[3 * i + k] = SphP[i].VelPred[k];
        }
        h_mass[i] = P[i].Mass;
        h_u[i] = SphP[i].InternalEnergyPred;
        h_hsml[i] = P[i].Hsml;
        h_cn[i] = SphP[i].ConditionNumber;
        h_vsig[i] = SphP[i].MaxSignalVel;
        h_num_ngb[i] = P[i].NumNgb;
        h_dhsml[i] = P[i].DhsmlNgbFactor;
        h_divvel[i] = P[i].Particle_DivVel;
        h_rho[i] = SphP[i].Density;
        h_pres[i] = SphP[i].Pressure;
        h_fce[i] = SphP[i].FaceClosureError;
        memcpy(h_nvt + 9 * (size_t)i, SphP[i].NV_T, 9 * sizeof(double));
        memcpy(h_grho + 3 * (size_t)i, SphP[i].Gradients.Density, 3 * sizeof(double));
        memcpy(h_gp + 3 * (size_t)i, SphP[i].Gradients.Pressure, 3 * sizeof(double));
        memcpy(h_gv + 9 * (size_t)i, SphP[i].Gradients.Velocity, 9 * sizeof(double));
        if (g_method) {
            /* GET_PARTICLE_INTEGERTIME: 2^TimeBin, 0 for bin 0 (allvars.h:909-912); a
             * negative bin is density()'s mark of a converged particle */
            const int bin = (P[i].TimeBin < 0) ? -P[i].TimeBin - 1 : P[i].TimeBin;
            h_ti[i] = bin ? ((int64_t)1 << bin) : 0;
            memcpy(h_pvel + 3 * (size_t)i, SphP[i].ParticleVel, 3 * sizeof(double));
            memcpy(h_gw + 3 * (size_t)i, SphP[i].GravWorkTerm, 3 * sizeof(double));
            h_dtm[i] = SphP[i].DtMass;
            h_dm[i] = SphP[i].dMass;
        }
    }
    gzm_host_view v;
    memset(&v, 0, sizeof(v));
    v.n = n;
    v.pos = h_pos; v.vel = h_vel; v.mass = h_mass; v.u = h_u;
    v.hsml = h_hsml; v.cn = h_cn; v.vsig = h_vsig;
    v.num_ngb = h_num_ngb; v.dhsml = h_dhsml; v.divvel = h_divvel; v.fce = h_fce;
    v.density = h_rho; v.pressure = h_pres; v.nvt = h_nvt;
    v.grad_rho = h_grho; v.grad_p = h_gp; v.grad_v = h_gv;
    if (g_method) {
        v.ti_step = h_ti; v.particle_vel = h_pvel; v.dt_mass = h_dtm; v.d_mass = h_dm; v.grav_work = h_gw;
    }
    int s = gzm_upload(g_ctx, &v);
    if (s) return fail(s, "gzm_upload");
    return 0;
}

enum { DL_DENSITY = 1, DL_GRADIENTS = 2, DL_FORCE = 4 };

/* device -> staging: the arrays a stage wrote (out2particle's role) */
static int download(int what)
{
    gzm_host_view v;
    memset(&v, 0, sizeof(v));
    v.n = N_gas;
    if (what & DL_DENSITY) {
        v.hsml = h_hsml; v.cn = h_cn; v.num_ngb = h_num_ngb; v.dhsml = h_dhsml; v.divvel = h_divvel;
        v.density = h_rho; v.pressure = h_pres; v.nvt = h_nvt; v.fce = h_fce;
        if (g_method) v.particle_vel = h_pvel;
    }
    if (what & DL_GRADIENTS) { v.grad_rho = h_grho; v.grad_p = h_gp; v.grad_v = h_gv; }
    if (what & DL_FORCE) {
        v.acc = h_acc; v.dtu = h_dtu; v.vsig = h_vsig; v.wakeup = h_wakeup;
        if (g_method) { v.dt_mass = h_dtm; v.d_mass = h_dm; v.grav_work = h_gw; }
    }
    int s = gzm_download(g_ctx, &v);
    if (s) return fail(s, "gzm_download");
    return 0;
}

void density(void)
{
    g_status = 0;
    if (N_gas <= 0) return;
    if (ensure_staging(N_gas)) { fail(GZM_ERR_NOMEM, "host staging"); endrun(g_status); return; }
    int na = 0;
    if (collect_active(1, &na)) { endrun(g_status); return; }
    if (ensure_ctx() || upload()) { endrun(g_status); return; }
    g_resident = 1;
    gzm_pass_stats st;
    memset(&st, 0, sizeof(st));
    int s = gzm_density(g_ctx, h_active, na, &st);
    if (s) { fail(s, "gzm_density"); endrun(s); return; }
    if (download(DL_DENSITY)) { endrun(g_status); return; }
    for (int a = 0; a < na; a++) {
        int i = h_active[a];
        if (!(P[i].Mass > 0)) continue;
        P[i].Hsml = h_hsml[i];
        P[i].NumNgb = h_num_ngb[i];
        P[i].DhsmlNgbFactor = h_dhsml[i];
        P[i].Particle_DivVel = h_divvel[i];
        SphP[i].Density = h_rho[i];
        SphP[i].Pressure = h_pres[i];
        SphP[i].ConditionNumber = h_cn[i];
        SphP[i].FaceClosureError = h_fce[i];
        memcpy(SphP[i].NV_T, h_nvt + 9 * (size_t)i, 9 * sizeof(double));
        if (g_method) memcpy(SphP[i].ParticleVel, h_pvel + 3 * (size_t)i, 3 * sizeof(double));
    }
    /* "mark as active again" (density.c:869-872): a particle that entered
     * with a negative TimeBin was left out of the iteration */
    int steps = 0;
    for (int i = FirstActiveParticle; i >= 0 && i < NumPart && steps <= NumPart; i = NextActiveParticle[i], steps++)
        if (P[i].TimeBin < 0) P[i].TimeBin = -P[i].TimeBin - 1;
}

/* The gradient and force stages run on the particle set density() uploaded;
 * called on their own (a caller that set P/SphP from elsewhere), they upload
 * it first. */
static int ensure_resident(void)
{
    if (ensure_staging(N_gas)) return fail(GZM_ERR_NOMEM, "host staging");
    if (g_resident && g_ctx) return 0;
    if (ensure_ctx() || upload()) return g_status;
    g_resident = 1;
    return 0;
}

void hydro_gradient_calc(void)
{
    g_status = 0;
    if (N_gas <= 0) return;
    int na = 0;
    if (ensure_resident() || collect_active(0, &na)) { endrun(g_status); return; }
    gzm_pass_stats st;
    memset(&st, 0, sizeof(st));
    int s = gzm_gradients(g_ctx, h_active, na, &st);
    if (s) { fail(s, "gzm_gradients"); endrun(s); return; }
    if (download(DL_GRADIENTS)) { endrun(g_status); return; }
    for (int a = 0; a < na; a++) {
        int i = h_active[a];
        memcpy(SphP[i].Gradients.Density, h_grho + 3 * (size_t)i, 3 * sizeof(double));
        memcpy(SphP[i].Gradients.Pressure, h_gp + 3 * (size_t)i, 3 * sizeof(double));
        memcpy(SphP[i].Gradients.Velocity, h_gv + 9 * (size_t)i, 9 * sizeof(double));
    }
}

void hydro_force(void)
{
    g_status = 0;
    if (N_gas <= 0) return;
    int na = 0;
    if (ensure_resident() || collect_active(0, &na)) { endrun(g_status); return; }
    gzm_pass_stats st;
    memset(&st, 0, sizeof(st));
    int s = gzm_hydro_force(g_ctx, h_active, na, &st);
    if (s) { fail(s, "gzm_hydro_force"); endrun(s); return; }
    if (download(DL_FORCE)) { endrun(g_status); return; }
    for (int a = 0; a < na; a++) {
        int i = h_active[a];
        for (int k = 0; k < 3; k++) SphP[i].HydroAccel[k] = h_acc[3 * (size_t)i + k];
        SphP[i].DtInternalEnergy = h_dtu[i];
        SphP[i].MaxSignalVel = h_vsig[i];
        if (g_method) {
            SphP[i].DtMass = h_dtm[i];
            memcpy(SphP[i].GravWorkTerm, h_gw + 3 * (size_t)i, 3 * sizeof(double));
        }
    }
    /* dMass: active particles' own sums, and the shares active neighbours
     * gave to any particle (SphP[j].dMass -= ..., hydro_evaluate.h:384-392) */
    if (g_method)
        for (int i = 0; i < N_gas; i++) SphP[i].dMass = h_dm[i];
    for (int i = 0; i < N_gas; i++)
        if (h_wakeup[i]) SphP[i].wakeup = 1;
    /* the next pass starts from P/SphP again (positions drift between passes) */
    g_resident = 0;
}

void compute_hydro_densities_and_forces(void)
{
    /* accel.c:57-111: density, force_update_hmax (the library derives the
     * per-cell search bound from the current hsml), gradients, force */
    density();
    if (g_status) return;
    hydro_gradient_calc();
    if (g_status) return;
    hydro_force();
}
