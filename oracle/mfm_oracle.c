/*
 * mfm_oracle.c -- FP64 CPU restatement of GIZMO's MFM hydro pass.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library and
 * the CPU baseline timed by bench.py; the product never links or calls it.
 *
 * Compile-time branch selection restated here (the reference's defaults for a
 * pure-hydro MFM run, scripts/test_problems/sedov.params:1-9):
 *   HYDRO_MESHLESS_FINITE_MASS (allvars.h:85-91), BOX_PERIODIC, SELFGRAVITY_OFF,
 *   KERNEL_FUNCTION=3 cubic spline (kernel.h:10-12), NUMDIMS=3,
 *   SLOPE_LIMITER_TOLERANCE=1 (allvars.h:68-74), WAKEUP=4.1 (allvars.h:76-81),
 *   CONDITION_NUMBER_DANGER=1e3 (allvars.h:1109-1113), no COOLING/GALSF/MAGNETIC,
 *   DOUBLEPRECISION so MyFloat = MyDouble = double (allvars.h:42,1321-1327).
 * Non-comoving, so every All.cf_* factor is 1 (gravity/cosmology.c:48).
 *
 * Build: gcc -O2 -ffp-contract=off -fopenmp -fPIC -shared (see oracle/Makefile).
 * Operation order follows the reference expressions so that equal summation
 * order gives bit-identical results.
 */
#include "mfm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* constants                                                                  */
/* ------------------------------------------------------------------------- */
#define NUMDIMS 3
#define NORM_COEFF 4.188790204786          /* allvars.h:1443 (literal, not 4*pi/3) */
#define KERNEL_NORM (8.0 / M_PI)           /* kernel.h:63 */
#define CONDITION_NUMBER_DANGER 1.0e3      /* allvars.h:1112 */
#define MIN_REAL_NUMBER 1e-56              /* allvars.h:1106-1107 */
#define WAKEUP 4.1                         /* allvars.h:78 */
#define TOL_ITER 1.e-6                     /* reimann.h:12 */
#define NMAX_ITER 1000                     /* reimann.h:13 */

static inline double DMAX(double a, double b) { return (a > b) ? a : b; } /* proto.h:123 */
static inline double DMIN(double a, double b) { return (a < b) ? a : b; } /* proto.h:124 */

/* ------------------------------------------------------------------------- */
/* kernel (kernel.h:182-197, 205-242, 373-374), cubic spline                 */
/* ------------------------------------------------------------------------- */
static inline void kernel_hinv(double h, double *hinv, double *hinv3, double *hinv4)
{
    *hinv = 1.0 / h;
    *hinv3 = *hinv * *hinv * *hinv;
    *hinv4 = *hinv3 * *hinv;
}

/* mode 0: wk and dwk, -1: wk only, +1: dwk only (kernel.h:199-203) */
static inline void kernel_main(double u, double hinv3, double hinv4, double *wk, double *dwk, int mode)
{
    if (u >= 1) { *wk = 0; *dwk = 0; return; }
    if (u < 0.5) {
        if (mode >= 0) *dwk = u * (18.0 * u - 12.0);
        if (mode <= 0) *wk = (1.0 + 6.0 * (u - 1.0) * u * u);
    } else {
        double t1 = (1.0 - u);
        double t2 = t1 * t1;
        if (mode >= 0) *dwk = -6.0 * t2;
        if (mode <= 0) *wk = 2.0 * t2 * t1;
    }
    if (mode >= 0) *dwk *= KERNEL_NORM * hinv4;
    if (mode <= 0) *wk *= KERNEL_NORM * hinv3;
}

void orc_kernel(double u, double h, double *wk, double *dwk)
{
    double hinv, hinv3, hinv4;
    kernel_hinv(h, &hinv, &hinv3, &hinv4);
    kernel_main(u, hinv3, hinv4, wk, dwk, 0);
}

/* ------------------------------------------------------------------------- */
/* 3x3 cofactor inverse and Frobenius condition number (system/system.c:355-391) */
/* ------------------------------------------------------------------------- */
static double matrix_invert_ndims(double T[3][3], double Tinv[3][3])
{
    int j, k;
    double FrobNorm = 0, FrobNorm_inv = 0, detT = 0;
    for (j = 0; j < 3; j++) for (k = 0; k < 3; k++) Tinv[j][k] = 0;
    for (j = 0; j < 3; j++) for (k = 0; k < 3; k++) FrobNorm += T[j][k] * T[j][k];
    detT = T[0][0] * T[1][1] * T[2][2]
         + T[0][1] * T[1][2] * T[2][0]
         + T[0][2] * T[1][0] * T[2][1]
         - T[0][2] * T[1][1] * T[2][0]
         - T[0][1] * T[1][0] * T[2][2]
         - T[0][0] * T[1][2] * T[2][1];
    if ((detT != 0) && !isnan(detT)) {
        Tinv[0][0] = (T[1][1] * T[2][2] - T[1][2] * T[2][1]) / detT;
        Tinv[0][1] = (T[0][2] * T[2][1] - T[0][1] * T[2][2]) / detT;
        Tinv[0][2] = (T[0][1] * T[1][2] - T[0][2] * T[1][1]) / detT;
        Tinv[1][0] = (T[1][2] * T[2][0] - T[1][0] * T[2][2]) / detT;
        Tinv[1][1] = (T[0][0] * T[2][2] - T[0][2] * T[2][0]) / detT;
        Tinv[1][2] = (T[0][2] * T[1][0] - T[0][0] * T[1][2]) / detT;
        Tinv[2][0] = (T[1][0] * T[2][1] - T[1][1] * T[2][0]) / detT;
        Tinv[2][1] = (T[0][1] * T[2][0] - T[0][0] * T[2][1]) / detT;
        Tinv[2][2] = (T[0][0] * T[1][1] - T[0][1] * T[1][0]) / detT;
    }
    for (j = 0; j < 3; j++) for (k = 0; k < 3; k++) FrobNorm_inv += Tinv[j][k] * Tinv[j][k];
    return DMAX(sqrt(FrobNorm * FrobNorm_inv) / NUMDIMS, 1);
}

double orc_matrix_invert(const double Tin[9], double Tout[9])
{
    double T[3][3], Ti[3][3];
    memcpy(T, Tin, sizeof(T));
    double cn = matrix_invert_ndims(T, Ti);
    memcpy(Tout, Ti, sizeof(Ti));
    return cn;
}

/* ------------------------------------------------------------------------- */
/* neighbour search: uniform cell grid (or brute force) + the exact FP64     */
/* predicates of density.c:277-281 and hydro_evaluate.h:170-178.             */
/* ------------------------------------------------------------------------- */
/* Octree search mode (orc_params.tree = 1), the CPU search the reference
 * runs (ngb_treefind_variable_threads / ngb_treefind_pairs_threads, ngb.c:102-130,
 * with the node opening of system/ngb_codeblock_checknode.h:38-60 and node
 * hsml maxima of force_update_hmax, forcetree_update.c:348-476): a cubic
 * root of the largest box side, eight children per node, leaves of at most
 * OT_BUCKET particles.  A node is opened when the particle lies within the
 * search radius of its cube on every axis (nearest periodic image), with
 * radius h_i for the density set and max(h_i, node hmax) for the pair set;
 * every candidate then passes the exact predicates of nb_test, so the sets
 * are those of the grid and of brute force. */
#define OT_BUCKET 8
typedef struct onode {
    double c[3], half, hmax;
    int child[8];  /* -1: none */
    int first, count, leaf;
} onode;

typedef struct otree {
    onode *node;
    int nn, cap;
    int *idx;      /* particles, grouped by leaf */
} otree;

typedef struct grid_t {
    int nc[3];
    double inv_s[3];
    int *start;  /* ncell+1 */
    int *idx;    /* n, ascending particle index inside every cell */
    otree *tree; /* octree mode: the search structure (the grid is not built) */
} grid_t;

static int ot_new(otree *t, const double c[3], double half, int first, int count)
{
    if (t->nn == t->cap) {
        t->cap = t->cap ? 2 * t->cap : 1024;
        t->node = (onode *)realloc(t->node, sizeof(onode) * (size_t)t->cap);
    }
    onode *o = &t->node[t->nn];
    for (int a = 0; a < 3; a++) o->c[a] = c[a];
    o->half = half;
    o->hmax = 0;
    for (int k = 0; k < 8; k++) o->child[k] = -1;
    o->first = first;
    o->count = count;
    o->leaf = 1;
    return t->nn++;
}

/* split node `id` (its particles idx[first, first+count)) into octants */
static void ot_split(otree *t, int id, const double *pos, int *tmp, int depth)
{
    onode o = t->node[id];
    if (o.count <= OT_BUCKET || depth >= 40) return;
    int cnt[8] = {0}, off[9];
    for (int k = 0; k < o.count; k++) {
        const double *x = pos + 3 * (size_t)t->idx[o.first + k];
        int oct = (x[0] >= o.c[0]) * 4 + (x[1] >= o.c[1]) * 2 + (x[2] >= o.c[2]);
        cnt[oct]++;
    }
    off[0] = 0;
    for (int q = 0; q < 8; q++) off[q + 1] = off[q] + cnt[q];
    int fill[8];
    for (int q = 0; q < 8; q++) fill[q] = off[q];
    for (int k = 0; k < o.count; k++) {
        int j = t->idx[o.first + k];
        const double *x = pos + 3 * (size_t)j;
        int oct = (x[0] >= o.c[0]) * 4 + (x[1] >= o.c[1]) * 2 + (x[2] >= o.c[2]);
        tmp[fill[oct]++] = j;
    }
    memcpy(t->idx + o.first, tmp, sizeof(int) * (size_t)o.count);
    t->node[id].leaf = 0;
    for (int q = 0; q < 8; q++) {
        if (!cnt[q]) continue;
        double c[3] = {o.c[0] + ((q & 4) ? 0.5 : -0.5) * o.half, o.c[1] + ((q & 2) ? 0.5 : -0.5) * o.half,
                       o.c[2] + ((q & 1) ? 0.5 : -0.5) * o.half};
        int ch = ot_new(t, c, 0.5 * o.half, o.first + off[q], cnt[q]);
        t->node[id].child[q] = ch;
        ot_split(t, ch, pos, tmp, depth + 1);
    }
}

static double ot_hmax(otree *t, int id, const double *hsml)
{
    onode *o = &t->node[id];
    double m = 0;
    if (o->leaf) {
        for (int k = 0; k < o->count; k++) m = DMAX(m, hsml[t->idx[o->first + k]]);
    } else {
        for (int q = 0; q < 8; q++)
            if (t->node[id].child[q] >= 0) m = DMAX(m, ot_hmax(t, t->node[id].child[q], hsml));
    }
    t->node[id].hmax = m;
    return m;
}

static otree *ot_build(const orc_params *p, int n, const double *pos, const double *hsml)
{
    otree *t = (otree *)calloc(1, sizeof(otree));
    t->idx = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *tmp = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; i++) t->idx[i] = i;
    double L = DMAX(p->box[0], DMAX(p->box[1], p->box[2]));
    double c[3] = {0.5 * L, 0.5 * L, 0.5 * L};
    int root = ot_new(t, c, 0.5 * L, 0, n);
    ot_split(t, root, pos, tmp, 0);
    if (hsml) ot_hmax(t, root, hsml);
    free(tmp);
    return t;
}

static void ot_free(otree *t)
{
    if (!t) return;
    free(t->node);
    free(t->idx);
    free(t);
}

static void grid_build(grid_t *g, const orc_params *p, int n, const double *pos, const double *hsml)
{
    g->tree = NULL;
    if (p->tree && !p->brute_force) {
        g->tree = ot_build(p, n, pos, hsml);
        g->start = NULL;
        g->idx = NULL;
        return;
    }
    double vol = p->box[0] * p->box[1] * p->box[2];
    double s = cbrt(vol * 8.0 / (n > 0 ? n : 1)); /* ~8 particles per cell */
    long long ncell = 1;
    for (int a = 0; a < 3; a++) {
        int c = (int)(p->box[a] / s);
        if (c < 1) c = 1;
        if (c > 1024) c = 1024;
        g->nc[a] = c;
        g->inv_s[a] = c / p->box[a];
        ncell *= c;
    }
    g->start = (int *)calloc((size_t)ncell + 1, sizeof(int));
    g->idx = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *cell = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; i++) {
        int c[3];
        for (int a = 0; a < 3; a++) {
            c[a] = (int)floor(pos[3 * i + a] * g->inv_s[a]);
            if (c[a] < 0) c[a] = 0;
            if (c[a] >= g->nc[a]) c[a] = g->nc[a] - 1;
        }
        cell[i] = (c[0] * g->nc[1] + c[1]) * g->nc[2] + c[2];
        g->start[cell[i] + 1]++;
    }
    for (long long c = 0; c < ncell; c++) g->start[c + 1] += g->start[c];
    int *fill = (int *)malloc(sizeof(int) * (size_t)ncell);
    for (long long c = 0; c < ncell; c++) fill[c] = g->start[c];
    for (int i = 0; i < n; i++) g->idx[fill[cell[i]]++] = i; /* stable: ascending i */
    free(fill);
    free(cell);
}

static void grid_free(grid_t *g)
{
    free(g->start);
    free(g->idx);
    ot_free(g->tree);
}

/* signed nearest-image wrap, NEAREST_XYZ (allvars.h:1519,1550,1556-1559) */
static inline double wrap1(double x, double L)
{
    double half = 0.5 * L;
    return (x > half) ? (x - L) : ((x < -half) ? (x + L) : x);
}

typedef struct nb_t {
    int *j;
    double *dp; /* 3 per entry */
    double *r2;
    int cnt, cap;
} nb_t;

static void nb_push(nb_t *b, int j, const double dp[3], double r2)
{
    if (b->cnt == b->cap) {
        b->cap = b->cap ? 2 * b->cap : 256;
        b->j = (int *)realloc(b->j, sizeof(int) * (size_t)b->cap);
        b->dp = (double *)realloc(b->dp, sizeof(double) * 3 * (size_t)b->cap);
        b->r2 = (double *)realloc(b->r2, sizeof(double) * (size_t)b->cap);
    }
    b->j[b->cnt] = j;
    b->dp[3 * b->cnt + 0] = dp[0];
    b->dp[3 * b->cnt + 1] = dp[1];
    b->dp[3 * b->cnt + 2] = dp[2];
    b->r2[b->cnt] = r2;
    b->cnt++;
}

static void nb_free(nb_t *b)
{
    free(b->j);
    free(b->dp);
    free(b->r2);
    memset(b, 0, sizeof(*b));
}

/* test one candidate: mode 0 density set, mode 1 pair set */
static inline void nb_test(nb_t *out, const orc_params *p, const double *pos, const double *hsml,
                           int i, int j, double h2_i, int mode)
{
    double dp[3];
    dp[0] = pos[3 * i + 0] - pos[3 * j + 0];
    dp[1] = pos[3 * i + 1] - pos[3 * j + 1];
    dp[2] = pos[3 * i + 2] - pos[3 * j + 2];
    dp[1] = wrap1(dp[1], p->box[1]);
    dp[0] = wrap1(dp[0], p->box[0]);
    dp[2] = wrap1(dp[2], p->box[2]);
    double r2 = dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2];
    if (mode == 0) {
        if (r2 < h2_i) nb_push(out, j, dp, r2);
    } else {
        double h_j = hsml[j];
        if ((r2 >= h2_i) && (r2 >= h_j * h_j)) return;
        if (r2 <= 0) return;
        nb_push(out, j, dp, r2);
    }
}

/* collect the neighbour set of particle i for search radius R */
static void nb_collect(nb_t *out, const orc_params *p, const grid_t *g, int n, const double *pos,
                       const double *hsml, int i, double h_i, double R, int mode)
{
    out->cnt = 0;
    double h2_i = h_i * h_i;
    if (p->brute_force || g == NULL) {
        for (int j = 0; j < n; j++) nb_test(out, p, pos, hsml, i, j, h2_i, mode);
        return;
    }
    if (g->tree) {
        /* depth-first walk with an explicit stack */
        const otree *t = g->tree;
        int stack[8 * 64];
        int sp = 0;
        stack[sp++] = 0;
        const double *xi = pos + 3 * (size_t)i;
        while (sp > 0) {
            const onode *o = &t->node[stack[--sp]];
            const double r = (mode == 0) ? h_i : DMAX(h_i, o->hmax);
            int far = 0;
            for (int a = 0; a < 3 && !far; a++)
                if (fabs(wrap1(xi[a] - o->c[a], p->box[a])) - o->half > r) far = 1;
            if (far) continue;
            if (o->leaf) {
                for (int k = 0; k < o->count; k++) nb_test(out, p, pos, hsml, i, t->idx[o->first + k], h2_i, mode);
            } else {
                for (int q = 7; q >= 0; q--)
                    if (o->child[q] >= 0) stack[sp++] = o->child[q];
            }
        }
        return;
    }
    int lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        double x = pos[3 * i + a];
        lo[a] = (int)floor((x - R) * g->inv_s[a] - 1e-6);
        hi[a] = (int)floor((x + R) * g->inv_s[a] + 1e-6);
        if (hi[a] - lo[a] + 1 >= g->nc[a]) { lo[a] = 0; hi[a] = g->nc[a] - 1; }
    }
    for (int cx = lo[0]; cx <= hi[0]; cx++) {
        int wx = ((cx % g->nc[0]) + g->nc[0]) % g->nc[0];
        for (int cy = lo[1]; cy <= hi[1]; cy++) {
            int wy = ((cy % g->nc[1]) + g->nc[1]) % g->nc[1];
            for (int cz = lo[2]; cz <= hi[2]; cz++) {
                int wz = ((cz % g->nc[2]) + g->nc[2]) % g->nc[2];
                int c = (wx * g->nc[1] + wy) * g->nc[2] + wz;
                for (int k = g->start[c]; k < g->start[c + 1]; k++)
                    nb_test(out, p, pos, hsml, i, g->idx[k], h2_i, mode);
            }
        }
    }
}

static double max_hsml(int n, const double *hsml)
{
    double m = 0;
    for (int i = 0; i < n; i++) if (hsml[i] > m) m = hsml[i];
    return m;
}

static void set_threads(const orc_params *p)
{
#ifdef _OPENMP
    if (p->nthreads > 0) omp_set_num_threads(p->nthreads);
#else
    (void)p;
#endif
}

static int cmp_int(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

long long orc_neighbor_counts(const orc_params *p, int n, const double *pos, const double *hsml,
                              const int *active, int n_active, int mode, int *counts)
{
    grid_t g;
    grid_build(&g, p, n, pos, hsml);
    double hmax = max_hsml(n, hsml);
    long long total = 0;
    set_threads(p);
#pragma omp parallel reduction(+ : total)
    {
        nb_t b = {0};
#pragma omp for schedule(dynamic, 256)
        for (int a = 0; a < n_active; a++) {
            int i = active[a];
            double R = (mode == 0) ? hsml[i] : DMAX(hsml[i], hmax);
            nb_collect(&b, p, &g, n, pos, hsml, i, hsml[i], R, mode);
            counts[a] = b.cnt;
            total += b.cnt;
        }
        nb_free(&b);
    }
    grid_free(&g);
    return total;
}

long long orc_neighbor_lists(const orc_params *p, int n, const double *pos, const double *hsml,
                             const int *active, int n_active, int mode,
                             long long *offsets, int *idx, long long capacity)
{
    int *counts = (int *)malloc(sizeof(int) * (size_t)(n_active > 0 ? n_active : 1));
    long long total = orc_neighbor_counts(p, n, pos, hsml, active, n_active, mode, counts);
    offsets[0] = 0;
    for (int a = 0; a < n_active; a++) offsets[a + 1] = offsets[a] + counts[a];
    free(counts);
    if (total > capacity) return -total;
    grid_t g;
    grid_build(&g, p, n, pos, hsml);
    double hmax = max_hsml(n, hsml);
#pragma omp parallel
    {
        nb_t b = {0};
#pragma omp for schedule(dynamic, 256)
        for (int a = 0; a < n_active; a++) {
            int i = active[a];
            double R = (mode == 0) ? hsml[i] : DMAX(hsml[i], hmax);
            nb_collect(&b, p, &g, n, pos, hsml, i, hsml[i], R, mode);
            memcpy(idx + offsets[a], b.j, sizeof(int) * (size_t)b.cnt);
            qsort(idx + offsets[a], (size_t)b.cnt, sizeof(int), cmp_int);
        }
        nb_free(&b);
    }
    grid_free(&g);
    return total;
}

/* ------------------------------------------------------------------------- */
/* density(): hydro/density.c:252-347 (pair sums) and 451-1044 (iteration)    */
/* ------------------------------------------------------------------------- */
static inline double ncorr_from_cn(double cn)
{
    /* density.c:571-582: ncorr = sqrt(1+(cn-c0)/CND) above c0 = 0.1*CND, capped at 2 */
    double c0 = 0.1 * (double)CONDITION_NUMBER_DANGER;
    double ncorr_ngb = 1;
    if (cn > c0) ncorr_ngb = sqrt(1.0 + (cn - c0) / ((double)CONDITION_NUMBER_DANGER));
    if (ncorr_ngb > 2) ncorr_ngb = 2;
    return ncorr_ngb;
}

/* One hsml update of the iteration body (density.c:594-846) from the sweep's
 * normalised neighbour number and DhsmlNgbFactor, with the sweep's target
 * desnumngb and window half-width desnumngbdev (before its growth with the
 * sweep number, which this applies).  Updates h and
 * the brackets Left/Right; returns HSML_DONE (inside the window, or held at
 * Min/MaxHsml), HSML_BRACKETED (the bracket is narrower than 1e-3 of Left:
 * accepted as it is), HSML_COLLAPSED (no bracket at all: density.c:780-783
 * ends the run) or HSML_AGAIN (sweep again with the new h). */
enum { HSML_DONE = 0, HSML_AGAIN = 1, HSML_BRACKETED = 2, HSML_COLLAPSED = -1 };

static int hsml_step(double NumNgb, double Dhsml, double *h_io, double *Left_io, double *Right_io, int iter,
                     double desnumngb, double desnumngbdev, double minsoft, double maxsoft)
{
    double h = *h_io, Left = *Left_io, Right = *Right_io;
    int redo = 0, maxflag = 0, minflag = 0, result = HSML_AGAIN;
    /* the window grows with the sweep so that the iteration ends (density.c:594) */
    if (iter > 1)
        desnumngbdev = DMIN(0.25 * desnumngb, desnumngbdev * exp(0.1 * log(desnumngb / (16. * desnumngbdev)) * (double)iter));
    if ((NumNgb < (desnumngb - desnumngbdev) && h < 0.999 * maxsoft) ||
        (NumNgb > (desnumngb + desnumngbdev) && h > 1.001 * minsoft)) redo = 1;
    if ((h >= 0.999 * maxsoft) && (NumNgb < (desnumngb - desnumngbdev))) {
        redo = 0;
        if (h == maxsoft) maxflag = 0;
        else { redo = 1; h = maxsoft; maxflag = 1; }
    }
    if ((h <= 1.001 * minsoft) && (NumNgb > (desnumngb + desnumngbdev))) {
        redo = 0;
        if (h == minsoft) minflag = 0;
        else { redo = 1; h = minsoft; minflag = 1; }
    }
    if (!redo) return HSML_DONE;
    if (Left > 0 && Right > 0 && (Right - Left) < 1.0e-3 * Left) { result = HSML_BRACKETED; goto out; }
    if ((maxflag == 0) && (minflag == 0)) {
        if (NumNgb < (desnumngb - desnumngbdev)) Left = DMAX(h, Left);
        else {
            if (Right != 0) { if (h < Right) Right = h; }
            else Right = h;
        }
        if (Right > 0 && Left > 0) {
            double maxjump = 0;
            if (iter > 1) maxjump = 0.2 * log(Right / Left);
            if (NumNgb > 1) {
                double jumpvar = Dhsml * log(desnumngb / NumNgb) / NUMDIMS;
                if (iter > 1) { if (fabs(jumpvar) < maxjump) { if (jumpvar < 0) jumpvar = -maxjump; else jumpvar = maxjump; } }
                h *= exp(jumpvar);
            } else {
                h *= 2.0;
            }
            if ((h < Right) && (h > Left)) {
                if (iter > 1) {
                    double hfac = exp(maxjump);
                    if (h > Right / hfac) h = Right / hfac;
                    if (h < Left * hfac) h = Left * hfac;
                }
            } else {
                if (h > Right) h = Right;
                if (h < Left) h = Left;
                h = pow(h * Left * Right, 1.0 / 3.0);
            }
        } else {
            if (Right == 0 && Left == 0) { result = HSML_COLLAPSED; goto out; }
            if (Right == 0 && Left > 0) {
                double fac_lim, fac;
                if (NumNgb > 1) fac_lim = log(desnumngb / NumNgb) / NUMDIMS; else fac_lim = 1.4;
                if ((NumNgb < 2 * desnumngb) && (NumNgb > 0.1 * desnumngb)) {
                    double slope = Dhsml;
                    if (iter > 2 && slope < 1) slope = 0.5 * (slope + 1);
                    fac = fac_lim * slope;
                    if (iter >= 4) { if (Dhsml == 1) fac *= 10; }
                    if (fac < fac_lim + 0.231) h *= exp(fac);
                    else h *= exp(fac_lim + 0.231);
                } else h *= exp(fac_lim);
            }
            if (Right > 0 && Left == 0) {
                double fac_lim, fac;
                if (NumNgb > 1) fac_lim = log(desnumngb / NumNgb) / NUMDIMS; else fac_lim = 1.4;
                if (fac_lim < -1.535) fac_lim = -1.535;
                if ((NumNgb < 2 * desnumngb) && (NumNgb > 0.1 * desnumngb)) {
                    double slope = Dhsml;
                    if (iter > 2 && slope < 1) slope = 0.5 * (slope + 1);
                    fac = fac_lim * slope;
                    if (iter >= 4) { if (Dhsml == 1) fac *= 10; }
                    if (fac > fac_lim - 0.231) h *= exp(fac);
                    else h *= exp(fac_lim - 0.231);
                } else h *= exp(fac_lim);
            }
        }
    }
    if (h < minsoft) h = minsoft;
    if (minflag == 1) h = minsoft;
    if (h > maxsoft) h = maxsoft;
    if (maxflag == 1) h = maxsoft;
out:
    *h_io = h;
    *Left_io = Left;
    *Right_io = Right;
    return result;
}

int orc_hsml_step(double num_ngb, double dhsml, double *h, double *left, double *right, int iter,
                  double des_num_ngb, double des_num_ngb_dev, double min_hsml, double max_hsml)
{
    return hsml_step(num_ngb, dhsml, h, left, right, iter, des_num_ngb, des_num_ngb_dev, min_hsml, max_hsml);
}

int orc_density(const orc_params *p, orc_state *s, const int *active, int n_active, orc_stats *st)
{
    const int n = s->n;
    const double *pos = s->pos, *vel = s->vel, *mass = s->mass;
    double *hsml = s->hsml;
    double *Left = (double *)calloc((size_t)n + 1, sizeof(double));
    double *Right = (double *)calloc((size_t)n + 1, sizeof(double));
    unsigned char *done = (unsigned char *)calloc((size_t)n + 1, 1);
    double maxsoft = p->max_hsml, minsoft = p->min_hsml;
    int status = 0;
    const int mfv = (p->method == 1) && (s->particle_vel != NULL);
    set_threads(p);

    /* density.c:470-490: active gas with mass>0, clamp the initial hsml */
    for (int a = 0; a < n_active; a++) {
        int i = active[a];
        if (!(mass[i] > 0)) { done[i] = 1; continue; }
        Left[i] = Right[i] = 0;
        if ((hsml[i] < 0) || !isfinite(hsml[i]) || (hsml[i] > 0.99 * maxsoft)) hsml[i] = 0.99 * maxsoft;
    }

    grid_t g;
    grid_build(&g, p, n, pos, hsml);
    int iter = 0;
    long long dens_int = 0;
    long long ntot;
    do {
        long long npleft = 0, sweep_int = 0;
        int bad = 0;
#pragma omp parallel reduction(+ : npleft, sweep_int, bad)
        {
            nb_t b = {0};
#pragma omp for schedule(dynamic, 128)
            for (int a = 0; a < n_active; a++) {
                int i = active[a];
                if (done[i]) continue;
                /* ---- density_evaluate (density.c:252-347) ---- */
                double h = hsml[i], hinv, hinv3, hinv4;
                kernel_hinv(h, &hinv, &hinv3, &hinv4);
                nb_collect(&b, p, &g, n, pos, hsml, i, h, h, 0);
                double Ngb = 0, Rho = 0, DhsmlNgb = 0, DivVel = 0, NVT[3][3] = {{0}};
                double divvel_l1 = 0, f_l1[3] = {0, 0, 0};
                double PVel[3] = {0, 0, 0}, pvel_l1[3] = {0, 0, 0}; /* MFV: sum of m_j W_ij v_j */
                for (int q = 0; q < b.cnt; q++) {
                    int j = b.j[q];
                    if (mass[j] <= 0) continue;
                    const double *dp = b.dp + 3 * q;
                    double r = sqrt(b.r2[q]);
                    double u = r * hinv, wk, dwk;
                    kernel_main(u, hinv3, hinv4, &wk, &dwk, 0);
                    double mj_wk = mass[j] * wk;
                    sweep_int++;
                    Ngb += wk;
                    Rho += mj_wk;
                    DhsmlNgb += -(NUMDIMS * hinv * wk + u * dwk);
                    if (mfv && r == 0) { /* density.c:291-293 */
                        for (int kv = 0; kv < 3; kv++) {
                            PVel[kv] += mj_wk * vel[3 * j + kv];
                            pvel_l1[kv] += fabs(mj_wk * vel[3 * j + kv]);
                        }
                    }
                    if (r > 0) {
                        NVT[0][0] += wk * dp[0] * dp[0];
                        NVT[0][1] += wk * dp[0] * dp[1];
                        NVT[0][2] += wk * dp[0] * dp[2];
                        NVT[1][1] += wk * dp[1] * dp[1];
                        NVT[1][2] += wk * dp[1] * dp[2];
                        NVT[2][2] += wk * dp[2] * dp[2];
                        NVT[1][0] += wk * dp[0];
                        NVT[2][0] += wk * dp[1];
                        NVT[2][1] += wk * dp[2];
                        double dv0 = vel[3 * i + 0] - vel[3 * j + 0];
                        double dv1 = vel[3 * i + 1] - vel[3 * j + 1];
                        double dv2 = vel[3 * i + 2] - vel[3 * j + 2];
                        if (mfv) { /* density.c:331-334: v_i - (v_i - v_j) */
                            const double dvv[3] = {dv0, dv1, dv2};
                            for (int kv = 0; kv < 3; kv++) {
                                PVel[kv] += mj_wk * (vel[3 * i + kv] - dvv[kv]);
                                pvel_l1[kv] += fabs(mj_wk * (vel[3 * i + kv] - dvv[kv]));
                            }
                        }
                        DivVel -= dwk * (dp[0] * dv0 + dp[1] * dv1 + dp[2] * dv2) / r;
                        divvel_l1 += fabs(dwk * (dp[0] * dv0 + dp[1] * dv1 + dp[2] * dv2) / r);
                        f_l1[0] += fabs(wk * dp[0]);
                        f_l1[1] += fabs(wk * dp[1]);
                        f_l1[2] += fabs(wk * dp[2]);
                    }
                }
                /* ---- iteration body (density.c:505-850) ---- */
                double desnumngb = p->des_num_ngb, desnumngbdev = p->max_num_ngb_dev;
                if (p->first_step) { if (p->max_num_ngb_dev > 0.05) desnumngbdev = 0.05; }
                double desnumngbdev_0 = desnumngbdev, ConditionNumber = 0;
                double NumNgb = Ngb, Dhsml = DhsmlNgb, Tinv[3][3];
                if (NumNgb > 0) {
                    divvel_l1 /= NumNgb;
                    Dhsml *= h / (NUMDIMS * NumNgb);
                    DivVel /= NumNgb;
                    NumNgb *= NORM_COEFF * pow(h, NUMDIMS);
                } else {
                    NumNgb = Dhsml = DivVel = 0;
                }
                if (Dhsml > -0.9) Dhsml = 1 / (1 + Dhsml); else Dhsml = 1;
                DivVel *= Dhsml;
                divvel_l1 *= fabs(Dhsml);
                double FCE, FCE_l1 = 0;
                {
                    double V_i = NORM_COEFF * pow(h, NUMDIMS) / NumNgb;
                    double dx_i = sqrt(V_i * (NVT[0][0] + NVT[1][1] + NVT[2][2]));
                    double fin[3] = {NVT[1][0], NVT[2][0], NVT[2][1]}, fout[3] = {0, 0, 0};
                    NVT[1][0] = NVT[0][1]; NVT[2][0] = NVT[0][2]; NVT[2][1] = NVT[1][2];
                    double norm = pow(h, 2 - NUMDIMS);
                    for (int k1 = 0; k1 < 3; k1++) for (int k2 = 0; k2 < 3; k2++) NVT[k1][k2] /= norm;
                    double cn_thresh = 10. * CONDITION_NUMBER_DANGER;
                    double trace = NVT[0][0] + NVT[1][1] + NVT[2][2];
                    double add = 1.05 * (trace / NUMDIMS) / cn_thresh;
                    while (1) {
                        ConditionNumber = matrix_invert_ndims(NVT, Tinv);
                        if (ConditionNumber < cn_thresh) break;
                        for (int k1 = 0; k1 < NUMDIMS; k1++) NVT[k1][k1] += add;
                        add *= 1.2;
                    }
                    for (int k1 = 0; k1 < 3; k1++) for (int k2 = 0; k2 < 3; k2++) NVT[k1][k2] = Tinv[k1][k2] / norm;
                    for (int k1 = 0; k1 < 3; k1++) for (int k2 = 0; k2 < 3; k2++)
                        fout[k1] += 2. * V_i * NVT[k1][k2] * fin[k2];
                    double leak = 0;
                    for (int k1 = 0; k1 < 3; k1++) leak += fabs(fout[k1]) / NUMDIMS;
                    FCE = leak / (2. * NUMDIMS * pow(dx_i, NUMDIMS - 1));
                    double leak_l1 = 0;
                    for (int k1 = 0; k1 < 3; k1++) {
                        double o = 0;
                        for (int k2 = 0; k2 < 3; k2++) o += fabs(2. * V_i * NVT[k1][k2]) * f_l1[k2];
                        leak_l1 += o / NUMDIMS;
                    }
                    FCE_l1 = leak_l1 / (2. * NUMDIMS * pow(dx_i, NUMDIMS - 1));
                }
                /* neighbour-number target (density.c:565-595) */
                double cn_stored = s->cn[i];
                if ((iter == 0) && (ConditionNumber > cn_stored) && (cn_stored > 0)) {
                    double nc = ncorr_from_cn(cn_stored);
                    double dn_ngb = fabs(NumNgb - p->des_num_ngb * nc) / (desnumngbdev_0 * nc);
                    nc = ncorr_from_cn(ConditionNumber);
                    double dn_ngb_alt = fabs(NumNgb - p->des_num_ngb * nc) / (desnumngbdev_0 * nc);
                    dn_ngb = DMIN(dn_ngb, dn_ngb_alt);
                    if (dn_ngb < 10.0) cn_stored = ConditionNumber;
                }
                double ncorr_ngb = ncorr_from_cn(cn_stored);
                {
                    double d00 = 0.35;
                    if (FCE > d00) ncorr_ngb = DMAX(ncorr_ngb, DMIN(FCE / d00, 2.));
                }
                desnumngb = p->des_num_ngb * ncorr_ngb;
                desnumngbdev = desnumngbdev_0 * ncorr_ngb;

                {
                    const int r = hsml_step(NumNgb, Dhsml, &h, &Left[i], &Right[i], iter, desnumngb, desnumngbdev,
                                            minsoft, maxsoft);
                    if (r == HSML_DONE || r == HSML_BRACKETED) {
                        done[i] = 1;
                        cn_stored = ConditionNumber; /* density.c:705-719, 737-738 */
                    } else if (r == HSML_COLLAPSED) {
                        npleft++;
                        bad++;
                    } else {
                        npleft++;
                    }
                }
                /* out2particle assignments of this sweep (density.c:177-244) plus the
                 * in-place updates of the iteration body */
                hsml[i] = h;
                s->num_ngb[i] = NumNgb;
                s->dhsml[i] = Dhsml;
                s->divvel[i] = DivVel;
                s->density[i] = Rho;
                memcpy(s->nvt + 9 * (size_t)i, NVT, 9 * sizeof(double));
                s->fce[i] = FCE;
                s->cn[i] = cn_stored;
                if (s->divvel_l1) s->divvel_l1[i] = divvel_l1;
                if (s->fce_l1) s->fce_l1[i] = FCE_l1;
                if (mfv) {
                    for (int kv = 0; kv < 3; kv++) s->particle_vel[3 * (size_t)i + kv] = PVel[kv];
                    if (s->pvel_l1) for (int kv = 0; kv < 3; kv++) s->pvel_l1[3 * (size_t)i + kv] = pvel_l1[kv];
                }
            }
            nb_free(&b);
        }
        dens_int += sweep_int;
        if (bad) { status = -1; break; }
        ntot = npleft;
        if (ntot > 0) {
            iter++;
            if (iter > p->maxiter) { status = 1155; break; }
        }
    } while (ntot > 0);
    grid_free(&g);

    /* finalize (density.c:880-1031) */
    for (int a = 0; a < n_active; a++) {
        int i = active[a];
        if (!(mass[i] > 0)) continue;
        if (mfv && (s->density[i] > 0)) {
            /* mesh motion 5 (density.c:886-894): 0.7 VelPred + 0.3 the kernel-weighted
             * mean velocity, the weights' sum being the density sum still held here */
            const double eps_pvel = 0.3;
            for (int k = 0; k < 3; k++) {
                double *w = s->particle_vel + 3 * (size_t)i + k;
                if (s->pvel_l1)
                    s->pvel_l1[3 * (size_t)i + k] = (1. - eps_pvel) * fabs(vel[3 * i + k]) +
                                                    eps_pvel * s->pvel_l1[3 * (size_t)i + k] / s->density[i];
                *w = (1. - eps_pvel) * vel[3 * i + k] + eps_pvel * *w / s->density[i];
            }
        }
        double h = hsml[i];
        if ((h > 0) && (s->num_ngb[i] > 0)) s->density[i] = mass[i] * s->num_ngb[i] / (NORM_COEFF * pow(h, NUMDIMS));
        else if (h <= 0) s->density[i] = 0;
        else s->density[i] = mass[i] / (NORM_COEFF * pow(h, NUMDIMS));
        s->pressure[i] = (p->gamma - 1) * s->u[i] * s->density[i]; /* eos.c:58-61 */
        if (s->num_ngb[i] > 0) s->num_ngb[i] = pow(s->num_ngb[i], 1. / NUMDIMS); else s->num_ngb[i] = 0;
    }
    if (st) {
        st->iterations = iter + 1;
        st->dens_interactions = dens_int;
        st->status = status;
    }
    free(Left);
    free(Right);
    free(done);
    return status;
}

/* ------------------------------------------------------------------------- */
/* hydro_gradient_calc(): gradients.c:537-572 (limiter, construction),       */
/* 1052-1253 (finalize), 1477-2089 (pair loop, MINMAX else-if semantics)      */
/* ------------------------------------------------------------------------- */
#define MINMAX_CHECK(x, xmin, xmax) ((x < xmin) ? (xmin = x) : ((x > xmax) ? (xmax = x) : (1)))

static void local_slopelimiter(double *grad, double valmax, double valmin, double alim, double h,
                               double shoot_tol, int pos_preserve, double d_max, double val_cen)
{
    int k;
    double d_abs = 0.0;
    for (k = 0; k < 3; k++) d_abs += grad[k] * grad[k];
    if (d_abs > 0) {
        d_abs = sqrt(d_abs);
        double cfac = 1 / (alim * h * d_abs);
        double fabs_max = fabs(valmax), fabs_min = fabs(valmin), abs_max = fabs_max, abs_min = fabs_min, f_corr_overshoot;
        if (abs_max < abs_min) { abs_max = fabs_min; abs_min = fabs_max; }
        f_corr_overshoot = DMIN(abs_min + shoot_tol * abs_max, abs_max);
        cfac *= f_corr_overshoot;
        if (pos_preserve == 1) {
            double fmin = DMIN(val_cen, DMAX(0, DMAX(MIN_REAL_NUMBER * val_cen, DMIN(0.5 * (val_cen + valmin), val_cen - f_corr_overshoot))));
            cfac = DMIN((((val_cen - fmin) / d_max) / d_abs), cfac);
        }
        if (cfac < 1) { for (k = 0; k < 3; k++) grad[k] *= cfac; }
    }
}

void orc_slope_limiter(double grad[3], double valmax, double valmin, double alim, double h,
                       double shoot_tol, int pos_preserve, double d_max, double val_cen)
{
    local_slopelimiter(grad, valmax, valmin, alim, h, shoot_tol, pos_preserve, d_max, val_cen);
}

static inline void construct_gradient(double *grad, const double *nvt)
{
    double v_tmp[3];
    for (int k = 0; k < 3; k++) v_tmp[k] = grad[k];
    for (int k = 0; k < 3; k++) grad[k] = nvt[3 * k + 0] * v_tmp[0] + nvt[3 * k + 1] * v_tmp[1] + nvt[3 * k + 2] * v_tmp[2];
}

static inline int gasgrad_isactive(const orc_state *s, int i)
{
    /* gradients.c:66-75 */
    if (s->mass[i] <= 0) return 0;
    if (s->density[i] <= 0 || s->hsml[i] <= 0) return 0;
    return 1;
}

int orc_gradients(const orc_params *p, orc_state *s, const int *active, int n_active, orc_stats *st)
{
    const int n = s->n;
    grid_t g;
    grid_build(&g, p, n, s->pos, s->hsml);
    double hmax = max_hsml(n, s->hsml);
    long long nint = 0;
    set_threads(p);
#pragma omp parallel reduction(+ : nint)
    {
        nb_t b = {0};
#pragma omp for schedule(dynamic, 128)
        for (int a = 0; a < n_active; a++) {
            int i = active[a];
            double gr[3] = {0, 0, 0}, gp[3] = {0, 0, 0}, gv[3][3] = {{0}};
            double maxd = 0, mx_rho = 0, mn_rho = 0, mx_p = 0, mn_p = 0, mx_v[3] = {0, 0, 0}, mn_v[3] = {0, 0, 0};
            double l1[15] = {0}, in1[15] = {0};
            if (gasgrad_isactive(s, i)) {
                double rho_i = s->density[i], p_i = s->pressure[i];
                const double *v_i = s->vel + 3 * i;
                double h_i = s->hsml[i], hinv, hinv3, hinv4;
                kernel_hinv(h_i, &hinv, &hinv3, &hinv4);
                nb_collect(&b, p, &g, n, s->pos, s->hsml, i, h_i, DMAX(h_i, hmax), 1);
                for (int q = 0; q < b.cnt; q++) {
                    int j = b.j[q];
                    if (!gasgrad_isactive(s, j)) continue;
                    nint++;
                    const double *dp = b.dp + 3 * q;
                    double r = sqrt(b.r2[q]), wk_i, dwk_i;
                    if (r < h_i) { double u = r * hinv; kernel_main(u, hinv3, hinv4, &wk_i, &dwk_i, -1); }
                    else wk_i = 0;
                    if (r > maxd) maxd = r;
                    double d_rho = s->density[j] - rho_i;
                    MINMAX_CHECK(d_rho, mn_rho, mx_rho);
                    double dpp = s->pressure[j] - p_i;
                    MINMAX_CHECK(dpp, mn_p, mx_p);
                    double dv[3];
                    for (int k = 0; k < 3; k++) dv[k] = s->vel[3 * j + k] - v_i[k];
                    for (int k = 0; k < 3; k++) MINMAX_CHECK(dv[k], mn_v[k], mx_v[k]);
                    if (r < h_i) {
                        for (int k = 0; k < 3; k++) {
                            double wk_xyz_i = -wk_i * dp[k];
                            gr[k] += wk_xyz_i * d_rho;
                            gp[k] += wk_xyz_i * dpp;
                            for (int k2 = 0; k2 < 3; k2++) gv[k2][k] += wk_xyz_i * dv[k2];
                            l1[k] += fabs(wk_xyz_i * d_rho);
                            l1[3 + k] += fabs(wk_xyz_i * dpp);
                            for (int k2 = 0; k2 < 3; k2++) l1[6 + 3 * k2 + k] += fabs(wk_xyz_i * dv[k2]);
                            in1[k] += fabs(wk_xyz_i) * (fabs(s->density[j]) + fabs(rho_i));
                            in1[3 + k] += fabs(wk_xyz_i) * (fabs(s->pressure[j]) + fabs(p_i));
                            for (int k2 = 0; k2 < 3; k2++)
                                in1[6 + 3 * k2 + k] += fabs(wk_xyz_i) * (fabs(s->vel[3 * j + k2]) + fabs(v_i[k2]));
                        }
                    }
                }
            }
            /* finalize (gradients.c:1052-1058, 1219-1253) */
            const double *nvt = s->nvt + 9 * (size_t)i;
            if (s->grad_l1) {
                double *o = s->grad_l1 + 15 * (size_t)i;
                for (int q = 0; q < 5; q++)
                    for (int k = 0; k < 3; k++)
                        o[3 * q + k] = fabs(nvt[3 * k + 0]) * l1[3 * q + 0] + fabs(nvt[3 * k + 1]) * l1[3 * q + 1] +
                                       fabs(nvt[3 * k + 2]) * l1[3 * q + 2];
            }
            if (s->grad_in) {
                double *o = s->grad_in + 15 * (size_t)i;
                for (int q = 0; q < 5; q++)
                    for (int k = 0; k < 3; k++)
                        o[3 * q + k] = fabs(nvt[3 * k + 0]) * in1[3 * q + 0] + fabs(nvt[3 * k + 1]) * in1[3 * q + 1] +
                                       fabs(nvt[3 * k + 2]) * in1[3 * q + 2];
            }
            construct_gradient(gr, nvt);
            construct_gradient(gp, nvt);
            for (int k = 0; k < 3; k++) construct_gradient(gv[k], nvt);
            double h = s->hsml[i];
            double stol = 0.1;                   /* SELFGRAVITY_OFF, no MAGNETIC/GALSF */
            double d_max = DMAX(h, maxd);
            double h_lim = h;
            double cn = s->cn[i];
            double a_limiter = 0.25;
            if (cn > 100) a_limiter = DMIN(0.5, 0.25 + 0.25 * (cn - 100) / 100);
            local_slopelimiter(gr, mx_rho, mn_rho, a_limiter, h_lim, 0, 1, d_max, s->density[i]);
            local_slopelimiter(gp, mx_p, mn_p, a_limiter, h_lim, stol, 1, d_max, s->pressure[i]);
            for (int k = 0; k < 3; k++) local_slopelimiter(gv[k], mx_v[k], mn_v[k], a_limiter, h_lim, stol, 0, 0, 0);
            memcpy(s->grad_rho + 3 * (size_t)i, gr, sizeof(gr));
            memcpy(s->grad_p + 3 * (size_t)i, gp, sizeof(gp));
            memcpy(s->grad_v + 9 * (size_t)i, gv, sizeof(gv));
        }
        nb_free(&b);
    }
    grid_free(&g);
    if (st) st->grad_interactions = nint;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Riemann problem (reimann.h:1-13, 129-248, 286-593, 685-1138), MFM, hydro   */
/* ------------------------------------------------------------------------- */
typedef struct gconst { double G0, G1, G2, G3, G4, G5, G6, G7, G8, G9; } gconst;

static gconst make_gconst(double gamma)
{
    gconst c;
    c.G0 = gamma;
    c.G1 = ((c.G0 - 1.0) / (2.0 * c.G0));
    c.G2 = ((c.G0 + 1.0) / (2.0 * c.G0));
    c.G3 = ((2.0 * c.G0 / (c.G0 - 1.0)));
    c.G4 = (2.0 / (c.G0 - 1.0));
    c.G5 = (2.0 / (c.G0 + 1.0));
    c.G6 = ((c.G0 - 1.0) / (c.G0 + 1.0));
    c.G7 = (0.5 * (c.G0 - 1.0));
    c.G8 = (1.0 / c.G0);
    c.G9 = (c.G0 - 1.0);
    return c;
}

typedef struct rstate { double rho, p, v[3], u, cs; } rstate;
typedef struct rout { double P_M, S_M, Fv[3], Fp; } rout;

static void reconstruct_face_states(double Q_i, const double Grad_Q_i[3], double Q_j, const double Grad_Q_j[3],
                                    const double distance_from_i[3], const double distance_from_j[3],
                                    double *Q_L, double *Q_R, int mode)
{
    if (mode == 0) { *Q_R = Q_i; *Q_L = Q_j; return; }
    if (Q_i == Q_j) { *Q_L = *Q_R = Q_i; return; }
    *Q_R = Q_i + Grad_Q_i[0] * distance_from_i[0] + Grad_Q_i[1] * distance_from_i[1] + Grad_Q_i[2] * distance_from_i[2];
    *Q_L = Q_j + Grad_Q_j[0] * distance_from_j[0] + Grad_Q_j[1] * distance_from_j[1] + Grad_Q_j[2] * distance_from_j[2];
    double Qmin, Qmax, Qmed, Qmax_eff, Qmin_eff, fac, Qmed_max, Qmed_min;
    double fac_minmax = 0.5, fac_meddev = 0.375;
    Qmed = 0.5 * (Q_i + Q_j);
    if (Q_i < Q_j) { Qmax = Q_j; Qmin = Q_i; } else { Qmax = Q_i; Qmin = Q_j; }
    fac = fac_minmax * (Qmax - Qmin);
    Qmax_eff = Qmax + fac;
    Qmin_eff = Qmin - fac;
    if (mode > 0) {
        if (Qmax < 0) { if (Qmax_eff > 0) Qmax_eff = Qmax * Qmax / (Qmax - (Qmax_eff - Qmax)); }
        if (Qmin > 0) { if (Qmin_eff < 0) Qmin_eff = Qmin * Qmin / (Qmin + (Qmin - Qmin_eff)); }
    }
    fac = fac_meddev * (Qmax - Qmin);
    Qmed_max = Qmed + fac;
    Qmed_min = Qmed - fac;
    if (Qmed_max > Qmax_eff) Qmed_max = Qmax_eff;
    if (Qmed_min < Qmin_eff) Qmed_min = Qmin_eff;
    if (Q_i < Q_j) {
        if (*Q_R < Qmin_eff) *Q_R = Qmin_eff;
        if (*Q_R > Qmed_max) *Q_R = Qmed_max;
        if (*Q_L > Qmax_eff) *Q_L = Qmax_eff;
        if (*Q_L < Qmed_min) *Q_L = Qmed_min;
    } else {
        if (*Q_R > Qmax_eff) *Q_R = Qmax_eff;
        if (*Q_R < Qmed_min) *Q_R = Qmed_min;
        if (*Q_L < Qmin_eff) *Q_L = Qmin_eff;
        if (*Q_L > Qmed_max) *Q_L = Qmed_max;
    }
}

void orc_reconstruct(double qi, const double gi[3], double qj, const double gj[3],
                     const double di[3], const double dj[3], double *ql, double *qr, int mode)
{
    reconstruct_face_states(qi, gi, qj, gj, di, dj, ql, qr, mode);
}

/* The three star-state estimates of get_wavespeeds_and_pressure_star
 * (reimann.h:527-593), each with the MIN_REAL_NUMBER floor the reference
 * applies after it.  S_L/S_R are the outer wave speeds the estimate used. */

/* Gaburov's HLLC estimate with the Davis wave speeds (reimann.h:541-547) */
static void pstar_hllc(const rstate *L, const rstate *R, rout *o, double v_line_L, double v_line_R,
                       double cs_L, double cs_R, double *S_L_out, double *S_R_out)
{
    double PT_L = L->p, PT_R = R->p;
    double S_L = DMIN(v_line_L, v_line_R) - DMAX(cs_L, cs_R);
    double S_R = DMAX(v_line_L, v_line_R) + DMAX(cs_L, cs_R);
    double rho_wt_L = L->rho * (S_L - v_line_L);
    double rho_wt_R = R->rho * (S_R - v_line_R);
    o->S_M = ((PT_R - PT_L) + rho_wt_L * v_line_L - rho_wt_R * v_line_R) / (rho_wt_L - rho_wt_R);
    o->P_M = (PT_L * rho_wt_R - PT_R * rho_wt_L + rho_wt_L * rho_wt_R * (v_line_R - v_line_L)) / (rho_wt_R - rho_wt_L);
    if (o->P_M <= MIN_REAL_NUMBER) { o->P_M = MIN_REAL_NUMBER; o->S_M = S_L = S_R = 0; }
    *S_L_out = S_L; *S_R_out = S_R;
}

/* Roe-averaged wave speeds (reimann.h:551-576) */
static void pstar_roe(const rstate *L, const rstate *R, rout *o, const gconst *c, const double n_unit[3],
                      double v_line_L, double v_line_R, double cs_L, double cs_R, double h_L, double h_R,
                      double *S_L_out, double *S_R_out)
{
    double PT_L = L->p, PT_R = R->p;
    double sqrt_rho_L = sqrt(L->rho);
    double sqrt_rho_R = sqrt(R->rho);
    double sqrt_rho_inv = 1 / (sqrt_rho_L + sqrt_rho_R);
    double vx_roe = (sqrt_rho_L * L->v[0] + sqrt_rho_R * R->v[0]) * sqrt_rho_inv;
    double vy_roe = (sqrt_rho_L * L->v[1] + sqrt_rho_R * R->v[1]) * sqrt_rho_inv;
    double vz_roe = (sqrt_rho_L * L->v[2] + sqrt_rho_R * R->v[2]) * sqrt_rho_inv;
    double v_line_roe = vx_roe * n_unit[0] + vy_roe * n_unit[1] + vz_roe * n_unit[2];
    double h_roe = (sqrt_rho_L * h_L + sqrt_rho_R * h_R) * sqrt_rho_inv;
    double cs_roe = sqrt(DMAX(1.e-30, c->G9 * (h_roe - 0.5 * (vx_roe * vx_roe + vy_roe * vy_roe + vz_roe * vz_roe))));
    double S_R = DMAX(v_line_R + cs_R, v_line_roe + cs_roe);
    double S_L = DMIN(v_line_L - cs_L, v_line_roe - cs_roe);
    double rho_wt_R = R->rho * (S_R - v_line_R);
    double rho_wt_L = -L->rho * (S_L - v_line_L);
    o->S_M = ((rho_wt_R * v_line_R + rho_wt_L * v_line_L) + (PT_L - PT_R)) / (rho_wt_R + rho_wt_L);
    o->P_M = L->rho * (v_line_L - S_L) * (v_line_L - o->S_M) + PT_L;
    if (o->P_M <= MIN_REAL_NUMBER) { o->P_M = MIN_REAL_NUMBER; o->S_M = S_L = S_R = 0; }
    *S_L_out = S_L; *S_R_out = S_R;
}

/* primitive-variable (PVRS) estimate, contact speed clamped to the fastest
 * signal (reimann.h:582-587) */
static void pstar_pvrs(const rstate *L, const rstate *R, rout *o, double v_line_L, double v_line_R,
                       double cs_L, double cs_R, double *S_L_out, double *S_R_out)
{
    double PT_L = L->p, PT_R = R->p;
    o->P_M = 0.5 * ((PT_L + PT_R) + (v_line_L - v_line_R) * 0.25 * (L->rho + R->rho) * (cs_L + cs_R));
    o->S_M = 0.5 * (v_line_R + v_line_L) + 2.0 * (PT_L - PT_R) / ((L->rho + R->rho) * (cs_L + cs_R));
    double S_plus = DMAX(DMAX(fabs(v_line_L - cs_L), fabs(v_line_R - cs_R)), DMAX(fabs(v_line_L + cs_L), fabs(v_line_R + cs_R)));
    double S_L = -S_plus, S_R = S_plus;
    if (o->S_M < S_L) o->S_M = S_L;
    if (o->S_M > S_R) o->S_M = S_R;
    if (o->P_M <= MIN_REAL_NUMBER) { o->P_M = MIN_REAL_NUMBER; o->S_M = S_L = S_R = 0; }
    *S_L_out = S_L; *S_R_out = S_R;
}

/* get_wavespeeds_and_pressure_star (reimann.h:527-593): vacuum test, then the
 * estimates in turn while P_M is <= 0, NaN or above the limiter.  Returns the
 * estimate that stands: 0 vacuum, 1 HLLC, 2 Roe, 3 PVRS. */
static int get_wavespeeds_and_pressure_star(const rstate *L, const rstate *R, rout *o, const gconst *c,
                                            const double n_unit[3],
                                            double v_line_L, double v_line_R, double cs_L, double cs_R,
                                            double h_L, double h_R, double press_tot_limiter,
                                            double *S_L, double *S_R)
{
    if ((v_line_R - v_line_L) > DMAX(cs_L, cs_R)) {
        o->P_M = MIN_REAL_NUMBER; o->S_M = *S_L = *S_R = 0;
        return 0;
    }
    pstar_hllc(L, R, o, v_line_L, v_line_R, cs_L, cs_R, S_L, S_R);
    if (!((o->P_M <= 0) || (isnan(o->P_M)) || (o->P_M > press_tot_limiter))) return 1;
    pstar_roe(L, R, o, c, n_unit, v_line_L, v_line_R, cs_L, cs_R, h_L, h_R, S_L, S_R);
    if (!((o->P_M <= 0) || (isnan(o->P_M)) || (o->P_M > press_tot_limiter))) return 2;
    pstar_pvrs(L, R, o, v_line_L, v_line_R, cs_L, cs_R, S_L, S_R);
    return 3;
}

static void riemann_kt(const rstate *L, const rstate *R, rout *o, const double n[3],
                       double v_line_L, double v_line_R, double cs_L, double cs_R, double h_L, double h_R)
{
    /* reimann.h:484-519, SLOPE_LIMITER_TOLERANCE=1 switch alpha */
    double S_L, S_R, S_M, alpha;
    int k;
    double delta_threshold = 0.001 * 0.5 * (L->rho + R->rho) * 0.5 * (cs_L + cs_R);
    double dv2 = 0, dv[3];
    for (k = 0; k < 3; k++) { dv[k] = R->rho * R->v[k] - L->rho * L->v[k]; dv2 += dv[k] * dv[k]; }
    double dvL = fabs(R->rho * v_line_R - L->rho * v_line_L);
    alpha = dvL / sqrt(delta_threshold * delta_threshold + dv2);
    S_L = alpha * cs_L; if (v_line_L > 0) S_L += v_line_L; else S_L -= v_line_L;
    S_R = alpha * cs_R; if (v_line_R > 0) S_R += v_line_R; else S_R -= v_line_R;
    S_M = DMAX(S_L, S_R);
    o->S_M = DMAX(cs_L + fabs(v_line_L), cs_R + fabs(v_line_R));
    double f_rho_left = L->rho * (v_line_L + S_M);
    double f_rho_right = R->rho * (v_line_R - S_M);
    double denom = 1 / (L->rho * v_line_L - R->rho * v_line_R + S_M * (L->rho + R->rho));
    o->P_M = (f_rho_left * R->p - f_rho_right * L->p) * denom;
    for (k = 0; k < 3; k++) o->Fv[k] = (R->v[k] - L->v[k]) * f_rho_left * f_rho_right * denom + (o->P_M) * n[k];
    o->Fp = (S_M * (f_rho_left * R->p + f_rho_right * L->p) + (h_R - h_L) * f_rho_left * f_rho_right) * denom;
}

static double guess_for_pressure(const rstate *L, const rstate *R, const gconst *c,
                                 double v_line_L, double v_line_R, double cs_L, double cs_R)
{
    double pmin, pmax;
    double pv = 0.5 * (L->p + R->p) - 0.125 * (v_line_R - v_line_L) * (L->p + R->p) * (cs_L + cs_R);
    pmin = DMIN(L->p, R->p);
    pmax = DMAX(L->p, R->p);
    if (pmin <= 0) return 0.5 * (pmin + pmax);
    double qrat = pmax / pmin;
    if (qrat <= 2.0 && (pmin <= pv && pv <= pmax)) return pv;
    if (pv < pmin) {
        double pnu = (cs_L + cs_R) - c->G7 * (v_line_R - v_line_L);
        double pde = cs_L / pow(L->p, c->G1) + cs_R / pow(R->p, c->G1);
        return pow(pnu / pde, c->G3);
    } else {
        double gel = sqrt((c->G5 / L->rho) / (c->G6 * L->p + pv));
        double ger = sqrt((c->G5 / R->rho) / (c->G6 * R->p + pv));
        double x = (gel * L->p + ger * R->p - (v_line_R - v_line_L)) / (gel + ger);
        if (x < pmin || x > pmax) x = pmin;
        return x;
    }
}

static int iterative_Riemann_solver(const rstate *L, const rstate *R, rout *o, const gconst *c,
                                    double v_line_L, double v_line_R, double cs_L, double cs_R)
{
    double Pg, Pg_prev, W_L, W_R, Z_L, Z_R, tol, pratio;
    int niter_Riemann = 0;
    double a0, a1, a2, dvel, check_vel;
    dvel = v_line_R - v_line_L;
    check_vel = c->G4 * (cs_R + cs_L) - dvel;
    if (check_vel < 0) return 0;
    tol = 100.0;
    Pg = guess_for_pressure(L, R, c, v_line_L, v_line_R, cs_L, cs_R);
    W_L = W_R = 0;
    while ((tol > TOL_ITER) && (niter_Riemann < NMAX_ITER)) {
        Pg_prev = Pg;
        if (Pg > L->p) {
            a0 = c->G5 / L->rho;
            a1 = c->G6 * L->p;
            a2 = sqrt(a0 / (Pg + a1));
            W_L = (Pg - L->p) * a2;
            Z_L = a2 * (1.0 - 0.5 * (Pg - L->p) / (a1 + Pg));
        } else {
            pratio = Pg / L->p;
            W_L = c->G4 * cs_L * (pow(pratio, c->G1) - 1);
            Z_L = 1 / (L->rho * cs_L) * pow(Pg / L->p, -c->G2);
        }
        if (Pg > R->p) {
            a0 = c->G5 / R->rho;
            a1 = c->G6 * R->p;
            a2 = sqrt(a0 / (Pg + a1));
            W_R = (Pg - R->p) * a2;
            Z_R = a2 * (1.0 - 0.5 * (Pg - R->p) / (a1 + Pg));
        } else {
            pratio = Pg / R->p;
            W_R = c->G4 * cs_R * (pow(pratio, c->G1) - 1);
            Z_R = 1 / (R->rho * cs_R) * pow(pratio, -c->G2);
        }
        if (niter_Riemann < NMAX_ITER / 2) Pg -= (W_L + W_R + dvel) / (Z_L + Z_R);
        else Pg -= 0.5 * (W_L + W_R + dvel) / (Z_L + Z_R);
        if (Pg < 0.1 * Pg_prev) Pg = 0.1 * Pg_prev;
        tol = 2.0 * fabs((Pg - Pg_prev) / (Pg + Pg_prev));
        niter_Riemann++;
    }
    if (niter_Riemann < NMAX_ITER) {
        o->P_M = Pg;
        o->S_M = 0.5 * (v_line_L + v_line_R) + 0.5 * (W_R - W_L);
        return 1;
    }
    return 0;
}

/* Riemann_solver_exact for MFM: fluxes are left untouched, only P_M/S_M
 * change (sample_reimann_standard returns at once, reimann.h:873-877). */
static int riemann_exact(const rstate *L, const rstate *R, rout *o, const gconst *c,
                         double v_line_L, double v_line_R, double cs_L, double cs_R)
{
    if ((L->p == 0 && R->p == 0) || (L->rho == 0 && R->rho == 0)) { o->P_M = o->S_M = 0; return 0; }
    if ((L->rho > 0) && (R->rho > 0)) {
        if (!iterative_Riemann_solver(L, R, o, c, v_line_L, v_line_R, cs_L, cs_R)) {
            /* sample_reimann_vaccum_internal at S=0 (reimann.h:818-846) */
            double S_L = v_line_L + c->G4 * cs_L;
            double S_R = v_line_R - c->G4 * cs_R;
            if (0.0 <= S_L) { o->P_M = 0; o->S_M = S_L; }
            else if (0.0 >= S_R) { o->P_M = 0; o->S_M = S_R; }
            else { o->P_M = 0; o->S_M = 0.0; }
        }
    } else {
        if ((L->rho < 0) || (R->rho < 0)) return 1234545;
        if (L->rho > 0) { o->P_M = 0; o->S_M = v_line_L + c->G4 * cs_L; }
        if (R->rho > 0) { o->P_M = 0; o->S_M = v_line_R - c->G4 * cs_R; }
    }
    return 0;
}

/* Riemann_solver (reimann.h:286-388) for MFM hydro.  L/R are by value as in
 * the reference; returns a stage mask (1 HLLC, 2 KT, 4 exact). */
static int riemann_solver(rstate L, rstate R, rout *o, const double n_unit[3], double press_tot_limiter,
                          const gconst *c, int *err)
{
    if ((L.p < 0 && R.p < 0) || (L.rho < 0) || (R.rho < 0)) { o->P_M = 0; return 0; }
    L.cs = sqrt(c->G0 * L.p / L.rho);
    R.cs = sqrt(c->G0 * R.p / R.rho);
    L.u = L.p / (c->G9 * L.rho);
    R.u = R.p / (c->G9 * R.rho);
    double cs_L = L.cs, cs_R = R.cs;
    double h_L = L.p / L.rho + L.u + 0.5 * (L.v[0] * L.v[0] + L.v[1] * L.v[1] + L.v[2] * L.v[2]);
    double h_R = R.p / R.rho + R.u + 0.5 * (R.v[0] * R.v[0] + R.v[1] * R.v[1] + R.v[2] * R.v[2]);
    double v_line_L = L.v[0] * n_unit[0] + L.v[1] * n_unit[1] + L.v[2] * n_unit[2];
    double v_line_R = R.v[0] * n_unit[0] + R.v[1] * n_unit[1] + R.v[2] * n_unit[2];
    int mask = 1;
    /* HLLC_Riemann_solver (reimann.h:400-425), MFM branch */
    double S_L, S_R;
    get_wavespeeds_and_pressure_star(&L, &R, o, c, n_unit, v_line_L, v_line_R, cs_L, cs_R, h_L, h_R, press_tot_limiter,
                                     &S_L, &S_R);
    for (int k = 0; k < 3; k++) o->Fv[k] = o->P_M * n_unit[k];
    o->Fp = o->P_M * o->S_M;
    if ((o->P_M < 0) || (isnan(o->P_M))) {
        riemann_kt(&L, &R, o, n_unit, v_line_L, v_line_R, cs_L, cs_R, h_L, h_R);
        mask |= 2;
    }
    if ((o->P_M < 0) || (isnan(o->P_M)) || (o->P_M > press_tot_limiter)) {
        int e = riemann_exact(&L, &R, o, c, v_line_L, v_line_R, cs_L, cs_R);
        if (e && err) *err = e;
        mask |= 4;
    }
    return mask;
}

/* Face-frame quantities that Riemann_solver derives from the two states
 * (reimann.h:327-345): sound speeds, enthalpies, normal velocities. */
typedef struct rprep { double cs_L, cs_R, h_L, h_R, v_line_L, v_line_R; } rprep;

static rprep riemann_prep(rstate *L, rstate *R, const double n_unit[3], const gconst *c)
{
    rprep q;
    L->cs = sqrt(c->G0 * L->p / L->rho);
    R->cs = sqrt(c->G0 * R->p / R->rho);
    L->u = L->p / (c->G9 * L->rho);
    R->u = R->p / (c->G9 * R->rho);
    q.cs_L = L->cs; q.cs_R = R->cs;
    q.h_L = L->p / L->rho + L->u + 0.5 * (L->v[0] * L->v[0] + L->v[1] * L->v[1] + L->v[2] * L->v[2]);
    q.h_R = R->p / R->rho + R->u + 0.5 * (R->v[0] * R->v[0] + R->v[1] * R->v[1] + R->v[2] * R->v[2]);
    q.v_line_L = L->v[0] * n_unit[0] + L->v[1] * n_unit[1] + L->v[2] * n_unit[2];
    q.v_line_R = R->v[0] * n_unit[0] + R->v[1] * n_unit[1] + R->v[2] * n_unit[2];
    return q;
}

static void load_states(const double Lin[5], const double Rin[5], rstate *L, rstate *R)
{
    memset(L, 0, sizeof(*L));
    memset(R, 0, sizeof(*R));
    L->rho = Lin[0]; L->p = Lin[1]; L->v[0] = Lin[2]; L->v[1] = Lin[3]; L->v[2] = Lin[4];
    R->rho = Rin[0]; R->p = Rin[1]; R->v[0] = Rin[2]; R->v[1] = Rin[3]; R->v[2] = Rin[4];
}

void orc_riemann_kt(double gamma, const double Lin[5], const double Rin[5], const double n[3], double out[6])
{
    gconst c = make_gconst(gamma);
    rstate L, R;
    load_states(Lin, Rin, &L, &R);
    rprep q = riemann_prep(&L, &R, n, &c);
    rout o;
    memset(&o, 0, sizeof(o));
    riemann_kt(&L, &R, &o, n, q.v_line_L, q.v_line_R, q.cs_L, q.cs_R, q.h_L, q.h_R);
    out[0] = o.P_M; out[1] = o.S_M; out[2] = o.Fv[0]; out[3] = o.Fv[1]; out[4] = o.Fv[2]; out[5] = o.Fp;
}

void orc_pstar_estimate(double gamma, const double Lin[5], const double Rin[5], const double n[3], int which,
                        double out[4])
{
    gconst c = make_gconst(gamma);
    rstate L, R;
    load_states(Lin, Rin, &L, &R);
    rprep q = riemann_prep(&L, &R, n, &c);
    rout o;
    memset(&o, 0, sizeof(o));
    double S_L = 0, S_R = 0;
    if (which == 0) pstar_hllc(&L, &R, &o, q.v_line_L, q.v_line_R, q.cs_L, q.cs_R, &S_L, &S_R);
    else if (which == 1)
        pstar_roe(&L, &R, &o, &c, n, q.v_line_L, q.v_line_R, q.cs_L, q.cs_R, q.h_L, q.h_R, &S_L, &S_R);
    else pstar_pvrs(&L, &R, &o, q.v_line_L, q.v_line_R, q.cs_L, q.cs_R, &S_L, &S_R);
    out[0] = o.P_M; out[1] = o.S_M; out[2] = S_L; out[3] = S_R;
}

int orc_wavespeeds(double gamma, const double Lin[5], const double Rin[5], const double n[3],
                   double press_tot_limiter, double out[4])
{
    gconst c = make_gconst(gamma);
    rstate L, R;
    load_states(Lin, Rin, &L, &R);
    rprep q = riemann_prep(&L, &R, n, &c);
    rout o;
    memset(&o, 0, sizeof(o));
    double S_L = 0, S_R = 0;
    int which = get_wavespeeds_and_pressure_star(&L, &R, &o, &c, n, q.v_line_L, q.v_line_R, q.cs_L, q.cs_R, q.h_L,
                                                 q.h_R, press_tot_limiter, &S_L, &S_R);
    out[0] = o.P_M; out[1] = o.S_M; out[2] = S_L; out[3] = S_R;
    return which;
}

int orc_riemann(double gamma, const double Lin[5], const double Rin[5], const double n[3],
                double press_tot_limiter, double out[6])
{
    gconst c = make_gconst(gamma);
    rstate L = {Lin[0], Lin[1], {Lin[2], Lin[3], Lin[4]}, 0, 0};
    rstate R = {Rin[0], Rin[1], {Rin[2], Rin[3], Rin[4]}, 0, 0};
    rout o;
    memset(&o, 0, sizeof(o));
    int err = 0;
    int mask = riemann_solver(L, R, &o, n, press_tot_limiter, &c, &err);
    out[0] = o.P_M; out[1] = o.S_M; out[2] = o.Fv[0]; out[3] = o.Fv[1]; out[4] = o.Fv[2]; out[5] = o.Fp;
    return err ? -err : mask;
}

/* ------------------------------------------------------------------------- */
/* Riemann problem for MFV hydro (HYDRO_MESHLESS_FINITE_VOLUME): the solver   */
/* returns the fluxes of mass, momentum and total energy through a face at    */
/* rest in the frame of the states, reimann.h:286-388 with the MFV branches   */
/* of 400-425, 484-519, 603-675, 685-724, 731-1030, 1146-1166                 */
/* ------------------------------------------------------------------------- */
typedef struct routv { double P_M, S_M, Frho, Fv[3], Fp; } routv;

/* HLLC_fluxes (reimann.h:603-675): the four-wave HLLC flux sampled at the
 * face (frame speed 0) from the star estimate's P_M, S_M and outer speeds */
static void hllc_fluxes_mfv(const rstate *L, const rstate *R, routv *o, const double n[3], double vL, double vR,
                            double hL, double hR, double S_L, double S_R)
{
    const double P_M = o->P_M, S_M = o->S_M, vf = 0;
    if ((P_M <= 0) || isnan(P_M)) return;
    if ((S_M == S_L) || (S_M == S_R) || (S_M == vf)) {
        o->Frho = 0;
        o->Fp = P_M * S_M;
        for (int k = 0; k < 3; k++) o->Fv[k] = P_M * n[k];
        return;
    }
    if (vf < S_L) { /* left state */
        o->Frho = L->rho * (vL - vf);
        o->Fp = L->rho * hL * (vL - vf) + L->p * vf;
        for (int k = 0; k < 3; k++) o->Fv[k] = o->Frho * L->v[k] + L->p * n[k];
    } else if ((S_L <= vf) && (vf <= S_M)) { /* left star state */
        double nfac = L->rho * (S_L - vL) / (S_L - S_M);
        if (nfac < 0) nfac = 0;
        o->Frho = L->rho * (vL - S_L) + nfac * (S_L - vf);
        double eK = L->rho * hL - L->p;
        o->Fp = (L->rho * hL * vL - eK * S_L) +
                (S_L - vf) * nfac * (eK / L->rho + (S_M - vL) * (S_M + L->p / (L->rho * (S_L - vL))));
        double pfac = nfac * (S_L - vf) * (S_M - vL) + L->p;
        for (int k = 0; k < 3; k++) o->Fv[k] = o->Frho * L->v[k] + pfac * n[k];
    } else if ((S_M <= vf) && (vf <= S_R)) { /* right star state */
        double nfac = R->rho * (S_R - vR) / (S_R - S_M);
        if (nfac < 0) nfac = 0;
        o->Frho = R->rho * (vR - S_R) + nfac * (S_R - vf);
        double eK = R->rho * hR - R->p;
        o->Fp = (R->rho * hR * vR - eK * S_R) +
                (S_R - vf) * nfac * (eK / R->rho + (S_M - vR) * (S_M + R->p / (R->rho * (S_R - vR))));
        double pfac = nfac * (S_R - vf) * (S_M - vR) + R->p;
        for (int k = 0; k < 3; k++) o->Fv[k] = o->Frho * R->v[k] + pfac * n[k];
    } else { /* right state */
        o->Frho = R->rho * (vR - vf);
        o->Fp = R->rho * hR * (vR - vf) + R->p * vf;
        for (int k = 0; k < 3; k++) o->Fv[k] = o->Frho * R->v[k] + R->p * n[k];
    }
}

/* Kurganov-Tadmor/PWK flux, MFV branch (reimann.h:484-502, 503-507) */
static void riemann_kt_mfv(const rstate *L, const rstate *R, routv *o, const double n[3], double vL, double vR,
                           double cL, double cR, double hL, double hR)
{
    double thr = 0.001 * 0.5 * (L->rho + R->rho) * 0.5 * (cL + cR);
    double dv2 = 0;
    for (int k = 0; k < 3; k++) {
        double d = R->rho * R->v[k] - L->rho * L->v[k];
        dv2 += d * d;
    }
    double alpha = fabs(R->rho * vR - L->rho * vL) / sqrt(thr * thr + dv2);
    double S_L = alpha * cL; if (vL > 0) S_L += vL; else S_L -= vL;
    double S_R = alpha * cR; if (vR > 0) S_R += vR; else S_R -= vR;
    double S_M = DMAX(S_L, S_R);
    o->S_M = DMAX(cL + fabs(vL), cR + fabs(vR));
    double frl = L->rho * (vL + S_M);
    double frr = R->rho * (vR - S_M);
    o->P_M = 0.5 * (R->p + L->p);
    o->Frho = 0.5 * (frl + frr);
    for (int k = 0; k < 3; k++) o->Fv[k] = 0.5 * (frl * L->v[k] + frr * R->v[k]) + (o->P_M) * n[k];
    o->Fp = 0.5 * (frl * hL + frr * hR + S_M * (R->p - L->p));
}

/* sample_reimann_vaccum_left (reimann.h:731-771): vacuum on the left, the
 * right state's rarefaction sampled at speed S; Frho/Fp/Fv hold the face
 * state (density, pressure, velocity) until face_to_flux_mfv */
static void sample_vacuum_left(double S, const rstate *R, routv *o, const gconst *c, const double n[3], double vR,
                               double cR)
{
    double S_R = vR - c->G4 * cR;
    if (S_R > S) {
        o->P_M = 0; o->S_M = S_R; o->Frho = 0;
    } else if (vR + cR > S) {
        double C_eff = c->G5 * (cR - c->G7 * (vR - S));
        o->P_M = R->p * pow(C_eff / cR, c->G3);
        o->S_M = c->G5 * (-cR + c->G7 * vR + S);
        o->Frho = R->rho * pow(C_eff / cR, c->G4);
    } else {
        o->P_M = R->p; o->S_M = vR; o->Frho = R->rho;
    }
    o->Fp = o->P_M;
    for (int k = 0; k < 3; k++) o->Fv[k] = R->v[k] + (o->S_M - vR) * n[k];
}

/* sample_reimann_vaccum_right (reimann.h:779-820), mirror of the above */
static void sample_vacuum_right(double S, const rstate *L, routv *o, const gconst *c, const double n[3], double vL,
                                double cL)
{
    double S_L = vL + c->G4 * cL;
    if (S_L < S) {
        o->P_M = 0; o->S_M = S_L; o->Frho = 0;
    } else if (vL - cL < S) {
        double C_eff = c->G5 * (cL + c->G7 * (vL - S));
        o->P_M = L->p * pow(C_eff / cL, c->G3);
        o->S_M = c->G5 * (cL + c->G7 * vL + S);
        o->Frho = L->rho * pow(C_eff / cL, c->G4);
    } else {
        o->P_M = L->p; o->S_M = vL; o->Frho = L->rho;
    }
    o->Fp = o->P_M;
    for (int k = 0; k < 3; k++) o->Fv[k] = L->v[k] + (o->S_M - vL) * n[k];
}

/* sample_reimann_vaccum_internal (reimann.h:828-858): vacuum created inside
 * the fan.  The velocity between the fans keeps the reference's (1 - n_k)
 * weighting of the tangential part. */
static void sample_vacuum_internal(double S, const rstate *L, const rstate *R, routv *o, const gconst *c,
                                   const double n[3], double vL, double vR, double cL, double cR)
{
    double S_L = vL + c->G4 * cL;
    double S_R = vR - c->G4 * cR;
    if (S <= S_L) sample_vacuum_right(S, L, o, c, n, vL, cL);
    else if (S >= S_R) sample_vacuum_left(S, R, o, c, n, vR, cR);
    else {
        o->P_M = 0; o->S_M = S; o->Frho = 0; o->Fp = o->P_M;
        for (int k = 0; k < 3; k++)
            o->Fv[k] = (L->v[k] + (R->v[k] - L->v[k]) * (S - S_L) / (S_R - S_L)) * (1 - n[k]) + S * n[k];
    }
}

/* sample_reimann_standard (reimann.h:866-1030): the exact solution's state
 * at speed S given the star pressure P_M and contact speed S_M */
static void sample_standard(double S, const rstate *L, const rstate *R, routv *o, const gconst *c, const double n[3],
                            double vL, double vR, double cL, double cR)
{
    if (S <= o->S_M) { /* left of the contact */
        if (o->P_M <= L->p) { /* left rarefaction */
            if (S <= vL - cL) {
                o->Fp = L->p; o->Frho = L->rho;
                for (int k = 0; k < 3; k++) o->Fv[k] = L->v[k];
                return;
            }
            double C_eff_L = cL * pow(o->P_M / L->p, c->G1);
            if (S > o->S_M - C_eff_L) { /* left star state */
                o->Frho = L->rho * pow(o->P_M / L->p, c->G8);
                o->Fp = o->P_M;
                for (int k = 0; k < 3; k++) o->Fv[k] = L->v[k] + (o->S_M - vL) * n[k];
                return;
            }
            double S_eff = c->G5 * (cL + c->G7 * vL + S); /* inside the fan */
            double C_eff = c->G5 * (cL + c->G7 * (vL - S));
            o->Frho = L->rho * pow(C_eff / cL, c->G4);
            o->Fp = L->p * pow(C_eff / cL, c->G3);
            for (int k = 0; k < 3; k++) o->Fv[k] = L->v[k] + (S_eff - vL) * n[k];
            return;
        }
        if (L->p > 0) { /* left shock */
            double pml = o->P_M / L->p;
            double S_L = vL - cL * sqrt(c->G2 * pml + c->G1);
            if (S <= S_L) {
                o->Fp = L->p; o->Frho = L->rho;
                for (int k = 0; k < 3; k++) o->Fv[k] = L->v[k];
                return;
            }
            o->Frho = L->rho * (pml + c->G6) / (pml * c->G6 + 1.0);
        } else {
            o->Frho = L->rho / c->G6;
        }
        o->Fp = o->P_M;
        for (int k = 0; k < 3; k++) o->Fv[k] = L->v[k] + (o->S_M - vL) * n[k];
        return;
    }
    /* right of the contact */
    if (o->P_M > R->p) { /* right shock */
        if (R->p > 0) {
            double pmr = o->P_M / R->p;
            double S_R = vR + cR * sqrt(c->G2 * pmr + c->G1);
            if (S >= S_R) {
                o->Fp = R->p; o->Frho = R->rho;
                for (int k = 0; k < 3; k++) o->Fv[k] = R->v[k];
                return;
            }
            o->Frho = R->rho * (pmr + c->G6) / (pmr * c->G6 + 1.0);
        } else {
            o->Frho = R->rho / c->G6;
        }
        o->Fp = o->P_M;
        for (int k = 0; k < 3; k++) o->Fv[k] = R->v[k] + (o->S_M - vR) * n[k];
        return;
    }
    if (S >= vR + cR) { /* right rarefaction: right data state */
        o->Fp = R->p; o->Frho = R->rho;
        for (int k = 0; k < 3; k++) o->Fv[k] = R->v[k];
        return;
    }
    double C_eff_R = cR * pow(o->P_M / R->p, c->G1);
    if (S <= o->S_M + C_eff_R) { /* right star state */
        o->Frho = R->rho * pow(o->P_M / R->p, c->G8);
        o->Fp = o->P_M;
        for (int k = 0; k < 3; k++) o->Fv[k] = R->v[k] + (o->S_M - vR) * n[k];
        return;
    }
    double S_eff = c->G5 * (-cR + c->G7 * vR + S); /* inside the fan */
    double C_eff = c->G5 * (cR - c->G7 * (vR - S));
    o->Frho = R->rho * pow(C_eff / cR, c->G4);
    o->Fp = R->p * pow(C_eff / cR, c->G3);
    for (int k = 0; k < 3; k++) o->Fv[k] = R->v[k] + (S_eff - vR) * n[k];
}

/* convert_face_to_flux (reimann.h:1146-1166): face state (rho, P, v) to the
 * Euler fluxes through a face at rest */
static void face_to_flux_mfv(routv *o, const double n[3], const gconst *c)
{
    double rho = o->Frho, P = o->Fp, v[3], v_line = 0, h = 0;
    for (int k = 0; k < 3; k++) {
        v[k] = o->Fv[k];
        v_line += v[k] * n[k];
        h += v[k] * v[k];
    }
    v_line -= 0;
    h *= 0.5 * rho;
    h += (c->G0 / c->G9) * P;
    o->Fp = h * v_line;
    o->Frho = rho * v_line;
    for (int k = 0; k < 3; k++) o->Fv[k] = o->Frho * v[k] + P * n[k];
}

/* Riemann_solver_exact with the MFV sampling (reimann.h:685-724) */
static int riemann_exact_mfv(const rstate *L, const rstate *R, routv *o, const gconst *c, const double n[3],
                             double vL, double vR, double cL, double cR)
{
    if ((L->p == 0 && R->p == 0) || (L->rho == 0 && R->rho == 0)) {
        o->P_M = o->S_M = 0;
        o->Frho = o->Fp = o->Fv[0] = o->Fv[1] = o->Fv[2] = 0;
        return 0;
    }
    if ((L->rho > 0) && (R->rho > 0)) {
        rout q;
        q.P_M = o->P_M; q.S_M = o->S_M;
        if (iterative_Riemann_solver(L, R, &q, c, vL, vR, cL, cR)) {
            o->P_M = q.P_M; o->S_M = q.S_M;
            sample_standard(0.0, L, R, o, c, n, vL, vR, cL, cR);
        } else {
            sample_vacuum_internal(0.0, L, R, o, c, n, vL, vR, cL, cR);
        }
    } else {
        if ((L->rho < 0) || (R->rho < 0)) return 1234545;
        if (L->rho > 0) sample_vacuum_right(0.0, L, o, c, n, vL, cL);
        if (R->rho > 0) sample_vacuum_left(0.0, R, o, c, n, vR, cR);
    }
    face_to_flux_mfv(o, n, c);
    return 0;
}

/* Riemann_solver (reimann.h:286-388) for MFV hydro: HLLC with the HLLC
 * fluxes, the KT flux on a failed estimate, the exact solver past the
 * limiter.  Returns the stage mask (1 HLLC, 2 KT, 4 exact). */
static int riemann_solver_mfv(rstate L, rstate R, routv *o, const double n[3], double press_tot_limiter,
                              const gconst *c, int *err)
{
    if ((L.p < 0 && R.p < 0) || (L.rho < 0) || (R.rho < 0)) { o->P_M = 0; return 0; }
    rprep q = riemann_prep(&L, &R, n, c);
    int mask = 1;
    rout e;
    e.P_M = o->P_M; e.S_M = o->S_M;
    double S_L, S_R;
    get_wavespeeds_and_pressure_star(&L, &R, &e, c, n, q.v_line_L, q.v_line_R, q.cs_L, q.cs_R, q.h_L, q.h_R,
                                     press_tot_limiter, &S_L, &S_R);
    o->P_M = e.P_M; o->S_M = e.S_M;
    if ((o->P_M >= 0) && !isnan(o->P_M))
        hllc_fluxes_mfv(&L, &R, o, n, q.v_line_L, q.v_line_R, q.h_L, q.h_R, S_L, S_R);
    if ((o->P_M < 0) || isnan(o->P_M)) {
        riemann_kt_mfv(&L, &R, o, n, q.v_line_L, q.v_line_R, q.cs_L, q.cs_R, q.h_L, q.h_R);
        mask |= 2;
    }
    if ((o->P_M < 0) || isnan(o->P_M) || (o->P_M > press_tot_limiter)) {
        int ee = riemann_exact_mfv(&L, &R, o, c, n, q.v_line_L, q.v_line_R, q.cs_L, q.cs_R);
        if (ee && err) *err = ee;
        mask |= 4;
    }
    return mask;
}

/* element exports: out = P_M, S_M, F_rho, F_v[3], F_p */
int orc_riemann_mfv(double gamma, const double Lin[5], const double Rin[5], const double n[3],
                    double press_tot_limiter, double out[7])
{
    gconst c = make_gconst(gamma);
    rstate L, R;
    load_states(Lin, Rin, &L, &R);
    routv o;
    memset(&o, 0, sizeof(o));
    int err = 0;
    int mask = riemann_solver_mfv(L, R, &o, n, press_tot_limiter, &c, &err);
    out[0] = o.P_M; out[1] = o.S_M; out[2] = o.Frho;
    out[3] = o.Fv[0]; out[4] = o.Fv[1]; out[5] = o.Fv[2]; out[6] = o.Fp;
    return err ? -err : mask;
}

void orc_riemann_kt_mfv(double gamma, const double Lin[5], const double Rin[5], const double n[3], double out[7])
{
    gconst c = make_gconst(gamma);
    rstate L, R;
    load_states(Lin, Rin, &L, &R);
    rprep q = riemann_prep(&L, &R, n, &c);
    routv o;
    memset(&o, 0, sizeof(o));
    riemann_kt_mfv(&L, &R, &o, n, q.v_line_L, q.v_line_R, q.cs_L, q.cs_R, q.h_L, q.h_R);
    out[0] = o.P_M; out[1] = o.S_M; out[2] = o.Frho;
    out[3] = o.Fv[0]; out[4] = o.Fv[1]; out[5] = o.Fv[2]; out[6] = o.Fp;
}

/* The exact solver alone (reimann.h:685-724 with 1146-1166): the Godunov
 * flux of the exact solution sampled at the face */
int orc_riemann_exact_mfv(double gamma, const double Lin[5], const double Rin[5], const double n[3], double out[7])
{
    gconst c = make_gconst(gamma);
    rstate L, R;
    load_states(Lin, Rin, &L, &R);
    rprep q = riemann_prep(&L, &R, n, &c);
    routv o;
    memset(&o, 0, sizeof(o));
    int e = riemann_exact_mfv(&L, &R, &o, &c, n, q.v_line_L, q.v_line_R, q.cs_L, q.cs_R);
    out[0] = o.P_M; out[1] = o.S_M; out[2] = o.Frho;
    out[3] = o.Fv[0]; out[4] = o.Fv[1]; out[5] = o.Fv[2]; out[6] = o.Fp;
    return e;
}

/* ------------------------------------------------------------------------- */
/* hydro_force(): hydro_toplevel.c:961-1057 (preloop, finalize 652-767),      */
/* hydro_evaluate.h:22-516, hydro_core_meshless.h:8-385,                      */
/* compute_finitevol_faces.h:14-106                                           */
/* ------------------------------------------------------------------------- */
typedef struct force_i_t {
    int i;
    double m, rho, P, sound, h, hinv, hinv3, hinv4, V, cn, cnumcrit2, fce, dhsml;
    const double *v, *T, *gr, *gp, *gv;
} force_i_t;

static void force_i_setup(const orc_params *p, const orc_state *s, int i, force_i_t *fi)
{
    fi->i = i;
    fi->m = s->mass[i];
    fi->rho = s->density[i];
    fi->P = s->pressure[i];
    fi->sound = sqrt(p->gamma * fi->P / fi->rho); /* eos.c:245-254 */
    fi->h = s->hsml[i];
    kernel_hinv(fi->h, &fi->hinv, &fi->hinv3, &fi->hinv4);
    fi->V = fi->m / fi->rho;
    fi->cn = s->cn[i];
    fi->cnumcrit2 = ((double)CONDITION_NUMBER_DANGER) * ((double)CONDITION_NUMBER_DANGER) - fi->cn * fi->cn;
    fi->fce = s->fce[i];
    fi->dhsml = s->dhsml[i];
    fi->v = s->vel + 3 * (size_t)i;
    fi->T = s->nvt + 9 * (size_t)i;
    fi->gr = s->grad_rho + 3 * (size_t)i;
    fi->gp = s->grad_p + 3 * (size_t)i;
    fi->gv = s->grad_v + 9 * (size_t)i;
}

/* one (i,j) interaction of hydro_force_evaluate: hydro_evaluate.h:170-298 and
 * the MFM core.  Writes Fv[3], Fp, vsig; returns status (0, 1234, 1234545).
 * `ro` is the per-i Riemann output that the reference keeps across pairs. */
static int near_tie(double a, double b)
{
    double m = fabs(a) > fabs(b) ? fabs(a) : fabs(b);
    return fabs(a - b) <= 4.0 * (nextafter(m, INFINITY) - m);
}

static int force_pair(const orc_params *p, const orc_state *s, const force_i_t *fi, const gconst *c,
                      int j, const double dp[3], double r2, rout *ro, double Fv_out[3], double *Fp_out,
                      double *vsig_out, int *mask_out, double *Fp_abs_out)
{
    double Fp_abs = 0; /* sum of |terms| that make up Fluxes_p (error scale only) */
    const double epsilon_entropic_eos_big = 0.5, epsilon_entropic_eos_small = 1.e-3; /* hydro_evaluate.h:45-54 */
    int status = 0;
    const double *v_i = fi->v;
    double rho_i = fi->rho;
    double h_j = s->hsml[j];
    double r = sqrt(r2);
    double rinv = 1 / r;
    const double *v_j = s->vel + 3 * (size_t)j;
    double dv[3] = {v_i[0] - v_j[0], v_i[1] - v_j[1], v_i[2] - v_j[2]};
    double rho_j = s->density[j], P_j_loc = s->pressure[j];
    double sound_i = fi->sound;
    double sound_j = sqrt(p->gamma * P_j_loc / rho_j);
    double vsig = sound_i + sound_j;
    double vdotr2 = dp[0] * dv[0] + dp[1] * dv[1] + dp[2] * dv[2];
    if (vdotr2 < 0) vsig -= 1.0 * vdotr2 * rinv; /* fac_mu = 1 */
    double wk_i, dwk_i, wk_j, dwk_j, u, hinv_j, hinv3_j, hinv4_j;
    if (r < fi->h) { u = r * fi->hinv; kernel_main(u, fi->hinv3, fi->hinv4, &wk_i, &dwk_i, 0); }
    else { dwk_i = 0; wk_i = 0; }
    if (r < h_j) {
        kernel_hinv(h_j, &hinv_j, &hinv3_j, &hinv4_j);
        u = r * hinv_j;
        kernel_main(u, hinv3_j, hinv4_j, &wk_j, &dwk_j, 0);
    } else { dwk_j = 0; wk_j = 0; }

    /* ---- hydro_core_meshless.h ---- */
    double Fluxes_p = 0, Fluxes_v[3] = {0, 0, 0};
    double leak_vs_tol = 0.5 * (fi->fce + s->fce[j]);
    double face_area_dot_vel = 0, face_vel_i = 0, face_vel_j = 0, Face_Area_Norm = 0;
    double Pressure_i = fi->P, Pressure_j = P_j_loc;
    double V_i = fi->V, V_j = s->mass[j] / rho_j;
    const double *T_i = fi->T, *T_j = s->nvt + 9 * (size_t)j;
    double cn_j = s->cn[j];
    int mask = 0;
    /* compute_finitevol_faces.h:14-106 */
    double Face_Area_Vec[3];
    double wt_i = V_i, wt_j = V_j;
    if ((fabs(V_i - V_j) / DMIN(V_i, V_j)) / NUMDIMS > 1.25) {
        wt_i = wt_j = V_i * V_j * (wk_i + wk_j) / (V_i * wk_i + V_j * wk_j);
    } else { wt_i = V_i; wt_j = V_j; }
    double facenormal_dot_dp = 0;
    for (int k = 0; k < 3; k++) {
        Face_Area_Vec[k] = wk_i * wt_i * (T_i[3 * k + 0] * dp[0] + T_i[3 * k + 1] * dp[1] + T_i[3 * k + 2] * dp[2])
                         + wk_j * wt_j * (T_j[3 * k + 0] * dp[0] + T_j[3 * k + 1] * dp[1] + T_j[3 * k + 2] * dp[2]);
        Face_Area_Norm += Face_Area_Vec[k] * Face_Area_Vec[k];
        facenormal_dot_dp += Face_Area_Vec[k] * dp[k];
    }
    if ((cn_j * cn_j > 1.0e12 + fi->cnumcrit2) || (facenormal_dot_dp < 0)) {
        Face_Area_Norm = -(wt_i * V_i * dwk_i + wt_j * V_j * dwk_j) / r;
        Face_Area_Vec[0] = Face_Area_Norm * dp[0];
        Face_Area_Vec[1] = Face_Area_Norm * dp[1];
        Face_Area_Vec[2] = Face_Area_Norm * dp[2];
        Face_Area_Norm = Face_Area_Norm * Face_Area_Norm * r2;
    }
    Face_Area_Norm = sqrt(Face_Area_Norm);
    if (Face_Area_Norm != 0) {
        double n_unit[3];
        for (int k = 0; k < 3; k++) n_unit[k] = Face_Area_Vec[k] / Face_Area_Norm;
        double s_i = 0.5 * r, s_j = -0.5 * r, s_star_ij = 0;
        s_i = s_star_ij - s_i;
        s_j = s_star_ij - s_j;
        double distance_from_i[3], distance_from_j[3], v_frame[3];
        distance_from_i[0] = dp[0] * rinv; distance_from_i[1] = dp[1] * rinv; distance_from_i[2] = dp[2] * rinv;
        for (int k = 0; k < 3; k++) { distance_from_j[k] = distance_from_i[k] * s_j; distance_from_i[k] *= s_i; }
        for (int k = 0; k < 3; k++) v_frame[k] = rinv * (-s_i * v_j[k] + s_j * v_i[k]);
        for (int k = 0; k < 3; k++) { face_vel_i += v_i[k] * n_unit[k]; face_vel_j += v_j[k] * n_unit[k]; }
        face_area_dot_vel = rinv * (-s_i * face_vel_j + s_j * face_vel_i);
        double v2_approach = 0;
        double vdotr2_phys = vdotr2;
        vdotr2_phys *= 1 / (r * 1.0);
        if (vdotr2_phys < 0) v2_approach = vdotr2_phys * vdotr2_phys;
        double vdotf2_phys = face_vel_i - face_vel_j;
        if (vdotf2_phys < 0) v2_approach = DMAX(v2_approach, vdotf2_phys * vdotf2_phys);
        int recon_mode = 1;
        if (leak_vs_tol > 1) recon_mode = 0;
        rstate RL, RR;
        memset(&RL, 0, sizeof(RL));
        memset(&RR, 0, sizeof(RR));
        reconstruct_face_states(rho_i, fi->gr, rho_j, s->grad_rho + 3 * (size_t)j,
                                distance_from_i, distance_from_j, &RL.rho, &RR.rho, recon_mode);
        reconstruct_face_states(Pressure_i, fi->gp, Pressure_j, s->grad_p + 3 * (size_t)j,
                                distance_from_i, distance_from_j, &RL.p, &RR.p, recon_mode);
        for (int k = 0; k < 3; k++) {
            reconstruct_face_states(v_i[k], fi->gv + 3 * k, v_j[k], s->grad_v + 9 * (size_t)j + 3 * k,
                                    distance_from_i, distance_from_j, &RL.v[k], &RR.v[k], recon_mode);
            RL.v[k] -= v_frame[k];
            RR.v[k] -= v_frame[k];
        }
        double press_i_tot = Pressure_i + rho_i * v2_approach;
        double press_j_tot = Pressure_j + rho_j * v2_approach;
        double press_tot_limiter = 1.1 * 1.0 * DMAX(press_i_tot, press_j_tot);
        if (recon_mode == 0)
            press_tot_limiter = DMAX(press_tot_limiter, DMAX(DMAX(Pressure_i, Pressure_j), 2. * DMAX(rho_i, rho_j) * v2_approach));
        int err = 0;
        mask = riemann_solver(RL, RR, ro, n_unit, press_tot_limiter, c, &err);
        if ((ro->P_M < 0) || (isnan(ro->P_M)) || (ro->P_M > 1.4 * press_tot_limiter)) {
            mask |= 8; /* first-order retry */
            RR.p = Pressure_i; RL.p = Pressure_j;
            RR.rho = rho_i; RL.rho = rho_j;
            for (int k = 0; k < 3; k++) { RR.v[k] = v_i[k] - v_frame[k]; RL.v[k] = v_j[k] - v_frame[k]; }
            mask |= riemann_solver(RL, RR, ro, n_unit, 1.4 * press_tot_limiter, c, &err);
            if ((ro->P_M < 0) || (isnan(ro->P_M))) {
                RR.p = Pressure_i; RL.p = Pressure_j;
                RR.rho = rho_i; RL.rho = rho_j;
                for (int k = 0; k < 3; k++) { RR.v[k] = 0; RL.v[k] = 0; }
                mask |= riemann_solver(RL, RR, ro, n_unit, 2.0 * press_tot_limiter, c, &err);
                if ((ro->P_M < 0) || (isnan(ro->P_M))) status = 1234; /* hydro_core_meshless.h:262 */
            }
        }
        if (err) status = err;
        if ((ro->P_M > 0) && (!isnan(ro->P_M))) {
            Fp_abs = fabs(ro->Fp);
            for (int k = 0; k < 3; k++) { ro->Fp += v_frame[k] * ro->Fv[k]; Fp_abs += fabs(v_frame[k] * ro->Fv[k]); }
            Fluxes_p = Face_Area_Norm * ro->Fp;
            Fp_abs *= Face_Area_Norm;
            for (int k = 0; k < 3; k++) Fluxes_v[k] = Face_Area_Norm * ro->Fv[k];
            /* MFM adiabatic face correction (hydro_core_meshless.h:335-374) */
            double SM_over_ceff = fabs(ro->S_M) / DMIN(sound_i, sound_j);
            if ((SM_over_ceff < epsilon_entropic_eos_big) || (leak_vs_tol > 1)) {
                int use_entropic_energy_equation = 1;
                double facenorm_pm = ro->P_M * Face_Area_Norm;
                double PdV_fac = ro->P_M * vdotr2_phys / 1.0;
                double PdV_i = dwk_i * V_i * V_i * fi->dhsml * PdV_fac;
                double PdV_j = dwk_j * V_j * V_j * s->dhsml[j] * PdV_fac;
                double du_old = facenorm_pm * (ro->S_M + face_area_dot_vel);
                double du_new = 0.5 * (PdV_i - PdV_j + facenorm_pm * (face_vel_i + face_vel_j));
                double cnum2 = cn_j * cn_j;
                if (SM_over_ceff > epsilon_entropic_eos_small && cnum2 < fi->cnumcrit2) {
                    if (near_tie(Pressure_i / rho_i, Pressure_j / rho_j)) mask |= 16;
                    if (Pressure_i / rho_i != Pressure_j / rho_j) {
                        if (Pressure_i / rho_i > Pressure_j / rho_j) {
                            double dtoj = -du_old + facenorm_pm * face_vel_j;
                            if (dtoj > 0) use_entropic_energy_equation = 0;
                            else { if (dtoj < 0) { if (dtoj > -du_new + facenorm_pm * face_vel_j) use_entropic_energy_equation = 0; } }
                        } else {
                            double dtoi = +du_old - facenorm_pm * face_vel_i;
                            if (dtoi > 0) use_entropic_energy_equation = 0;
                            else { if (dtoi < 0) { if (dtoi > +du_new - facenorm_pm * face_vel_i) use_entropic_energy_equation = 0; } }
                        }
                    }
                }
                if (cnum2 >= fi->cnumcrit2) use_entropic_energy_equation = 1;
                if (use_entropic_energy_equation) {
                    Fluxes_p += du_new - du_old;
                    Fp_abs += fabs(0.5 * PdV_i) + fabs(0.5 * PdV_j) + fabs(0.5 * facenorm_pm * (face_vel_i + face_vel_j)) +
                              fabs(du_old);
                }
            }
        } else {
            Fluxes_p = 0; Fluxes_v[0] = Fluxes_v[1] = Fluxes_v[2] = 0;
        }
    }
    Fv_out[0] = Fluxes_v[0]; Fv_out[1] = Fluxes_v[1]; Fv_out[2] = Fluxes_v[2];
    *Fp_out = Fluxes_p;
    *vsig_out = vsig;
    if (mask_out) *mask_out = mask;
    if (Fp_abs_out) *Fp_abs_out = Fp_abs;
    return status;
}

/* One (i,j) interaction of the MFV force loop: hydro_evaluate.h:170-298,
 * compute_finitevol_faces.h:14-106 and hydro_core_meshless.h:8-334 with the
 * HYDRO_MESHLESS_FINITE_VOLUME branches: the face moves with
 * rinv (-s_i v_j + s_j w_i), w_i = ParticleVel of i and v_j = VelPred of j
 * (hydro_evaluate.h:195-196 sets the neighbour's element velocity to its
 * VelPred), the signal speed takes 3 |v.r|/r, the limiter is doubled, the
 * fluxes carry mass and are boosted back from the face frame, and there is
 * no entropic face correction.  Writes Fv[3], Fp, Frho (times the face area),
 * vsig; Fv_abs / Fp_abs (may be NULL): the sums of |terms| behind them. */
static int force_pair_mfv(const orc_params *p, const orc_state *s, const force_i_t *fi, const gconst *c, int j,
                          const double dp[3], double r2, routv *ro, const double w_i[3], double Fv_out[3],
                          double *Fp_out, double *Frho_out, double *vsig_out, int *mask_out, double Fv_abs[3],
                          double *Fp_abs_out, double *Frho_abs_out)
{
    int status = 0, mask = 0;
    const double *v_i = fi->v;
    double rho_i = fi->rho;
    double h_j = s->hsml[j];
    double r = sqrt(r2);
    double rinv = 1 / r;
    const double *v_j = s->vel + 3 * (size_t)j;
    double dv[3] = {v_i[0] - v_j[0], v_i[1] - v_j[1], v_i[2] - v_j[2]};
    double rho_j = s->density[j], P_j = s->pressure[j];
    double sound_i = fi->sound;
    double sound_j = sqrt(p->gamma * P_j / rho_j);
    double vsig = sound_i + sound_j;
    double vdotr2 = dp[0] * dv[0] + dp[1] * dv[1] + dp[2] * dv[2];
    if (vdotr2 < 0) vsig -= 3 * 1.0 * vdotr2 * rinv; /* hydro_evaluate.h:240-242, fac_mu = 1 */
    double wk_i, dwk_i, wk_j, dwk_j, u, hinv_j, hinv3_j, hinv4_j;
    if (r < fi->h) { u = r * fi->hinv; kernel_main(u, fi->hinv3, fi->hinv4, &wk_i, &dwk_i, 0); }
    else { dwk_i = 0; wk_i = 0; }
    if (r < h_j) {
        kernel_hinv(h_j, &hinv_j, &hinv3_j, &hinv4_j);
        u = r * hinv_j;
        kernel_main(u, hinv3_j, hinv4_j, &wk_j, &dwk_j, 0);
    } else { dwk_j = 0; wk_j = 0; }

    double Fluxes_rho = 0, Fluxes_p = 0, Fluxes_v[3] = {0, 0, 0}, Fp_abs = 0, Fva[3] = {0, 0, 0}, Frho_abs = 0;
    double leak_vs_tol = 0.5 * (fi->fce + s->fce[j]);
    double face_vel_i = 0, face_vel_j = 0, Face_Area_Norm = 0;
    double V_i = fi->V, V_j = s->mass[j] / rho_j;
    const double *T_i = fi->T, *T_j = s->nvt + 9 * (size_t)j;
    double cn_j = s->cn[j];
    double Face_Area_Vec[3];
    double wt_i = V_i, wt_j = V_j;
    if ((fabs(V_i - V_j) / DMIN(V_i, V_j)) / NUMDIMS > 1.25) {
        wt_i = wt_j = V_i * V_j * (wk_i + wk_j) / (V_i * wk_i + V_j * wk_j);
    }
    double facenormal_dot_dp = 0;
    for (int k = 0; k < 3; k++) {
        Face_Area_Vec[k] = wk_i * wt_i * (T_i[3 * k + 0] * dp[0] + T_i[3 * k + 1] * dp[1] + T_i[3 * k + 2] * dp[2])
                         + wk_j * wt_j * (T_j[3 * k + 0] * dp[0] + T_j[3 * k + 1] * dp[1] + T_j[3 * k + 2] * dp[2]);
        Face_Area_Norm += Face_Area_Vec[k] * Face_Area_Vec[k];
        facenormal_dot_dp += Face_Area_Vec[k] * dp[k];
    }
    if ((cn_j * cn_j > 1.0e12 + fi->cnumcrit2) || (facenormal_dot_dp < 0)) {
        Face_Area_Norm = -(wt_i * V_i * dwk_i + wt_j * V_j * dwk_j) / r;
        Face_Area_Vec[0] = Face_Area_Norm * dp[0];
        Face_Area_Vec[1] = Face_Area_Norm * dp[1];
        Face_Area_Vec[2] = Face_Area_Norm * dp[2];
        Face_Area_Norm = Face_Area_Norm * Face_Area_Norm * r2;
    }
    Face_Area_Norm = sqrt(Face_Area_Norm);
    if (Face_Area_Norm != 0) {
        double n_unit[3];
        for (int k = 0; k < 3; k++) n_unit[k] = Face_Area_Vec[k] / Face_Area_Norm;
        double s_i = 0.5 * r, s_j = -0.5 * r, s_star_ij = 0;
        s_i = s_star_ij - s_i;
        s_j = s_star_ij - s_j;
        double distance_from_i[3], distance_from_j[3], v_frame[3];
        distance_from_i[0] = dp[0] * rinv; distance_from_i[1] = dp[1] * rinv; distance_from_i[2] = dp[2] * rinv;
        for (int k = 0; k < 3; k++) { distance_from_j[k] = distance_from_i[k] * s_j; distance_from_i[k] *= s_i; }
        for (int k = 0; k < 3; k++) v_frame[k] = rinv * (-s_i * v_j[k] + s_j * w_i[k]);
        for (int k = 0; k < 3; k++) { face_vel_i += v_i[k] * n_unit[k]; face_vel_j += v_j[k] * n_unit[k]; }
        double v2_approach = 0;
        double vdotr2_phys = vdotr2;
        vdotr2_phys *= 1 / (r * 1.0);
        if (vdotr2_phys < 0) v2_approach = vdotr2_phys * vdotr2_phys;
        double vdotf2_phys = face_vel_i - face_vel_j;
        if (vdotf2_phys < 0) v2_approach = DMAX(v2_approach, vdotf2_phys * vdotf2_phys);
        int recon_mode = 1;
        if (leak_vs_tol > 1) recon_mode = 0;
        rstate RL, RR;
        memset(&RL, 0, sizeof(RL));
        memset(&RR, 0, sizeof(RR));
        reconstruct_face_states(rho_i, fi->gr, rho_j, s->grad_rho + 3 * (size_t)j,
                                distance_from_i, distance_from_j, &RL.rho, &RR.rho, recon_mode);
        reconstruct_face_states(fi->P, fi->gp, P_j, s->grad_p + 3 * (size_t)j,
                                distance_from_i, distance_from_j, &RL.p, &RR.p, recon_mode);
        for (int k = 0; k < 3; k++) {
            reconstruct_face_states(v_i[k], fi->gv + 3 * k, v_j[k], s->grad_v + 9 * (size_t)j + 3 * k,
                                    distance_from_i, distance_from_j, &RL.v[k], &RR.v[k], recon_mode);
            RL.v[k] -= v_frame[k];
            RR.v[k] -= v_frame[k];
        }
        double press_i_tot = fi->P + rho_i * v2_approach;
        double press_j_tot = P_j + rho_j * v2_approach;
        double press_tot_limiter = 1.1 * 1.0 * DMAX(press_i_tot, press_j_tot);
        press_tot_limiter *= 2.0; /* hydro_core_meshless.h:194-196 */
        if (recon_mode == 0)
            press_tot_limiter = DMAX(press_tot_limiter, DMAX(DMAX(fi->P, P_j), 2. * DMAX(rho_i, rho_j) * v2_approach));
        int err = 0;
        mask = riemann_solver_mfv(RL, RR, ro, n_unit, press_tot_limiter, c, &err);
        if ((ro->P_M < 0) || (isnan(ro->P_M)) || (ro->P_M > 1.4 * press_tot_limiter)) {
            mask |= 8;
            RR.p = fi->P; RL.p = P_j;
            RR.rho = rho_i; RL.rho = rho_j;
            for (int k = 0; k < 3; k++) { RR.v[k] = v_i[k] - v_frame[k]; RL.v[k] = v_j[k] - v_frame[k]; }
            mask |= riemann_solver_mfv(RL, RR, ro, n_unit, 1.4 * press_tot_limiter, c, &err);
            if ((ro->P_M < 0) || (isnan(ro->P_M))) {
                RR.p = fi->P; RL.p = P_j;
                RR.rho = rho_i; RL.rho = rho_j;
                for (int k = 0; k < 3; k++) { RR.v[k] = 0; RL.v[k] = 0; }
                mask |= riemann_solver_mfv(RL, RR, ro, n_unit, 2.0 * press_tot_limiter, c, &err);
                if ((ro->P_M < 0) || (isnan(ro->P_M))) status = 1234;
            }
        }
        if (err) status = err;
        if ((ro->P_M > 0) && (!isnan(ro->P_M))) {
            /* error scale of the mass flux: the terms rho (v - S) it is made of,
             * ~rho (|v| + c) in the face frame.  Where the flux cancels to
             * rounding (a uniform flow), |F_rho| itself is no scale at all. */
            double dvi = 0, dvj = 0;
            for (int k = 0; k < 3; k++) { dvi += fabs(v_i[k] - v_frame[k]); dvj += fabs(v_j[k] - v_frame[k]); }
            const double frho_abs = fabs(ro->Frho) + DMAX(rho_i, rho_j) * (dvi + dvj + sound_i + sound_j);
            Frho_abs = Face_Area_Norm * frho_abs;
            /* de-boost from the face frame (hydro_core_meshless.h:285-295) */
            Fp_abs = fabs(ro->Fp);
            for (int k = 0; k < 3; k++) {
                Fva[k] = fabs(ro->Fv[k]) + fabs(v_frame[k]) * frho_abs;
                Fp_abs += fabs(v_frame[k] * ro->Fv[k]) + (0.5 * v_frame[k] * v_frame[k]) * frho_abs;
                ro->Fp += v_frame[k] * ro->Fv[k];
                ro->Fp += (0.5 * v_frame[k] * v_frame[k]) * ro->Frho;
                ro->Fv[k] += v_frame[k] * ro->Frho;
            }
            Fluxes_rho = Face_Area_Norm * ro->Frho;
            Fluxes_p = Face_Area_Norm * ro->Fp;
            for (int k = 0; k < 3; k++) Fluxes_v[k] = Face_Area_Norm * ro->Fv[k];
            Fp_abs *= Face_Area_Norm;
            for (int k = 0; k < 3; k++) Fva[k] *= Face_Area_Norm;
        }
    }
    Fv_out[0] = Fluxes_v[0]; Fv_out[1] = Fluxes_v[1]; Fv_out[2] = Fluxes_v[2];
    *Fp_out = Fluxes_p;
    *Frho_out = Fluxes_rho;
    *vsig_out = vsig;
    if (mask_out) *mask_out = mask;
    if (Fv_abs) for (int k = 0; k < 3; k++) Fv_abs[k] = Fva[k];
    if (Fp_abs_out) *Fp_abs_out = Fp_abs;
    if (Frho_abs_out) *Frho_abs_out = Frho_abs;
    return status;
}

int orc_force_pair(const orc_params *p, const orc_state *s, int i, int j, double out[6])
{
    const gconst c = make_gconst(p->gamma);
    force_i_t fi;
    force_i_setup(p, s, i, &fi);
    double dp[3];
    for (int a = 0; a < 3; a++) dp[a] = s->pos[3 * i + a] - s->pos[3 * j + a];
    dp[1] = wrap1(dp[1], p->box[1]);
    dp[0] = wrap1(dp[0], p->box[0]);
    dp[2] = wrap1(dp[2], p->box[2]);
    double r2 = dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2];
    rout ro;
    memset(&ro, 0, sizeof(ro));
    int mask = 0;
    int stt = force_pair(p, s, &fi, &c, j, dp, r2, &ro, out, out + 3, out + 4, &mask, NULL);
    out[5] = mask;
    return stt;
}

/* One active particle's MFV force loop and final operations
 * (hydro_evaluate.h:146-498 with the mass exchange of 379-401,
 * hydro_toplevel.c:652-767 with the MFV terms of 661-665 and 742-752). */
static int mfv_particle(const orc_params *p, orc_state *s, const gconst *c, const grid_t *g, double hmax, nb_t *b,
                        int i, long long *nint, long long *nretry, long long *nexact)
{
    const int n = s->n;
    int status = 0;
    double Acc[3] = {0, 0, 0}, DtE = 0, DtMass = 0, dMass = 0, GravWork[3] = {0, 0, 0};
    double MaxSignalVel = MIN_REAL_NUMBER;
    double acc_l1[3] = {0, 0, 0}, dte_l1 = 0, dtmass_l1 = 0, dmass_l1 = 0, gw_l1[3] = {0, 0, 0};
    double m_i = s->mass[i], rho_i = s->density[i];
    const double *w_i = s->particle_vel + 3 * (size_t)i;
    const long long ti_i = s->ti_step ? s->ti_step[i] : 1;
    const double dt_i = (double)ti_i * p->timebase_interval; /* hydro_evaluate.h:82 */
    force_i_t fi;
    force_i_setup(p, s, i, &fi);
    if ((m_i > 0) && (rho_i > 0)) {
        double out_vsig = fi.sound;
        routv ro;
        memset(&ro, 0, sizeof(ro));
        nb_collect(b, p, g, n, s->pos, s->hsml, i, fi.h, DMAX(fi.h, hmax), 1);
        for (int q = 0; q < b->cnt; q++) {
            int j = b->j[q];
            if (s->mass[j] <= 0) continue;
            if (s->density[j] <= 0) continue;
            (*nint)++;
            const double *dp = b->dp + 3 * q;
            double Fv[3], Fp, Frho, vsig, Fva[3], Fpa, Frhoa;
            int mask = 0;
            int e = force_pair_mfv(p, s, &fi, c, j, dp, b->r2[q], &ro, w_i, Fv, &Fp, &Frho, &vsig, &mask, Fva, &Fpa,
                                   &Frhoa);
            if (e > status) status = e;
            if (mask & 8) (*nretry)++;
            if (mask & 4) (*nexact)++;
            /* mass exchange (hydro_evaluate.h:379-392), j_is_active_for_fluxes = 0 */
            const long long ti_j = s->ti_step ? s->ti_step[j] : 1;
            double dmass_holder = Frho * dt_i, dmass_limiter;
            if (dmass_holder > 0) dmass_limiter = s->mass[j]; else dmass_limiter = m_i;
            dmass_limiter *= 0.1;
            if (fabs(dmass_holder) > dmass_limiter) dmass_holder *= dmass_limiter / fabs(dmass_holder);
            if (ti_i < ti_j) {
                dMass += dmass_holder;
#pragma omp atomic
                s->d_mass[j] -= dmass_holder;
                dmass_l1 += Frhoa * dt_i;
                if (s->dmass_l1) {
#pragma omp atomic
                    s->dmass_l1[j] += Frhoa * dt_i;
                }
            }
            if (ti_i == ti_j) {
                dMass += 0.5 * dmass_holder;
#pragma omp atomic
                s->d_mass[j] -= 0.5 * dmass_holder;
                dmass_l1 += 0.5 * Frhoa * dt_i;
                if (s->dmass_l1) {
#pragma omp atomic
                    s->dmass_l1[j] += 0.5 * Frhoa * dt_i;
                }
            }
            DtMass += Frho;
            dtmass_l1 += Frhoa;
            for (int k = 0; k < 3; k++) {
                GravWork[k] += Frho * dp[k];
                gw_l1[k] += Frhoa * fabs(dp[k]);
            }
            for (int k = 0; k < 3; k++) Acc[k] += Fv[k];
            DtE += Fp;
            for (int k = 0; k < 3; k++) acc_l1[k] += Fva[k];
            dte_l1 += Fpa;
            if (vsig > out_vsig) out_vsig = vsig;
            if (s->wakeup && !s->is_active[j]) {
                if (vsig > WAKEUP * s->vsig[j]) {
#pragma omp atomic write
                    s->wakeup[j] = 1;
                }
            }
        }
        if (MaxSignalVel < out_vsig) MaxSignalVel = out_vsig;
    }
    if (m_i > 0) {
        /* hydro_final_operations_and_cleanup (hydro_toplevel.c:661-665, 738-764) */
        const double *v = fi.v;
        double v2_p = MaxSignalVel * MaxSignalVel;
        for (int k = 0; k < 3; k++) v2_p += (v[k] - w_i[k]) * (v[k] - w_i[k]);
        MaxSignalVel = sqrt(v2_p);
        for (int k = 0; k < 3; k++) {
            DtE -= (v[k] / 1.0) * Acc[k];
            DtE += 0.5 * (v[k] / 1.0) * (v[k] / 1.0) * DtMass;
            Acc[k] -= (v[k] / 1.0) * DtMass;
            Acc[k] /= m_i;
            dte_l1 += fabs(v[k]) * acc_l1[k] + 0.5 * v[k] * v[k] * dtmass_l1;
            acc_l1[k] += fabs(v[k]) * dtmass_l1;
            acc_l1[k] /= m_i;
        }
        DtE -= s->u[i] * DtMass;
        dte_l1 += fabs(s->u[i]) * dtmass_l1;
        DtE /= m_i;
        dte_l1 /= m_i;
        if (fi.h >= 0.99 * p->max_hsml) DtE = 0;
    }
    s->acc[3 * (size_t)i + 0] = Acc[0];
    s->acc[3 * (size_t)i + 1] = Acc[1];
    s->acc[3 * (size_t)i + 2] = Acc[2];
    s->dtu[i] = DtE;
    s->vsig[i] = MaxSignalVel;
    s->dt_mass[i] = DtMass;
#pragma omp atomic
    s->d_mass[i] += dMass;
    for (int k = 0; k < 3; k++) s->grav_work[3 * (size_t)i + k] = GravWork[k];
    if (s->acc_l1) for (int k = 0; k < 3; k++) s->acc_l1[3 * (size_t)i + k] = acc_l1[k];
    if (s->dtu_l1) s->dtu_l1[i] = dte_l1;
    if (s->dtmass_l1) s->dtmass_l1[i] = dtmass_l1;
    if (s->dmass_l1) {
#pragma omp atomic
        s->dmass_l1[i] += dmass_l1;
    }
    if (s->gravwork_l1) for (int k = 0; k < 3; k++) s->gravwork_l1[3 * (size_t)i + k] = gw_l1[k];
    if (s->tie_flag) s->tie_flag[i] = 0;
    return status;
}

int orc_hydro_force(const orc_params *p, orc_state *s, const int *active, int n_active, orc_stats *st)
{
    const int n = s->n;
    const gconst c = make_gconst(p->gamma);
    grid_t g;
    grid_build(&g, p, n, s->pos, s->hsml);
    double hmax = max_hsml(n, s->hsml);
    long long nint = 0, nretry = 0, nexact = 0;
    int status = 0;
    const int mfv = (p->method == 1);
    if (mfv) {
        /* hydro_force_initial_operations_preloop (hydro_toplevel.c:968-983):
         * DtMass, dMass, GravWorkTerm of the active particles start at zero;
         * dMass of every neighbour, active or not, then receives its share */
        for (int a = 0; a < n_active; a++) {
            int i = active[a];
            s->dt_mass[i] = 0;
            s->d_mass[i] = 0;
            for (int k = 0; k < 3; k++) s->grav_work[3 * (size_t)i + k] = 0;
            if (s->dmass_l1) s->dmass_l1[i] = 0;
        }
    }
    set_threads(p);
#pragma omp parallel reduction(+ : nint, nretry, nexact) reduction(max : status)
    {
        nb_t b = {0};
#pragma omp for schedule(dynamic, 64)
        for (int a = 0; a < n_active; a++) {
            int i = active[a];
            if (mfv) {
                int e = mfv_particle(p, s, &c, &g, hmax, &b, i, &nint, &nretry, &nexact);
                if (e > status) status = e;
                continue;
            }
            /* preloop (hydro_toplevel.c:968-976) */
            double Acc[3] = {0, 0, 0}, DtE = 0, MaxSignalVel = MIN_REAL_NUMBER;
            double acc_l1[3] = {0, 0, 0}, dte_l1 = 0;
            int tie = 0;
            double m_i = s->mass[i], rho_i = s->density[i];
            if ((m_i > 0) && (rho_i > 0)) {
                force_i_t fi;
                force_i_setup(p, s, i, &fi);
                double out_vsig = fi.sound;
                rout ro;
                memset(&ro, 0, sizeof(ro));
                nb_collect(&b, p, &g, n, s->pos, s->hsml, i, fi.h, DMAX(fi.h, hmax), 1);
                for (int q = 0; q < b.cnt; q++) {
                    int j = b.j[q];
                    if (s->mass[j] <= 0) continue;
                    if (s->density[j] <= 0) continue;
                    nint++;
                    double Fv[3], Fp, vsig, Fp_abs;
                    int mask = 0;
                    int e = force_pair(p, s, &fi, &c, j, b.dp + 3 * q, b.r2[q], &ro, Fv, &Fp, &vsig, &mask, &Fp_abs);
                    if (e > status) status = e;
                    if (mask & 8) nretry++;
                    if (mask & 16) tie = 1;
                    if (mask & 4) nexact++;
                    for (int k = 0; k < 3; k++) Acc[k] += Fv[k];
                    DtE += Fp;
                    for (int k = 0; k < 3; k++) acc_l1[k] += fabs(Fv[k]);
                    dte_l1 += Fp_abs;
                    if (vsig > out_vsig) out_vsig = vsig;
                    if (s->wakeup && !s->is_active[j]) {
                        if (vsig > WAKEUP * s->vsig[j]) {
#pragma omp atomic write
                            s->wakeup[j] = 1;
                        }
                    }
                }
                if (MaxSignalVel < out_vsig) MaxSignalVel = out_vsig;
                /* hydro_final_operations_and_cleanup (hydro_toplevel.c:738-764) */
                for (int k = 0; k < 3; k++) {
                    DtE -= (fi.v[k] / 1.0) * Acc[k];
                    Acc[k] /= m_i;
                    dte_l1 += fabs(fi.v[k]) * acc_l1[k];
                    acc_l1[k] /= m_i;
                }
                DtE /= m_i;
                dte_l1 /= m_i;
                if (fi.h >= 0.99 * p->max_hsml) DtE = 0;
            } else if (m_i > 0) {
                /* rho <= 0: the pair loop returns at once; the final operations still run */
                DtE /= m_i;
            }
            s->acc[3 * (size_t)i + 0] = Acc[0];
            s->acc[3 * (size_t)i + 1] = Acc[1];
            s->acc[3 * (size_t)i + 2] = Acc[2];
            s->dtu[i] = DtE;
            s->vsig[i] = MaxSignalVel;
            if (s->acc_l1) for (int k = 0; k < 3; k++) s->acc_l1[3 * (size_t)i + k] = acc_l1[k];
            if (s->dtu_l1) s->dtu_l1[i] = dte_l1;
            if (s->tie_flag) s->tie_flag[i] = (unsigned char)tie;
        }
        nb_free(&b);
    }
    grid_free(&g);
    if (st) {
        st->force_interactions = nint;
        st->riemann_retries = nretry;
        st->exact_solver_calls = nexact;
        st->status = status;
    }
    return status;
}
