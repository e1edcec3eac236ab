/*
 * gizmo_io.c -- HDF5 IC / snapshot layout of the gas of one rank over the
 * host mirror's P[] / SphP[].  See include/gizmo_io.h for the contract.
 * Built against the HDF5 C library (1.10 API).
 */
#include "gizmo_io.h"

#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hdf5.h>

#include "gizmo_mfm_host.h"

static char g_err[512];

const char *gizmo_io_last_error(void) { return g_err; }

static int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

static void quiet_hdf5(void) { H5Eset_auto2(H5E_DEFAULT, NULL, NULL); }

/* ------------------------------------------------------------------------- */
/* header attributes                                                          */
/* ------------------------------------------------------------------------- */
static int read_attr(hid_t grp, const char *name, hid_t memtype, void *buf, hssize_t want, int required)
{
    if (H5Aexists(grp, name) <= 0) return required ? -1 : 1;
    hid_t a = H5Aopen(grp, name, H5P_DEFAULT);
    if (a < 0) return -1;
    hid_t sp = H5Aget_space(a);
    hssize_t npts = H5Sget_simple_extent_npoints(sp);
    H5Sclose(sp);
    int rc = -1;
    if (npts == want) rc = H5Aread(a, memtype, buf) < 0 ? -1 : 0;
    H5Aclose(a);
    return rc;
}

static int write_attr(hid_t grp, const char *name, hid_t memtype, const void *buf, hsize_t count)
{
    hid_t sp = (count == 1) ? H5Screate(H5S_SCALAR) : H5Screate_simple(1, &count, NULL);
    if (sp < 0) return -1;
    hid_t a = H5Acreate2(grp, name, memtype, sp, H5P_DEFAULT, H5P_DEFAULT);
    int rc = (a < 0) ? -1 : (H5Awrite(a, memtype, buf) < 0 ? -1 : 0);
    if (a >= 0) H5Aclose(a);
    H5Sclose(sp);
    return rc;
}

/* read_header_attributes_in_hdf5 (read_ic.c:1379-1411) plus the attributes a
 * snapshot carries besides them */
static int read_header(hid_t f, gizmo_io_header *h)
{
    memset(h, 0, sizeof(*h));
    if (H5Lexists(f, "Header", H5P_DEFAULT) <= 0) return fail(GIZMO_IO_ERR_FORMAT, "no /Header group");
    hid_t g = H5Gopen2(f, "Header", H5P_DEFAULT);
    if (g < 0) return fail(GIZMO_IO_ERR_FORMAT, "cannot open /Header");
    struct { const char *name; hid_t t; void *buf; hssize_t n; int required; } a[] = {
        {"NumPart_ThisFile", H5T_NATIVE_INT, h->npart, 6, 1},
        {"NumPart_Total", H5T_NATIVE_UINT, h->npart_total, 6, 1},
        {"NumPart_Total_HighWord", H5T_NATIVE_UINT, h->npart_total_high, 6, 1},
        {"MassTable", H5T_NATIVE_DOUBLE, h->mass_table, 6, 1},
        {"Time", H5T_NATIVE_DOUBLE, &h->time, 1, 1},
        {"NumFilesPerSnapshot", H5T_NATIVE_INT, &h->num_files, 1, 1},
        {"Flag_DoublePrecision", H5T_NATIVE_INT, &h->flag_double, 1, 1},
        {"Redshift", H5T_NATIVE_DOUBLE, &h->redshift, 1, 0},
        {"BoxSize", H5T_NATIVE_DOUBLE, &h->box_size, 1, 0},
        {"HubbleParam", H5T_NATIVE_DOUBLE, &h->hubble_param, 1, 0},
        {"ComovingIntegrationOn", H5T_NATIVE_INT, &h->comoving, 1, 0},
    };
    int rc = 0;
    for (size_t k = 0; k < sizeof(a) / sizeof(a[0]) && !rc; k++)
        if (read_attr(g, a[k].name, a[k].t, a[k].buf, a[k].n, a[k].required) < 0)
            rc = fail(GIZMO_IO_ERR_FORMAT, "Header attribute %s missing or of the wrong size", a[k].name);
    H5Gclose(g);
    return rc;
}

/* ------------------------------------------------------------------------- */
/* datasets                                                                   */
/* ------------------------------------------------------------------------- */
/* Dataset `name` of `grp` with n rows of `width` values into out[n * width]
 * (double), read in the input precision (float or double memory type, as
 * read_file reads MyInputFloat blocks, read_ic.c:1136-1150) and widened.
 * Returns 0, 1 when the dataset is absent and !required, or an error code. */
static int read_block(hid_t grp, const char *name, long n, int width, int input_double, double *out, int required)
{
    if (H5Lexists(grp, name, H5P_DEFAULT) <= 0)
        return required ? fail(GIZMO_IO_ERR_FORMAT, "PartType0/%s missing", name) : 1;
    hid_t d = H5Dopen2(grp, name, H5P_DEFAULT);
    if (d < 0) return fail(GIZMO_IO_ERR_FORMAT, "cannot open PartType0/%s", name);
    hid_t sp = H5Dget_space(d);
    int rank = H5Sget_simple_extent_ndims(sp);
    hsize_t dims[2] = {0, 0};
    H5Sget_simple_extent_dims(sp, dims, NULL);
    H5Sclose(sp);
    const int ok_shape = (width == 1) ? (rank == 1 && dims[0] == (hsize_t)n)
                                      : (rank == 2 && dims[0] == (hsize_t)n && dims[1] == (hsize_t)width);
    if (!ok_shape) {
        H5Dclose(d);
        return fail(GIZMO_IO_ERR_FORMAT, "PartType0/%s has the wrong shape for %ld particles", name, n);
    }
    const size_t cnt = (size_t)n * (size_t)width;
    herr_t e;
    if (input_double) {
        e = H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out);
    } else {
        float *tmp = malloc(cnt * sizeof(float) + 1);
        if (!tmp) { H5Dclose(d); return fail(GIZMO_IO_ERR_NOMEM, "no memory for PartType0/%s", name); }
        e = H5Dread(d, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, tmp);
        for (size_t k = 0; k < cnt; k++) out[k] = (double)tmp[k];
        free(tmp);
    }
    H5Dclose(d);
    return (e < 0) ? fail(GIZMO_IO_ERR_FORMAT, "reading PartType0/%s failed", name) : 0;
}

static int write_block(hid_t grp, const char *name, long n, int width, int output_double, const double *in)
{
    hsize_t dims[2] = {(hsize_t)n, (hsize_t)width};
    hid_t sp = H5Screate_simple(width == 1 ? 1 : 2, dims, NULL);
    hid_t ft = output_double ? H5T_NATIVE_DOUBLE : H5T_NATIVE_FLOAT;
    hid_t d = H5Dcreate2(grp, name, ft, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    int rc = (d < 0) ? -1 : (H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, in) < 0 ? -1 : 0);
    if (d >= 0) H5Dclose(d);
    H5Sclose(sp);
    return rc ? fail(GIZMO_IO_ERR_FORMAT, "writing PartType0/%s failed", name) : 0;
}

/* periodic wrap into [0, L) (the BOX_PERIODIC position wrap applied to every
 * particle, predict.c / domain.c); positions further than 1e6 boxes out are
 * rejected rather than walked back */
static int wrap_into_box(double *x, double L)
{
    if (!isfinite(*x) || fabs(*x) > 1e6 * L) return -1;
    while (*x < 0) *x += L;
    while (*x >= L) *x -= L;
    return 0;
}

int gizmo_read_ic(const char *fname, int input_double, gizmo_io_header *hdr_out)
{
    quiet_hdf5();
    g_err[0] = 0;
    if (!fname) return fail(GIZMO_IO_ERR_FORMAT, "no file name");
    if (!(All.BoxSize > 0)) return fail(GIZMO_IO_ERR_FORMAT, "All.BoxSize must be set (> 0) before reading");
    hid_t f = H5Fopen(fname, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return fail(GIZMO_IO_ERR_OPEN, "can't open file `%s' for reading initial conditions", fname);
    gizmo_io_header h;
    int rc = read_header(f, &h);
    if (!rc) {
        /* read_ic.c:820-833 */
        if (input_double && h.flag_double == 0)
            rc = fail(GIZMO_IO_ERR_PRECISION_SINGLE, "input read in double precision, but the file is in single precision");
        else if (!input_double && h.flag_double)
            rc = fail(GIZMO_IO_ERR_PRECISION_DOUBLE, "input read in single precision, but the file is in double precision");
        else if (h.num_files > 1)
            rc = fail(GIZMO_IO_ERR_MULTIFILE, "NumFilesPerSnapshot=%d: read each file of the set", h.num_files);
    }
    if (!rc && h.num_files <= 1) {
        /* read_ic.c:800-806: a single file holds the totals */
        for (int t = 0; t < 6; t++) { h.npart_total[t] = (unsigned int)h.npart[t]; h.npart_total_high[t] = 0; }
    }
    const long n = h.npart[0];
    if (!rc && n <= 0) rc = fail(GIZMO_IO_ERR_FORMAT, "the file holds no gas (NumPart_ThisFile[0]=%d)", h.npart[0]);
    hid_t g = -1;
    if (!rc) {
        if (H5Lexists(f, "PartType0", H5P_DEFAULT) <= 0 || (g = H5Gopen2(f, "PartType0", H5P_DEFAULT)) < 0)
            rc = fail(GIZMO_IO_ERR_FORMAT, "no /PartType0 group for %ld gas particles", n);
    }
    double *buf = NULL;
    unsigned long long *ids = NULL;
    if (!rc) {
        buf = malloc(sizeof(double) * 3 * (size_t)n);
        ids = malloc(sizeof(unsigned long long) * (size_t)n);
        if (!buf || !ids || gizmo_host_allocate((int)n)) rc = fail(GIZMO_IO_ERR_NOMEM, "no memory for %ld particles", n);
    }
    /* Coordinates, Velocities (read_ic.c:187-211) */
    if (!rc && !(rc = read_block(g, "Coordinates", n, 3, input_double, buf, 1))) {
        for (long i = 0; i < n && !rc; i++)
            for (int k = 0; k < 3 && !rc; k++) {
                P[i].Pos[k] = buf[3 * i + k];
                if (wrap_into_box(&P[i].Pos[k], All.BoxSize))
                    rc = fail(GIZMO_IO_ERR_FORMAT, "particle %ld: position %g cannot be placed in the box", i, buf[3 * i + k]);
            }
    }
    if (!rc && !(rc = read_block(g, "Velocities", n, 3, input_double, buf, 1)))
        for (long i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) P[i].Vel[k] = SphP[i].VelPred[k] = buf[3 * i + k];
    /* ParticleIDs (read_ic.c:213-215), read as 64-bit IDs (LONGIDS) */
    if (!rc) {
        hid_t d = (H5Lexists(g, "ParticleIDs", H5P_DEFAULT) > 0) ? H5Dopen2(g, "ParticleIDs", H5P_DEFAULT) : -1;
        if (d < 0) rc = fail(GIZMO_IO_ERR_FORMAT, "PartType0/ParticleIDs missing");
        else {
            hid_t sp = H5Dget_space(d);
            hssize_t npts = H5Sget_simple_extent_npoints(sp);
            H5Sclose(sp);
            if (npts != n || H5Dread(d, H5T_NATIVE_ULLONG, H5S_ALL, H5S_ALL, H5P_DEFAULT, ids) < 0)
                rc = fail(GIZMO_IO_ERR_FORMAT, "PartType0/ParticleIDs unreadable or of the wrong size");
            H5Dclose(d);
            for (long i = 0; i < n && !rc; i++) P[i].ID = ids[i];
        }
    }
    /* Masses only for a zero MassTable entry (read_ic.c:105, 229-231) */
    if (!rc) {
        if (h.mass_table[0] != 0) {
            for (long i = 0; i < n; i++) P[i].Mass = h.mass_table[0];
        } else if (!(rc = read_block(g, "Masses", n, 1, input_double, buf, 1))) {
            for (long i = 0; i < n; i++) P[i].Mass = buf[i];
        }
    }
    /* InternalEnergy (read_ic.c:267-269), InternalEnergyPred (init.c:503) */
    if (!rc && !(rc = read_block(g, "InternalEnergy", n, 1, input_double, buf, 1)))
        for (long i = 0; i < n; i++) SphP[i].InternalEnergy = SphP[i].InternalEnergyPred = buf[i];
    /* SmoothingLength when present (INPUT_READ_HSML / restart from a snapshot) */
    if (!rc) {
        int s = read_block(g, "SmoothingLength", n, 1, input_double, buf, 0);
        if (s == 0) for (long i = 0; i < n; i++) P[i].Hsml = buf[i];
        else if (s == 1) for (long i = 0; i < n; i++) P[i].Hsml = 0;
        else rc = s;
    }
    if (!rc)
        for (long i = 0; i < n; i++) { P[i].Type = 0; P[i].TimeBin = 0; }
    free(buf);
    free(ids);
    if (g >= 0) H5Gclose(g);
    H5Fclose(f);
    if (rc && P) gizmo_host_free();
    if (!rc && hdr_out) *hdr_out = h;
    return rc;
}

int gizmo_write_snapshot(const char *fname, int output_double)
{
    quiet_hdf5();
    g_err[0] = 0;
    if (!fname) return fail(GIZMO_IO_ERR_FORMAT, "no file name");
    if (N_gas <= 0 || !P || !SphP) return fail(GIZMO_IO_ERR_FORMAT, "no gas particles to write");
    hid_t f = H5Fcreate(fname, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return fail(GIZMO_IO_ERR_OPEN, "can't open file `%s' for writing snapshot", fname);
    int rc = 0;
    /* write_file header fill (io.c:4112-4135) and attributes (io.c:4496-4566) */
    hid_t h = H5Gcreate2(f, "Header", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    int npart[6] = {N_gas, 0, 0, 0, 0, 0};
    unsigned int ntot[6] = {(unsigned int)N_gas, 0, 0, 0, 0, 0}, nhigh[6] = {0, 0, 0, 0, 0, 0};
    double mass_table[6] = {0, 0, 0, 0, 0, 0};
    double t = All.Time, z = 0, box = All.BoxSize, hubble = 1.0, desngb = All.DesNumNgb;
    int one = 1, zero = 0, flag_double = output_double ? 1 : 0, kernel_id = 3;
    if (h < 0) rc = -1;
    rc = rc ? rc : write_attr(h, "NumPart_ThisFile", H5T_NATIVE_INT, npart, 6);
    rc = rc ? rc : write_attr(h, "NumPart_Total", H5T_NATIVE_UINT, ntot, 6);
    rc = rc ? rc : write_attr(h, "NumPart_Total_HighWord", H5T_NATIVE_UINT, nhigh, 6);
    rc = rc ? rc : write_attr(h, "MassTable", H5T_NATIVE_DOUBLE, mass_table, 6);
    rc = rc ? rc : write_attr(h, "Time", H5T_NATIVE_DOUBLE, &t, 1);
    rc = rc ? rc : write_attr(h, "Redshift", H5T_NATIVE_DOUBLE, &z, 1);
    rc = rc ? rc : write_attr(h, "BoxSize", H5T_NATIVE_DOUBLE, &box, 1);
    rc = rc ? rc : write_attr(h, "NumFilesPerSnapshot", H5T_NATIVE_INT, &one, 1);
    rc = rc ? rc : write_attr(h, "ComovingIntegrationOn", H5T_NATIVE_INT, &zero, 1);
    rc = rc ? rc : write_attr(h, "HubbleParam", H5T_NATIVE_DOUBLE, &hubble, 1);
    rc = rc ? rc : write_attr(h, "Flag_Sfr", H5T_NATIVE_INT, &zero, 1);
    rc = rc ? rc : write_attr(h, "Flag_Cooling", H5T_NATIVE_INT, &zero, 1);
    rc = rc ? rc : write_attr(h, "Flag_StellarAge", H5T_NATIVE_INT, &zero, 1);
    rc = rc ? rc : write_attr(h, "Flag_Metals", H5T_NATIVE_INT, &zero, 1);
    rc = rc ? rc : write_attr(h, "Flag_Feedback", H5T_NATIVE_INT, &zero, 1);
    rc = rc ? rc : write_attr(h, "Flag_DoublePrecision", H5T_NATIVE_INT, &flag_double, 1);
    rc = rc ? rc : write_attr(h, "Flag_IC_Info", H5T_NATIVE_INT, &zero, 1);
    rc = rc ? rc : write_attr(h, "Kernel_Function_ID", H5T_NATIVE_INT, &kernel_id, 1);
    rc = rc ? rc : write_attr(h, "Effective_Kernel_NeighborNumber", H5T_NATIVE_DOUBLE, &desngb, 1);
    if (h >= 0) H5Gclose(h);
    if (rc) { H5Fclose(f); return fail(GIZMO_IO_ERR_FORMAT, "writing the Header of `%s' failed", fname); }
    /* PartType0 blocks in write_file order (io.c:4067-4095) */
    const long n = N_gas;
    hid_t g = H5Gcreate2(f, "PartType0", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    double *buf = malloc(sizeof(double) * 3 * (size_t)n);
    unsigned long long *ids = malloc(sizeof(unsigned long long) * (size_t)n);
    if (g < 0 || !buf || !ids) rc = fail(GIZMO_IO_ERR_NOMEM, "cannot create PartType0 of `%s'", fname);
    if (!rc) {
        for (long i = 0; i < n; i++) for (int k = 0; k < 3; k++) buf[3 * i + k] = P[i].Pos[k];
        rc = write_block(g, "Coordinates", n, 3, output_double, buf);
    }
    if (!rc) {
        for (long i = 0; i < n; i++) for (int k = 0; k < 3; k++) buf[3 * i + k] = P[i].Vel[k];
        rc = write_block(g, "Velocities", n, 3, output_double, buf);
    }
    if (!rc) {
        for (long i = 0; i < n; i++) ids[i] = P[i].ID;
        hsize_t dims = (hsize_t)n;
        hid_t sp = H5Screate_simple(1, &dims, NULL);
        hid_t d = H5Dcreate2(g, "ParticleIDs", H5T_NATIVE_ULLONG, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        if (d < 0 || H5Dwrite(d, H5T_NATIVE_ULLONG, H5S_ALL, H5S_ALL, H5P_DEFAULT, ids) < 0)
            rc = fail(GIZMO_IO_ERR_FORMAT, "writing PartType0/ParticleIDs failed");
        if (d >= 0) H5Dclose(d);
        H5Sclose(sp);
    }
    struct { const char *name; int what; } blocks[] = {
        {"Masses", 0}, {"InternalEnergy", 1}, {"Density", 2}, {"SmoothingLength", 3}};
    for (size_t b = 0; b < sizeof(blocks) / sizeof(blocks[0]) && !rc; b++) {
        for (long i = 0; i < n; i++) {
            switch (blocks[b].what) {
            case 0: buf[i] = P[i].Mass; break;
            case 1: buf[i] = SphP[i].InternalEnergy; break;
            case 2: buf[i] = SphP[i].Density; break;
            default: buf[i] = P[i].Hsml; break;
            }
        }
        rc = write_block(g, blocks[b].name, n, 1, output_double, buf);
    }
    free(buf);
    free(ids);
    if (g >= 0) H5Gclose(g);
    H5Fclose(f);
    return rc;
}

/* ------------------------------------------------------------------------- */
/* setup_smoothinglengths (init.c:1067-1110)                                  */
/* ------------------------------------------------------------------------- */
typedef struct { uint64_t key; int i; } keyed_t;

static int cmp_keyed(const void *a, const void *b)
{
    const keyed_t *x = a, *y = b;
    return (x->key < y->key) ? -1 : (x->key > y->key) ? 1 : (x->i - y->i);
}

/* 21 bits per axis, interleaved with x highest: the cells of level L are the
 * runs of equal key >> (3 (21 - L)) */
static uint64_t morton21(const double pos[3], double L)
{
    uint64_t key = 0;
    uint32_t c[3];
    for (int k = 0; k < 3; k++) {
        double s = pos[k] / L * 2097152.0;
        c[k] = (s <= 0) ? 0u : (s >= 2097151.0 ? 2097151u : (uint32_t)s);
    }
    for (int b = 20; b >= 0; b--)
        for (int k = 0; k < 3; k++) key = (key << 1) | ((c[k] >> b) & 1u);
    return key;
}

int gizmo_setup_smoothinglengths(void)
{
    g_err[0] = 0;
    const long n = N_gas;
    if (n <= 0 || !P) return fail(GIZMO_IO_ERR_FORMAT, "no gas particles");
    if (!(All.BoxSize > 0)) return fail(GIZMO_IO_ERR_FORMAT, "All.BoxSize must be > 0");
    keyed_t *k = malloc(sizeof(keyed_t) * (size_t)n);
    double *cum = malloc(sizeof(double) * (size_t)(n + 1));
    if (!k || !cum) { free(k); free(cum); return fail(GIZMO_IO_ERR_NOMEM, "no memory"); }
    for (long i = 0; i < n; i++) { k[i].key = morton21(P[i].Pos, All.BoxSize); k[i].i = (int)i; }
    qsort(k, (size_t)n, sizeof(keyed_t), cmp_keyed);
    cum[0] = 0;
    for (long s = 0; s < n; s++) cum[s + 1] = cum[s] + P[k[s].i].Mass;
    for (long s = 0; s < n; s++) {
        const int i = k[s].i;
        const double m = P[i].Mass;
        double h = 0.05 * All.BoxSize;
        /* from the finest cell upwards: the first cell with 2 DesNumNgb m */
        for (int L = 21; L >= 0; L--) {
            const int shift = 3 * (21 - L);
            const uint64_t cell = (shift >= 64) ? 0 : (k[s].key >> shift);
            long lo = s, hi = s;  /* the run of equal cell key around s */
            long a = 0, b = s;
            while (a < b) { long mid = (a + b) / 2; if (((shift >= 64) ? 0 : (k[mid].key >> shift)) < cell) a = mid + 1; else b = mid; }
            lo = a;
            a = s; b = n;
            while (a < b) { long mid = (a + b) / 2; if (((shift >= 64) ? 0 : (k[mid].key >> shift)) <= cell) a = mid + 1; else b = mid; }
            hi = a;
            const double M = cum[hi] - cum[lo];
            if (!(2 * All.DesNumNgb * m > M) || L == 0) {
                const double len = All.BoxSize / (double)(1L << L);
                if (M > 0 && len > 0) h = pow(3.0 / (4 * M_PI) * All.DesNumNgb * m / M, 0.333333) * len;
                break;
            }
        }
        P[i].Hsml = h;
    }
    free(k);
    free(cum);
    return 0;
}
