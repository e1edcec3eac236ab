/*
 * gizmo_mfm_host.h -- C host that keeps GIZMO's hydro entry points and
 * particle structures for one rank and runs them on the MI355X library
 * (gzm.h).  Drop-in for the three calls of compute_hydro_densities_and_forces()
 * (accel.c:57-111):
 *
 *   void density(void);              proto.h:723   (hydro/density.c:451)
 *   void hydro_gradient_calc(void);  proto.h:962   (hydro/gradients.c:577)
 *   void hydro_force(void);          proto.h:749   (hydro/hydro_toplevel.c:1045)
 *
 * They take no arguments and work on the globals P[], SphP[], All, NumPart,
 * N_gas and the active list FirstActiveParticle / NextActiveParticle[]
 * (allvars.h:1652-1654), exactly as in the reference.  Only the fields the
 * MFM pass reads or writes are declared; names and meaning follow
 * allvars.h:2450-2778 (particle_data) and 2807-3235 (gas_cell_data).
 *
 * Error behaviour follows the reference: a failure calls endrun(code) with
 * the reference's code (1155 hsml not converged, 1234 no positive Riemann
 * pressure, 1234545 negative density into the exact solver).  endrun()
 * prints and exits unless a handler is installed with
 * gizmo_host_set_endrun_handler().
 *
 * Particles 0..N_gas-1 are gas (Type 0), as in GIZMO's memory layout; the
 * box is periodic with side All.BoxSize.  P[]/SphP[] must not change between
 * density() and hydro_force() of one pass, as in accel.c:57-111.
 */
#ifndef GIZMO_MFM_HOST_H
#define GIZMO_MFM_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

typedef double MyDouble; /* DOUBLEPRECISION (allvars.h:42, 1321-1327) */
typedef double MyFloat;

struct particle_data {
    MyDouble Pos[3];
    MyDouble Vel[3];
    MyDouble Mass;
    MyFloat Hsml;
    MyFloat NumNgb;          /* NumNgb^(1/3) after density() (density.c:1020) */
    MyFloat DhsmlNgbFactor;
    MyFloat Particle_DivVel;
    int Type;
    int TimeBin;
    unsigned long long ID;
};

struct gas_cell_data {
    MyDouble Density;
    MyDouble Pressure;
    MyDouble InternalEnergy;
    MyDouble InternalEnergyPred;
    MyDouble VelPred[3];
    MyDouble HydroAccel[3];
    MyDouble DtInternalEnergy;
    MyDouble MaxSignalVel;
    MyDouble ConditionNumber;
    MyDouble FaceClosureError;
    MyDouble NV_T[3][3];
    struct {
        MyDouble Density[3];
        MyDouble Pressure[3];
        MyDouble Velocity[3][3];
    } Gradients;
    int wakeup;              /* PPPZ[i].wakeup (hydro_evaluate.h:485-495) */
};

struct global_data_all_processes {
    double BoxSize;
    double DesNumNgb;
    double MaxNumNgbDeviation;
    double MinHsml;
    double MaxHsml;
    double Time;
    double TimeBegin;
    int MaxPart;
};

extern struct particle_data *P;
extern struct gas_cell_data *SphP;
extern struct global_data_all_processes All;
extern int NumPart, N_gas;
extern int FirstActiveParticle;
extern int *NextActiveParticle;

/* the three entry points (reference semantics) */
void density(void);
void hydro_gradient_calc(void);
void hydro_force(void);
/* accel.c:57-111 order: density, hmax update (implicit here), gradients, force */
void compute_hydro_densities_and_forces(void);

/* endrun (allvars.h:1635-1636): default prints and exits with the code */
typedef void (*gizmo_endrun_handler)(int code, const char *where);
void gizmo_host_set_endrun_handler(gizmo_endrun_handler handler);
void endrun(int code);

/* helpers for embedding and tests (not in the reference) */
int  gizmo_host_allocate(int n_gas);        /* P, SphP, NextActiveParticle for n gas particles */
void gizmo_host_free(void);
void gizmo_host_all_active(void);           /* active list = every particle, in index order */
int  gizmo_host_set_active(const int *idx, int n); /* active list = idx[0..n-1], in that order */
int  gizmo_host_set_device(int device);     /* GPU ordinal used by the next pass */
int  gizmo_host_last_status(void);          /* status of the last library call */
const char *gizmo_host_last_error(void);
void gizmo_host_release(void);              /* destroy the device context */

#ifdef __cplusplus
}
#endif
#endif
