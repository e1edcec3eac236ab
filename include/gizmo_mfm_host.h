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
 * box is periodic with sides All.BoxSize * All.BoxLong[k].  P[]/SphP[] must
 * not change between density() and hydro_force() of one pass, as in
 * accel.c:57-111.
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
    /* HYDRO_MESHLESS_FINITE_VOLUME fields (allvars.h gas_cell_data), used
     * after gizmo_host_set_method(1): density() sets ParticleVel, hydro_force()
     * DtMass and GravWorkTerm of the active particles and adds to dMass of
     * every particle an active neighbour exchanged mass with */
    MyDouble ParticleVel[3];
    MyDouble DtMass;
    MyDouble dMass;
    MyDouble GravWorkTerm[3];
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
    /* time integration (kick / drift / timestep) */
    double CourantFac;       /* All.CourantFac (timestep.c:485-487) */
    double MaxSizeTimestep;  /* All.MaxSizeTimestep (timestep.c:260) */
    double MinEgySpec;       /* All.MinEgySpec: specific-energy floor, 0 = none (eos.c:221-231) */
    double BoxLong[3];       /* BOX_LONG_X/Y/Z: box side k is BoxSize * BoxLong[k]; 0 counts as 1
                                (allvars.h:1518-1560) */
    double GasGamma;         /* adiabatic index, GAMMA_DEFAULT / EOS_GAMMA (allvars.h:1092); 0 = 5/3 */
    double Timebase_interval; /* physical time of one integer time unit (UNIT_INTEGERTIME_IN_PHYSICAL,
                                allvars.h:908, without comoving integration); the MFV mass exchange
                                uses it with each particle's integer step 2^TimeBin */
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

/* time integration of the gas with equal timesteps for every particle
 * (FORCE_EQUAL_TIMESTEPS), on top of the hydro pass: */
/* kicks.c:255-402 (do_the_kick, gas part): u += DtInternalEnergy dt, halved
 * instead where that would fall below u/2, floored at MinEgySpec;
 * Vel += HydroAccel dt; mode 1 (second half kick) also sets VelPred = Vel and
 * InternalEnergyPred = InternalEnergy */
void gizmo_kick_gas(double dt, int mode);
/* predict.c:120-240 (drift_particle, gas part): Pos += Vel dt, wrapped into
 * the periodic box; VelPred += HydroAccel dt; Density *= exp(-DivVel dt) with
 * DivVel dt clamped to +-0.3; InternalEnergyPred += DtInternalEnergy dt with
 * the same halving floor and MinEgySpec; Pressure = (gamma - 1) rho u_pred
 * (set_eos_pressure, eos.c:58-61) */
void gizmo_drift_gas(double dt);
/* timestep.c:482-491: CourantFac * Get_Particle_Size(i) / (0.5 MaxSignalVel),
 * with Get_Particle_Size = 1.61199 Hsml / NumNgb^(1/3) (predict.c:381-395),
 * at most MaxSizeTimestep */
double gizmo_gas_timestep(int i);
/* the smallest of them over the gas (FORCE_EQUAL_TIMESTEPS, timestep.c:71-77) */
double gizmo_global_timestep(void);
/* The main loop of run.c with equal timesteps: hydro forces at the start,
 * then per step dt = min(global timestep, t_end - Time), first half kick,
 * drift, Time += dt, hydro forces, second half kick; until Time reaches t_end
 * or max_steps steps.  Returns the number of steps taken, or the negative
 * status of a failed pass (endrun is called as well). */
int gizmo_run_until(double t_end, int max_steps);

/* helpers for embedding and tests (not in the reference) */
int  gizmo_host_allocate(int n_gas);        /* P, SphP, NextActiveParticle for n gas particles */
void gizmo_host_free(void);
void gizmo_host_all_active(void);           /* active list = every particle, in index order */
int  gizmo_host_set_active(const int *idx, int n); /* active list = idx[0..n-1], in that order */
int  gizmo_host_set_device(int device);     /* GPU ordinal used by the next pass */
/* hydro method of the next passes: 0 meshless finite mass (default), 1 meshless
 * finite volume with HYDRO_FIX_MESH_MOTION = 5 (FP64); the reference fixes it at
 * compile time (HYDRO_MESHLESS_FINITE_MASS / _VOLUME).  The kick / drift /
 * timestep helpers below integrate MFM runs only.  Returns 0, or -1 for
 * another value. */
int  gizmo_host_set_method(int method);
int  gizmo_host_last_status(void);          /* status of the last library call */
const char *gizmo_host_last_error(void);
double gizmo_box_side(int k);               /* periodic box side k: BoxSize * BoxLong[k] */
double gizmo_gamma(void);                   /* adiabatic index of the gas: GasGamma, or 5/3 */
void gizmo_host_release(void);              /* destroy the device context */

#ifdef __cplusplus
}
#endif
#endif
