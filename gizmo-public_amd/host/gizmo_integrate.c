/*
 * gizmo_integrate.c -- kick, drift and timestep of the gas over P[]/SphP[]
 * around the hydro pass, with one timestep for every particle
 * (FORCE_EQUAL_TIMESTEPS).  See include/gizmo_mfm_host.h for the contract.
 *
 * These are the gas parts of do_the_kick (kicks.c:255-402), drift_particle
 * (predict.c:120-240), get_timestep (timestep.c:482-491) and the step of
 * run.c for a non-comoving, gravity-free MFM run: every All.cf_* factor is 1
 * and the gravitational terms vanish.  They are O(N) host loops beside the
 * O(N * neighbours) pass on the device.
 */
#include "gizmo_mfm_host.h"

#include <math.h>

/* the halving floor of kicks.c:329 and predict.c:215: a step that would take
 * the specific energy below half its value halves it instead */
static inline double energy_step(double u, double du)
{
    const double e = u + du;
    return (e < 0.5 * u) ? 0.5 * u : e;
}

void gizmo_kick_gas(double dt, int mode)
{
    for (int i = 0; i < N_gas; i++) {
        if (P[i].Type != 0 || !(P[i].Mass > 0)) continue;
        SphP[i].InternalEnergy = energy_step(SphP[i].InternalEnergy, SphP[i].DtInternalEnergy * dt);
        /* check_particle_for_temperature_minimum (eos.c:221-231) */
        if (All.MinEgySpec && SphP[i].InternalEnergy < All.MinEgySpec) {
            SphP[i].InternalEnergy = All.MinEgySpec;
            SphP[i].DtInternalEnergy = 0;
        }
        for (int k = 0; k < 3; k++) P[i].Vel[k] += SphP[i].HydroAccel[k] * dt;
        if (mode == 1) {
            /* kicks.c:386-392: after a full step the predicted values restart
             * from the current ones */
            for (int k = 0; k < 3; k++) SphP[i].VelPred[k] = P[i].Vel[k];
            SphP[i].InternalEnergyPred = SphP[i].InternalEnergy;
        }
    }
}

void gizmo_drift_gas(double dt)
{
    const double gm1 = gizmo_gamma() - 1.0;
    double L[3];
    for (int k = 0; k < 3; k++) L[k] = gizmo_box_side(k);
    for (int i = 0; i < N_gas; i++) {
        for (int k = 0; k < 3; k++) {
            double x = P[i].Pos[k] + P[i].Vel[k] * dt;
            /* BOX_PERIODIC wrap (a particle moves less than a box per step) */
            while (x < 0) x += L[k];
            while (x >= L[k]) x -= L[k];
            P[i].Pos[k] = x;
        }
        if (P[i].Type != 0 || !(P[i].Mass > 0)) continue;
        double divv_fac = P[i].Particle_DivVel * dt;
        if (divv_fac > 0.3) divv_fac = 0.3;
        if (divv_fac < -0.3) divv_fac = -0.3;
        for (int k = 0; k < 3; k++) SphP[i].VelPred[k] += SphP[i].HydroAccel[k] * dt;
        SphP[i].Density *= exp(-divv_fac);
        SphP[i].InternalEnergyPred = energy_step(SphP[i].InternalEnergyPred, SphP[i].DtInternalEnergy * dt);
        if (SphP[i].InternalEnergyPred < All.MinEgySpec) SphP[i].InternalEnergyPred = All.MinEgySpec;
        SphP[i].Pressure = gm1 * SphP[i].InternalEnergyPred * SphP[i].Density; /* eos.c:58-61 */
    }
}

double gizmo_gas_timestep(int i)
{
    double dt = All.MaxSizeTimestep > 0 ? All.MaxSizeTimestep : HUGE_VAL;
    if (P[i].Type == 0 && P[i].Mass > 0) {
        const double csnd = 0.5 * SphP[i].MaxSignalVel;
        const double L_particle = 1.61199 * P[i].Hsml / P[i].NumNgb; /* Get_Particle_Size */
        const double dt_courant = All.CourantFac * L_particle / csnd;
        if (dt_courant < dt) dt = dt_courant;
    }
    return dt;
}

double gizmo_global_timestep(void)
{
    double dt = HUGE_VAL;
    for (int i = 0; i < N_gas; i++) {
        const double d = gizmo_gas_timestep(i);
        if (d < dt) dt = d;
    }
    return dt;
}

int gizmo_run_until(double t_end, int max_steps)
{
    if (N_gas <= 0) return 0;
    gizmo_host_all_active();
    compute_hydro_densities_and_forces();
    if (gizmo_host_last_status()) return -gizmo_host_last_status();
    int steps = 0;
    while (All.Time < t_end && steps < max_steps) {
        double dt = gizmo_global_timestep();
        if (!(dt > 0) || !isfinite(dt)) {
            endrun(1); /* no usable timestep (the reference stops on a zero step, timestep.c) */
            return -1;
        }
        const int last = (All.Time + dt >= t_end);
        if (last) dt = t_end - All.Time;
        gizmo_kick_gas(0.5 * dt, 0);
        gizmo_drift_gas(dt);
        All.Time = last ? t_end : All.Time + dt;
        compute_hydro_densities_and_forces();
        if (gizmo_host_last_status()) return -gizmo_host_last_status();
        gizmo_kick_gas(0.5 * dt, 1);
        steps++;
    }
    return steps;
}
