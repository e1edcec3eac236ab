/*
 * mfm_oracle.h -- CPU FP64 restatement of GIZMO's meshless-finite-mass hydro
 * pass (density / gradients / force).  TEST INFRASTRUCTURE ONLY: this code is
 * the checker that the HIP library is compared against, and the CPU baseline
 * that bench.py times beside it.  Nothing in the product links or calls it.
 *
 * Every function cites the reference lines it restates.  Arithmetic is kept
 * in the reference's operation order so that, for the same summation order,
 * results agree to the last bit (compile with -ffp-contract=off).
 *
 * Array conventions (all arrays indexed by particle, caller order):
 *   pos[n][3], vel[n][3]           FP64 positions / VelPred
 *   nvt[n][9]                      row-major 3x3 (NV_T after inversion)
 *   grad_rho[n][3], grad_p[n][3]   gradients
 *   grad_v[n][3][3]                grad_v[i][k][d] = d v_k / d x_d
 *                                  (reference SphP.Gradients.Velocity[k][d])
 */
#ifndef MFM_ORACLE_H
#define MFM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    double box[3];          /* boxSize_X/Y/Z (allvars.h:1462) */
    double gamma;           /* EOS_GAMMA (allvars.h:1092) */
    double des_num_ngb;     /* All.DesNumNgb */
    double max_num_ngb_dev; /* All.MaxNumNgbDeviation */
    double min_hsml;        /* All.MinHsml */
    double max_hsml;        /* All.MaxHsml */
    int first_step;         /* All.Time == All.TimeBegin (density.c:509) */
    int maxiter;            /* MAXITER (allvars.h:1294) */
    int nthreads;           /* OpenMP threads (<=0: runtime default) */
    int brute_force;        /* 1: O(N^2) neighbour search (truth), 0: cell grid or octree */
    int tree;               /* 1: octree walk with node hsml maxima, the reference's search
                               (ngb.c:102-130); 0: cell grid */
    int method;             /* 0: meshless finite mass (HYDRO_MESHLESS_FINITE_MASS); 1: meshless
                               finite volume (HYDRO_MESHLESS_FINITE_VOLUME) with the smoothed
                               mesh motion HYDRO_FIX_MESH_MOTION = 5 (allvars.h:94-98) */
    double timebase_interval; /* MFV: physical time of one integer time unit,
                               UNIT_INTEGERTIME_IN_PHYSICAL (allvars.h:908) */
} orc_params;

typedef struct orc_state {
    int n;
    /* inputs */
    const double *pos;
    const double *vel;
    const double *mass;
    const double *u;          /* InternalEnergyPred */
    /* density outputs (hsml and cn are in/out) */
    double *hsml;
    double *num_ngb;          /* NumNgb^(1/NUMDIMS) after the pass */
    double *dhsml;            /* DhsmlNgbFactor */
    double *divvel;           /* Particle_DivVel */
    double *density;
    double *pressure;
    double *nvt;              /* [n][9] */
    double *cn;               /* ConditionNumber: previous value in, new out */
    double *fce;              /* FaceClosureError */
    /* gradient outputs */
    double *grad_rho;
    double *grad_p;
    double *grad_v;
    /* force outputs (vsig is in/out: previous MaxSignalVel feeds WAKEUP) */
    double *acc;
    double *dtu;
    double *vsig;
    unsigned char *wakeup;    /* may be NULL */
    unsigned char *is_active; /* [n] 1 if particle is in the active list */
    /* optional (may be NULL): sums of the absolute values of the pair terms
     * behind each summed output, after the same normalisation.  They bound
     * what a different summation order can change (n * eps * sum|terms|)
     * and serve as the error scale of parity tests. */
    double *divvel_l1;        /* [n] */
    double *fce_l1;           /* [n] */
    double *grad_l1;          /* [n][15]: rho(3), P(3), v(3x3) */
    double *acc_l1;           /* [n][3] */
    double *dtu_l1;           /* [n] */
    /* optional: 1 where the force loop compared P_i/rho_i with P_j/rho_j
     * (hydro_core_meshless.h:349-351) for values within 4 ulp of each other.
     * Such a branch is decided by rounding, so any change of summation order
     * (in the reference as well) can flip it. */
    unsigned char *tie_flag;  /* [n] */
    /* optional: sum over the gradient pairs of |w_ij dp_ij| (|Q_i| + |Q_j|),
     * through |NV_T| like grad_l1: the scale of the error that rounding of
     * the inputs rho, P, v themselves (not of this stage's sums) puts into
     * the differences Q_j - Q_i.  Where Q is uniform up to rounding (an
     * unjittered lattice) the differences are pure rounding and grad_l1 is
     * itself rounding-sized. */
    double *grad_in;          /* [n][15] */
    /* MFV only (method 1): ParticleVel (density output, force input) [n][3];
     * DtMass, dMass (in/out: inactive particles accumulate their share),
     * GravWorkTerm [n][3]; ti_step = P[].dt_step, the integer timestep
     * (NULL: equal for every particle) */
    double *particle_vel;
    double *dt_mass;
    double *d_mass;
    double *grav_work;
    const long long *ti_step;
    /* optional: sums of |terms| of the MFV outputs (error scales; a pair's
     * mass-flux term counts as rho (|v - v_face| + c) times the face area,
     * the size of the terms the flux is computed from, so that a flux that
     * cancels to rounding in a uniform flow has a meaningful scale) */
    double *pvel_l1;          /* [n][3] */
    double *dtmass_l1;        /* [n] */
    double *dmass_l1;         /* [n] */
    double *gravwork_l1;      /* [n][3] */
} orc_state;

typedef struct orc_stats {
    int iterations;               /* number of density sweeps */
    long long dens_interactions;  /* sum over sweeps of |{j: r2 < h_i^2}| */
    long long grad_interactions;  /* sum over active i of gradient pair-set size */
    long long force_interactions; /* sum over active i of force pair-set size */
    int status;                   /* 0 ok, else the reference endrun code */
    long long riemann_retries;    /* pairs that needed the first-order retry */
    long long exact_solver_calls; /* pairs that reached the exact solver */
} orc_stats;

/* ---- element functions (exported for closed-form tests) ---- */
void   orc_kernel(double u, double h, double *wk, double *dwk);
double orc_matrix_invert(const double T[9], double Tinv[9]);
void   orc_reconstruct(double qi, const double gi[3], double qj, const double gj[3],
                       const double di[3], const double dj[3], double *ql, double *qr, int mode);
void   orc_slope_limiter(double grad[3], double valmax, double valmin, double alim, double h,
                         double shoot_tol, int pos_preserve, double d_max, double val_cen);
/* One call of the reference Riemann_solver() (reimann.h:286-388) for an MFM run
 * with the given adiabatic index.  L/R: rho, p, v[3].  Outputs P_M, S_M and the
 * MFM fluxes F_v[3], F_p.  Returns a bit mask of the stages used
 * (1 HLLC, 2 KT, 4 exact). */
int    orc_riemann(double gamma, const double L[5], const double R[5], const double n[3],
                   double press_tot_limiter, double out[6]);
/* The Kurganov-Tadmor/PWK flux alone (reimann.h:484-519), MFM branch:
 * out = P_M, S_M, F_v[3], F_p. */
void   orc_riemann_kt(double gamma, const double L[5], const double R[5], const double n[3], double out[6]);
/* Riemann_solver() for an MFV run (reimann.h:286-388 with the MFV branches):
 * out = P_M, S_M, F_rho, F_v[3], F_p in the frame of the states; returns the
 * stage mask (1 HLLC with the HLLC fluxes, 2 KT, 4 exact) or -1234545. */
int    orc_riemann_mfv(double gamma, const double L[5], const double R[5], const double n[3],
                       double press_tot_limiter, double out[7]);
/* The KT flux alone, MFV branch (reimann.h:484-507): out as above. */
void   orc_riemann_kt_mfv(double gamma, const double L[5], const double R[5], const double n[3], double out[7]);
/* The exact solver alone with the MFV sampling at the face (reimann.h:685-724,
 * 731-1030, 1146-1166): out as above; returns 0 or 1234545. */
int    orc_riemann_exact_mfv(double gamma, const double L[5], const double R[5], const double n[3],
                             double out[7]);
/* One star-state estimate of get_wavespeeds_and_pressure_star
 * (reimann.h:527-593): which 0 Gaburov HLLC (Davis wave speeds), 1 Roe
 * averages, 2 PVRS; out = P_M, S_M, S_L, S_R. */
void   orc_pstar_estimate(double gamma, const double L[5], const double R[5], const double n[3], int which,
                          double out[4]);
/* The whole estimate cascade with its limiter tests: out = P_M, S_M, S_L,
 * S_R; returns the estimate that stands (0 vacuum, 1 HLLC, 2 Roe, 3 PVRS). */
int    orc_wavespeeds(double gamma, const double L[5], const double R[5], const double n[3],
                      double press_tot_limiter, double out[4]);

/* One hsml update of density()'s iteration (density.c:594-846) from the
 * sweep's normalised neighbour number and DhsmlNgbFactor, for the target
 * des_num_ngb and window half-width des_num_ngb_dev (grown with the sweep
 * number iter here, as the reference does).  Updates h and the
 * brackets; returns 0 converged (or held at Min/MaxHsml), 1 sweep again,
 * 2 accepted by the bracket-width test, -1 no bracket (the run ends). */
int    orc_hsml_step(double num_ngb, double dhsml, double *h, double *left, double *right, int iter,
                     double des_num_ngb, double des_num_ngb_dev, double min_hsml, double max_hsml);

/* ---- the three stages (density.c:451, gradients.c:577, hydro_toplevel.c:1045) ---- */
int  orc_density(const orc_params *p, orc_state *s, const int *active, int n_active, orc_stats *st);
int  orc_gradients(const orc_params *p, orc_state *s, const int *active, int n_active, orc_stats *st);
int  orc_hydro_force(const orc_params *p, orc_state *s, const int *active, int n_active, orc_stats *st);

/* One (i,j) interaction of the force loop with the state in `s` (gradients etc.
 * already filled): out = F_v[3], F_p (before the mass division), v_sig, stage mask. */
int  orc_force_pair(const orc_params *p, const orc_state *s, int i, int j, double out[6]);

/* ---- neighbour sets for a fixed hsml (bit-exact comparisons) ----
 * mode 0: density set {j : r2 < h_i^2} (self included)                 density.c:281
 * mode 1: pair set    {j : r2 > 0 and (r2 < h_i^2 or r2 < h_j^2)}       hydro_evaluate.h:177-178
 * counts[i] for every i in active; if offsets != NULL, also a CSR list
 * (offsets[n_active+1], idx) with each row sorted ascending. */
long long orc_neighbor_counts(const orc_params *p, int n, const double *pos, const double *hsml,
                              const int *active, int n_active, int mode, int *counts);
long long orc_neighbor_lists(const orc_params *p, int n, const double *pos, const double *hsml,
                             const int *active, int n_active, int mode,
                             long long *offsets, int *idx, long long capacity);

#ifdef __cplusplus
}
#endif
#endif
