/*
 * gizmo_io.h -- GIZMO's HDF5 initial-condition and snapshot layout for the
 * gas of one rank, over the host mirror's P[] / SphP[] (gizmo_mfm_host.h).
 *
 * File layout (the reference's ICFormat 3 / SnapFormat 3, as
 * scripts/make_IC.py:96-165 writes an IC):
 *   /Header attributes    NumPart_ThisFile int[6], NumPart_Total uint[6],
 *                         NumPart_Total_HighWord uint[6], MassTable double[6],
 *                         Time, Redshift, BoxSize, HubbleParam double,
 *                         NumFilesPerSnapshot, Flag_DoublePrecision,
 *                         ComovingIntegrationOn int
 *   /PartType0 datasets   Coordinates [N][3], Velocities [N][3], ParticleIDs [N],
 *                         Masses [N] (only when MassTable[0] == 0),
 *                         InternalEnergy [N]; snapshots add Density [N] and
 *                         SmoothingLength [N]
 *
 *   gizmo_read_ic()               read_ic() + read_file() for one HDF5 file
 *                                 (read_ic.c:33, 698, 1379-1411) and the gas
 *                                 part of init() (init.c:503-507)
 *   gizmo_write_snapshot()        savepositions() / write_file() for the gas
 *                                 (io.c:35, 4067; header attributes io.c:4496-4566)
 *   gizmo_setup_smoothinglengths() the initial kernel length of init.c:1067-1110
 *
 * Errors follow the reference: the functions return the code the reference
 * would pass to endrun() (10 / 11 for a precision flag that does not match the
 * build's input precision, read_ic.c:820-833; 123 for a file that cannot be
 * opened, read_ic.c:718-722) or GIZMO_IO_ERR_FORMAT for a file that is not in
 * this layout; the message is in gizmo_io_last_error().  Nothing aborts.
 */
#ifndef GIZMO_IO_H
#define GIZMO_IO_H

#ifdef __cplusplus
extern "C" {
#endif

#define GIZMO_IO_ERR_PRECISION_DOUBLE 10   /* file in double, build reads single (read_ic.c:829) */
#define GIZMO_IO_ERR_PRECISION_SINGLE 11   /* file in single, build reads double (read_ic.c:822) */
#define GIZMO_IO_ERR_OPEN 123              /* cannot open the file (read_ic.c:720) */
#define GIZMO_IO_ERR_FORMAT 3001           /* missing or malformed group, dataset or attribute */
#define GIZMO_IO_ERR_NOMEM 3002            /* host allocation failed */
#define GIZMO_IO_ERR_MULTIFILE 3003        /* NumFilesPerSnapshot > 1: read each file of the set */

typedef struct gizmo_io_header {
    int npart[6];                      /* NumPart_ThisFile */
    unsigned int npart_total[6];       /* NumPart_Total */
    unsigned int npart_total_high[6];  /* NumPart_Total_HighWord */
    double mass_table[6];              /* MassTable */
    double time;                       /* Time */
    double redshift;                   /* Redshift (written; not read by read_ic) */
    double box_size;                   /* BoxSize (written; the run's box comes from All.BoxSize) */
    double hubble_param;               /* HubbleParam (written) */
    int num_files;                     /* NumFilesPerSnapshot */
    int flag_double;                   /* Flag_DoublePrecision */
    int comoving;                      /* ComovingIntegrationOn (written) */
} gizmo_io_header;

/* Read the gas (PartType0) of one HDF5 IC or snapshot file into P[]/SphP[]:
 * allocates them for NumPart_ThisFile[0] particles (gizmo_host_allocate),
 * then Pos, Vel, ID, Mass (from the dataset, or MassTable[0] when that is
 * non-zero), InternalEnergy; VelPred = Vel, InternalEnergyPred =
 * InternalEnergy, Type 0, all particles active (init.c:503-507).
 * SmoothingLength is read into Hsml when present (INPUT_READ_HSML,
 * restarts from a snapshot), else Hsml = 0.  Positions are wrapped into the
 * periodic box [0, All.BoxSize).
 * input_double selects the input precision the reference fixes at compile
 * time (INPUT_IN_DOUBLEPRECISION): 0 reads every floating-point dataset as
 * float (MyInputFloat, allvars.h:1335-1348) and needs Flag_DoublePrecision 0;
 * 1 reads double and needs Flag_DoublePrecision 1.  Other particle types in
 * the file are left out (the hydro pass sees gas only); their counts stay in
 * *hdr.  hdr may be NULL. */
int gizmo_read_ic(const char *fname, int input_double, gizmo_io_header *hdr);

/* Write the gas of P[]/SphP[] (N_gas particles) as an HDF5 snapshot:
 * Header attributes as write_header_attributes_in_hdf5 (io.c:4496-4566)
 * with Time = All.Time, BoxSize = All.BoxSize, MassTable zero (masses per
 * particle), and PartType0 Coordinates, Velocities, ParticleIDs, Masses,
 * InternalEnergy, Density, SmoothingLength.  output_double: 1 writes
 * floating-point blocks as double (OUTPUT_IN_DOUBLEPRECISION), 0 as float;
 * Flag_DoublePrecision records the choice. */
int gizmo_write_snapshot(const char *fname, int output_double);

/* Initial kernel lengths of the gas (init.c:1067-1110): for each particle,
 * the smallest cell of the octree over the periodic box (cells of side
 * BoxSize / 2^L) that holds at least 2 * DesNumNgb times the particle's mass,
 * and Hsml = (3 DesNumNgb m / (4 pi M_cell))^(1/3) * side.  A cell without
 * mass (impossible for a particle's own cell) gives 0.05 BoxSize. */
int gizmo_setup_smoothinglengths(void);

const char *gizmo_io_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
