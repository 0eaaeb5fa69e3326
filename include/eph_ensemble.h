/*
 * eph_ensemble.h -- batched ensembles: many independent fix eph systems
 * (replicas) advanced together on one device.
 *
 * USER-EPH runs one system per LAMMPS instance; an ensemble of cascades is a
 * loop of separate runs (one FixEPH each, fix_eph.cpp:61-241).  Here every
 * replica keeps its own atoms, seed, T_e grid and energy tally, but all
 * replicas live in one device system and every phase of the MD step is one
 * launch for all of them:
 *   - atoms of all replicas are concatenated (locals first, ghosts after);
 *     ghosts and neighbour lists are built on the device per replica box,
 *     with the same order of images and neighbours a replica gets alone;
 *   - the fix sweeps run over chunks of <= 256 atoms of one replica, so each
 *     replica's noise stream (seed, tag, step), T_e cell lookup and energy
 *     partial sums are those of a separate run;
 *   - the T_e grids are concatenated and advanced by one stencil launch per
 *     sub-step (replicas with fewer sub-steps copy); with temperature-
 *     dependent C_e(T), kappa(T) the counts are found on the device each step
 *     and the host reads the largest before the stencil launches;
 *   - the whole step can be recorded once as a hipGraph and replayed.
 * Neighbour rebuilds are collective: when any atom of any replica has moved
 * more than skin/2, every replica rebuilds (the steps are logged, so a
 * separate run forced to rebuild at the same steps reproduces a replica bit
 * for bit).
 */
#ifndef EPH_ENSEMBLE_H
#define EPH_ENSEMBLE_H

#include "eph_md.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct eph_ensemble eph_ensemble;

int eph_ensemble_create(eph_ensemble** out, int device);
void eph_ensemble_destroy(eph_ensemble* e);
/* message of the last failure on e (or of the last failed create when e is NULL) */
const char* eph_ensemble_last_error(const eph_ensemble* e);

/* Add a replica: the harness system of cfg and a fix built from the fix
 * command words arg (as eph_create takes them) on that system's box.  All
 * replicas must share the .beta file, flags, model, dt and masses; each has
 * its own seed and T_e grid (C_e and kappa constant or temperature-dependent,
 * as its grid file says).  masses has ntypes + 1
 * entries in the layout of eph_set_masses (LAMMPS atom->mass, [0] unused). */
int eph_ensemble_add(eph_ensemble* e, const eph_md_config* cfg, int narg, const char* const* arg,
                     int ntypes, double dt, double boltz, double ftm2v, const double* masses);

/* Concatenate the replicas on the device, build ghosts and lists, compute the
 * initial pair forces.  The first step run is first_step. */
int eph_ensemble_setup(eph_ensemble* e, long long first_step);

/* nsteps Verlet steps of every replica (use_graph: record the step once as a
 * hipGraph and replay it).  *ms receives the device time of the steps;
 * rebuilt[k] (nsteps entries, may be NULL) is 1 when step first + k rebuilt. */
int eph_ensemble_run(eph_ensemble* e, int nsteps, int use_graph, double* ms, int* rebuilt);

/* Forms of the density and force sweeps, numbered as eph_set_sweep_forms
 * numbers them: 0 one chunk of a replica per 256-thread block, tables read
 * through the caches; 1, 2, 3 persistent blocks of 1024, 512, 256 threads with
 * the tables in LDS, walking tiles of 4, 2, 1 chunks; 4, 5, 6 form 0 held to
 * 4, 3, 2 blocks per CU.  Every form gives the same bits. */
#define EPH_ENSEMBLE_FORMS 7

/* Time every form of the two sweeps on the current system and keep the fastest
 * of each for later steps.  variants[2] = chosen forms (density, force);
 * ms[2 * EPH_ENSEMBLE_FORMS] = average time of each density form, then each
 * force form, -1 for a form that cannot run here (may be NULL).  The state of
 * the replicas is unchanged. */
int eph_ensemble_autotune(eph_ensemble* e, int* variants, double* ms);

/* pin the forms instead (0 .. EPH_ENSEMBLE_FORMS - 1 each) */
int eph_ensemble_set_variants(eph_ensemble* e, int density, int force);

/* replicas, total local atoms, total T_e cells */
int eph_ensemble_size(const eph_ensemble* e, int* nrep, long long* natoms, long long* ncells);
/* replica r: local atoms, T_e cells, FDM sub-steps per MD step */
int eph_ensemble_info(const eph_ensemble* e, int r, int* nlocal, int* ncells, int* substeps);

/* which: 0 x, 1 v, 2 f ([nlocal][3]); 3 T_e ([ncells]); 4 (E_local, Ee);
 * 5 the candidate lists at the last density sweep of the whole batch (r only
 * has to be valid): waves that scanned their full lists, waves in all, moved
 * atoms of all replicas, retaken (0 or 1).  Meaningful after a chunk-form
 * sweep; the LDS-table forms always walk the full lists and leave them as
 * they were. */
int eph_ensemble_get(eph_ensemble* e, int r, int which, double* out);

#ifdef __cplusplus
}
#endif

#endif /* EPH_ENSEMBLE_H */
