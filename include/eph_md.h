/*
 * eph_md.h -- minimal molecular-dynamics harness standing in for the LAMMPS
 * pieces FixEPH relies on (no counterpart in USER-EPH itself):
 *   create_atoms fcc + velocity create  -> eph_md_create_fcc
 *   periodic ghost images (Comm::borders, x_ghost = x_owner + shift)
 *   full neighbour list with ghosts at r_cutoff + skin (NeighList); the fix's
 *   list in distance order (squared distance quantized at a power-of-two
 *   scale, ties in stencil order), host and device builds alike
 *   pair lj/cut (Newton off, full list)
 *   neighbor check yes: rebuild when an atom moved more than skin / 2
 *   Verlet step order: initial_integrate, comm/neighbor, pair, post_force,
 *   final_integrate, end_of_step
 * Host arrays are LAMMPS-layout ([n][3] doubles, 1-based types, tags 1..N).
 */
#ifndef EPH_MD_H
#define EPH_MD_H

#include "eph_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct eph_md eph_md;

typedef struct eph_md_config {
  int cells;              /* fcc cells per box edge                            */
  double a;               /* lattice constant (A)                              */
  int ntypes;             /* species, assigned equiatomic at random if > 1     */
  unsigned int seed;      /* seeds species, velocities, displacements, PKA dir */
  double T;               /* Maxwell temperature (K), momentum zeroed, rescaled */
  double disp_sigma;      /* Gaussian lattice displacement (A), 0 = perfect    */
  double pka_eV;          /* kinetic energy of the atom nearest the centre     */
  double mass;            /* g/mol, all species                                */
  double cut_fix;         /* fix neighbour cutoff (r_cutoff of the .beta)      */
  double cut_lj;          /* pair cutoff                                       */
  double skin;            /* neighbour skin (A)                                */
  double lj_eps, lj_sigma;
} eph_md_config;

int eph_md_create_fcc(eph_md** out, const eph_md_config* cfg);
void eph_md_destroy(eph_md* md);
const char* eph_md_last_error(const eph_md* md);

/* sizes: nlocal, nghost, fix-list entries, lj-list entries; box lo/hi */
int eph_md_info(const eph_md* md, int* nlocal, int* nghost, long long* nfix, long long* nlj,
                double* boxlo, double* boxhi);
/* copies of the host state: x, v [nall][3]; type, mask, tag [nall];
 * ghost owner [nghost], shift [nghost][3]; fix list CSR off [nlocal+1], idx */
int eph_md_get(const eph_md* md, double* x, double* v, int* type, int* mask, int* tag,
               int* ghost_owner, double* ghost_shift, int* fix_off, int* fix_idx);

/* host-side LAMMPS stand-ins (for driving a fix through host arrays):
 * x_ghost = x_owner + shift, v_ghost = v_owner for arrays of length nall */
int eph_md_update_ghosts(const eph_md* md, double* x, double* v);
/* pair lj/cut over the lj list: f[nlocal][3] = sum of pair forces */
int eph_md_lj_host(const eph_md* md, const double* x, double* f);
/* 1 if an atom moved more than skin/2 since the last build */
int eph_md_needs_rebuild(const eph_md* md, const double* x);
/* remap locals into the box, rebuild ghosts and both lists from x[nlocal][3],
 * v[nlocal][3] (ghost entries of the caller's arrays are ignored) */
int eph_md_rebuild(eph_md* md, const double* x, const double* v);

/* ---- device-resident run (the benchmark path) ---- */
/* upload the system into a fix context (atoms, ghost map, fix list) */
int eph_md_attach(eph_md* md, eph_context* ctx);
/* nsteps full Verlet steps on the device: fix hooks + ghost update + LJ pair;
 * rebuilds go through the host when the skin check trips.  Timings in ms,
 * summed over the steps, are returned when the pointers are non-NULL:
 *   t[0] fix hooks (initial_integrate + post_force + final_integrate +
 *        end_of_step), t[1] whole steps, t[2] pair + ghost update,
 *   t[3] host rebuilds (count in *nrebuild). */
int eph_md_run(eph_md* md, eph_context* ctx, int nsteps, long long first_step, double* t,
               int* nrebuild);
/* rebuilds on the device (1, default; falls back to the host when a capacity
 * is exceeded) or through the host (0); both give the same ghosts and lists */
int eph_md_set_device_rebuild(eph_md* md, int on);
/* rebuild at the next step whatever the displacements */
int eph_md_force_rebuild(eph_md* md);
/* copy the device-resident system (atoms, ghost map, both lists) back into the
 * harness's host arrays, so eph_md_info / eph_md_get describe it */
int eph_md_pull(eph_md* md, eph_context* ctx);
/* Full-precision binary checkpoint of a device-resident run: atoms, forces,
 * ghosts, both lists, the rebuild reference positions, the density sweep's
 * candidate lists and the fix's state (eph_checkpoint_save).  Load into a
 * harness made with the same configuration and a fix made with the same
 * command, both attached; the next eph_md_run then goes on bit for bit, down
 * to which waves walk their candidates.  path NULL: the harness keeps the
 * checkpoint in host memory instead (one at a time; load with NULL restores it,
 * also into the same harness and fix, which replays a segment of a run).
 * Files of another system or another fix are refused before anything changes;
 * a file damaged past its headers can leave the harness partly restored.
 * Errors are reported through eph_last_error(ctx). */
int eph_md_checkpoint_save(eph_md* md, eph_context* ctx, const char* path);
int eph_md_checkpoint_load(eph_md* md, eph_context* ctx, const char* path);

#ifdef __cplusplus
}
#endif

#endif /* EPH_MD_H */
