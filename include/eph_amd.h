/*
 * eph_amd.h -- C ABI of the MI355X-native `fix eph` hot path.
 *
 * One opaque context per fix instance.  The context owns every device buffer
 * (atoms, neighbour list, spline tables, T_e grid, scratch) and a HIP stream;
 * the caller owns host arrays.  Every call returns an int status (0 = ok,
 * negative = enum eph_status) and never aborts the host; the message of the
 * last failure is available from eph_last_error().
 *
 * Two ways to drive the hooks:
 *   host arrays  -- eph_initial_integrate / eph_post_force / eph_final_integrate /
 *                   eph_end_of_step take LAMMPS-layout host arrays ([n][3]
 *                   doubles), upload, run the kernels and download, like the
 *                   reference GPU path (fix_eph_gpu.cpp:91-205).
 *   device state -- eph_dev_* run the same kernels on the atoms already
 *                   resident in the context (used by the device-resident MD
 *                   harness of eph_md.h and by the benchmark).
 *
 * References are USER-EPH file:line.
 */
#ifndef EPH_AMD_H
#define EPH_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct eph_context eph_context;
typedef struct eph_grid eph_grid;

enum eph_status {
  EPH_OK = 0,
  EPH_ERR_ARGS = -1,        /* bad argument / illegal fix command            */
  EPH_ERR_FILE = -2,        /* file missing or unreadable                    */
  EPH_ERR_FORMAT = -3,      /* malformed .beta / grid / parameter file       */
  EPH_ERR_OOM = -4,         /* device allocation failed                      */
  EPH_ERR_HIP = -5,         /* HIP runtime error                             */
  EPH_ERR_STATE = -6,       /* call out of order (e.g. no atoms uploaded)    */
  EPH_ERR_NONFINITE = -7,   /* T_e grid became non-finite (FDM blow-up)      */
  EPH_ERR_UNSUPPORTED = -8, /* model 3 (PRLCM), or a feature outside a style  */
  EPH_ERR_NO_DEVICE = -9    /* no HIP device                                 */
};

/* eph_flag bits, fix_eph.h:43-50 */
enum eph_flag {
  EPH_FRICTION = 0x01, EPH_RANDOM = 0x02, EPH_FDM = 0x04,
  EPH_NOINT = 0x08, EPH_NOFRICTION = 0x10, EPH_NORANDOM = 0x20
};

/* hooks a host must call, fix_eph.cpp:293-302 (FixEPH::setmask) */
enum eph_hook {
  EPH_HOOK_INITIAL_INTEGRATE = 1, EPH_HOOK_POST_FORCE = 2,
  EPH_HOOK_FINAL_INTEGRATE = 4, EPH_HOOK_END_OF_STEP = 8
};

/* ---------------------------------------------------------------- lifetime */

/* FixEPH::FixEPH, fix_eph.cpp:61-241.  arg/narg are the fix command words
 * exactly as LAMMPS passes them (arg[0] = fix ID ... arg[17+] = element per
 * type).  boxlo/boxhi = domain->boxlo/boxhi, dt = update->dt, boltz and
 * ftm2v = force->boltz / force->ftm2v, ntypes = atom->ntypes.  device = HIP
 * device ordinal.
 * arg[2] selects the style: "eph" (models 1 TTM, 2 PRB, 4 PRL, -1 TESTING in
 * arg[5]; model 3 is rejected: the reference indexes its rho table by
 * jtype - i, fix_eph.cpp:600), "eph/coloured" (fix_eph_coloured.cpp: PRL),
 * "eph/coloured/exp" (fix_eph_coloured_exp.cpp: arg[5] = tau0, the applied
 * f_EPH and f_RNG are exponentially filtered) and "eph/coloured/exp/v1"
 * (fix_eph_coloured_exp_v1.cpp: arg[5] = tau0, the w sum sees filtered
 * velocities and the random sum filtered noise).  "eph/atomic"
 * (fix_eph_atomic.cpp) has its own layout: ID group eph/atomic seed flags T_e
 * T_file inner_loops output beta_file kappa_file element... (narg >= 12);
 * per-atom electronic energies replace the T_e grid (flags: 1 friction,
 * 2 random, 4 heat diffusion, 8 no integration, 16/32 forces not applied),
 * and boxlo/boxhi are not used. */
int eph_create(eph_context** out, int narg, const char* const* arg, int ntypes,
               const double boxlo[3], const double boxhi[3], double dt, double boltz,
               double ftm2v, int device);
/* FixEPH::~FixEPH, fix_eph.cpp:244-261 */
void eph_destroy(eph_context* ctx);
/* message of the last failure on ctx (or of the last failed eph_create when ctx is NULL) */
const char* eph_last_error(const eph_context* ctx);

/* FixEPH::init, fix_eph.cpp:263-287: rejects 2-d, non-periodic and triclinic
 * boxes; *cutoff receives the r_cutoff the neighbour request must use. */
int eph_init(eph_context* ctx, int dimension, int nonperiodic, int triclinic, double* cutoff);
/* FixEPH::setmask, fix_eph.cpp:293-302: returns the eph_hook bits */
int eph_setmask(void);
/* FixEPH::reset_dt, fix_eph.cpp:909-916 (and zeta, fix_eph_coloured_exp.cpp:683-690) */
int eph_reset_dt(eph_context* ctx, double dt);
/* Filter state of the coloured styles, [nlocal][3] by local index: which 0 is
 * f_dis (eph/coloured/exp) or zv (.../v1), which 1 is f_sto or zi.  The
 * reference carries these with the atoms (pack_exchange / copy_arrays of
 * fix_eph_coloured_exp.cpp); a caller that reorders atoms reads them before
 * and writes them back after eph_set_atoms. */
int eph_get_filter_state(eph_context* ctx, int which, double* out);
int eph_set_filter_state(eph_context* ctx, int which, const double* in);
/* atom->mass[1..ntypes] (element 0 unused, as in LAMMPS) */
int eph_set_masses(eph_context* ctx, const double* mass, int ntypes);
/* Fix::groupbit; atoms with (mask & groupbit) belong to the fix group */
int eph_set_groupbit(eph_context* ctx, int groupbit);
/* flags, model, seed, r_cutoff, rho_cutoff, eta_factor, dtf (read-only view) */
int eph_get_params(const eph_context* ctx, int* flags, int* model, int* seed, double* r_cutoff,
                   double* rho_cutoff, double* eta_factor, double* dtf);

/* ------------------------------------------------------------------- atoms */

/* atom->x, atom->v ([nall][3]), atom->type (1-based), atom->mask, atom->tag
 * for nall = nlocal + nghost; replaces cpu_to_device_EPH_GPU of
 * lib/eph_gpu.h:75 and the host staging of fix_eph_gpu.cpp:110-112. */
int eph_set_atoms(eph_context* ctx, int nlocal, int nghost, const double* x, const double* v,
                  const int* type, const int* mask, const int* tag);
/* On one process the forward communication of fix_eph.cpp:951-1009 is
 * ghost <- owner: owner[g] is the local index whose image is ghost nlocal+g,
 * shift[g][3] the periodic offset (x_ghost = x_owner + shift). */
int eph_set_ghost_map(eph_context* ctx, int nghost, const int* owner, const double* shift);
/* NeighList numneigh/firstneigh (full list incl. ghosts) in CSR form:
 * neighbours of i are idx[off[i] .. off[i+1]), each masked with neighmask
 * (NEIGHMASK) as fix_eph.cpp:451-452 does.  Replaces transfer_neighbour_list
 * of fix_eph_gpu.cpp:241-280 / EPH_GPU::grow_neigh of lib/eph_gpu.h:67.
 * A masked index outside [0, nlocal + nghost) returns EPH_ERR_ARGS and drops
 * the list: post_force then returns EPH_ERR_STATE until a valid list is set. */
int eph_set_neighbors(eph_context* ctx, int nlocal, const int* off, const int* idx, int neighmask);
/* Inject the Gaussian noise xi ([nlocal][3]) used by the next post_force for
 * group atoms, instead of drawing it; NULL returns to the built-in stream.
 * Replaces the RanMars draws of fix_eph.cpp:854-862. */
int eph_set_noise(eph_context* ctx, const double* xi);
/* the built-in stream draws xi_i = N(0,1)^3 from Philox4x32-10 keyed by
 * (seed, tag_i, ntimestep); this sets ntimestep for the next post_force. */
int eph_set_timestep(eph_context* ctx, long long ntimestep);

/* ------------------------------------------------------- hooks, host arrays */

/* FixEPH::initial_integrate, fix_eph.cpp:305-328; x, v updated in place */
int eph_initial_integrate(eph_context* ctx, double* x, double* v, const double* f);
/* FixEPH::post_force, fix_eph.cpp:841-907 (+ calculate_environment :431-466,
 * force_prl :687-837); x, v are the current [nall][3] arrays, f [nlocal][3]
 * is updated in place (f += f_EPH, f += f_RNG per the flags). */
int eph_post_force(eph_context* ctx, const double* x, const double* v, double* f);
/* post_force of a run over several MPI ranks, split where FixEPH exchanges
 * ghost values (comm->forward_comm with states XI, RHO, WI, fix_eph.cpp:863-871
 * and pack/unpack_forward_comm :951-1009).  Those exchanges feed only force_prl's
 * stages 2 and 3, so one exchange of EPH_HALO_COLS values per atom carries all
 * three: eph_post_force_local draws xi and computes rho, alpha and w of the local
 * atoms and writes halo[nlocal][EPH_HALO_COLS] = (xi_x, xi_y, xi_z, rho, w_x,
 * w_y, w_z); the caller sends the rows of the owners to their ghosts, and
 * eph_post_force_ghosts takes halo_ghost[nghost][EPH_HALO_COLS] in ghost order,
 * finishes the forces and updates f[nlocal][3].  x, v are [nall][3].  Styles eph,
 * eph/coloured and eph/coloured/exp with model 4 (else EPH_ERR_UNSUPPORTED). */
#define EPH_HALO_COLS 7
int eph_post_force_local(eph_context* ctx, const double* x, const double* v, double* halo);
int eph_post_force_ghosts(eph_context* ctx, const double* halo_ghost, double* f);
/* end_of_step of a run over several ranks, split at EPH_FDM's sync_before
 * (eph_fdm.h:479-491) and the tally's MPI_Allreduce (fix_eph.cpp:402): the first
 * call deposits this rank's atoms and returns the grid's dT_e[G] and this rank's
 * E_local; the caller sums both over the ranks and passes the sums to the second,
 * which solves the grid (every rank the same bits, so no broadcast is needed)
 * and adds the energy to Ee.  write_output: write the heat map (one rank). */
int eph_end_of_step_local(eph_context* ctx, const double* x, const double* v, double* dT,
                          double* E_local);
int eph_end_of_step_reduced(eph_context* ctx, const double* dT, double E, long long ntimestep,
                            int write_output);
/* FixEPH::final_integrate, fix_eph.cpp:330-348 */
int eph_final_integrate(eph_context* ctx, double* v, const double* f);
/* FixEPH::end_of_step, fix_eph.cpp:350-429 (+ EPH_FDM::solve, eph_fdm.h:267-400) */
int eph_end_of_step(eph_context* ctx, const double* x, const double* v, long long ntimestep);
/* FixEPH::post_run, fix_eph.cpp:1019-1021: writes the T_e state file */
int eph_post_run(eph_context* ctx);
/* Full-precision binary checkpoint of the fix's own state: the T_e grid, the
 * energy tally, the last step and the per-atom state of the styles that keep
 * one (the filter states of eph/coloured/exp*, E_i of eph/atomic), by atom tag.
 * No counterpart in the reference, whose T.restart (eph_post_run) keeps six
 * digits of T_e only; a context created with the same command, given the same
 * atoms and list, that loads this file goes on bit for bit.  Load after
 * eph_set_atoms (every local tag must be in the file); the file must come from
 * the same style, model, flags and grid (EPH_ERR_FORMAT otherwise).  The whole
 * file is read and checked before the context changes: a refused file leaves
 * it as it was. */
int eph_checkpoint_save(eph_context* ctx, const char* path);
int eph_checkpoint_load(eph_context* ctx, const char* path);
/* FixEPH::compute_vector, fix_eph.cpp:940-948: 0 -> Ee, 1 -> mean T_e
 * (eph/atomic, fix_eph_atomic.cpp:855-864: sum and mean over the group's atoms) */
int eph_compute_vector(eph_context* ctx, int i, double* out);
/* per-atom array of fix_eph.cpp:406-428: [nlocal][8] = rho, beta(rho), f_EPH, f_RNG;
 * eph/atomic (fix_eph_atomic.cpp:416-448): [nlocal][12] = the same, then rho_a,
 * E, dE, T */
int eph_get_peratom(eph_context* ctx, double* array);
/* columns of that array: 8, or 12 for eph/atomic */
int eph_peratom_cols(const eph_context* ctx);
/* eph/atomic: the per-atom electronic energies E_i, [nlocal] by local index.
 * The reference carries them with the atoms (pack_exchange / copy_arrays,
 * fix_eph_atomic.cpp:950-964); a caller that reorders atoms reads them before
 * and writes them back after eph_set_atoms.  Setting them also replaces the
 * set-up from T_e that the first eph_set_atoms performs. */
int eph_get_atomic_energy(eph_context* ctx, double* E);
int eph_set_atomic_energy(eph_context* ctx, const double* E);
/* internal per-atom state after post_force, for checking:
 * which 0 rho [nall], 1 w [nall][3], 2 xi [nall][3], 3 f_EPH [nlocal][3],
 * 4 f_RNG [nlocal][3], 5 alpha [nall], 6 active neighbours (r < r_cutoff)
 * per local atom [nlocal], 7 list entries per local atom [nlocal];
 * eph/atomic also 8 rho_a [nall], 9 E [nall], 10 dE [nlocal], 11 T [nall];
 * 12: 4 values about the density sweep's candidate list at its last run:
 * waves that scanned the full list, waves in all, atoms found moved, 1 if the
 * candidates of all atoms were retaken */
int eph_get_state(eph_context* ctx, int which, double* out);
/* energy deposited in the last end_of_step (E_local of fix_eph.cpp:357-402) */
int eph_last_energy(eph_context* ctx, double* E_local);
/* non-zero once get_beta/get_alpha saw rho > rho_cutoff (eph_beta.h:174-194) */
int eph_rho_cutoff_warning(eph_context* ctx);

/* ---------------------------------------------------- hooks, device-resident */

int eph_dev_initial_integrate(eph_context* ctx);
int eph_dev_post_force(eph_context* ctx);
int eph_dev_final_integrate(eph_context* ctx);
int eph_dev_end_of_step(eph_context* ctx, long long ntimestep);
/* copy resident x, v ([nall][3]) / f ([nlocal][3]) back to the host */
int eph_dev_get_atoms(eph_context* ctx, double* x, double* v, double* f);
/* wait for all work queued on the context's stream */
int eph_sync(eph_context* ctx);
/* the context's HIP stream (hipStream_t), for timing with HIP events */
void* eph_stream(eph_context* ctx);

/* Kernel forms of the two neighbour sweeps (no counterpart in the reference;
 * every form gives the same bits).  0: one 256-thread block per 256 atoms,
 * spline tables read through the caches (the force sweep's waves fetch their
 * neighbours' records together, 8 lanes per record, through LDS); 1, 2, 3:
 * persistent blocks of 1024, 512 or 256 threads (one per CU: 4, 2 or 1 waves
 * per SIMD) with the rho(r^2) tables staged in LDS (only when they fit); -1:
 * chosen by system size (the default: form 1 for the density sweep of large
 * systems, form 0 otherwise).  eph_autotune times every form of each sweep on
 * the resident atoms (after eph_dev_post_force or the harness's attach) and
 * keeps the fastest; forms[2] receives the choice (density, force) and
 * ms[2 * EPH_SWEEP_FORMS] the times (density forms, then force forms; 0 for
 * a form the device cannot run).  f is not changed.  Forms 4, 5, 6 are form 0
 * held to 4, 3 or 2 blocks per CU (fewer neighbour records in flight per L2,
 * less latency hidden). */
/* Candidate list of the density sweep (no counterpart in the reference; the
 * results are the same bits either way).  In form 0 (and 4-6) the rho loop of a
 * wave walks the entries of the full list that were within r_cutoff + delta
 * when the candidates were last taken, unless an atom (local or ghost) that
 * has moved 0.45 delta since then lies within r_cutoff + 0.1 delta of the
 * wave's atoms; such waves scan the full list.  Any change of the lists,
 * ghost map, atoms, types, masks or group retakes the candidates of all atoms
 * in the next sweep, and so do more than 512 moved atoms.  Ghost positions
 * uploaded at every hook are checked as they are (a box that changes between
 * list uploads is allowed).  delta < 0: 0.1 r_cutoff (the default), 0: always
 * the full list. */
int eph_set_candidate_margin(eph_context* ctx, double delta);

#define EPH_SWEEP_FORMS 7
int eph_set_sweep_forms(eph_context* ctx, int density, int force);
int eph_autotune(eph_context* ctx, int* forms, double* ms);

/* Per-kernel timing with HIP events on the context's stream (no counterpart
 * in the reference; used to measure kernel durations).  eph_timing_report with
 * k < 0 returns the number of kernels recorded; with k >= 0 it fills the name,
 * summed milliseconds and launch count of kernel k. */
int eph_timing_enable(eph_context* ctx, int on);
int eph_timing_report(eph_context* ctx, int k, char* name, int namelen, double* total_ms,
                      long long* count);

/* Plain device-to-device copy of `bytes` bytes (double4 loads/stores), run
 * `reps` times on a fresh buffer pair; *ms receives the average kernel time.
 * No counterpart in the reference: it calibrates the HBM byte counters of a
 * profiled run and measures the achievable copy bandwidth. */
int eph_copy_probe(long long bytes, int reps, int device, double* ms);

/* ------------------------------------------------ host-side table checks */
/* Host construction only, no device needed.
 * EPH_Spline(dx, y) of eph_spline.h:31-132: coeffs [n][4] = a, b, c, d of the
 * global-x cubic per knot. */
int eph_spline_build(double dx, const double* y, int n, double* coeffs);
/* EPH_Beta(file) of eph_beta.h:39-128: table `which` (0 rho(r), 1 rho(r^2),
 * 2 beta(rho), 3 alpha(rho)) of element `elem`.  Returns the knot count (or a
 * negative status); coeffs [n][4] is filled when non-NULL. */
int eph_beta_table(const char* path, int elem, int which, double* coeffs, double* inv_dx,
                   double* r_cutoff, double* rho_cutoff);

/* EPH_kappa(file) of eph_kappa.h:55-150 (fix eph/atomic): element count, the
 * reference's pair-slot count (n + 1)(n - 1) / 2 (1 for one element), r_cutoff
 * and T_max. */
int eph_kappa_info(const char* path, int* n_elem, int* n_pairs, double* r_cutoff, double* T_max);
/* table `which` of a .kappa file: 0 rho(r), 1 rho(r^2) of element `index`
 * (out [n][4] coefficients, *dx = 1 / knot spacing); 2 E(T) of element
 * `index`, 3 K(T) of pair slot `index` (out [n][2] = y, forward slope with 0
 * at the top knot; *dx = knot spacing).  Returns n or a negative status. */
int eph_kappa_table(const char* path, int index, int which, double* out, double* dx);

/* --------------------------------------------------------------- T_e grid */
/* EPH_FDM counterpart (eph_fdm.h); also usable standalone. */

/* EPH_FDM(nx, ny, nz, box, T_e, C_e, rho_e, kappa_e), eph_fdm.h:28-46 */
int eph_grid_create(eph_grid** out, int nx, int ny, int nz, const double box[6], double T_e,
                    double C_e, double rho_e, double kappa_e, int device);
/* EPH_FDM(const char*), eph_fdm.h:48-119 (grid file + optional parameter file) */
int eph_grid_load(eph_grid** out, const char* path, int device);
void eph_grid_destroy(eph_grid* g);
const char* eph_grid_last_error(const eph_grid* g);
/* the grid owned by a fix context (not to be destroyed by the caller) */
eph_grid* eph_context_grid(eph_context* ctx);
int eph_grid_set_dt(eph_grid* g, double dt);          /* eph_fdm.h:154-157 */
int eph_grid_set_steps(eph_grid* g, int steps);       /* eph_fdm.h:159-162 */
/* one cell's T, S, rho_e, C_e, kappa, flag (0 constant, 1 dynamic, 2 zero
 * derivative) and T-dependence flag, as a grid-file line sets them */
int eph_grid_set_cell(eph_grid* g, int i, int j, int k, double T, double S, double rho,
                      double C, double kappa, int flag, int tdyn);
/* EPH_FDM::insert_energy (eph_fdm.h:172-179) for n points, in array order */
int eph_grid_insert_energy(eph_grid* g, int n, const double* x, const double* y,
                           const double* z, const double* E);
/* EPH_FDM::solve, eph_fdm.h:267-400; *substeps receives the sub-step count */
int eph_grid_solve(eph_grid* g, int* substeps);
/* which: 0 T, 1 dT (deposit), 2 C_e, 3 rho_e, 4 kappa, 5 S */
int eph_grid_get(eph_grid* g, int which, double* out);
int eph_grid_set(eph_grid* g, int which, const double* in);
/* EPH_FDM::get_index / get_T for n points (eph_fdm.h:182-187, 494-509) */
int eph_grid_lookup(eph_grid* g, int n, const double* x, const double* y, const double* z,
                    int* index, double* T);
int eph_grid_info(const eph_grid* g, int* dims, double* box, double* last_r);
int eph_grid_T_mean(eph_grid* g, double* out);        /* eph_fdm.h:189-196 */
int eph_grid_save_state(eph_grid* g, const char* path);                  /* :226-265 */
int eph_grid_save_temperature(eph_grid* g, const char* base, int n);     /* :198-224 */

/* No counterpart in the reference.  eph_grid_insert_energy for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): x, y, z, E are n values each in
 * device memory; the deposit's scratch is one fixed array per GPU of
 * eph_noalloc_scratch_bytes() bytes (20 MiB), enough for 16 n + 20 G bytes
 * plus about 2 KB (G = grid cells); larger inputs return EPH_ERR_ARGS.  The
 * work is enqueued on the grid's stream and the call returns without waiting
 * (eph_grid_get and eph_grid_solve order after it).  All calls on one GPU
 * share that scratch: calls that overlap in time on one GPU race. */
int eph_grid_insert_energy_noalloc(eph_grid* g, int n, const double* x, const double* y,
                                   const double* z, const double* E);
long long eph_noalloc_scratch_bytes(void);

#ifdef __cplusplus
}
#endif

#endif /* EPH_AMD_H */
