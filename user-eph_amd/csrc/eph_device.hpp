// eph_device.hpp -- device data layout and the kernel launch interface.
//
// Layout in HBM (FP64 throughout):
//   pos[nall]  double4 (x, y, z, element index)   gathered per neighbour
//   vel[nall]  double4 (vx, vy, vz, 0)             gathered in the w sweep
//   pv[nall]   PosVel  (pos, vel), 64 B            gathered in the w sweep
//   rec[nall]  ForceRec (pos, (w, rho), (xi, alpha)), one 128-B line per
//              atom: the force sweep gathers one line per neighbour
//   f, f_EPH, f_RNG: structure of arrays [3][nlocal] (streamed, not gathered)
//   neighbour lists: transposed ELL, entry k of atom i at [k * stride + i],
//     so lane i of a wavefront reads consecutive words (one thread per atom,
//     neighbours summed in list order exactly as the reference loops do).
//   tables: per element, knots of (a, b, c, d) as double4.
//   T_e grid: flat double arrays in i + j nx + k nx ny order (eph_fdm.h:110).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace eph {

struct TablesView {
  int n_elem = 0;
  int n_rho = 0;
  double inv_dr2 = 0;         // 1 / d(r^2)
  const double4* rho_r2 = nullptr;  // [n_rho][n_elem]: knot k of element e at k * n_elem + e
  int n_beta = 0;
  double inv_drho = 0;
  const double4* alpha = nullptr;   // [n_elem][n_beta]
  const double4* beta = nullptr;    // [n_elem][n_beta]
  double rc2 = 0;
  double rho_cut = 0;
  const double4* rho_r = nullptr;   // rho(r) tables [n_elem][n_rho] (model 2)
  double inv_dr = 0;                // 1 / dr
};

struct GridView {
  int nx = 1, ny = 1, nz = 1;
  double x0 = 0, y0 = 0, z0 = 0;
  double dx = 1, dy = 1, dz = 1;
  double dV = 1;
  double* T = nullptr;      // current T_e
  double* Tn = nullptr;     // ping-pong buffer
  double* dT = nullptr;     // deposited power density (dT_e)
  const double* S = nullptr;
  const double* rho = nullptr;
  double* C = nullptr;
  double* kappa = nullptr;
  const int8_t* flag = nullptr;
  const int8_t* tdyn = nullptr;
  // temperature-dependent parameters
  const double4* C_T = nullptr; int n_CT = 0; double inv_dT_C = 0;
  const double4* kappa_T = nullptr; int n_kT = 0; double inv_dT_k = 0;
  const double* E_y = nullptr; const double* E_dy = nullptr; int n_E = 0; double E_dx = 1;
};

struct PosVel;

struct AtomsView {
  int nlocal = 0, nghost = 0;
  int stride = 0;             // ELL column stride (>= nlocal, multiple of 64)
  double4* pos = nullptr;
  double4* vel = nullptr;
  double* f = nullptr;        // [3][fstride]
  int fstride = 0;
  const int* type = nullptr;  // 1-based LAMMPS type
  const int* mask = nullptr;
  const int* tag = nullptr;
  int groupbit = 1;
  // when set, the kernels that move atoms before post_force (initial
  // integrate, ghost update, remap) also keep the packed (position,
  // velocity) records of the w sweep current
  PosVel* pv = nullptr;
};

struct StepParams {
  int flags = 0;
  double eta = 0;        // sqrt(2 k_B / dt)
  double dtv = 0, dtf = 0;
  double dt = 0;
  const double* mass = nullptr;  // [ntypes + 1]
  uint32_t seed = 0;
  uint64_t step = 0;
  int xi_mode = 0;       // 0: zero, 1: injected [nlocal][3], 2: Philox stream
  const double* xi_in = nullptr;
  int model = 4;         // fix eph model: 1 TTM, 2 PRB, 4 PRL, -1 TESTING
  int colour = 0;        // fix style: 0 eph, 1 eph/coloured, 2 .../exp, 3 .../exp/v1
  double zeta = 0;       // exponential filter weight 1 - exp(-dt / tau0) (styles 2, 3)
};

// Per-atom record of the force sweep: position (with element), (w, rho) and
// (xi, alpha) in one 128-B cache line; the fourth slot is unused.
struct ForceRec {
  double4 p, rw, xa, pad;
};
// Position and velocity side by side (64 B), gathered by the w sweep.
struct PosVel {
  double4 p, v;
};

// Work buffers of one post_force / end_of_step
struct WorkView {
  const int* nbr = nullptr;   // full list, ELL
  const int* nnbr = nullptr;
  int* act = nullptr;         // active (r^2 < rc^2) entries, ELL, written by the density sweep
  int* nact = nullptr;
  ForceRec* rec = nullptr;    // written by the density sweep (locals) and the ghost copy
  PosVel* pv = nullptr;       // packed from pos, vel before the density sweep
  double* fE = nullptr;       // [3][fstride]
  double* fR = nullptr;       // [3][fstride]
  int* flags_out = nullptr;   // bit 0: rho > rho_cutoff seen
  const int* ghost_owner = nullptr;
  const double4* ghost_shift = nullptr;
  // filter states of the coloured styles: (f_dis, f_sto) for eph/coloured/exp,
  // (zv, zi) for eph/coloured/exp/v1; .w unused
  double4* cA = nullptr;
  double4* cB = nullptr;
  const double4* vw = nullptr;  // velocities the w sweep uses (vel, or zv)
  // Candidate list of the cooperative density sweep: the entries of the full
  // list with r^2 < rcand2 = (rc + delta)^2 at the positions cref, taken for
  // every local atom in one sweep.  Until the next such sweep, an atom (local
  // or ghost) whose distance from its cref reaches 0.45 delta is "moved"; the
  // check before each sweep lists the current positions of the moved atoms.
  // An entry j left out of atom i's candidates was at least rc + delta away,
  // so it can be within rc now only if i or j moved, and then a moved atom
  // lies within rc of i.  A wave none of whose atoms has a moved atom within
  // `reach` (> rc) walks its candidates: the same active entries in the same
  // order.  Any other wave scans the full list.  More than moved_cap moved
  // atoms retakes the candidates of all atoms.  rcand2 = 0: not used.
  int* cand = nullptr;          // ELL, same stride and width as nbr
  int* ncand = nullptr;
  double4* cref = nullptr;      // [nlocal + nghost]
  double4* moved = nullptr;     // [moved_cap] positions of the moved atoms
  // [0] candidates valid, [1] waves that scanned the full list, [2] moved
  // count; [3], [4] the same two of the last sweep, [5] 1 if it retook them
  int* cstate = nullptr;
  double rcand2 = 0.0;
  double half_delta2 = 0.0;     // (0.45 delta)^2
  double reach = 0.0;           // rc + 0.1 delta
  int moved_cap = 0;
};
// moved atoms listed per step before the candidates of all atoms are retaken
constexpr int kMovedCap = 512;

// Ordered deposit of per-atom values into grid cells (bit-identical to a
// serial loop over atoms): cell of each item, counts, offsets and the
// per-cell item lists.
struct DepositView {
  int* cell = nullptr;        // [n]
  int* count = nullptr;       // [G + 1]
  int* start = nullptr;       // [G + 1]
  int* fill = nullptr;        // [G]
  int* items = nullptr;       // [n]
  double* v1 = nullptr;       // first value per item (friction part)
  double* v2 = nullptr;       // second value per item (random part), may be null
  double* partial = nullptr;  // [2 * nblocks] energy partial sums
  int* scan_tmp = nullptr;    // block sums for the scan
  int* maxcount = nullptr;    // [1]
};

// Per-cell buckets of the deposit for grids with a few atoms per cell
// (launch_deposit_bucket): count and kBucket atom slots per cell, and the
// queue of cells that received more atoms than that.
constexpr int kBucket = 16;
struct BucketView {
  int* count = nullptr;   // [G], zero between steps
  int* items = nullptr;   // [G][kBucket]
  int* queue = nullptr;   // [G]
  int* nqueue = nullptr;  // [1]
};

constexpr int kBlock = 256;
constexpr int kCellConstant = 0, kCellDynamic = 1, kCellZeroDerivative = 2;
constexpr int kMaxSortInCell = 64;
// block size of the persistent LDS-table forms of the ensemble's sweeps
constexpr int kBigBlock = 1024;
// kernel forms of the two neighbour sweeps: 0 reads the rho(r^2) tables
// through the caches; the others stage them in LDS, with 1024, 512 or 256
// threads per block (one block per CU)
// kChunkOcc4/3/2: the cache-table form held to 4, 3 or 2 blocks per CU (its
// LDS request set accordingly), which trades latency hiding for fewer
// neighbour records in flight per L2
constexpr int kLdsForm1024 = 1, kLdsForm512 = 2, kLdsForm256 = 3;
constexpr int kChunkOcc4 = 4, kChunkOcc3 = 5, kChunkOcc2 = 6, kSweepForms = 7;
// rho(r^2) tables up to this size are staged in LDS (160 KB per CU on gfx950)
constexpr size_t kLdsTableMax = 128 * 1024;
constexpr int kLdsMinAtoms = 32768;

// consecutive atoms [start, start + count) of replica `rep`, count <= kBlock
struct Chunk {
  int rep, start, count, pad;
};

// ---- launch wrappers (eph_kernels.hip) ----
hipError_t launch_ghost_xv(const AtomsView& a, const int* owner, const double4* shift, hipStream_t s);
// indices outside [0, nall) set bit 0 of *bad (and are stored as 0); a row
// with more than `width` entries sets bit 1 and is stored empty
hipError_t launch_csr_to_ell(int nlocal, const int* off, const int* idx, int neighmask, int nall,
                             int* ell, int* nn, int stride, int width, int* bad, hipStream_t s);
// form: 0 one block per 256 atoms with tables read from global memory;
// kLdsForm1024/512/256 persistent blocks of that many threads with the
// tables in LDS (when they fit); -1 by size (LDS with 1024 threads for the
// density sweep, 512 for the force sweep)
// ev (optional): two events recorded around the sweep kernel alone (not the
// candidate check before it and the flag update after it)
hipError_t launch_density_w(const AtomsView& a, const TablesView& t, const StepParams& p,
                            const WorkView& w, hipStream_t s, int form = -1, hipEvent_t* ev = nullptr);
// the tables fit in one workgroup's LDS on the current device
bool lds_tables_possible(const TablesView& t);
hipError_t launch_pack_pv(const AtomsView& a, const WorkView& w, hipStream_t s);
hipError_t launch_ghost_rec(const AtomsView& a, const WorkView& w, hipStream_t s);
// multi-rank halo: the locals' (xi, rho, w) rows out, the ghosts' rows in
hipError_t launch_halo_pack(const ForceRec* rec, int n, double* out, hipStream_t s);
hipError_t launch_ghost_halo(const AtomsView& a, const TablesView& t, const WorkView& w,
                             const double* in, hipStream_t s);
hipError_t launch_prl_forces(const AtomsView& a, const TablesView& t, const GridView& g,
                             const StepParams& p, const WorkView& w, hipStream_t s, int form = -1);
hipError_t launch_integrate(const AtomsView& a, const StepParams& p, bool initial, hipStream_t s);
// models 1 (TTM) and 2 (PRB) of fix eph, after the density sweep
hipError_t launch_ttm_forces(const AtomsView& a, const TablesView& t, const GridView& g,
                             const StepParams& p, const WorkView& w, hipStream_t s);
hipError_t launch_prb_forces(const AtomsView& a, const TablesView& t, const GridView& g,
                             const StepParams& p, const WorkView& w, hipStream_t s);
// eph/coloured/exp/v1: zv <- zv (1 - zeta) + zeta v for locals, then ghosts <- owners
hipError_t launch_colour_v(const AtomsView& a, const StepParams& p, const WorkView& w, hipStream_t s);
hipError_t launch_fill_int(int* p, int n, int v, hipStream_t s);
hipError_t launch_fill_double(double* p, long n, double v, hipStream_t s);
// deposit of the fix (energy of friction / random forces at end of step)
hipError_t launch_deposit_atoms(const AtomsView& a, const GridView& g, const StepParams& p,
                                const WorkView& w, const DepositView& d, int nblocks, hipStream_t s);
hipError_t launch_deposit_points(int n, const double* x, const double* y, const double* z,
                                 const double* E, double dt, const GridView& g, const DepositView& d,
                                 hipStream_t s);
// qbuf: 2 + 2 G ints of queue space for cells holding more than
// kMaxSortInCell items.  d.count must hold the per-cell counts.
hipError_t launch_deposit_cells(const GridView& g, const DepositView& d, int n_items, int* qbuf,
                                hipStream_t s);
// the whole deposit and tally through per-cell buckets (fric / rand: which
// parts); with kick, the final velocity-Verlet kick of the group's atoms
// first, in the same pass.  b.count must be zero on entry and is left zero
hipError_t launch_deposit_bucket(const AtomsView& a, const GridView& g, const StepParams& p,
                                 const WorkView& w, const DepositView& d, const BucketView& b,
                                 bool fric, bool rand, bool kick, double* E_local, double* Ee,
                                 hipStream_t s);
hipError_t launch_energy_reduce(const DepositView& d, int nblocks, double* E_local, double* Ee,
                                hipStream_t s);
// grid
hipError_t launch_grid_tdyn_update(const GridView& g, hipStream_t s);
hipError_t launch_grid_minmax(const GridView& g, double* out3, double* scratch, hipStream_t s);
hipError_t launch_fdm_substep(const GridView& g, const double* Tin, double* Tout, double inner_dt,
                              bool any_tdyn, hipStream_t s);
hipError_t launch_grid_lookup(const GridView& g, int n, const double* x, const double* y,
                              const double* z, int* idx, double* T, hipStream_t s);
hipError_t launch_sum(const double* v, int n, double* out, double* scratch, hipStream_t s);
hipError_t launch_nonfinite(const double* v, int n, int* flag, hipStream_t s);
hipError_t launch_peratom(const AtomsView& a, const TablesView& t, const WorkView& w,
                          double* out, hipStream_t s);
// host-layout staging
hipError_t launch_rows_into_d4(const double* rows, int n, double4* out, hipStream_t s);
// whole records: w = species[type - 1] (species given) or 0
hipError_t launch_rows_into_d4_w(const double* rows, int n, const int* type, const int* species,
                                 double4* out, hipStream_t s);
// in[i * stride4] for i < n (stride4 = 4 reads one field of ForceRec)
hipError_t launch_d4_into_rows(const double4* in, int n, double* rows, hipStream_t s, int stride4 = 1);
hipError_t launch_rows_into_soa(const double* rows, int n, double* soa, int stride, hipStream_t s);
hipError_t launch_soa_into_rows(const double* soa, int n, int stride, double* rows, hipStream_t s);
hipError_t launch_d4_component(const double4* in, int n, int comp, double* out, hipStream_t s,
                               int stride4 = 1);
// harness
hipError_t launch_lj(const AtomsView& a, const int* nbr, const int* nn, int stride, double eps,
                     double sigma, double rc2, hipStream_t s);
hipError_t launch_max_disp2(const AtomsView& a, const double4* ref, double* partial, int nblocks,
                            hipStream_t s);
hipError_t launch_copy_probe(double4* dst, const double4* src, long n, hipStream_t s);
// batched replicas (one launch per phase for all replicas)
// variant 0: one chunk per 256-thread block, tables from global memory;
// variant 1: persistent 1024-thread blocks with the tables in LDS (falls back
// to 0 when the tables do not fit).  Both give the same bits.
bool batch_lds_possible(const TablesView& t);
// ng: the device ghost count (the ghost slots of a hold a.nghost = capacity);
// rep: the replica of each of those atoms.  With both the cache-table forms
// walk candidate lists, without them full lists.
hipError_t launch_density_w_batch(const AtomsView& a, const TablesView& t, const StepParams& p,
                                  const WorkView& w, const Chunk* chunks, int nchunks,
                                  const uint32_t* seeds, const uint64_t* step, hipStream_t s,
                                  int variant = 0, const int* ng = nullptr,
                                  const int* rep = nullptr);
hipError_t launch_prl_forces_batch(const AtomsView& a, const TablesView& t, const GridView* grids,
                                   const StepParams& p, const WorkView& w, const Chunk* chunks,
                                   int nchunks, hipStream_t s, int variant = 0);
hipError_t launch_deposit_atoms_batch(const AtomsView& a, const GridView* grids, const int* cell_off,
                                      const StepParams& p, const WorkView& w, const DepositView& d,
                                      const Chunk* chunks, int nchunks, hipStream_t s);
hipError_t launch_energy_reduce_batch(const double* partial, const int* chunk0, const int* nchunks,
                                      int nrep, double* E_local, double* Ee, hipStream_t s);
hipError_t launch_fdm_substep_batch(const GridView* grids, const int* cell_rep, const int* cell_off,
                                    const double* Tin, double* Tout, const double* inner_dt,
                                    const int* steps, int n, int G, bool tdyn, hipStream_t s);
hipError_t launch_grid_tdyn_update_batch(const GridView* grids, const int* cell_rep, const int* cell_off,
                                         int G, hipStream_t s);
hipError_t launch_fdm_substeps_batch(const GridView* grids, const int8_t* rep_tdyn, const double* grid_dt,
                                     const int* grid_steps, int* sub_n, double* inner_dt, int* max_sub,
                                     int R, hipStream_t s);
hipError_t launch_copy_double(const double* src, double* dst, int n, hipStream_t s);
hipError_t launch_ghost_xv_dev(const AtomsView& a, const int* owner, const double4* shift,
                               const int* ng, int cap, hipStream_t s);
hipError_t launch_pack_pv_dev(const AtomsView& a, const WorkView& w, const int* ng, int cap,
                              hipStream_t s);
hipError_t launch_ghost_rec_dev(const AtomsView& a, const WorkView& w, const int* ng, int cap,
                                hipStream_t s);
hipError_t launch_step_advance(uint64_t* step, hipStream_t s);
hipError_t launch_step_log(const int* flag, const uint64_t* step, const uint64_t* first, int* log,
                           int nlog, hipStream_t s);

}  // namespace eph
