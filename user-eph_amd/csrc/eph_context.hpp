// eph_context.hpp -- internal definition of the opaque handles of eph_amd.h.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "eph_device.hpp"
#include "eph_devarray.hpp"
#include "eph_host.hpp"
#include "eph_atomic.hpp"
#include "eph_stage.hpp"

extern "C" {
#include "eph_amd.h"
}

namespace eph {

// Scratch of the ordered deposit (per item and per cell arrays).
struct DepositBuffers {
  DevArray<int> cell, count, start, fill, items, scan_tmp, qbuf;
  DevArray<double> v1, v2, partial;
  hipError_t reserve(size_t n_items, size_t G, size_t nblocks);
  DepositView view(bool with_v2) {
    DepositView d;
    d.cell = cell.p; d.count = count.p; d.start = start.p; d.fill = fill.p; d.items = items.p;
    d.v1 = v1.p; d.v2 = with_v2 ? v2.p : nullptr; d.partial = partial.p; d.scan_tmp = scan_tmp.p;
    return d;
  }
};

// Runs the error-returning HIP call and turns a failure into a status.
#define EPH_HIP(obj, call)                                                        \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) return (obj)->fail_hip(e_, #call);                      \
  } while (0)

}  // namespace eph

struct eph_grid {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string err;
  eph::GridHost h;
  eph::GridView v;
  eph::DevArray<double> T, Tn, dT, S, rho, C, kappa, sum_tmp, minmax, E_y, E_dy, pts;
  eph::DevArray<int8_t> flag, tdyn;
  eph::DevArray<double4> C_T, kappa_T;
  eph::DevArray<int> nonfinite, lookup_idx;
  eph::DepositBuffers dep;
  // per-cell buckets of the deposit (count kept at zero between steps)
  eph::DevArray<int> bcount, bitems, bqueue, bnq;
  int bucket_bytes_ok(size_t G);  // reserves the buckets; 0 when they are ready
  bool any_tdyn = false;
  bool minmax_valid = false;
  double cmin = 0, rmin = 0, kmax = 0;
  double last_r = 0;
  int last_steps = 0;

  int fail(int code, const std::string& m) { err = m; return code; }
  int fail_hip(hipError_t e, const char* what) {
    err = std::string("HIP error ") + hipGetErrorString(e) + " in " + what;
    return e == hipErrorOutOfMemory ? EPH_ERR_OOM : EPH_ERR_HIP;
  }
  int upload();            // host description -> device (all arrays)
  void refresh_minmax();   // host min/max of the stability estimate
  int solve(int* substeps);
  void substeps(double* r, unsigned* steps, double* inner_dt) const;
  int insert_points(int n, const double* x, const double* y, const double* z, const double* E);
  int download(int which, std::vector<double>& out);
  int sync_params_from_device();  // C, kappa after temperature-dependent updates
};

struct eph_context {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // fix parameters (fix_eph.cpp:61-241)
  int seed = 0, flags = 0, model = 0, ntypes = 0, T_freq = 0, groupbit = 1;
  // fix style: 0 eph, 1 eph/coloured, 2 eph/coloured/exp, 3 eph/coloured/exp/v1,
  // 4 eph/atomic (per-atom electronic energies, no T_e grid; state in `at`)
  int colour = 0;
  std::unique_ptr<eph::AtomicState> at;
  bool atomic() const { return colour == 4; }
  double tau0 = 0, zeta = 0;
  std::string T_out, T_state;
  double dt = 0, boltz = 0, ftm2v = 0, eta = 0, dtv = 0, dtf = 0;
  std::vector<int> type_map;
  std::vector<double> mass_host;
  eph::BetaModel beta;
  eph::TablesView tabs;
  eph::DevArray<double4> t_rho_r2, t_alpha, t_beta, t_rho_r;
  eph_grid* grid = nullptr;
  // atoms
  int nlocal = 0, nghost = 0, stride = 0, fstride = 0, maxnbr = 0;
  bool have_atoms = false, have_list = false, have_forces = false;
  eph::DevArray<double4> pos, vel, ghost_shift, stage4;
  // filter states of the coloured styles, kept across steps (and across
  // set_atoms calls: entries beyond the old size start at zero)
  eph::DevArray<double4> colA, colB;
  eph::DevArray<eph::ForceRec> rec;
  eph::DevArray<eph::PosVel> pv;
  // field f of every ForceRec (0 position, 1 (w, rho), 2 (xi, alpha)), stride 4 double4
  const double4* rec_field(int f) const { return reinterpret_cast<const double4*>(rec.p) + f; }
  eph::DevArray<double> f, fE, fR, xi_in, stage, mass, scal;  // scal: [0] E_local [1] Ee
  eph::DevArray<int> type, mask, tag, ghost_owner, nbr, nnbr, act, nact, csr_off, csr_idx,
      flags_dev, list_bad, species_dev;  // species_dev: element of each LAMMPS type
  // candidate list of the cooperative density sweep (WorkView::cand);
  // cand_delta < 0: 0.1 r_cutoff, 0: none
  eph::DevArray<int> cand, ncand, cstate;
  eph::DevArray<double4> cref, moved;
  double cand_delta = -1.0;
  // after any change of the lists, ghosts, types, masks or group
  void drop_candidates();
  int xi_mode_next = 0;  // 1 once eph_set_noise injected values
  // the packed (position, velocity) records match pos and vel: the kernels
  // that move atoms keep them so; anything else that writes pos or vel
  // (uploads, restores) clears this and the next post_force repacks
  bool pv_fresh = false;
  // the multi-rank hooks: the deposit leaves Ee alone (the caller adds the
  // sum over ranks), and only the rank that writes files writes the heat map
  bool defer_tally = false, write_output = true;
  std::vector<double> halo_host;  // [nall][7] staging of the halo hooks
  // kernel forms of the two sweeps (-1 chosen by system size; see eph_autotune)
  int form_density = -1, form_force = -1;
  long long step = 0;
  double* pinned = nullptr;  // small pinned host scratch
  // copies of the host-array hooks (pinned chunks, worker threads), made on
  // first use
  std::unique_ptr<eph::HostStager> stager;
  bool stage_failed = false;  // no pinned memory: the runtime copies directly
  // optional per-kernel timing
  bool timing = false;
  // own_a: a goes back to the pool with this entry (false when an earlier
  // entry's b is this entry's a)
  struct TimedLaunch { const char* name; hipEvent_t a, b; bool own_a = true; };
  std::vector<TimedLaunch> pending;
  std::vector<hipEvent_t> event_pool;
  struct KernelStat { std::string name; double ms = 0; long count = 0; };
  std::vector<KernelStat> stats;

  int fail(int code, const std::string& m) { err = m; return code; }
  int fail_hip(hipError_t e, const char* what) {
    err = std::string("HIP error ") + hipGetErrorString(e) + " in " + what;
    return e == hipErrorOutOfMemory ? EPH_ERR_OOM : EPH_ERR_HIP;
  }
  eph::AtomsView atoms_view();
  eph::StepParams step_params();
  eph::WorkView work_view();
  int reserve_atoms(int nl, int ng);
  int set_atoms(int nl, int ng, const double* x, const double* v, const int* ty, const int* mk,
                const int* tg);
  int set_ghost_map(int ng, const int* owner, const double* shift);
  int set_neighbors(int nl, const int* off, const int* idx, int neighmask);
  int dev_initial_integrate();
  int dev_post_force();
  int dev_post_force_local();   // post_force up to the ghost exchange
  int dev_post_force_ghosts();  // the rest, with the ghosts' halo values in place
  int dev_final_integrate();
  int dev_end_of_step(long long ntimestep);
  int dev_final_integrate_end_of_step(long long ntimestep);
  int end_of_step_after_deposit(long long ntimestep);
  int dev_deposit(bool kick, bool* kicked);
  bool bucket_deposit() const;
  int autotune(int* forms, double* ms);
  int grow_filter_state(size_t nall);
  int check_flags();
  // caller's host memory <-> device buffers on the context's stream: h2d
  // returns once src may be reused, d2h once dst holds the data
  int h2d(void* dst, const void* src, size_t bytes);
  int d2h(void* dst, const void* src, size_t bytes);
  // timing helpers
  hipEvent_t take_event();
  int timed_density(const eph::AtomsView& a, const eph::StepParams& p, const eph::WorkView& w);
  void timed_begin(const char* name, hipEvent_t* a);
  void timed_end(const char* name, hipEvent_t a);
  int resolve_timing();
};
