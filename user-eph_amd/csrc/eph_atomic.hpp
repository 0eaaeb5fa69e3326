// eph_atomic.hpp -- the `fix eph/atomic` style (fix_eph_atomic.cpp): per-atom
// electronic energies E_i with temperatures T_i = T(E_i) from the .kappa
// file's E(T) table, the PRL friction and random forces with per-atom T, the
// per-atom energy deposit and the neighbour heat diffusion of heat_solve.
#pragma once

#include <string>
#include <vector>

#include "eph_devarray.hpp"
#include "eph_device.hpp"
#include "eph_host.hpp"

struct eph_context;

namespace eph {

struct AtomicView {
  double* E = nullptr;      // [nall]   E_a_i[.][0] (ghosts: copies of their owners)
  double* E1 = nullptr;     // [nlocal] E_a_i[.][1], the diffusion's output
  double* dE = nullptr;     // [nlocal] deposit of the step
  double* rho_a = nullptr;  // [nall]   locality density of the .kappa file
  double* T = nullptr;      // [nall]   T(E)
  double* K = nullptr;      // [nall]   K(T) of the atom's slot
  const double4* krho = nullptr;  // rho(r^2) tables of the .kappa file [n_kelem][k_nr]
  int k_nr = 0;
  double k_inv_dr2 = 0, k_rc2 = 0;
  const double* ET_y = nullptr;   // E(T) per element [n_kelem][nT]
  const double* ET_dy = nullptr;  // forward slopes, [nT - 1] used
  const double* K_y = nullptr;    // K(T) per slot [n_pairs][nT]
  const double* K_dy = nullptr;
  int nT = 0;
  double dT = 1;
  const int* kmap = nullptr;      // kappa element (and K slot) of LAMMPS type t at [t - 1]
};

hipError_t launch_atomic_sync(const AtomsView& a, const WorkView& w, const AtomicView& k,
                              bool with_K, hipStream_t s);
hipError_t launch_atomic_env(const AtomsView& a, const TablesView& t, const StepParams& p,
                             const WorkView& w, const AtomicView& k, hipStream_t s);
hipError_t launch_atomic_ghost(const AtomsView& a, const WorkView& w, const AtomicView& k,
                               hipStream_t s);
hipError_t launch_atomic_forces(const AtomsView& a, const TablesView& t, const StepParams& p,
                                const WorkView& w, const AtomicView& k, hipStream_t s);
hipError_t launch_heat_add(const AtomsView& a, const AtomicView& k, double scaling, hipStream_t s);
hipError_t launch_heat_diffuse(const AtomsView& a, const WorkView& w, const AtomicView& k,
                               double dt, hipStream_t s);
hipError_t launch_heat_copy(const AtomsView& a, const AtomicView& k, hipStream_t s);
// masked staging for the tallies: out[i] = group ? v[i] : 0 (locals)
hipError_t launch_group_values(const AtomsView& a, const double* v, double* out, hipStream_t s);
hipError_t launch_peratom_atomic(const AtomsView& a, const TablesView& t, const WorkView& w,
                                 const AtomicView& k, double* out, hipStream_t s);

// host state of one eph/atomic context
struct AtomicState {
  KappaModel kap;
  int inner_loops = 0;
  double T0 = 0;
  bool init_from_T = true;  // arg[6] == "NULL": E_i = E(T_e) at set-up
  bool ready = false;       // energies set up from the first atoms
  std::vector<int> kmap_host;
  int ngroup = 0;           // local atoms in the group (for the mean T)
  double Ee = 0, Te = 0;    // compute_vector (fix_eph_atomic.cpp:855-864)
  DevArray<double4> krho;
  DevArray<double> ET_y, ET_dy, K_y, K_dy, E, E1, dE, rho_a, T, K, st1, st2, sums;
  DevArray<int> kmap;
  AtomicView view();
};

// style set-up from the fix command (fix_eph_atomic.cpp:59-269); fills ctx
// fields and returns a status, err set on failure
int atomic_parse(eph_context* c, int narg, const char* const* arg, int ntypes, std::string& err);
int atomic_upload(eph_context* c);
int atomic_reserve(eph_context* c, size_t nall);
// energies from the first atoms (host arrays of the locals)
int atomic_setup_energies(eph_context* c, const int* type, const int* mask);
int atomic_post_force(eph_context* c);
int atomic_end_of_step(eph_context* c);

}  // namespace eph
