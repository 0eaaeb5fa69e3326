/* fix_eph_amd.h -- LAMMPS adaptor: `fix ID group eph/amd ...` runs the fix eph
 * hot path on an MI355X through the C ABI of eph_amd.h.  Copy this file and
 * fix_eph_amd.cpp into a LAMMPS src/ tree (with include/eph_amd.h on the
 * include path and libeph_amd.so on the link line).  The style name without
 * "/amd" is the reference style the context runs, with that style's command
 * words: eph, eph/coloured, eph/coloured/exp, eph/coloured/exp/v1
 * (fix_eph.cpp:61-241, fix_eph_coloured_exp.cpp:43) and eph/atomic
 * (fix_eph_atomic.cpp:59-269).  The exp-filtered coloured styles and
 * eph/atomic keep per-atom state from step to step (filter states, electronic
 * energies); the adaptor holds it in LAMMPS per-atom rows while atoms are
 * sorted or exchanged, as the reference fixes' grow/copy callbacks do.  Over
 * several MPI ranks (styles eph, eph/coloured, eph/coloured/exp, model 4) each
 * rank's context runs its own atoms; the ghosts' xi, rho and w come through
 * comm->forward_comm and the grid's deposits and the tally through
 * MPI_Allreduce, as FixEPH and EPH_FDM exchange them. */
#ifdef FIX_CLASS
// clang-format off
FixStyle(eph/amd,FixEPHAMD)
FixStyle(eph/coloured/amd,FixEPHAMD)
FixStyle(eph/coloured/exp/amd,FixEPHAMD)
FixStyle(eph/coloured/exp/v1/amd,FixEPHAMD)
FixStyle(eph/atomic/amd,FixEPHAMD)
// clang-format on
#else

#ifndef LMP_FIX_EPH_AMD_H
#define LMP_FIX_EPH_AMD_H

#include <vector>

#include "fix.h"

struct eph_context;

namespace LAMMPS_NS {

class FixEPHAMD : public Fix {
 public:
  FixEPHAMD(class LAMMPS *, int, char **);
  ~FixEPHAMD() override;
  int setmask() override;
  void init() override;
  void init_list(int, class NeighList *) override;
  void reset_dt() override;
  void initial_integrate(int) override;
  void post_force(int) override;
  void final_integrate() override;
  void end_of_step() override;
  double compute_vector(int) override;
  void post_run() override;
  void grow_arrays(int) override;
  void pre_exchange() override;
  void copy_arrays(int, int, int) override;
  int pack_forward_comm(int, int *, double *, int, int *) override;
  void unpack_forward_comm(int, int, double *) override;

 private:
  eph_context *ctx;
  class NeighList *list;
  int ncols;
  double **array;
  bigint list_step;  // neighbor->lastcall of the list on the device, -1 none
  std::vector<int> off, idx, owner;
  // per-atom state carried by LAMMPS while atoms move: 6 columns (the two
  // filter states of eph/coloured/exp*), 1 (E_i of eph/atomic) or none
  int nstate;
  double **state;
  bool state_on_host;  // the rows hold the context's state (since pre_exchange)
  std::vector<double> buf;
  // several ranks: the halo rows of the locals and the ghosts (EPH_HALO_COLS
  // each), the grid's deposits being summed, and its size
  bool multi;
  std::vector<double> halo, halo_ghost, dT;
  void state_to_host();
  void state_to_device();

  void check(int rc);
  void upload_atoms();
  void upload_list();
};

}  // namespace LAMMPS_NS

#endif
#endif
