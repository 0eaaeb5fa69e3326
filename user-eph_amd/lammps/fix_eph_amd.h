/* fix_eph_amd.h -- LAMMPS adaptor: `fix ID group eph/amd ...` runs the fix eph
 * hot path on an MI355X through the C ABI of eph_amd.h.  Copy this file and
 * fix_eph_amd.cpp into a LAMMPS src/ tree (with include/eph_amd.h on the
 * include path and libeph_amd.so on the link line).  The command words after
 * the style name are those of `fix eph` (fix_eph.cpp:61-241).  The style name
 * without "/amd" is the reference style the context runs: eph and
 * eph/coloured.  The exp-filtered coloured styles and eph/atomic keep per-atom
 * state from step to step; carrying it with migrating atoms needs the
 * grow/copy/pack_exchange callbacks (eph_get/set_filter_state,
 * eph_get/set_atomic_energy), which this adaptor does not implement. */
#ifdef FIX_CLASS
// clang-format off
FixStyle(eph/amd,FixEPHAMD)
FixStyle(eph/coloured/amd,FixEPHAMD)
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

 private:
  eph_context *ctx;
  class NeighList *list;
  int ncols;
  double **array;
  bigint list_step;  // neighbor->lastcall of the list on the device, -1 none
  std::vector<int> off, idx, owner;

  void check(int rc);
  void upload_atoms();
  void upload_list();
};

}  // namespace LAMMPS_NS

#endif
#endif
