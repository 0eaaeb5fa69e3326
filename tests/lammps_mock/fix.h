// Test harness: the Fix base class (hooks and output flags used by fixes).
#pragma once
#include <string>
#include "pointers.h"
namespace LAMMPS_NS {
namespace FixConst {
enum {
  INITIAL_INTEGRATE = 1 << 0,
  POST_INTEGRATE = 1 << 1,
  PRE_EXCHANGE = 1 << 2,
  PRE_NEIGHBOR = 1 << 3,
  POST_NEIGHBOR = 1 << 4,
  PRE_FORCE = 1 << 5,
  PRE_REVERSE = 1 << 6,
  POST_FORCE = 1 << 7,
  FINAL_INTEGRATE = 1 << 8,
  END_OF_STEP = 1 << 9,
};
}  // namespace FixConst
class NeighList;
class Fix : protected Pointers {
 public:
  char *id = nullptr;
  int groupbit = 1;
  int vector_flag = 0, size_vector = 0, global_freq = 0, extvector = 0;
  int peratom_flag = 0, size_peratom_cols = 0, peratom_freq = 0;
  int comm_forward = 0;  // values per atom of pack/unpack_forward_comm
  double **array_atom = nullptr;
  Fix(LAMMPS *lmp, int narg, char **arg) : Pointers(lmp) {
    id_store = narg > 0 ? arg[0] : "";
    id = const_cast<char *>(id_store.c_str());
  }
  ~Fix() override = default;
  virtual int setmask() = 0;
  virtual void init() {}
  virtual void init_list(int, NeighList *) {}
  virtual void reset_dt() {}
  virtual void initial_integrate(int) {}
  virtual void post_force(int) {}
  virtual void final_integrate() {}
  virtual void end_of_step() {}
  virtual double compute_vector(int) { return 0.0; }
  virtual void post_run() {}
  virtual void grow_arrays(int) {}
  virtual void pre_exchange() {}
  virtual void copy_arrays(int, int, int) {}
  virtual int pack_forward_comm(int, int *, double *, int, int *) { return 0; }
  virtual void unpack_forward_comm(int, int, double *) {}

 private:
  std::string id_store;
};
}  // namespace LAMMPS_NS
