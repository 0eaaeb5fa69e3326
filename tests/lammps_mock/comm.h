// Test harness: rank count and the ghost-velocity flag.
#pragma once
#include "pointers.h"
namespace LAMMPS_NS {
class Comm {
 public:
  int me = 0, nprocs = 1;
  int ghost_velocity = 0;
};
}  // namespace LAMMPS_NS
