// Test harness: timestep size and counter.
#pragma once
#include "pointers.h"
namespace LAMMPS_NS {
class Update {
 public:
  double dt = 0.001;
  bigint ntimestep = 0;
};
}  // namespace LAMMPS_NS
