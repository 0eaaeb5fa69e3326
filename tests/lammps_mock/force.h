// Test harness: the unit constants of the fix (metal units).
#pragma once
#include "pointers.h"
namespace LAMMPS_NS {
class Force {
 public:
  double boltz = 8.617343e-5;
  double ftm2v = 1.0 / 1.0364269e-4;
};
}  // namespace LAMMPS_NS
