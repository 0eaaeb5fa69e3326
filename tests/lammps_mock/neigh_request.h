// Test harness: a list request; the cutoff it asks for is recorded.
#pragma once
#include "pointers.h"
namespace LAMMPS_NS {
class NeighRequest {
 public:
  int flags = 0;
  double cutoff = -1.0;
  void set_cutoff(double c) { cutoff = c; }
};
}  // namespace LAMMPS_NS
