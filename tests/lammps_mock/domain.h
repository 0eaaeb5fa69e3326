// Test harness: an orthogonal periodic box.
#pragma once
#include "pointers.h"
namespace LAMMPS_NS {
class Domain {
 public:
  double boxlo[3] = {0, 0, 0}, boxhi[3] = {0, 0, 0};
  int dimension = 3, nonperiodic = 0, triclinic = 0;
};
}  // namespace LAMMPS_NS
