// Test harness: a full neighbour list, LAMMPS layout.
#pragma once
#include "pointers.h"
namespace LAMMPS_NS {
class NeighList {
 public:
  int inum = 0;
  int *ilist = nullptr;
  int *numneigh = nullptr;
  int **firstneigh = nullptr;
};
}  // namespace LAMMPS_NS
