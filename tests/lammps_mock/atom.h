// Test harness: per-atom arrays ([n][3] rows), types 1..ntypes, tag -> local map.
#pragma once
#include <unordered_map>
#include "pointers.h"
namespace LAMMPS_NS {
class Atom {
 public:
  enum { GROW = 0, RESTART = 1, BORDER = 2 };
  enum { MAP_NONE = 0, MAP_ARRAY = 1, MAP_HASH = 2 };
  double **x = nullptr, **v = nullptr, **f = nullptr;
  double *mass = nullptr;  // [ntypes + 1], element 0 unused
  int *type = nullptr, *mask = nullptr;
  tagint *tag = nullptr;
  int nlocal = 0, nghost = 0, ntypes = 0, nmax = 0;
  bigint natoms = 0;
  int map_style = MAP_ARRAY;
  std::unordered_map<tagint, int> tag_map;  // tag -> local index
  int map(tagint t) {
    auto it = tag_map.find(t);
    return it == tag_map.end() ? -1 : it->second;
  }
  void add_callback(int) {}
  void delete_callback(const char *, int) {}
};
}  // namespace LAMMPS_NS
