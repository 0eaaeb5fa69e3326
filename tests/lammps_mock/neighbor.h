// Test harness: list requests and the step of the last list build.
#pragma once
#include "neigh_request.h"
#include "pointers.h"
namespace LAMMPS_NS {
namespace NeighConst {
enum { REQ_DEFAULT = 0, REQ_FULL = 1 << 0, REQ_GHOST = 1 << 1 };
}
class Fix;
class Neighbor {
 public:
  NeighRequest request;
  bigint lastcall = -1;
  int ago = -1;
  NeighRequest *add_request(Fix *, int flags) {
    request.flags = flags;
    return &request;
  }
};
}  // namespace LAMMPS_NS
