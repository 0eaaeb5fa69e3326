// Integer types and the neighbour-mask of a default LAMMPS build (test harness).
#pragma once
#include <cstdint>
namespace LAMMPS_NS {
typedef int tagint;
typedef int64_t bigint;
}  // namespace LAMMPS_NS
#define NEIGHMASK 0x1FFFFFFF
