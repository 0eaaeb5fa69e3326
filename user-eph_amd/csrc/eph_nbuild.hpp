// eph_nbuild.hpp -- device rebuild of ghosts and neighbour lists (harness).
#pragma once

#include <cmath>
#include <cstdint>

#include "eph_device.hpp"

namespace eph {

// Order of every fix list of the harness (host and device builds alike): the
// entries are first found in stencil order (found position t), then sorted by
// (q(r^2), t), q(r^2) = r^2 x scale truncated, r^2 as at the build.  scale is
// the power of two that keeps cf2 x scale at or below 2^23, so q fits 23 bits
// and is exact arithmetic on both sides; t breaks ties.  Close pairs come
// first, and entry k of neighbouring atoms lies at similar distances, which
// the sweeps' table loads gain from.  The device sorts keys q x 2^9 + t and
// takes lists of at most kListSortMax entries (a longer list is an error).
constexpr int kListSortMax = 512;
inline double list_sort_scale(double cf2) {
  return std::ldexp(1.0, 23 - static_cast<int>(std::ceil(std::log2(cf2 > 1.0 ? cf2 : 1.0))));
}
__host__ __device__ inline uint32_t list_sort_q(double r2, double scale) {
  return static_cast<uint32_t>(r2 * scale);
}

// per-replica boxes and bin grids (arrays of 3 per replica unless noted)
struct NbBoxes {
  const double* lo = nullptr;
  const double* hi = nullptr;
  const double* L = nullptr;
  const double* glo = nullptr;   // bin grid origin = lo - cmax
  const int* nb = nullptr;       // bins per dimension
  const int* bin_base = nullptr; // [replicas] first bin of each replica
  double bs = 1.0;               // bin size = cmax / 2
  double cg = 1.0;               // ghost cutoff = cmax
};

struct NbAtoms {
  PosVel* pv = nullptr;  // packed records kept current by the remap (may be null)
  double4* pos = nullptr;
  double4* vel = nullptr;
  int* type = nullptr;
  int* mask = nullptr;
  int* tag = nullptr;
  int* rep = nullptr;
  int* owner = nullptr;          // [ghost_cap]
  double4* shift = nullptr;      // [ghost_cap]
};

struct NbLists {
  int* nbr_f = nullptr; int* nn_f = nullptr; int cap_f = 0; double cf2 = 0;  // fix list
  double sort_scale = 0;  // list_sort_scale(cf2): the fix lists' order (see list_sort_key)
  int* nbr_l = nullptr; int* nn_l = nullptr; int cap_l = 0; double cl2 = 0;  // pair list
  int stride = 0;
  // the density sweep's candidates (entries of the fix list with r^2 < cand2,
  // in list order, capacity cap_f), written with the lists when cand is set
  int* cand = nullptr; int* ncand = nullptr; double cand2 = 0;
};

struct NbBuild {
  int nlocal = 0, ghost_cap = 0, nbins = 0;
  NbAtoms at;
  NbBoxes box;
  NbLists lists;
  int* go = nullptr;        // [1] rebuild flag
  int* nghost = nullptr;    // [4] ghosts after each pass; [3] is the result
  int* overflow = nullptr;  // [1] bit 0 ghosts, bit 1 lists
  int* nrebuild = nullptr;  // [1] count of device rebuilds
  // [1] the density sweep's candidate state: a rebuild that writes the
  // candidates (lists.cand) marks them valid with cref = the positions of the
  // rebuild; one that does not clears it
  int* cand_valid = nullptr;
  double4* cref = nullptr;  // [nlocal + ghost_cap]
  int* cnt = nullptr;       // [nlocal + ghost_cap]
  int* start = nullptr;     // [nlocal + ghost_cap + 1]
  int* bin_of = nullptr;    // [nlocal + ghost_cap]
  int* bcount = nullptr;    // [nbins + 1]
  int* bstart = nullptr;    // [nbins + 1]
  int* bfill = nullptr;     // [nbins + 1]
  int* items = nullptr;     // [nlocal + ghost_cap]
  int* scan_tmp = nullptr;  // [(nlocal + ghost_cap + nbins) / 4096 + 2]
  double4* xref = nullptr;  // [nlocal]
};

hipError_t launch_nb_check(int nlocal, const double4* pos, const double4* ref, double lim, int* go,
                           hipStream_t s);
hipError_t launch_nb_force(int* go, hipStream_t s);
hipError_t launch_nb_build(const NbBuild& w, hipStream_t s);

}  // namespace eph
