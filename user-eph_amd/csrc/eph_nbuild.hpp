// eph_nbuild.hpp -- device rebuild of ghosts and neighbour lists (harness).
#pragma once

#include "eph_device.hpp"

namespace eph {

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
