// eph_harness.hpp -- internal definition of the MD harness handle (eph_md.h).
#pragma once

#include <string>
#include <vector>

#include "eph_context.hpp"
#include "eph_nbuild.hpp"

extern "C" {
#include "eph_md.h"
}

namespace eph {
struct CSR {
  std::vector<int> off, idx;
};
}  // namespace eph

struct eph_md {
  std::string err;
  eph_md_config cfg{};
  int nlocal = 0, nghost = 0;
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, L[3] = {0, 0, 0};
  std::vector<double> x, v;          // [nall][3]
  std::vector<int> type, mask, tag;  // [nall]
  std::vector<int> owner;            // [nghost]
  std::vector<double> shift;         // [nghost][3]
  std::vector<double> x_build;       // [nlocal][3]
  eph::CSR fix, lj;
  // device side
  eph::DevArray<int> lj_nbr, lj_nn, lj_off, lj_idx;
  eph::DevArray<double4> xref;
  eph::DevArray<double> partial;
  int lj_stride = 0;
  // device rebuild (eph_nbuild.hip)
  bool device_rebuild = true;
  bool force_next = false;
  int ghost_cap = 0, cap_f = 0, cap_l = 0, nbins = 0;
  eph::DevArray<int> rep, flags, cnt, start, bin_of, bcount, bstart, bfill, items, scan_tmp, nbi;
  eph::DevArray<double> box_d;
  eph::NbBuild nbw;
  int* pinned = nullptr;  // go, nghost, overflow
  std::vector<char> ckpt;  // eph_md_checkpoint_save(md, ctx, NULL) keeps it here

  int fail(int code, const std::string& m) { err = m; return code; }
  int fail_hip(hipError_t e, const char* what) {
    err = std::string("HIP error ") + hipGetErrorString(e) + " in " + what;
    return EPH_ERR_HIP;
  }
  ~eph_md() {
    if (pinned) (void)hipHostFree(pinned);
  }
  void wrap();
  void build_ghosts();
  void build_lists();
  void rebuild_all() {
    wrap();
    build_ghosts();
    build_lists();
    x_build.assign(x.begin(), x.begin() + 3 * (size_t)nlocal);
  }
  int upload(eph_context* ctx);
};
