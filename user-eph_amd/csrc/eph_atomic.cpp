// eph_atomic.cpp -- host side of the fix eph/atomic style: command parsing,
// the .kappa tables on the device, the energies' set-up and the step hooks.
#include "eph_atomic.hpp"

#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "eph_context.hpp"

namespace eph {

namespace {

// EPH_Linear::reverse_lookup (eph_linear.h:52-61) on the host
double linear_reverse(const LinearTable& L, double v) {
  size_t lo = 0, hi = L.y.size();
  while (lo < hi) {
    const size_t mid = (lo + hi) / 2;
    if (v < L.y[mid]) hi = mid; else lo = mid + 1;
  }
  if (lo >= L.y.size() || lo == 0) return 0.0;
  const size_t idx = lo - 1;
  return idx * L.dx + 1. / L.dy[idx] * (v - L.y[idx]);
}

}  // namespace

AtomicView AtomicState::view() {
  AtomicView k;
  k.E = E.p; k.E1 = E1.p; k.dE = dE.p; k.rho_a = rho_a.p; k.T = T.p; k.K = K.p;
  k.krho = krho.p; k.k_nr = kap.n_r;
  k.k_inv_dr2 = kap.rho_r_sq.empty() ? 0.0 : kap.rho_r_sq[0].inv_dx;
  k.k_rc2 = kap.r_cutoff_sq;
  k.ET_y = ET_y.p; k.ET_dy = ET_dy.p; k.K_y = K_y.p; k.K_dy = K_dy.p;
  k.nT = kap.n_T; k.dT = kap.dT;
  k.kmap = kmap.p;
  return k;
}

// fix_eph_atomic.cpp:59-269.  arg: ID group eph/atomic seed flags T_e T_file
// inner_loops output beta_file kappa_file element...
int atomic_parse(eph_context* c, int narg, const char* const* arg, int ntypes, std::string& err) {
  if (narg < 12) { err = "fix_eph_atomic: too few arguments"; return EPH_ERR_ARGS; }
  auto* at = c->at.get();
  c->seed = std::atoi(arg[3]);
  c->flags = (int)std::strtol(arg[4], nullptr, 0);
  c->model = 4;
  at->T0 = std::atof(arg[5]);
  at->init_from_T = std::strcmp("NULL", arg[6]) == 0;
  at->inner_loops = std::atoi(arg[7]);
  if (at->inner_loops < 1) at->inner_loops = 0;
  c->ntypes = ntypes;
  if (ntypes < 1 || ntypes > narg - 11) {
    err = "fix_eph_atomic: number of types larger than provided in fix";
    return EPH_ERR_ARGS;
  }
  if (!load_beta(arg[9], c->beta, err))
    return err.rfind("cannot open", 0) == 0 ? EPH_ERR_FILE : EPH_ERR_FORMAT;
  if (!load_kappa(arg[10], at->kap, err))
    return err.rfind("cannot open", 0) == 0 ? EPH_ERR_FILE : EPH_ERR_FORMAT;
  c->type_map.assign(ntypes, INT_MAX);
  at->kmap_host.assign(ntypes, INT_MAX);
  for (int i = 0; i < ntypes; ++i) {
    for (size_t j = 0; j < c->beta.names.size(); ++j)
      if (c->beta.names[j] == arg[11 + i]) { c->type_map[i] = (int)j; break; }
    for (size_t j = 0; j < at->kap.names.size(); ++j)
      if (at->kap.names[j] == arg[11 + i]) { at->kmap_host[i] = (int)j; break; }
    if (c->type_map[i] == INT_MAX || at->kmap_host[i] == INT_MAX) {
      err = "fix_eph_atomic: elements not found in input file";
      return EPH_ERR_ARGS;
    }
    // K(T) is looked up by the element index among n_pairs slots
    // (fix_eph_atomic.cpp:707, 734); the reference's slot count is smaller
    // than the element count for two elements
    if (at->kmap_host[i] >= at->kap.n_pairs) {
      err = "fix_eph_atomic: K(T) slot of an element beyond the .kappa file's pair count";
      return EPH_ERR_FORMAT;
    }
  }
  return EPH_OK;
}

int atomic_upload(eph_context* c) {
  auto* at = c->at.get();
  const KappaModel& m = at->kap;
  const int ne = (int)m.names.size(), nr = m.n_r, nT = m.n_T, np = m.n_pairs;
  std::vector<Cubic> kr((size_t)ne * nr);
  std::vector<double> ey((size_t)ne * nT), edy((size_t)ne * nT, 0.0), ky((size_t)np * nT),
      kdy((size_t)np * nT, 0.0);
  for (int e = 0; e < ne; ++e) {
    std::copy(m.rho_r_sq[e].knots.begin(), m.rho_r_sq[e].knots.end(), kr.begin() + (size_t)e * nr);
    std::copy(m.E_T[e].y.begin(), m.E_T[e].y.end(), ey.begin() + (size_t)e * nT);
    std::copy(m.E_T[e].dy.begin(), m.E_T[e].dy.end(), edy.begin() + (size_t)e * nT);
  }
  for (int k = 0; k < np; ++k) {
    std::copy(m.K_T[k].y.begin(), m.K_T[k].y.end(), ky.begin() + (size_t)k * nT);
    std::copy(m.K_T[k].dy.begin(), m.K_T[k].dy.end(), kdy.begin() + (size_t)k * nT);
  }
  EPH_HIP(c, at->krho.reserve(kr.size()));
  EPH_HIP(c, at->ET_y.reserve(ey.size()));
  EPH_HIP(c, at->ET_dy.reserve(edy.size()));
  EPH_HIP(c, at->K_y.reserve(ky.size()));
  EPH_HIP(c, at->K_dy.reserve(kdy.size()));
  EPH_HIP(c, at->kmap.reserve(at->kmap_host.size()));
  EPH_HIP(c, at->sums.reserve(512));
  hipStream_t s = c->stream;
  EPH_HIP(c, hipMemcpyAsync(at->krho.p, kr.data(), kr.size() * sizeof(Cubic), hipMemcpyHostToDevice, s));
  EPH_HIP(c, hipMemcpyAsync(at->ET_y.p, ey.data(), ey.size() * sizeof(double), hipMemcpyHostToDevice, s));
  EPH_HIP(c, hipMemcpyAsync(at->ET_dy.p, edy.data(), edy.size() * sizeof(double), hipMemcpyHostToDevice, s));
  EPH_HIP(c, hipMemcpyAsync(at->K_y.p, ky.data(), ky.size() * sizeof(double), hipMemcpyHostToDevice, s));
  EPH_HIP(c, hipMemcpyAsync(at->K_dy.p, kdy.data(), kdy.size() * sizeof(double), hipMemcpyHostToDevice, s));
  EPH_HIP(c, hipMemcpyAsync(at->kmap.p, at->kmap_host.data(), at->kmap_host.size() * sizeof(int),
                            hipMemcpyHostToDevice, s));
  EPH_HIP(c, hipStreamSynchronize(s));
  return EPH_OK;
}

// per-atom arrays for nall atom slots; E keeps its values when it grows
// (grow_arrays keeps E_a_i), the rest is rewritten every step
int atomic_reserve(eph_context* c, size_t nall) {
  auto* at = c->at.get();
  if (at->E.p && at->E.n >= nall) return EPH_OK;
  const size_t want = nall + nall / 8 + 64;
  EPH_HIP(c, hipStreamSynchronize(c->stream));
  DevArray<double> e;
  EPH_HIP(c, e.reserve(want));
  // on the context's stream, finished before the old buffer is released
  if (at->E.p && at->E.n) {
    EPH_HIP(c, hipMemcpyAsync(e.p, at->E.p, at->E.n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    EPH_HIP(c, hipStreamSynchronize(c->stream));
  }
  at->E = std::move(e);
  for (DevArray<double>* a : {&at->E1, &at->dE, &at->rho_a, &at->T, &at->K, &at->st1, &at->st2})
    EPH_HIP(c, a->reserve(want));
  return EPH_OK;
}

// the constructor's set-up with the atoms present (fix_eph_atomic.cpp:242-268):
// E_i = E(T_e) for group atoms when no temperature file is given, the total
// Ee and the mean T_e, summed in atom order as the reference does
int atomic_setup_energies(eph_context* c, const int* type, const int* mask) {
  auto* at = c->at.get();
  const int n = c->nlocal;
  std::vector<double> E(n, 0.0);
  at->ngroup = 0;
  for (int i = 0; i < n; ++i) {
    if (!(mask[i] & c->groupbit)) continue;
    ++at->ngroup;
    if (at->init_from_T) E[i] = linear_forward(at->kap.E_T[at->kmap_host[type[i] - 1]], at->T0);
  }
  double Ee = 0.;
  for (int i = 0; i < n; ++i)
    if (mask[i] & c->groupbit) Ee += E[i];
  double Te = 0.0;
  for (int i = 0; i < n; ++i)
    if (mask[i] & c->groupbit) Te += linear_reverse(at->kap.E_T[at->kmap_host[type[i] - 1]], E[i]);
  if (at->ngroup > 0) Te /= static_cast<double>(at->ngroup);
  at->Ee = Ee;
  at->Te = Te / static_cast<double>(at->ngroup > 0 ? 1 : 0);
  if (n) EPH_HIP(c, hipMemcpyAsync(at->E.p, E.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  EPH_HIP(c, hipStreamSynchronize(c->stream));
  at->ready = true;
  return EPH_OK;
}

// FixEPHAtomic::post_force (fix_eph_atomic.cpp:764-824)
int atomic_post_force(eph_context* c) {
  auto* at = c->at.get();
  if (!at->ready) return c->fail(EPH_ERR_STATE, "atoms must be set");
  StepParams p = c->step_params();
  p.xi_mode = (c->flags & EPH_RANDOM) ? (c->xi_mode_next ? 1 : 2) : 0;
  const AtomsView a = c->atoms_view();
  const WorkView w = c->work_view();
  const AtomicView k = at->view();
  hipStream_t s = c->stream;
  hipEvent_t e;
  c->timed_begin("atomic_env", &e);
  EPH_HIP(c, launch_atomic_sync(a, w, k, false, s));
  EPH_HIP(c, launch_pack_pv(a, w, s));
  EPH_HIP(c, launch_atomic_env(a, c->tabs, p, w, k, s));
  EPH_HIP(c, launch_atomic_ghost(a, w, k, s));
  c->timed_end("atomic_env", e);
  c->timed_begin("atomic_forces", &e);
  EPH_HIP(c, launch_atomic_forces(a, c->tabs, p, w, k, s));
  c->timed_end("atomic_forces", e);
  c->xi_mode_next = 0;
  c->have_forces = true;
  return EPH_OK;
}

// FixEPHAtomic::end_of_step (fix_eph_atomic.cpp:379-414) with heat_solve
// (:660-762) when the HEAT flag is set
int atomic_end_of_step(eph_context* c) {
  auto* at = c->at.get();
  if (!at->ready) return c->fail(EPH_ERR_STATE, "atoms must be set");
  const AtomsView a = c->atoms_view();
  const WorkView w = c->work_view();
  const AtomicView k = at->view();
  hipStream_t s = c->stream;
  hipEvent_t e;
  if (c->flags & 0x04) {
    c->timed_begin("heat", &e);
    const int loops = at->inner_loops > 0 ? at->inner_loops : 1;
    const double scaling = 1.0 / static_cast<double>(loops);
    const double dt = c->dtv * scaling;
    for (int l = 0; l < loops; ++l) {
      EPH_HIP(c, launch_heat_add(a, k, scaling, s));
      EPH_HIP(c, launch_atomic_sync(a, w, k, true, s));
      EPH_HIP(c, launch_heat_diffuse(a, w, k, dt, s));
      EPH_HIP(c, launch_heat_copy(a, k, s));
    }
    c->timed_end("heat", e);
  }
  // Ee = sum of the group's E, Te = mean of their T (:384-410)
  EPH_HIP(c, launch_atomic_sync(a, w, k, false, s));
  EPH_HIP(c, launch_group_values(a, at->E.p, at->st1.p, s));
  EPH_HIP(c, launch_group_values(a, at->T.p, at->st2.p, s));
  EPH_HIP(c, launch_sum(at->st1.p, c->nlocal, at->sums.p + 256, at->sums.p, s));
  EPH_HIP(c, launch_sum(at->st2.p, c->nlocal, at->sums.p + 257, at->sums.p, s));
  double h[2] = {0, 0};
  EPH_HIP(c, hipMemcpyAsync(h, at->sums.p + 256, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  EPH_HIP(c, hipStreamSynchronize(s));
  at->Ee = h[0];
  double Tl = h[1];
  if (at->ngroup > 0) Tl /= static_cast<double>(at->ngroup);
  at->Te = Tl / static_cast<double>(at->ngroup > 0 ? 1 : 0);
  return EPH_OK;
}

}  // namespace eph
