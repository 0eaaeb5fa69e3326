// eph_context.cpp -- the C ABI of eph_amd.h: fix context, T_e grid, hooks.
#include "eph_context.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>

using namespace eph;

namespace {
thread_local std::string g_create_error;

inline int nblocks_of(long n) { return static_cast<int>((n + kBlock - 1) / kBlock); }
}  // namespace

// =================================================================== deposit
hipError_t DepositBuffers::reserve(size_t n_items, size_t G, size_t nblocks) {
  hipError_t e;
  if ((e = cell.reserve(n_items)) != hipSuccess) return e;
  if ((e = items.reserve(n_items)) != hipSuccess) return e;
  if ((e = v1.reserve(n_items)) != hipSuccess) return e;
  if ((e = v2.reserve(n_items)) != hipSuccess) return e;
  if ((e = count.reserve(G + 1)) != hipSuccess) return e;
  if ((e = start.reserve(G + 1)) != hipSuccess) return e;
  if ((e = fill.reserve(G + 1)) != hipSuccess) return e;
  if ((e = qbuf.reserve(2 + 2 * G)) != hipSuccess) return e;
  if ((e = scan_tmp.reserve(G / 4096 + 2)) != hipSuccess) return e;
  if ((e = partial.reserve(2 * nblocks + 2)) != hipSuccess) return e;
  return hipSuccess;
}

// ====================================================================== grid
int eph_grid::upload() {
  const size_t G = h.ntotal();
  EPH_HIP(this, T.reserve(G));
  EPH_HIP(this, Tn.reserve(G));
  EPH_HIP(this, dT.reserve(G));
  EPH_HIP(this, S.reserve(G));
  EPH_HIP(this, rho.reserve(G));
  EPH_HIP(this, C.reserve(G));
  EPH_HIP(this, kappa.reserve(G));
  EPH_HIP(this, flag.reserve(G));
  EPH_HIP(this, tdyn.reserve(G));
  EPH_HIP(this, sum_tmp.reserve(512));
  EPH_HIP(this, minmax.reserve(4));
  EPH_HIP(this, nonfinite.reserve(1));
  const size_t bd = G * sizeof(double);
  EPH_HIP(this, hipMemcpyAsync(T.p, h.T.data(), bd, hipMemcpyHostToDevice, stream));
  EPH_HIP(this, hipMemcpyAsync(Tn.p, h.T.data(), bd, hipMemcpyHostToDevice, stream));
  EPH_HIP(this, hipMemsetAsync(dT.p, 0, bd, stream));
  EPH_HIP(this, hipMemcpyAsync(S.p, h.S.data(), bd, hipMemcpyHostToDevice, stream));
  EPH_HIP(this, hipMemcpyAsync(rho.p, h.rho.data(), bd, hipMemcpyHostToDevice, stream));
  EPH_HIP(this, hipMemcpyAsync(C.p, h.C.data(), bd, hipMemcpyHostToDevice, stream));
  EPH_HIP(this, hipMemcpyAsync(kappa.p, h.kappa.data(), bd, hipMemcpyHostToDevice, stream));
  EPH_HIP(this, hipMemcpyAsync(flag.p, h.flag.data(), G, hipMemcpyHostToDevice, stream));
  EPH_HIP(this, hipMemcpyAsync(tdyn.p, h.tdyn.data(), G, hipMemcpyHostToDevice, stream));
  any_tdyn = std::any_of(h.tdyn.begin(), h.tdyn.end(), [](int8_t t) { return t != 0; });
  v = GridView();
  v.nx = (int)h.nx; v.ny = (int)h.ny; v.nz = (int)h.nz;
  v.x0 = h.x0; v.y0 = h.y0; v.z0 = h.z0;
  v.dx = h.dx; v.dy = h.dy; v.dz = h.dz; v.dV = h.dV;
  v.T = T.p; v.Tn = Tn.p; v.dT = dT.p; v.S = S.p; v.rho = rho.p; v.C = C.p; v.kappa = kappa.p;
  v.flag = flag.p; v.tdyn = tdyn.p;
  if (h.has_params) {
    const size_t nc = h.C_T.knots.size(), nk = h.kappa_T.knots.size(), ne = h.E_T.y.size();
    EPH_HIP(this, C_T.reserve(nc));
    EPH_HIP(this, kappa_T.reserve(nk));
    EPH_HIP(this, E_y.reserve(ne));
    EPH_HIP(this, E_dy.reserve(ne));
    EPH_HIP(this, hipMemcpyAsync(C_T.p, h.C_T.knots.data(), nc * sizeof(Cubic), hipMemcpyHostToDevice, stream));
    EPH_HIP(this, hipMemcpyAsync(kappa_T.p, h.kappa_T.knots.data(), nk * sizeof(Cubic), hipMemcpyHostToDevice, stream));
    EPH_HIP(this, hipMemcpyAsync(E_y.p, h.E_T.y.data(), ne * sizeof(double), hipMemcpyHostToDevice, stream));
    EPH_HIP(this, hipMemcpyAsync(E_dy.p, h.E_T.dy.data(), (ne - 1) * sizeof(double), hipMemcpyHostToDevice, stream));
    v.C_T = C_T.p; v.n_CT = (int)nc; v.inv_dT_C = h.C_T.inv_dx;
    v.kappa_T = kappa_T.p; v.n_kT = (int)nk; v.inv_dT_k = h.kappa_T.inv_dx;
    v.E_y = E_y.p; v.E_dy = E_dy.p; v.n_E = (int)ne; v.E_dx = h.E_T.dx;
  }
  minmax_valid = false;
  EPH_HIP(this, hipStreamSynchronize(stream));
  return EPH_OK;
}

void eph_grid::refresh_minmax() {
  const size_t G = h.ntotal();
  cmin = h.C[0]; rmin = h.rho[0]; kmax = h.kappa[0];
  for (size_t i = 1; i < G; ++i) {
    if (h.flag[i] != CELL_CONSTANT) {
      if (h.C[i] < cmin) cmin = h.C[i];
      if (h.rho[i] < rmin) rmin = h.rho[i];
      if (h.kappa[i] > kmax) kmax = h.kappa[i];
    }
  }
  minmax_valid = true;
}

// sub-step count and inner dt of EPH_FDM::solve (eph_fdm.h:302-314) from the
// stability extrema
void eph_grid::substeps(double* r_out, unsigned* steps_out, double* inner_out) const {
  const double steps_d = static_cast<double>(h.steps);
  double inner_dt = h.dt / steps_d;
  const double dtdxdydz = inner_dt * (1.0 / h.dx / h.dx + 1.0 / h.dy / h.dy + 1.0 / h.dz / h.dz);
  const double r = dtdxdydz / cmin / rmin * kmax;
  unsigned new_steps = static_cast<unsigned>(h.steps);
  if (r > 0.4) {
    inner_dt = 0.4 * inner_dt / r;
    new_steps = std::max(static_cast<unsigned>(h.dt / inner_dt), 1u);
    inner_dt = h.dt / new_steps;
  }
  *r_out = r;
  *steps_out = new_steps;
  *inner_out = inner_dt;
}

// EPH_FDM::solve (eph_fdm.h:267-400) on the device
int eph_grid::solve(int* substeps_out) {
  if (any_tdyn) {
    EPH_HIP(this, launch_grid_tdyn_update(v, stream));
    EPH_HIP(this, launch_grid_minmax(v, minmax.p, nullptr, stream));
    double mm[3];
    EPH_HIP(this, hipMemcpyAsync(mm, minmax.p, sizeof mm, hipMemcpyDeviceToHost, stream));
    EPH_HIP(this, hipStreamSynchronize(stream));
    cmin = mm[0]; rmin = mm[1]; kmax = mm[2];
  } else if (!minmax_valid) {
    refresh_minmax();
  }
  double r = 0, inner_dt = 0;
  unsigned new_steps = 0;
  substeps(&r, &new_steps, &inner_dt);
  last_r = r;
  for (unsigned n = 0; n < new_steps; ++n) {
    EPH_HIP(this, launch_fdm_substep(v, T.p, Tn.p, inner_dt, any_tdyn, stream));
    std::swap(T.p, Tn.p);
    v.T = T.p;
    v.Tn = Tn.p;
  }
  EPH_HIP(this, launch_fill_double(dT.p, (long)h.ntotal(), 0.0, stream));
  EPH_HIP(this, launch_nonfinite(T.p, (int)h.ntotal(), nonfinite.p, stream));
  last_steps = (int)new_steps;
  if (substeps_out) *substeps_out = (int)new_steps;
  return EPH_OK;
}

int eph_grid::insert_points(int n, const double* x, const double* y, const double* z,
                            const double* E) {
  if (n <= 0) return EPH_OK;
  const size_t G = h.ntotal();
  EPH_HIP(this, pts.reserve(4 * (size_t)n));
  EPH_HIP(this, dep.reserve(n, G, nblocks_of(n)));
  const size_t b = n * sizeof(double);
  EPH_HIP(this, hipMemcpyAsync(pts.p, x, b, hipMemcpyHostToDevice, stream));
  EPH_HIP(this, hipMemcpyAsync(pts.p + n, y, b, hipMemcpyHostToDevice, stream));
  EPH_HIP(this, hipMemcpyAsync(pts.p + 2 * n, z, b, hipMemcpyHostToDevice, stream));
  EPH_HIP(this, hipMemcpyAsync(pts.p + 3 * n, E, b, hipMemcpyHostToDevice, stream));
  DepositView d = dep.view(false);
  EPH_HIP(this, launch_fill_int(d.count, (int)G + 1, 0, stream));
  EPH_HIP(this, launch_deposit_points(n, pts.p, pts.p + n, pts.p + 2 * n, pts.p + 3 * n, h.dt, v, d, stream));
  EPH_HIP(this, launch_deposit_cells(v, d, n, dep.qbuf.p, stream));
  EPH_HIP(this, hipStreamSynchronize(stream));
  return EPH_OK;
}

int eph_grid::download(int which, std::vector<double>& out) {
  const size_t G = h.ntotal();
  out.resize(G);
  const double* src = which == 0 ? T.p : which == 1 ? dT.p : which == 2 ? C.p : which == 3 ? rho.p
                    : which == 4 ? kappa.p : S.p;
  EPH_HIP(this, hipMemcpyAsync(out.data(), src, G * sizeof(double), hipMemcpyDeviceToHost, stream));
  EPH_HIP(this, hipStreamSynchronize(stream));
  return EPH_OK;
}

int eph_grid::sync_params_from_device() {
  if (!any_tdyn) return EPH_OK;
  int rc;
  std::vector<double> tmp;
  if ((rc = download(2, tmp)) != EPH_OK) return rc;
  h.C = tmp;
  if ((rc = download(4, tmp)) != EPH_OK) return rc;
  h.kappa = tmp;
  return EPH_OK;
}

static int grid_init_device(eph_grid* g, int device, hipStream_t s) {
  g->device = device;
  if (hipSetDevice(device) != hipSuccess) return g->fail(EPH_ERR_NO_DEVICE, "no HIP device");
  if (s) {
    g->stream = s;
  } else {
    if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess)
      return g->fail(EPH_ERR_HIP, "stream creation failed");
    g->own_stream = true;
  }
  return g->upload();
}

// ================================================================== context
AtomsView eph_context::atoms_view() {
  AtomsView a;
  a.nlocal = nlocal; a.nghost = nghost; a.stride = stride;
  a.pos = pos.p; a.vel = vel.p; a.f = f.p; a.fstride = fstride;
  a.type = type.p; a.mask = mask.p; a.tag = tag.p; a.groupbit = groupbit;
  // eph/coloured/exp/v1 sweeps filtered velocities and eph/atomic packs its
  // own: only the other styles keep the records current as atoms move
  a.pv = (colour == 3 || atomic()) ? nullptr : pv.p;
  return a;
}

StepParams eph_context::step_params() {
  StepParams p;
  p.flags = flags; p.eta = eta; p.dtv = dtv; p.dtf = dtf; p.dt = dt;
  p.mass = mass.p; p.seed = static_cast<uint32_t>(seed); p.step = static_cast<uint64_t>(step);
  p.xi_mode = 0; p.xi_in = xi_in.p;
  p.model = model; p.colour = colour; p.zeta = zeta;
  return p;
}

void eph_context::drop_candidates() {
  if (cstate.p) (void)launch_fill_int(cstate.p, 8, 0, stream);
}

WorkView eph_context::work_view() {
  WorkView w;
  w.nbr = nbr.p; w.nnbr = nnbr.p; w.act = act.p; w.nact = nact.p;
  w.rec = rec.p; w.pv = pv.p; w.fE = fE.p; w.fR = fR.p; w.flags_out = flags_dev.p;
  w.ghost_owner = ghost_owner.p; w.ghost_shift = ghost_shift.p;
  w.cA = colA.p; w.cB = colB.p;
  w.vw = colour == 3 ? colA.p : vel.p;
  const double delta = cand_delta < 0.0 ? 0.1 * beta.r_cutoff : cand_delta;
  if (delta > 0.0 && cstate.p && cand.p && cand.n >= act.n && ncand.n >= nact.n &&
      cref.n >= static_cast<size_t>(nlocal) + nghost && moved.n >= static_cast<size_t>(kMovedCap)) {
    w.cand = cand.p; w.ncand = ncand.p; w.cref = cref.p; w.cstate = cstate.p; w.moved = moved.p;
    w.moved_cap = kMovedCap;
    const double rc = beta.r_cutoff + delta;
    w.rcand2 = rc * rc;
    // moved = 0.45 delta (not delta/2): an entry left out that neither atom
    // of the pair has moved is farther than rc + 0.1 delta, and a moved atom
    // is looked for within rc + 0.1 delta of a wave's atoms -- margins far
    // above the rounding of the distances
    w.half_delta2 = (0.45 * delta) * (0.45 * delta);
    w.reach = beta.r_cutoff + 0.1 * delta;
  }
  return w;
}

// Filter states persist from step to step, so growing them keeps the old
// entries (LAMMPS's grow_arrays) and zeroes the new ones.
int eph_context::grow_filter_state(size_t nall) {
  if (colour < 2 || (colA.p && colA.n >= nall)) return EPH_OK;
  const size_t want = nall + nall / 8 + 64;
  EPH_HIP(this, hipStreamSynchronize(stream));  // the old states are final
  for (eph::DevArray<double4>* a : {&colA, &colB}) {
    eph::DevArray<double4> b;
    EPH_HIP(this, b.reserve(want));  // zero-filled
    if (a->p && a->n) {  // on the stream, finished before the old buffer is released
      EPH_HIP(this, hipMemcpyAsync(b.p, a->p, a->n * sizeof(double4), hipMemcpyDeviceToDevice, stream));
      EPH_HIP(this, hipStreamSynchronize(stream));
    }
    *a = std::move(b);
  }
  return EPH_OK;
}

int eph_context::reserve_atoms(int nl, int ng) {
  const size_t nall = (size_t)nl + ng;
  const int st = ((nl + 63) / 64) * 64 + 64;
  EPH_HIP(this, pos.reserve(nall));
  EPH_HIP(this, vel.reserve(nall));
  EPH_HIP(this, rec.reserve(nall));
  EPH_HIP(this, pv.reserve(nall));
  EPH_HIP(this, type.reserve(nall));
  EPH_HIP(this, mask.reserve(nall));
  EPH_HIP(this, tag.reserve(nall));
  EPH_HIP(this, nnbr.reserve(st));
  EPH_HIP(this, nact.reserve(st));
  EPH_HIP(this, ncand.reserve(st));
  EPH_HIP(this, cref.reserve(nall));
  EPH_HIP(this, moved.reserve(kMovedCap));
  EPH_HIP(this, cstate.reserve(8));
  EPH_HIP(this, stage.reserve(static_cast<size_t>(EPH_HALO_COLS) * nall));  // rows, halos
  EPH_HIP(this, xi_in.reserve(3 * (size_t)st));
  if (st != stride || f.n < 3 * (size_t)st) {
    EPH_HIP(this, f.reserve(3 * (size_t)st));
    EPH_HIP(this, fE.reserve(3 * (size_t)st));
    EPH_HIP(this, fR.reserve(3 * (size_t)st));
    EPH_HIP(this, hipMemsetAsync(fE.p, 0, fE.n * sizeof(double), stream));
    EPH_HIP(this, hipMemsetAsync(fR.p, 0, fR.n * sizeof(double), stream));
    if (stride != st) have_list = false;  // ELL layout depends on the stride
  }
  stride = st;
  fstride = st;
  nlocal = nl;
  nghost = ng;
  pv_fresh = false;
  // the filter states (and the eph/atomic energies) cover every atom slot a
  // later device rebuild may fill
  if (atomic()) return eph::atomic_reserve(this, nall);
  return grow_filter_state(nall);
}

int eph_context::set_atoms(int nl, int ng, const double* x, const double* v, const int* ty,
                           const int* mk, const int* tg) {
  if (nl < 0 || ng < 0 || (nl + ng > 0 && (!x || !v || !ty || !mk))) return fail(EPH_ERR_ARGS, "bad atom arrays");
  const size_t nall = (size_t)nl + ng;
  for (size_t i = 0; i < nall; ++i)
    if (ty[i] < 1 || ty[i] > ntypes) return fail(EPH_ERR_ARGS, "atom type out of range");
  int rc = reserve_atoms(nl, ng);
  if (rc == EPH_OK) drop_candidates();
  if (rc != EPH_OK) return rc;
  if (species_dev.n < (size_t)ntypes) {
    EPH_HIP(this, species_dev.reserve(ntypes));
    EPH_HIP(this, hipMemcpyAsync(species_dev.p, type_map.data(), ntypes * sizeof(int),
                                 hipMemcpyHostToDevice, stream));
  }
  // the caller's [n][3] rows go up as they are and become (x, y, z, element)
  // and (v, 0) records on the device
  if ((rc = h2d(type.p, ty, nall * sizeof(int)))) return rc;
  if ((rc = h2d(mask.p, mk, nall * sizeof(int)))) return rc;
  if ((rc = h2d(stage.p, x, 3 * nall * sizeof(double)))) return rc;
  EPH_HIP(this, launch_rows_into_d4_w(stage.p, (int)nall, type.p, species_dev.p, pos.p, stream));
  if ((rc = h2d(stage.p, v, 3 * nall * sizeof(double)))) return rc;
  EPH_HIP(this, launch_rows_into_d4_w(stage.p, (int)nall, type.p, nullptr, vel.p, stream));
  if (atomic() && !at->ready && (rc = eph::atomic_setup_energies(this, ty, mk)) != EPH_OK) return rc;
  if (tg) {
    if ((rc = h2d(tag.p, tg, nall * sizeof(int)))) return rc;
  } else {
    std::vector<int> t(nall);
    for (size_t i = 0; i < nall; ++i) t[i] = (int)(i < (size_t)nl ? i + 1 : 0);
    if ((rc = h2d(tag.p, t.data(), nall * sizeof(int)))) return rc;
  }
  EPH_HIP(this, hipStreamSynchronize(stream));
  have_atoms = true;
  return EPH_OK;
}

int eph_context::set_ghost_map(int ng, const int* owner, const double* shift) {
  if (ng != nghost) return fail(EPH_ERR_ARGS, "ghost map size differs from nghost");
  drop_candidates();
  if (ng == 0) return EPH_OK;
  for (int g = 0; g < ng; ++g)
    if (owner[g] < 0 || owner[g] >= nlocal) return fail(EPH_ERR_ARGS, "ghost owner out of range");
  std::vector<double4> sh(ng);
  for (int g = 0; g < ng; ++g)
    sh[g] = shift ? make_double4(shift[3 * g], shift[3 * g + 1], shift[3 * g + 2], 0.0)
                  : make_double4(0.0, 0.0, 0.0, 0.0);
  EPH_HIP(this, ghost_owner.reserve(ng));
  EPH_HIP(this, ghost_shift.reserve(ng));
  EPH_HIP(this, hipMemcpyAsync(ghost_owner.p, owner, ng * sizeof(int), hipMemcpyHostToDevice, stream));
  EPH_HIP(this, hipMemcpyAsync(ghost_shift.p, sh.data(), ng * sizeof(double4), hipMemcpyHostToDevice, stream));
  EPH_HIP(this, hipStreamSynchronize(stream));
  return EPH_OK;
}

int eph_context::set_neighbors(int nl, const int* off, const int* idx, int neighmask) {
  if (!have_atoms) return fail(EPH_ERR_STATE, "atoms must be set before the neighbour list");
  if (nl != nlocal) return fail(EPH_ERR_ARGS, "neighbour list size differs from nlocal");
  drop_candidates();
  if (!off) return fail(EPH_ERR_ARGS, "CSR offsets missing");
  if (off[0] != 0) return fail(EPH_ERR_ARGS, "CSR offsets must start at 0");
  if (off[nl] > 0 && !idx) return fail(EPH_ERR_ARGS, "CSR indices missing");
  int mx = 0;
  const int nall = nlocal + nghost;
  for (int i = 0; i < nl; ++i) {
    const int n = off[i + 1] - off[i];
    if (n < 0) return fail(EPH_ERR_ARGS, "CSR offsets must be non-decreasing");
    mx = std::max(mx, n);
  }
  const long nnz = off[nl];
  maxnbr = mx;
  EPH_HIP(this, nbr.reserve((size_t)std::max(mx, 1) * stride));
  EPH_HIP(this, act.reserve((size_t)std::max(mx, 1) * stride));
  EPH_HIP(this, cand.reserve((size_t)std::max(mx, 1) * stride));
  EPH_HIP(this, csr_off.reserve(nl + 1));
  EPH_HIP(this, csr_idx.reserve(std::max<long>(nnz, 1)));
  EPH_HIP(this, hipMemcpyAsync(csr_off.p, off, (nl + 1) * sizeof(int), hipMemcpyHostToDevice, stream));
  if (nnz) {
    if (int rc = h2d(csr_idx.p, idx, nnz * sizeof(int))) return rc;
  }
  // the indices are range-checked on the device while they are converted
  EPH_HIP(this, list_bad.reserve(1));
  EPH_HIP(this, hipMemsetAsync(list_bad.p, 0, sizeof(int), stream));
  // rows are bounded by the width the host offsets give, so a device copy of
  // the offsets that differs from them cannot steer writes past the buffers
  EPH_HIP(this, launch_csr_to_ell(nl, csr_off.p, csr_idx.p, neighmask, nall, nbr.p, nnbr.p, stride,
                                  std::max(mx, 1), list_bad.p, stream));
  int bad = 0;
  EPH_HIP(this, hipMemcpyAsync(&bad, list_bad.p, sizeof(int), hipMemcpyDeviceToHost, stream));
  EPH_HIP(this, hipStreamSynchronize(stream));
  have_list = bad == 0;
  if (bad & 2) return fail(EPH_ERR_STATE, "neighbour offsets changed on their way to the device (the list is dropped)");
  if (bad) return fail(EPH_ERR_ARGS, "neighbour index out of range (the list is dropped)");
  return EPH_OK;
}

hipEvent_t eph_context::take_event() {
  if (!event_pool.empty()) {
    hipEvent_t e = event_pool.back();
    event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

void eph_context::timed_begin(const char*, hipEvent_t* a) {
  *a = nullptr;
  if (!timing) return;
  *a = take_event();
  (void)hipEventRecord(*a, stream);
}

void eph_context::timed_end(const char* name, hipEvent_t a) {
  if (!timing || !a) return;
  hipEvent_t b = take_event();
  (void)hipEventRecord(b, stream);
  pending.push_back({name, a, b});
}

int eph_context::resolve_timing() {
  if (pending.empty()) return EPH_OK;
  EPH_HIP(this, hipStreamSynchronize(stream));
  for (auto& t : pending) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, t.a, t.b);
    auto it = std::find_if(stats.begin(), stats.end(), [&](const KernelStat& s) { return s.name == t.name; });
    if (it == stats.end()) {
      stats.push_back({t.name, 0.0, 0});
      it = stats.end() - 1;
    }
    it->ms += ms;
    it->count += 1;
    if (t.own_a) event_pool.push_back(t.a);
    event_pool.push_back(t.b);
  }
  pending.clear();
  return EPH_OK;
}

int eph_context::dev_initial_integrate() {
  if (!have_atoms) return fail(EPH_ERR_STATE, "no atoms");
  if (flags & EPH_NOINT) return EPH_OK;
  hipEvent_t a;
  timed_begin("integrate", &a);
  EPH_HIP(this, launch_integrate(atoms_view(), step_params(), true, stream));
  timed_end("integrate", a);
  return EPH_OK;
}

int eph_context::dev_final_integrate() {
  if (!have_atoms) return fail(EPH_ERR_STATE, "no atoms");
  if (flags & EPH_NOINT) return EPH_OK;
  hipEvent_t a;
  timed_begin("integrate", &a);
  EPH_HIP(this, launch_integrate(atoms_view(), step_params(), false, stream));
  timed_end("integrate", a);
  return EPH_OK;
}

// The density sweep; with timing on, the sweep kernel alone is "density_w",
// the candidate check before it "cand_check" and the flag update after it
// "cand_flags".
int eph_context::timed_density(const AtomsView& a, const StepParams& p, const WorkView& w) {
  if (!timing) {
    EPH_HIP(this, launch_density_w(a, tabs, p, w, stream, form_density));
    return EPH_OK;
  }
  hipEvent_t e0 = take_event(), ev[2] = {take_event(), take_event()};
  (void)hipEventRecord(e0, stream);
  EPH_HIP(this, launch_density_w(a, tabs, p, w, stream, form_density, ev));
  hipEvent_t e1 = take_event();
  (void)hipEventRecord(e1, stream);
  pending.push_back({"cand_check", e0, ev[0], true});
  pending.push_back({"density_w", ev[0], ev[1], false});
  pending.push_back({"cand_flags", ev[1], e1, false});
  return EPH_OK;
}

int eph_context::dev_post_force() {
  if (!have_atoms || !have_list) return fail(EPH_ERR_STATE, "atoms and neighbour list must be set");
  if (atomic()) return eph::atomic_post_force(this);
  StepParams p = step_params();
  p.xi_mode = (flags & EPH_RANDOM) ? (xi_mode_next ? 1 : 2) : 0;
  const AtomsView a = atoms_view();
  const WorkView w = work_view();
  hipEvent_t e;
  if (colour == 3 && (flags & EPH_FRICTION)) EPH_HIP(this, launch_colour_v(a, p, w, stream));
  if (!pv_fresh || !a.pv) {
    timed_begin("pack_pv", &e);
    EPH_HIP(this, launch_pack_pv(a, w, stream));
    timed_end("pack_pv", e);
    pv_fresh = a.pv != nullptr;
  }
  // models other than PRL need rho (and xi) but no w sum
  StepParams pd = p;
  if (model != 4) pd.flags &= ~EPH_FRICTION;
  if (int rc = timed_density(a, pd, w)) return rc;
  if (model == 4) {
    timed_begin("ghost_rec", &e);
    EPH_HIP(this, launch_ghost_rec(a, w, stream));
    timed_end("ghost_rec", e);
    timed_begin("prl_forces", &e);
    EPH_HIP(this, launch_prl_forces(a, tabs, grid->v, p, w, stream, form_force));
    timed_end("prl_forces", e);
  } else if (model == 1) {
    timed_begin("ttm_forces", &e);
    EPH_HIP(this, launch_ttm_forces(a, tabs, grid->v, p, w, stream));
    timed_end("ttm_forces", e);
  } else if (model == 2) {
    timed_begin("prb_forces", &e);
    EPH_HIP(this, launch_prb_forces(a, tabs, grid->v, p, w, stream));
    timed_end("prb_forces", e);
  }
  xi_mode_next = 0;
  have_forces = true;
  return EPH_OK;
}

// post_force of a multi-rank run, first part: the noise, rho, alpha and w of
// the local atoms (fix_eph.cpp:852-862, 866-867 and force_prl's stage 1);
// their xi, rho and w then go to the other ranks' ghosts (the XI, RHO and WI
// exchanges of :863-871, which only the later stages read)
int eph_context::dev_post_force_local() {
  if (!have_atoms || !have_list) return fail(EPH_ERR_STATE, "atoms and neighbour list must be set");
  if (atomic() || colour == 3 || model != 4)
    return fail(EPH_ERR_UNSUPPORTED, "the split post_force serves the PRL model of eph, "
                                     "eph/coloured and eph/coloured/exp");
  StepParams p = step_params();
  p.xi_mode = (flags & EPH_RANDOM) ? (xi_mode_next ? 1 : 2) : 0;
  const AtomsView a = atoms_view();
  const WorkView w = work_view();
  if (!pv_fresh || !a.pv) {
    EPH_HIP(this, launch_pack_pv(a, w, stream));
    pv_fresh = a.pv != nullptr;
  }
  return timed_density(a, p, w);
}

// the second part, once the ghosts' records hold their exchanged values:
// stages 2 and 3 and the force update
int eph_context::dev_post_force_ghosts() {
  StepParams p = step_params();
  p.xi_mode = (flags & EPH_RANDOM) ? (xi_mode_next ? 1 : 2) : 0;
  const AtomsView a = atoms_view();
  const WorkView w = work_view();
  hipEvent_t e;
  timed_begin("prl_forces", &e);
  EPH_HIP(this, launch_prl_forces(a, tabs, grid->v, p, w, stream, form_force));
  timed_end("prl_forces", e);
  xi_mode_next = 0;
  have_forces = true;
  return EPH_OK;
}

// Times every form of each sweep on the resident atoms and keeps the fastest.
// The density sweep rewrites only what the next post_force rewrites anyway;
// the force sweep runs with both force additions off, so f is not touched.
int eph_context::autotune(int* forms, double* ms) {
  if (!have_atoms || !have_list) return fail(EPH_ERR_STATE, "atoms and neighbour list must be set");
  if (atomic() || model != 4) return fail(EPH_ERR_UNSUPPORTED, "the sweep forms belong to the PRL model");
  StepParams p = step_params();
  p.xi_mode = (flags & EPH_RANDOM) ? (xi_mode_next ? 1 : 2) : 0;
  const AtomsView a = atoms_view();
  const WorkView w = work_view();
  EPH_HIP(this, launch_pack_pv(a, w, stream));
  EPH_HIP(this, launch_density_w(a, tabs, p, w, stream, form_density));
  EPH_HIP(this, launch_ghost_rec(a, w, stream));
  StepParams pf = p;
  pf.flags |= EPH_NOFRICTION | EPH_NORANDOM;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t he = hipEventCreate(&e0);
  if (he == hipSuccess) he = hipEventCreate(&e1);
  constexpr int nf = kSweepForms;
  double t[2 * nf] = {};
  const bool lds_ok = lds_tables_possible(tabs);  // forms 1..3 need the tables in LDS
  auto runs = [&](int v) { return lds_ok || v == 0 || v >= kChunkOcc4; };
  for (int phase = 0; phase < 2 && he == hipSuccess; ++phase) {
    for (int v = 0; v < nf && he == hipSuccess; ++v) {
      if (!runs(v)) continue;
      for (int rep = 0; rep < 4 && he == hipSuccess; ++rep) {  // the first launch warms up
        he = hipEventRecord(e0, stream);
        if (he == hipSuccess)
          he = phase == 0 ? launch_density_w(a, tabs, p, w, stream, v)
                          : launch_prl_forces(a, tabs, grid->v, pf, w, stream, v);
        if (he == hipSuccess) he = hipEventRecord(e1, stream);
        if (he == hipSuccess) he = hipEventSynchronize(e1);
        float m = 0;
        if (he == hipSuccess) he = hipEventElapsedTime(&m, e0, e1);
        if (rep > 0) t[nf * phase + v] += m / 3.0;
      }
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (he != hipSuccess) return fail_hip(he, "autotune");
  auto best = [&](const double* tt) {
    int b = 0;
    for (int v = 1; v < nf; ++v)
      if (runs(v) && tt[v] < tt[b]) b = v;
    return b;
  };
  form_density = best(t);
  form_force = best(t + nf);
  if (forms) { forms[0] = form_density; forms[1] = form_force; }
  if (ms) std::copy(t, t + 2 * nf, ms);
  return EPH_OK;
}

// final_integrate then end_of_step with nothing in between (the harness's
// own Verlet step): where the bucketed deposit runs, one pass does both
int eph_context::dev_final_integrate_end_of_step(long long ntimestep) {
  if (atomic()) {
    int rc = dev_final_integrate();
    return rc != EPH_OK ? rc : dev_end_of_step(ntimestep);
  }
  int rc;
  if (bucket_deposit() && !(flags & EPH_NOINT)) {
    bool kicked = false;
    if ((rc = dev_deposit(true, &kicked)) != EPH_OK) return rc;
    if (!kicked) return fail(EPH_ERR_STATE, "the fused final kick did not run");
  } else {
    if ((rc = dev_final_integrate()) != EPH_OK) return rc;
    if ((rc = dev_deposit(false, nullptr)) != EPH_OK) return rc;
  }
  return end_of_step_after_deposit(ntimestep);
}

int eph_context::dev_end_of_step(long long ntimestep) {
  if (atomic()) return eph::atomic_end_of_step(this);
  int rc = dev_deposit(false, nullptr);
  if (rc != EPH_OK) return rc;
  return end_of_step_after_deposit(ntimestep);
}

// the FDM solve and the heat-map output of end_of_step (fix_eph.cpp:392-399)
int eph_context::end_of_step_after_deposit(long long ntimestep) {
  int rc = EPH_OK;
  hipEvent_t e;
  if (flags & EPH_FDM) {
    timed_begin("fdm_solve", &e);
    rc = grid->solve(nullptr);
    if (rc != EPH_OK) return fail(rc, grid->err);
    timed_end("fdm_solve", e);
  }
  if (write_output && T_freq > 0 && (ntimestep % T_freq) == 0) {
    std::vector<double> T;
    rc = grid->download(0, T);
    if (rc != EPH_OK) return fail(rc, grid->err);
    if (!save_grid_temperature(grid->h, T, T_out, (int)(ntimestep / T_freq)))
      return fail(EPH_ERR_FILE, "cannot write " + T_out);
  }
  return EPH_OK;
}

int eph_grid::bucket_bytes_ok(size_t G) {
  EPH_HIP(this, bcount.reserve(G));  // zero-filled when first made
  EPH_HIP(this, bitems.reserve(G * kBucket));
  EPH_HIP(this, bqueue.reserve(G));
  EPH_HIP(this, bnq.reserve(1));
  return EPH_OK;
}

// the deposit goes through per-cell buckets when there is one and the grid
// holds a few atoms per cell on average
bool eph_context::bucket_deposit() const {
  return (flags & (EPH_FRICTION | EPH_RANDOM)) &&
         static_cast<double>(nlocal) <= 6.0 * static_cast<double>(grid->h.ntotal());
}

// the deposit half of end_of_step (fix_eph.cpp:364-404): dT_e and the tally.
// kick: apply final_integrate's kick in the same pass where the bucketed
// deposit runs (*kicked tells whether it did)
int eph_context::dev_deposit(bool kick, bool* kicked_out) {
  if (!have_atoms) return fail(EPH_ERR_STATE, "no atoms");
  bool kicked = false;
  if (kicked_out) *kicked_out = false;
  if (flags & EPH_NOINT) kick = false;
  const bool fric = flags & EPH_FRICTION, rnd = flags & EPH_RANDOM;
  const int nb = nblocks_of(nlocal);
  const size_t G = grid->h.ntotal();
  hipEvent_t e;
  // a few atoms per cell on average: per-cell buckets (three launches)
  if (bucket_deposit()) {
    EPH_HIP(this, grid->dep.reserve(nlocal, 0, nb));
    if (int rc = grid->bucket_bytes_ok(G)) return fail(rc, grid->err);
    DepositView d = grid->dep.view(true);
    BucketView b;
    b.count = grid->bcount.p; b.items = grid->bitems.p; b.queue = grid->bqueue.p; b.nqueue = grid->bnq.p;
    timed_begin("deposit", &e);
    EPH_HIP(this, launch_deposit_bucket(atoms_view(), grid->v, step_params(), work_view(), d, b, fric, rnd,
                                        kick, scal.p, scal.p + (defer_tally ? 2 : 1), stream));
    timed_end(kick ? "final_integrate+deposit" : "deposit", e);
    kicked = kick;
    if (kicked_out) *kicked_out = kicked;
    return EPH_OK;
  }
  if (fric || rnd) {
    EPH_HIP(this, grid->dep.reserve(nlocal, G, nb));
    DepositView d = grid->dep.view(true);
    timed_begin("deposit", &e);
    EPH_HIP(this, launch_fill_int(d.count, (int)G + 1, 0, stream));
    EPH_HIP(this, launch_deposit_atoms(atoms_view(), grid->v, step_params(), work_view(), d, nb, stream));
    DepositView dc = d;
    if (!fric) dc.v1 = nullptr;
    if (!rnd) dc.v2 = nullptr;
    EPH_HIP(this, launch_deposit_cells(grid->v, dc, nlocal, grid->dep.qbuf.p, stream));
    EPH_HIP(this, launch_energy_reduce(d, nb, scal.p, scal.p + (defer_tally ? 2 : 1), stream));
    timed_end("deposit", e);
  } else {
    EPH_HIP(this, launch_fill_double(scal.p, 1, 0.0, stream));
  }
  return EPH_OK;
}

int eph_context::check_flags() {
  int fl = 0, gf = 0;
  EPH_HIP(this, hipMemcpyAsync(&fl, flags_dev.p, sizeof(int), hipMemcpyDeviceToHost, stream));
  if (grid) EPH_HIP(this, hipMemcpyAsync(&gf, grid->nonfinite.p, sizeof(int), hipMemcpyDeviceToHost, stream));
  EPH_HIP(this, hipStreamSynchronize(stream));
  if (gf & 2) return fail(EPH_ERR_NONFINITE, "T_e grid became non-finite (FDM instability)");
  return EPH_OK;
}

// ===================================================================== C ABI
extern "C" {

int eph_setmask(void) {
  return EPH_HOOK_INITIAL_INTEGRATE | EPH_HOOK_POST_FORCE | EPH_HOOK_FINAL_INTEGRATE |
         EPH_HOOK_END_OF_STEP;
}

const char* eph_last_error(const eph_context* ctx) {
  return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

int eph_create(eph_context** out, int narg, const char* const* arg, int ntypes,
               const double boxlo[3], const double boxhi[3], double dt, double boltz,
               double ftm2v, int device) {
  if (!out) return EPH_ERR_ARGS;
  *out = nullptr;
  auto fail = [&](int code, const std::string& m) {
    g_create_error = m;
    return code;
  };
  if (!arg || narg < 3) return fail(EPH_ERR_ARGS, "Illegal fix eph command: too few arguments");
  const bool atomic_style = arg[2] && std::strcmp(arg[2], "eph/atomic") == 0;
  if (!atomic_style && narg < 18) return fail(EPH_ERR_ARGS, "Illegal fix eph command: too few arguments");
  if (!(dt > 0)) return fail(EPH_ERR_ARGS, "timestep must be positive");
  auto* c = new eph_context();
  auto bail = [&](int code, const std::string& m) {
    g_create_error = m;
    eph_destroy(c);
    return code;
  };
  c->device = device;
  // arguments and input files are checked on the host first (no device needed)
  // fix_eph.cpp:85-132
  // fix styles with FixEPH's command layout: eph (fix_eph.cpp), eph/coloured
  // (fix_eph_coloured.cpp, PRL whatever arg[5] says, :652), eph/coloured/exp
  // and eph/coloured/exp/v1 (arg[5] = tau0, fix_eph_coloured_exp.cpp:113)
  const std::string style = arg[2] ? arg[2] : "";
  GridHost gh;
  std::string err;
  if (atomic_style) {
    c->colour = 4;
    c->at.reset(new eph::AtomicState());
    if (int rc = eph::atomic_parse(c, narg, arg, ntypes, err)) return bail(rc, err);
  } else {
  if (style == "eph" || style == "eph/gpu") c->colour = 0;
  else if (style == "eph/coloured") c->colour = 1;
  else if (style == "eph/coloured/exp") c->colour = 2;
  else if (style == "eph/coloured/exp/v1") c->colour = 3;
  else return bail(EPH_ERR_ARGS, "unknown fix style " + style);
  c->seed = std::atoi(arg[3]);
  c->flags = (int)std::strtol(arg[4], nullptr, 0);
  if (c->colour >= 2) {
    c->tau0 = std::atof(arg[5]);
    if (!(c->tau0 > 0)) return bail(EPH_ERR_ARGS, "tau0 must be positive");
    c->model = 4;
  } else {
    c->model = c->colour == 1 ? 4 : std::atoi(arg[5]);
  }
  if (c->model == 3)
    return bail(EPH_ERR_UNSUPPORTED,
                "model 3 (PRLCM) indexes rho by jtype - i in the reference (fix_eph.cpp:600): undefined");
  if (c->model != 1 && c->model != 2 && c->model != 4 && c->model != -1)
    return bail(EPH_ERR_ARGS, "invalid eph model");
  const double v_rho = std::atof(arg[6]), v_Ce = std::atof(arg[7]);
  const double v_kappa = std::atof(arg[8]), v_Te = std::atof(arg[9]);
  const int nx = std::atoi(arg[10]), ny = std::atoi(arg[11]), nz = std::atoi(arg[12]);
  // fix_eph.cpp:134-161
  if (std::strcmp("NULL", arg[13]) == 0) {
    if (nx < 1 || ny < 1 || nz < 1) return bail(EPH_ERR_ARGS, "FixEPH: non-positive grid values");
    const double box[6] = {boxlo[0], boxhi[0], boxlo[1], boxhi[1], boxlo[2], boxhi[2]};
    if (!make_grid_const(nx, ny, nz, box, v_Te, v_Ce, v_rho, v_kappa, gh, err)) return bail(EPH_ERR_ARGS, err);
    c->T_state = "T.restart";
  } else {
    if (!load_grid(arg[13], gh, err)) return bail(err.rfind("cannot open", 0) == 0 ? EPH_ERR_FILE : EPH_ERR_FORMAT, err);
    c->T_state = std::string(arg[13]) + ".restart";
  }
  c->T_freq = std::atoi(arg[14]);
  if (c->T_freq > 0) c->T_out = arg[15];
  gh.dt = dt;
  // fix_eph.cpp:171-197
  c->ntypes = ntypes;
  if (ntypes < 1 || ntypes > narg - 17)
    return bail(EPH_ERR_ARGS, "Fix eph: number of types larger than provided in fix");
  if (!load_beta(arg[16], c->beta, err))
    return bail(err.rfind("cannot open", 0) == 0 ? EPH_ERR_FILE : EPH_ERR_FORMAT, err);
  c->type_map.assign(ntypes, INT_MAX);
  for (int i = 0; i < ntypes; ++i) {
    for (size_t j = 0; j < c->beta.names.size(); ++j)
      if (c->beta.names[j] == arg[17 + i]) c->type_map[i] = (int)j;
    if (c->type_map[i] == INT_MAX) return bail(EPH_ERR_ARGS, "Fix eph: elements not found in input file");
  }
  }  // fix eph layout
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return bail(EPH_ERR_NO_DEVICE, "no HIP device");
  if (device < 0 || device >= ndev) return bail(EPH_ERR_NO_DEVICE, "HIP device ordinal out of range");
  if (hipSetDevice(device) != hipSuccess) return bail(EPH_ERR_NO_DEVICE, "hipSetDevice failed");
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(EPH_ERR_HIP, "stream creation failed");
  if (!c->atomic()) {
    c->grid = new eph_grid();
    c->grid->h = gh;
    if (int rc = grid_init_device(c->grid, device, c->stream)) return bail(rc, c->grid->err);
  }
  // fix_eph.cpp:199-204
  c->dt = dt;
  c->boltz = boltz;
  c->ftm2v = ftm2v;
  c->eta = std::sqrt(2.0 * boltz / dt);
  c->dtv = dt;
  c->dtf = 0.5 * dt * ftm2v;
  if (c->colour >= 2) c->zeta = 1.0 - std::exp(-dt / c->tau0);  // fix_eph_coloured_exp.cpp:190
  // tables
  const int ne = (int)c->beta.names.size();
  const int nr = c->beta.n_rho, nb = c->beta.n_beta;
  if (c->model == 2 && ntypes > ne)
    return bail(EPH_ERR_ARGS, "model 2 evaluates rho(r) of element jtype - 1: the .beta file needs "
                              "at least ntypes elements (fix_eph.cpp:537)");
  std::vector<Cubic> r2((size_t)ne * nr), al((size_t)ne * nb), be((size_t)ne * nb), r1((size_t)ne * nr);
  for (int e = 0; e < ne; ++e) {
    std::copy(c->beta.rho[e].knots.begin(), c->beta.rho[e].knots.end(), r1.begin() + (size_t)e * nr);
    // rho(r^2) knots interleaved by element: the elements' intervals at one r^2
    // share a cache line (knot k of element e at k * ne + e)
    for (int k = 0; k < nr; ++k) r2[(size_t)k * ne + e] = c->beta.rho_r_sq[e].knots[k];
    std::copy(c->beta.alpha[e].knots.begin(), c->beta.alpha[e].knots.end(), al.begin() + (size_t)e * nb);
    std::copy(c->beta.beta[e].knots.begin(), c->beta.beta[e].knots.end(), be.begin() + (size_t)e * nb);
  }
  if (c->t_rho_r2.reserve(r2.size()) || c->t_alpha.reserve(al.size()) || c->t_beta.reserve(be.size()) ||
      c->t_rho_r.reserve(r1.size()) ||
      c->scal.reserve(4) || c->flags_dev.reserve(1) || c->mass.reserve(ntypes + 1))
    return bail(EPH_ERR_OOM, "device allocation failed");
  c->mass_host.assign(ntypes + 1, 1.0);
  if (hipMemcpyAsync(c->t_rho_r2.p, r2.data(), r2.size() * sizeof(Cubic), hipMemcpyHostToDevice, c->stream) ||
      hipMemcpyAsync(c->t_alpha.p, al.data(), al.size() * sizeof(Cubic), hipMemcpyHostToDevice, c->stream) ||
      hipMemcpyAsync(c->t_beta.p, be.data(), be.size() * sizeof(Cubic), hipMemcpyHostToDevice, c->stream) ||
      hipMemcpyAsync(c->t_rho_r.p, r1.data(), r1.size() * sizeof(Cubic), hipMemcpyHostToDevice, c->stream) ||
      hipMemcpyAsync(c->mass.p, c->mass_host.data(), (ntypes + 1) * sizeof(double), hipMemcpyHostToDevice,
                     c->stream) ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return bail(EPH_ERR_HIP, "table upload failed");
  TablesView& t = c->tabs;
  t.n_elem = ne; t.n_rho = nr; t.inv_dr2 = c->beta.rho_r_sq[0].inv_dx; t.rho_r2 = c->t_rho_r2.p;
  t.n_beta = nb; t.inv_drho = c->beta.alpha[0].inv_dx; t.alpha = c->t_alpha.p; t.beta = c->t_beta.p;
  t.rc2 = c->beta.r_cutoff_sq; t.rho_cut = c->beta.rho_cutoff;
  t.rho_r = c->t_rho_r.p; t.inv_dr = c->beta.rho[0].inv_dx;
  if (c->atomic())
    if (int rc = eph::atomic_upload(c)) return bail(rc, c->err);
  *out = c;
  return EPH_OK;
}

void eph_destroy(eph_context* ctx) {
  if (!ctx) return;
  if (ctx->stream) {
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
  }
  for (auto& t : ctx->pending) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
  if (ctx->grid) eph_grid_destroy(ctx->grid);
  if (ctx->pinned) (void)hipHostFree(ctx->pinned);
  hipStream_t s = ctx->stream;
  delete ctx;
  if (s) (void)hipStreamDestroy(s);
}

int eph_init(eph_context* ctx, int dimension, int nonperiodic, int triclinic, double* cutoff) {
  if (!ctx) return EPH_ERR_ARGS;
  if (dimension == 2) return ctx->fail(EPH_ERR_ARGS, "Cannot use fix eph with 2d simulation");
  if (nonperiodic != 0) return ctx->fail(EPH_ERR_ARGS, "Cannot use nonperiodic boundares with fix eph");
  if (triclinic) return ctx->fail(EPH_ERR_ARGS, "Cannot use fix eph with triclinic box");
  if (cutoff) *cutoff = ctx->beta.r_cutoff;
  return eph_reset_dt(ctx, ctx->dt);
}

int eph_reset_dt(eph_context* ctx, double dt) {
  if (!ctx || !(dt > 0)) return EPH_ERR_ARGS;
  ctx->dt = dt;
  ctx->eta = std::sqrt(2.0 * ctx->boltz / dt);
  ctx->dtv = dt;
  ctx->dtf = 0.5 * dt * ctx->ftm2v;
  if (ctx->colour >= 2) ctx->zeta = 1.0 - std::exp(-dt / ctx->tau0);  // fix_eph_coloured_exp.cpp:685
  if (ctx->grid) ctx->grid->h.dt = dt;
  return EPH_OK;
}

int eph_get_filter_state(eph_context* ctx, int which, double* out) {
  if (!ctx || !out || which < 0 || which > 1) return EPH_ERR_ARGS;
  if (ctx->colour < 2) return ctx->fail(EPH_ERR_STATE, "the fix style keeps no filter state");
  const int n = ctx->nlocal;
  if (n == 0) return EPH_OK;
  std::vector<double4> h(n);
  const eph::DevArray<double4>& src = which == 0 ? ctx->colA : ctx->colB;
  EPH_HIP(ctx, hipMemcpyAsync(h.data(), src.p, n * sizeof(double4), hipMemcpyDeviceToHost, ctx->stream));
  EPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < n; ++i) { out[3 * i] = h[i].x; out[3 * i + 1] = h[i].y; out[3 * i + 2] = h[i].z; }
  return EPH_OK;
}

int eph_set_filter_state(eph_context* ctx, int which, const double* in) {
  if (!ctx || !in || which < 0 || which > 1) return EPH_ERR_ARGS;
  if (ctx->colour < 2) return ctx->fail(EPH_ERR_STATE, "the fix style keeps no filter state");
  const int n = ctx->nlocal;
  if (n == 0) return EPH_OK;
  int rc = ctx->grow_filter_state((size_t)n + ctx->nghost);
  if (rc != EPH_OK) return rc;
  std::vector<double4> h(n);
  for (int i = 0; i < n; ++i) h[i] = make_double4(in[3 * i], in[3 * i + 1], in[3 * i + 2], 0.0);
  eph::DevArray<double4>& dst = which == 0 ? ctx->colA : ctx->colB;
  EPH_HIP(ctx, hipMemcpyAsync(dst.p, h.data(), n * sizeof(double4), hipMemcpyHostToDevice, ctx->stream));
  EPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return EPH_OK;
}

int eph_set_masses(eph_context* ctx, const double* mass, int ntypes) {
  if (!ctx || !mass || ntypes != ctx->ntypes) return EPH_ERR_ARGS;
  ctx->mass_host.assign(mass, mass + ntypes + 1);
  EPH_HIP(ctx, hipMemcpyAsync(ctx->mass.p, ctx->mass_host.data(), (ntypes + 1) * sizeof(double),
                              hipMemcpyHostToDevice, ctx->stream));
  EPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return EPH_OK;
}

int eph_set_groupbit(eph_context* ctx, int groupbit) {
  if (!ctx) return EPH_ERR_ARGS;
  ctx->groupbit = groupbit;
  ctx->drop_candidates();
  return EPH_OK;
}

int eph_set_candidate_margin(eph_context* ctx, double delta) {
  if (!ctx || !(delta == delta)) return EPH_ERR_ARGS;
  ctx->cand_delta = delta;
  ctx->drop_candidates();
  return EPH_OK;
}

int eph_get_params(const eph_context* ctx, int* flags, int* model, int* seed, double* r_cutoff,
                   double* rho_cutoff, double* eta_factor, double* dtf) {
  if (!ctx) return EPH_ERR_ARGS;
  if (flags) *flags = ctx->flags;
  if (model) *model = ctx->model;
  if (seed) *seed = ctx->seed;
  if (r_cutoff) *r_cutoff = ctx->beta.r_cutoff;
  if (rho_cutoff) *rho_cutoff = ctx->beta.rho_cutoff;
  if (eta_factor) *eta_factor = ctx->eta;
  if (dtf) *dtf = ctx->dtf;
  return EPH_OK;
}

int eph_set_atoms(eph_context* ctx, int nlocal, int nghost, const double* x, const double* v,
                  const int* type, const int* mask, const int* tag) {
  if (!ctx) return EPH_ERR_ARGS;
  return ctx->set_atoms(nlocal, nghost, x, v, type, mask, tag);
}

int eph_set_ghost_map(eph_context* ctx, int nghost, const int* owner, const double* shift) {
  if (!ctx || (nghost > 0 && !owner)) return EPH_ERR_ARGS;
  return ctx->set_ghost_map(nghost, owner, shift);
}

int eph_set_neighbors(eph_context* ctx, int nlocal, const int* off, const int* idx, int neighmask) {
  if (!ctx || !off || (!idx && nlocal > 0 && off[nlocal] > 0)) return EPH_ERR_ARGS;
  return ctx->set_neighbors(nlocal, off, idx, neighmask);
}

int eph_set_noise(eph_context* ctx, const double* xi) {
  if (!ctx) return EPH_ERR_ARGS;
  if (!xi) {
    ctx->xi_mode_next = 0;
    return EPH_OK;
  }
  if (!ctx->have_atoms) return ctx->fail(EPH_ERR_STATE, "no atoms");
  if (int rc = ctx->h2d(ctx->xi_in.p, xi, 3 * (size_t)ctx->nlocal * sizeof(double))) return rc;
  ctx->xi_mode_next = 1;
  return EPH_OK;
}

int eph_set_timestep(eph_context* ctx, long long ntimestep) {
  if (!ctx) return EPH_ERR_ARGS;
  ctx->step = ntimestep;
  return EPH_OK;
}

// ---------------------------------------------------------- host-array hooks
namespace {
constexpr size_t kStageChunk = size_t(8) << 20;  // bytes per pinned staging chunk
// smaller copies go through the runtime directly: a context that only ever
// moves small arrays (the replicas of an ensemble) starts no staging threads
constexpr size_t kStageMin = size_t(4) << 20;
}

// the staging path for a copy of `bytes`, or null for the runtime's own copy
// (small copies, or pinned memory not available)
static HostStager* stager_for(eph_context* c, size_t bytes) {
  if (bytes < kStageMin || c->stage_failed) return nullptr;
  if (!c->stager) {
    auto s = std::make_unique<HostStager>();
    if (s->init(HostStager::default_threads(), kStageChunk) != hipSuccess) {
      (void)hipGetLastError();
      c->stage_failed = true;
      return nullptr;
    }
    c->stager = std::move(s);
  }
  return c->stager.get();
}

int eph_context::h2d(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return EPH_OK;
  if (HostStager* st = stager_for(this, bytes)) {
    EPH_HIP(this, st->to_device(dst, src, bytes, stream));
    return EPH_OK;
  }
  EPH_HIP(this, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
  EPH_HIP(this, hipStreamSynchronize(stream));
  return EPH_OK;
}

int eph_context::d2h(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return EPH_OK;
  if (HostStager* st = stager_for(this, bytes)) {
    EPH_HIP(this, st->to_host(dst, src, bytes, stream));
    return EPH_OK;
  }
  EPH_HIP(this, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream));
  EPH_HIP(this, hipStreamSynchronize(stream));
  return EPH_OK;
}

static int upload_rows(eph_context* c, const double* rows, int n, double4* dst) {
  c->pv_fresh = false;  // positions or velocities written behind the packed records
  if (n == 0) return EPH_OK;
  if (int rc = c->h2d(c->stage.p, rows, 3 * (size_t)n * sizeof(double))) return rc;
  EPH_HIP(c, launch_rows_into_d4(c->stage.p, n, dst, c->stream));
  return EPH_OK;
}

static int download_rows(eph_context* c, const double4* src, int n, double* rows, int stride4 = 1) {
  if (n == 0) return EPH_OK;
  EPH_HIP(c, launch_d4_into_rows(src, n, c->stage.p, c->stream, stride4));
  return c->d2h(rows, c->stage.p, 3 * (size_t)n * sizeof(double));
}

static int upload_force(eph_context* c, const double* f) {
  if (c->nlocal == 0) return EPH_OK;
  if (int rc = c->h2d(c->stage.p, f, 3 * (size_t)c->nlocal * sizeof(double))) return rc;
  EPH_HIP(c, launch_rows_into_soa(c->stage.p, c->nlocal, c->f.p, c->fstride, c->stream));
  return EPH_OK;
}

int eph_initial_integrate(eph_context* ctx, double* x, double* v, const double* f) {
  if (!ctx || !x || !v || !f) return EPH_ERR_ARGS;
  if (!ctx->have_atoms) return ctx->fail(EPH_ERR_STATE, "no atoms");
  if (ctx->flags & EPH_NOINT) return EPH_OK;
  int rc;
  if ((rc = upload_rows(ctx, x, ctx->nlocal, ctx->pos.p))) return rc;
  if ((rc = upload_rows(ctx, v, ctx->nlocal, ctx->vel.p))) return rc;
  if ((rc = upload_force(ctx, f))) return rc;
  if ((rc = ctx->dev_initial_integrate())) return rc;
  if ((rc = download_rows(ctx, ctx->pos.p, ctx->nlocal, x))) return rc;
  EPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if ((rc = download_rows(ctx, ctx->vel.p, ctx->nlocal, v))) return rc;
  EPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return EPH_OK;
}

int eph_post_force(eph_context* ctx, const double* x, const double* v, double* f) {
  if (!ctx || !x || !v || !f) return EPH_ERR_ARGS;
  if (!ctx->have_atoms) return ctx->fail(EPH_ERR_STATE, "no atoms");
  const int nall = ctx->nlocal + ctx->nghost;
  int rc;
  if ((rc = upload_rows(ctx, x, nall, ctx->pos.p))) return rc;
  if ((rc = upload_rows(ctx, v, nall, ctx->vel.p))) return rc;
  if ((rc = upload_force(ctx, f))) return rc;
  if ((rc = ctx->dev_post_force())) return rc;
  if (ctx->nlocal) {
    EPH_HIP(ctx, launch_soa_into_rows(ctx->f.p, ctx->nlocal, ctx->fstride, ctx->stage.p, ctx->stream));
    if ((rc = ctx->d2h(f, ctx->stage.p, 3 * (size_t)ctx->nlocal * sizeof(double)))) return rc;
  }
  EPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return ctx->check_flags();
}

int eph_final_integrate(eph_context* ctx, double* v, const double* f) {
  if (!ctx || !v || !f) return EPH_ERR_ARGS;
  if (!ctx->have_atoms) return ctx->fail(EPH_ERR_STATE, "no atoms");
  if (ctx->flags & EPH_NOINT) return EPH_OK;
  int rc;
  if ((rc = upload_rows(ctx, v, ctx->nlocal, ctx->vel.p))) return rc;
  if ((rc = upload_force(ctx, f))) return rc;
  if ((rc = ctx->dev_final_integrate())) return rc;
  if ((rc = download_rows(ctx, ctx->vel.p, ctx->nlocal, v))) return rc;
  EPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return EPH_OK;
}

int eph_end_of_step(eph_context* ctx, const double* x, const double* v, long long ntimestep) {
  if (!ctx || !x || !v) return EPH_ERR_ARGS;
  if (!ctx->have_atoms) return ctx->fail(EPH_ERR_STATE, "no atoms");
  int rc;
  if ((rc = upload_rows(ctx, x, ctx->nlocal, ctx->pos.p))) return rc;
  if ((rc = upload_rows(ctx, v, ctx->nlocal, ctx->vel.p))) return rc;
  if ((rc = ctx->dev_end_of_step(ntimestep))) return rc;
  EPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return ctx->check_flags();
}

int eph_post_force_local(eph_context* ctx, const double* x, const double* v, double* halo) {
  if (!ctx || !x || !v || (!halo && ctx->nlocal > 0)) return EPH_ERR_ARGS;
  if (!ctx->have_atoms) return ctx->fail(EPH_ERR_STATE, "no atoms");
  const int nall = ctx->nlocal + ctx->nghost;
  int rc;
  if ((rc = upload_rows(ctx, x, nall, ctx->pos.p))) return rc;
  if ((rc = upload_rows(ctx, v, nall, ctx->vel.p))) return rc;
  if ((rc = ctx->dev_post_force_local())) return rc;
  if (ctx->nlocal) {
    EPH_HIP(ctx, launch_halo_pack(ctx->rec.p, ctx->nlocal, ctx->stage.p, ctx->stream));
    if ((rc = ctx->d2h(halo, ctx->stage.p, (size_t)EPH_HALO_COLS * ctx->nlocal * sizeof(double)))) return rc;
  }
  EPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return ctx->check_flags();
}

int eph_post_force_ghosts(eph_context* ctx, const double* halo_ghost, double* f) {
  if (!ctx || !f || (!halo_ghost && ctx->nghost > 0)) return EPH_ERR_ARGS;
  if (!ctx->have_atoms) return ctx->fail(EPH_ERR_STATE, "no atoms");
  int rc;
  if (ctx->nghost) {
    // the ghosts' records from their exchanged xi, rho, w, with alpha from rho
    // as fix_eph.cpp:757 and :800 evaluate it for a neighbour; enqueued before
    // the force upload reuses the staging buffer
    if ((rc = ctx->h2d(ctx->stage.p, halo_ghost, (size_t)EPH_HALO_COLS * ctx->nghost * sizeof(double))))
      return rc;
    EPH_HIP(ctx, launch_ghost_halo(ctx->atoms_view(), ctx->tabs, ctx->work_view(), ctx->stage.p,
                                   ctx->stream));
  }
  if ((rc = upload_force(ctx, f))) return rc;
  if ((rc = ctx->dev_post_force_ghosts())) return rc;
  if (ctx->nlocal) {
    EPH_HIP(ctx, launch_soa_into_rows(ctx->f.p, ctx->nlocal, ctx->fstride, ctx->stage.p, ctx->stream));
    if ((rc = ctx->d2h(f, ctx->stage.p, 3 * (size_t)ctx->nlocal * sizeof(double)))) return rc;
  }
  EPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return ctx->check_flags();
}

int eph_end_of_step_local(eph_context* ctx, const double* x, const double* v, double* dT,
                          double* E_local) {
  if (!ctx || !x || !v || !dT || !E_local) return EPH_ERR_ARGS;
  if (!ctx->have_atoms) return ctx->fail(EPH_ERR_STATE, "no atoms");
  if (ctx->atomic()) return ctx->fail(EPH_ERR_UNSUPPORTED, "eph/atomic has no T_e grid");
  int rc;
  if ((rc = upload_rows(ctx, x, ctx->nlocal, ctx->pos.p))) return rc;
  if ((rc = upload_rows(ctx, v, ctx->nlocal, ctx->vel.p))) return rc;
  ctx->defer_tally = true;
  rc = ctx->dev_deposit(false, nullptr);
  ctx->defer_tally = false;
  if (rc) return rc;
  std::vector<double> d;
  if ((rc = ctx->grid->download(1, d))) return ctx->fail(rc, ctx->grid->err);
  std::copy(d.begin(), d.end(), dT);
  EPH_HIP(ctx, hipMemcpyAsync(E_local, ctx->scal.p, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  EPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return EPH_OK;
}

int eph_end_of_step_reduced(eph_context* ctx, const double* dT, double E, long long ntimestep,
                            int write_output) {
  if (!ctx || !dT) return EPH_ERR_ARGS;
  if (ctx->atomic()) return ctx->fail(EPH_ERR_UNSUPPORTED, "eph/atomic has no T_e grid");
  eph_grid* g = ctx->grid;
  const size_t G = g->h.ntotal();
  double sc[2] = {0.0, 0.0};
  EPH_HIP(ctx, hipMemcpyAsync(sc, ctx->scal.p, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  EPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  sc[0] = E;
  sc[1] += E;  // Ee += the sum over ranks (fix_eph.cpp:402-404)
  EPH_HIP(ctx, hipMemcpyAsync(ctx->scal.p, sc, 2 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  if (int rc = ctx->h2d(g->dT.p, dT, G * sizeof(double))) return rc;
  ctx->write_output = write_output != 0;
  const int rc = ctx->end_of_step_after_deposit(ntimestep);
  ctx->write_output = true;
  if (rc) return rc;
  EPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return ctx->check_flags();
}

int eph_post_run(eph_context* ctx) {
  if (!ctx) return EPH_ERR_ARGS;
  if (ctx->atomic()) return EPH_OK;  // fix_eph_atomic.cpp:946-948 writes nothing
  std::vector<double> T;
  int rc;
  if ((rc = ctx->grid->sync_params_from_device())) return ctx->fail(rc, ctx->grid->err);
  if ((rc = ctx->grid->download(0, T))) return ctx->fail(rc, ctx->grid->err);
  if (!save_grid_state(ctx->grid->h, T, ctx->T_state)) return ctx->fail(EPH_ERR_FILE, "cannot write " + ctx->T_state);
  return EPH_OK;
}

int eph_compute_vector(eph_context* ctx, int i, double* out) {
  if (!ctx || !out) return EPH_ERR_ARGS;
  if (ctx->atomic()) {  // fix_eph_atomic.cpp:855-864
    *out = i == 1 ? ctx->at->Te : ctx->at->Ee;
    return EPH_OK;
  }
  if (i == 1) return eph_grid_T_mean(ctx->grid, out);
  EPH_HIP(ctx, hipMemcpyAsync(out, ctx->scal.p + 1, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  EPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return EPH_OK;
}

int eph_last_energy(eph_context* ctx, double* E_local) {
  if (!ctx || !E_local) return EPH_ERR_ARGS;
  EPH_HIP(ctx, hipMemcpyAsync(E_local, ctx->scal.p, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  EPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return EPH_OK;
}

int eph_rho_cutoff_warning(eph_context* ctx) {
  if (!ctx) return EPH_ERR_ARGS;
  int fl = 0;
  EPH_HIP(ctx, hipMemcpyAsync(&fl, ctx->flags_dev.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  EPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return fl & 1;
}

int eph_get_peratom(eph_context* ctx, double* array) {
  if (!ctx || !array) return EPH_ERR_ARGS;
  if (!ctx->have_atoms) return ctx->fail(EPH_ERR_STATE, "no atoms");
  if (ctx->nlocal == 0) return EPH_OK;
  if (ctx->atomic()) {
    EPH_HIP(ctx, ctx->stage.reserve(std::max<size_t>(ctx->stage.n, 12 * (size_t)ctx->nlocal)));
    EPH_HIP(ctx, eph::launch_peratom_atomic(ctx->atoms_view(), ctx->tabs, ctx->work_view(),
                                            ctx->at->view(), ctx->stage.p, ctx->stream));
    return ctx->d2h(array, ctx->stage.p, 12 * (size_t)ctx->nlocal * sizeof(double));
  }
  EPH_HIP(ctx, ctx->stage.reserve(std::max<size_t>(ctx->stage.n, 8 * (size_t)ctx->nlocal)));
  EPH_HIP(ctx, launch_peratom(ctx->atoms_view(), ctx->tabs, ctx->work_view(), ctx->stage.p, ctx->stream));
  return ctx->d2h(array, ctx->stage.p, 8 * (size_t)ctx->nlocal * sizeof(double));
}

int eph_peratom_cols(const eph_context* ctx) {
  if (!ctx) return EPH_ERR_ARGS;
  return ctx->atomic() ? 12 : 8;
}

int eph_get_atomic_energy(eph_context* ctx, double* E) {
  if (!ctx || !E) return EPH_ERR_ARGS;
  if (!ctx->atomic()) return ctx->fail(EPH_ERR_STATE, "per-atom energies belong to eph/atomic");
  if (ctx->nlocal)
    EPH_HIP(ctx, hipMemcpyAsync(E, ctx->at->E.p, ctx->nlocal * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  EPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return EPH_OK;
}

int eph_set_atomic_energy(eph_context* ctx, const double* E) {
  if (!ctx || !E) return EPH_ERR_ARGS;
  if (!ctx->atomic()) return ctx->fail(EPH_ERR_STATE, "per-atom energies belong to eph/atomic");
  if (!ctx->have_atoms) return ctx->fail(EPH_ERR_STATE, "no atoms");
  if (ctx->nlocal) {
    EPH_HIP(ctx, hipMemcpyAsync(ctx->at->E.p, E, ctx->nlocal * sizeof(double), hipMemcpyHostToDevice,
                                ctx->stream));
    EPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  ctx->at->ready = true;
  return EPH_OK;
}

int eph_get_state(eph_context* ctx, int which, double* out) {
  if (!ctx || !out) return EPH_ERR_ARGS;
  if (!ctx->have_atoms) return ctx->fail(EPH_ERR_STATE, "no atoms");
  const int nall = ctx->nlocal + ctx->nghost;
  hipStream_t s = ctx->stream;
  switch (which) {
    case 0: EPH_HIP(ctx, launch_d4_component(ctx->rec_field(1), nall, 3, ctx->stage.p, s, 4));
      EPH_HIP(ctx, hipMemcpyAsync(out, ctx->stage.p, nall * sizeof(double), hipMemcpyDeviceToHost, s)); break;
    case 5: EPH_HIP(ctx, launch_d4_component(ctx->rec_field(2), nall, 3, ctx->stage.p, s, 4));
      EPH_HIP(ctx, hipMemcpyAsync(out, ctx->stage.p, nall * sizeof(double), hipMemcpyDeviceToHost, s)); break;
    case 1: { int rc = download_rows(ctx, ctx->rec_field(1), nall, out, 4); if (rc) return rc; break; }
    case 2: { int rc = download_rows(ctx, ctx->rec_field(2), nall, out, 4); if (rc) return rc; break; }
    case 3: case 4: {
      const double* src = which == 3 ? ctx->fE.p : ctx->fR.p;
      EPH_HIP(ctx, launch_soa_into_rows(src, ctx->nlocal, ctx->fstride, ctx->stage.p, s));
      EPH_HIP(ctx, hipMemcpyAsync(out, ctx->stage.p, 3 * (size_t)ctx->nlocal * sizeof(double), hipMemcpyDeviceToHost, s));
      break;
    }
    case 6: case 7: {
      std::vector<int> cnt(ctx->nlocal);
      const int* src = which == 6 ? ctx->nact.p : ctx->nnbr.p;
      if (ctx->nlocal)
        EPH_HIP(ctx, hipMemcpyAsync(cnt.data(), src, ctx->nlocal * sizeof(int), hipMemcpyDeviceToHost, s));
      EPH_HIP(ctx, hipStreamSynchronize(s));
      for (int i = 0; i < ctx->nlocal; ++i) out[i] = cnt[i];
      return EPH_OK;
    }
    case 8: case 9: case 10: case 11: {
      if (!ctx->atomic()) return ctx->fail(EPH_ERR_ARGS, "per-atom energies belong to eph/atomic");
      const eph::AtomicState& at = *ctx->at;
      const double* src = which == 8 ? at.rho_a.p : which == 9 ? at.E.p : which == 10 ? at.dE.p : at.T.p;
      const int n = which == 10 ? ctx->nlocal : nall;
      if (n) EPH_HIP(ctx, hipMemcpyAsync(out, src, n * sizeof(double), hipMemcpyDeviceToHost, s));
      break;
    }
    case 12: {  // candidate-list statistics of the last cooperative density sweep
      int st[8] = {};
      if (ctx->cstate.p) EPH_HIP(ctx, hipMemcpyAsync(st, ctx->cstate.p, sizeof st, hipMemcpyDeviceToHost, s));
      EPH_HIP(ctx, hipStreamSynchronize(s));
      const int nwaves = (ctx->nlocal + 63) / 64;
      out[0] = st[3]; out[1] = nwaves; out[2] = st[4]; out[3] = st[5];
      return EPH_OK;
    }
    default: return ctx->fail(EPH_ERR_ARGS, "unknown state selector");
  }
  EPH_HIP(ctx, hipStreamSynchronize(s));
  return EPH_OK;
}

// ------------------------------------------------------- device-resident hooks
int eph_dev_initial_integrate(eph_context* ctx) { return ctx ? ctx->dev_initial_integrate() : EPH_ERR_ARGS; }
int eph_dev_post_force(eph_context* ctx) { return ctx ? ctx->dev_post_force() : EPH_ERR_ARGS; }
int eph_dev_final_integrate(eph_context* ctx) { return ctx ? ctx->dev_final_integrate() : EPH_ERR_ARGS; }
int eph_dev_end_of_step(eph_context* ctx, long long ntimestep) {
  return ctx ? ctx->dev_end_of_step(ntimestep) : EPH_ERR_ARGS;
}

int eph_dev_get_atoms(eph_context* ctx, double* x, double* v, double* f) {
  if (!ctx) return EPH_ERR_ARGS;
  const int nall = ctx->nlocal + ctx->nghost;
  int rc;
  if (x) {
    if ((rc = download_rows(ctx, ctx->pos.p, nall, x))) return rc;
    EPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  if (v) {
    if ((rc = download_rows(ctx, ctx->vel.p, nall, v))) return rc;
    EPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  if (f && ctx->nlocal) {
    EPH_HIP(ctx, launch_soa_into_rows(ctx->f.p, ctx->nlocal, ctx->fstride, ctx->stage.p, ctx->stream));
    EPH_HIP(ctx, hipMemcpyAsync(f, ctx->stage.p, 3 * (size_t)ctx->nlocal * sizeof(double),
                                hipMemcpyDeviceToHost, ctx->stream));
    EPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  return EPH_OK;
}

int eph_sync(eph_context* ctx) {
  if (!ctx) return EPH_ERR_ARGS;
  EPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return ctx->check_flags();
}

void* eph_stream(eph_context* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int eph_set_sweep_forms(eph_context* ctx, int density, int force) {
  if (!ctx || density < -1 || density >= kSweepForms || force < -1 || force >= kSweepForms)
    return EPH_ERR_ARGS;
  if ((density >= 1 || force >= 1) && !lds_tables_possible(ctx->tabs))
    return ctx->fail(EPH_ERR_ARGS, "the spline tables do not fit in LDS on this device");
  ctx->form_density = density;
  ctx->form_force = force;
  return EPH_OK;
}

int eph_autotune(eph_context* ctx, int* forms, double* ms) {
  return ctx ? ctx->autotune(forms, ms) : EPH_ERR_ARGS;
}

int eph_timing_enable(eph_context* ctx, int on) {
  if (!ctx) return EPH_ERR_ARGS;
  int rc = ctx->resolve_timing();
  ctx->timing = on != 0;
  ctx->stats.clear();
  return rc;
}

int eph_timing_report(eph_context* ctx, int k, char* name, int namelen, double* total_ms,
                      long long* count) {
  if (!ctx) return EPH_ERR_ARGS;
  int rc = ctx->resolve_timing();
  if (rc) return rc;
  if (k < 0) return (int)ctx->stats.size();
  if (k >= (int)ctx->stats.size()) return EPH_ERR_ARGS;
  if (name && namelen > 0) {
    std::strncpy(name, ctx->stats[k].name.c_str(), namelen - 1);
    name[namelen - 1] = 0;
  }
  if (total_ms) *total_ms = ctx->stats[k].ms;
  if (count) *count = ctx->stats[k].count;
  return EPH_OK;
}

int eph_copy_probe(long long bytes, int reps, int device, double* ms) {
  if (bytes < 32 || reps < 1) return EPH_ERR_ARGS;
  if (hipSetDevice(device) != hipSuccess) return EPH_ERR_NO_DEVICE;
  const long n = static_cast<long>(bytes / 32);
  double4 *a = nullptr, *b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = EPH_OK;
  float t = 0.f;
  if (hipMalloc(reinterpret_cast<void**>(&a), n * 32) || hipMalloc(reinterpret_cast<void**>(&b), n * 32)) {
    rc = EPH_ERR_OOM;
  } else if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) || hipEventCreate(&e0) ||
             hipEventCreate(&e1) || hipMemsetAsync(a, 0, n * 32, s) ||
             launch_copy_probe(b, a, n, s) || hipEventRecord(e0, s)) {
    rc = EPH_ERR_HIP;
  } else {
    for (int r = 0; r < reps && rc == EPH_OK; ++r)
      if (launch_copy_probe(b, a, n, s)) rc = EPH_ERR_HIP;
    if (rc == EPH_OK && (hipEventRecord(e1, s) || hipEventSynchronize(e1) || hipEventElapsedTime(&t, e0, e1)))
      rc = EPH_ERR_HIP;
  }
  if (ms) *ms = t / reps;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}

// ========================================================= host table checks
int eph_spline_build(double dx, const double* y, int n, double* coeffs) {
  if (!y || !coeffs || n < 4) return EPH_ERR_ARGS;
  Spline s;
  std::string err;
  if (!build_spline(dx, std::vector<double>(y, y + n), s, err)) {
    g_create_error = err;
    return EPH_ERR_ARGS;
  }
  std::memcpy(coeffs, s.knots.data(), (size_t)n * sizeof(Cubic));
  return EPH_OK;
}

int eph_beta_table(const char* path, int elem, int which, double* coeffs, double* inv_dx,
                   double* r_cutoff, double* rho_cutoff) {
  if (!path || which < 0 || which > 3) return EPH_ERR_ARGS;
  BetaModel m;
  std::string err;
  if (!load_beta(path, m, err)) {
    g_create_error = err;
    return err.rfind("cannot open", 0) == 0 ? EPH_ERR_FILE : EPH_ERR_FORMAT;
  }
  if (elem < 0 || elem >= (int)m.names.size()) return EPH_ERR_ARGS;
  const Spline& s = which == 0 ? m.rho[elem] : which == 1 ? m.rho_r_sq[elem]
                  : which == 2 ? m.beta[elem] : m.alpha[elem];
  if (coeffs) std::memcpy(coeffs, s.knots.data(), s.knots.size() * sizeof(Cubic));
  if (inv_dx) *inv_dx = s.inv_dx;
  if (r_cutoff) *r_cutoff = m.r_cutoff;
  if (rho_cutoff) *rho_cutoff = m.rho_cutoff;
  return (int)s.knots.size();
}

int eph_kappa_info(const char* path, int* n_elem, int* n_pairs, double* r_cutoff, double* T_max) {
  if (!path) return EPH_ERR_ARGS;
  KappaModel m;
  std::string err;
  if (!load_kappa(path, m, err)) {
    g_create_error = err;
    return err.rfind("cannot open", 0) == 0 ? EPH_ERR_FILE : EPH_ERR_FORMAT;
  }
  if (n_elem) *n_elem = (int)m.names.size();
  if (n_pairs) *n_pairs = m.n_pairs;
  if (r_cutoff) *r_cutoff = m.r_cutoff;
  if (T_max) *T_max = m.T_max;
  return EPH_OK;
}

int eph_kappa_table(const char* path, int index, int which, double* out, double* dx) {
  if (!path || which < 0 || which > 3) return EPH_ERR_ARGS;
  KappaModel m;
  std::string err;
  if (!load_kappa(path, m, err)) {
    g_create_error = err;
    return err.rfind("cannot open", 0) == 0 ? EPH_ERR_FILE : EPH_ERR_FORMAT;
  }
  const int lim = which == 3 ? m.n_pairs : (int)m.names.size();
  if (index < 0 || index >= lim) return EPH_ERR_ARGS;
  if (which < 2) {
    const Spline& s = which == 0 ? m.rho_r[index] : m.rho_r_sq[index];
    if (out) std::memcpy(out, s.knots.data(), s.knots.size() * sizeof(Cubic));
    if (dx) *dx = s.inv_dx;
    return (int)s.knots.size();
  }
  const LinearTable& L = which == 2 ? m.E_T[index] : m.K_T[index];
  for (size_t i = 0; out && i < L.y.size(); ++i) {
    out[2 * i] = L.y[i];
    out[2 * i + 1] = i < L.dy.size() ? L.dy[i] : 0.0;
  }
  if (dx) *dx = L.dx;
  return (int)L.y.size();
}

// ================================================================ grid C ABI
int eph_grid_create(eph_grid** out, int nx, int ny, int nz, const double box[6], double T_e,
                    double C_e, double rho_e, double kappa_e, int device) {
  if (!out || !box) return EPH_ERR_ARGS;
  *out = nullptr;
  auto* g = new eph_grid();
  std::string err;
  if (!make_grid_const(nx, ny, nz, box, T_e, C_e, rho_e, kappa_e, g->h, err)) {
    g_create_error = err;
    delete g;
    return EPH_ERR_ARGS;
  }
  if (int rc = grid_init_device(g, device, nullptr)) {
    g_create_error = g->err;
    eph_grid_destroy(g);
    return rc;
  }
  *out = g;
  return EPH_OK;
}

int eph_grid_load(eph_grid** out, const char* path, int device) {
  if (!out || !path) return EPH_ERR_ARGS;
  *out = nullptr;
  auto* g = new eph_grid();
  std::string err;
  if (!load_grid(path, g->h, err)) {
    g_create_error = err;
    delete g;
    return err.rfind("cannot open", 0) == 0 ? EPH_ERR_FILE : EPH_ERR_FORMAT;
  }
  if (int rc = grid_init_device(g, device, nullptr)) {
    g_create_error = g->err;
    eph_grid_destroy(g);
    return rc;
  }
  *out = g;
  return EPH_OK;
}

void eph_grid_destroy(eph_grid* g) {
  if (!g) return;
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  hipStream_t s = g->own_stream ? g->stream : nullptr;
  delete g;
  if (s) (void)hipStreamDestroy(s);
}

const char* eph_grid_last_error(const eph_grid* g) { return g ? g->err.c_str() : g_create_error.c_str(); }

eph_grid* eph_context_grid(eph_context* ctx) { return ctx ? ctx->grid : nullptr; }

int eph_grid_set_dt(eph_grid* g, double dt) {
  if (!g) return EPH_ERR_ARGS;
  g->h.dt = dt;
  return EPH_OK;
}

int eph_grid_set_steps(eph_grid* g, int steps) {
  if (!g || steps < 1) return EPH_ERR_ARGS;
  g->h.steps = (size_t)steps;
  return EPH_OK;
}

int eph_grid_set_cell(eph_grid* g, int i, int j, int k, double T, double S, double rho, double C,
                      double kappa, int flag, int tdyn) {
  if (!g) return EPH_ERR_ARGS;
  if (i < 0 || j < 0 || k < 0 || (size_t)i >= g->h.nx || (size_t)j >= g->h.ny || (size_t)k >= g->h.nz)
    return g->fail(EPH_ERR_ARGS, "cell index out of range");
  if (tdyn && !g->h.has_params) return g->fail(EPH_ERR_ARGS, "temperature-dependent cell without parameter file");
  int rc;
  if ((rc = g->sync_params_from_device())) return rc;
  std::vector<double> Tcur;
  if ((rc = g->download(0, Tcur))) return rc;
  std::vector<double> dTcur;
  if ((rc = g->download(1, dTcur))) return rc;
  g->h.T = Tcur;
  const size_t c = i + j * g->h.nx + k * g->h.nx * g->h.ny;
  g->h.T[c] = T; g->h.S[c] = S; g->h.rho[c] = rho; g->h.C[c] = C; g->h.kappa[c] = kappa;
  g->h.flag[c] = (int8_t)flag; g->h.tdyn[c] = (int8_t)tdyn;
  if ((rc = g->upload())) return rc;
  EPH_HIP(g, hipMemcpyAsync(g->dT.p, dTcur.data(), dTcur.size() * sizeof(double), hipMemcpyHostToDevice, g->stream));
  EPH_HIP(g, hipStreamSynchronize(g->stream));
  return EPH_OK;
}

int eph_grid_insert_energy(eph_grid* g, int n, const double* x, const double* y, const double* z,
                           const double* E) {
  if (!g || n < 0 || (n > 0 && (!x || !y || !z || !E))) return EPH_ERR_ARGS;
  return g->insert_points(n, x, y, z, E);
}

int eph_grid_solve(eph_grid* g, int* substeps) {
  if (!g) return EPH_ERR_ARGS;
  int rc = g->solve(substeps);
  if (rc) return rc;
  int fl = 0;
  EPH_HIP(g, hipMemcpyAsync(&fl, g->nonfinite.p, sizeof(int), hipMemcpyDeviceToHost, g->stream));
  EPH_HIP(g, hipStreamSynchronize(g->stream));
  if (fl & 2) return g->fail(EPH_ERR_NONFINITE, "T_e grid became non-finite (FDM instability)");
  return EPH_OK;
}

int eph_grid_get(eph_grid* g, int which, double* out) {
  if (!g || !out || which < 0 || which > 5) return EPH_ERR_ARGS;
  std::vector<double> tmp;
  int rc = g->download(which, tmp);
  if (rc) return rc;
  std::copy(tmp.begin(), tmp.end(), out);
  return EPH_OK;
}

int eph_grid_set(eph_grid* g, int which, const double* in) {
  if (!g || !in || which < 0 || which > 5) return EPH_ERR_ARGS;
  const size_t G = g->h.ntotal();
  if (which == 1) {
    EPH_HIP(g, hipMemcpyAsync(g->dT.p, in, G * sizeof(double), hipMemcpyHostToDevice, g->stream));
    EPH_HIP(g, hipStreamSynchronize(g->stream));
    return EPH_OK;
  }
  int rc;
  if ((rc = g->sync_params_from_device())) return rc;
  std::vector<double> Tcur, dTcur;
  if ((rc = g->download(0, Tcur)) || (rc = g->download(1, dTcur))) return rc;
  g->h.T = Tcur;
  std::vector<double>& dst = which == 0 ? g->h.T : which == 2 ? g->h.C : which == 3 ? g->h.rho
                           : which == 4 ? g->h.kappa : g->h.S;
  std::copy(in, in + G, dst.begin());
  if ((rc = g->upload())) return rc;
  EPH_HIP(g, hipMemcpyAsync(g->dT.p, dTcur.data(), G * sizeof(double), hipMemcpyHostToDevice, g->stream));
  EPH_HIP(g, hipStreamSynchronize(g->stream));
  return EPH_OK;
}

int eph_grid_lookup(eph_grid* g, int n, const double* x, const double* y, const double* z,
                    int* index, double* T) {
  if (!g || n < 0) return EPH_ERR_ARGS;
  if (n == 0) return EPH_OK;
  EPH_HIP(g, g->pts.reserve(4 * (size_t)n));
  EPH_HIP(g, g->lookup_idx.reserve(n));
  const size_t b = n * sizeof(double);
  EPH_HIP(g, hipMemcpyAsync(g->pts.p, x, b, hipMemcpyHostToDevice, g->stream));
  EPH_HIP(g, hipMemcpyAsync(g->pts.p + n, y, b, hipMemcpyHostToDevice, g->stream));
  EPH_HIP(g, hipMemcpyAsync(g->pts.p + 2 * n, z, b, hipMemcpyHostToDevice, g->stream));
  EPH_HIP(g, launch_grid_lookup(g->v, n, g->pts.p, g->pts.p + n, g->pts.p + 2 * n, g->lookup_idx.p,
                                g->pts.p + 3 * n, g->stream));
  if (index) EPH_HIP(g, hipMemcpyAsync(index, g->lookup_idx.p, n * sizeof(int), hipMemcpyDeviceToHost, g->stream));
  if (T) EPH_HIP(g, hipMemcpyAsync(T, g->pts.p + 3 * n, b, hipMemcpyDeviceToHost, g->stream));
  EPH_HIP(g, hipStreamSynchronize(g->stream));
  return EPH_OK;
}

int eph_grid_info(const eph_grid* g, int* dims, double* box, double* last_r) {
  if (!g) return EPH_ERR_ARGS;
  if (dims) { dims[0] = (int)g->h.nx; dims[1] = (int)g->h.ny; dims[2] = (int)g->h.nz; }
  if (box) {
    box[0] = g->h.x0; box[1] = g->h.x1; box[2] = g->h.y0; box[3] = g->h.y1; box[4] = g->h.z0; box[5] = g->h.z1;
  }
  if (last_r) *last_r = g->last_r;
  return EPH_OK;
}

int eph_grid_T_mean(eph_grid* g, double* out) {
  if (!g || !out) return EPH_ERR_ARGS;
  const int G = (int)g->h.ntotal();
  EPH_HIP(g, launch_sum(g->T.p, G, g->sum_tmp.p + 256, g->sum_tmp.p, g->stream));
  double s = 0;
  EPH_HIP(g, hipMemcpyAsync(&s, g->sum_tmp.p + 256, sizeof(double), hipMemcpyDeviceToHost, g->stream));
  EPH_HIP(g, hipStreamSynchronize(g->stream));
  *out = s / G;
  return EPH_OK;
}

int eph_grid_save_state(eph_grid* g, const char* path) {
  if (!g || !path) return EPH_ERR_ARGS;
  std::vector<double> T;
  int rc;
  if ((rc = g->sync_params_from_device()) || (rc = g->download(0, T))) return rc;
  return save_grid_state(g->h, T, path) ? EPH_OK : g->fail(EPH_ERR_FILE, std::string("cannot write ") + path);
}

int eph_grid_save_temperature(eph_grid* g, const char* base, int n) {
  if (!g || !base) return EPH_ERR_ARGS;
  std::vector<double> T;
  int rc = g->download(0, T);
  if (rc) return rc;
  return save_grid_temperature(g->h, T, base, n) ? EPH_OK : g->fail(EPH_ERR_FILE, std::string("cannot write ") + base);
}

}  // extern "C"
