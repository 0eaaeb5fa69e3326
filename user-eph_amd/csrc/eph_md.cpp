// eph_md.cpp -- minimal MD harness standing in for the LAMMPS parts the fix
// relies on (see eph_md.h).  Host side builds the system, ghosts and lists;
// the device side runs the Verlet loop with the fix hooks.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>

#include "eph_harness.hpp"
#include "eph_rng.hpp"

using namespace eph;

namespace {

constexpr double kMvv2e = 1.0364269e-4;  // metal units: g/mol A^2/ps^2 -> eV
constexpr double kBoltz = 8.617343e-5;   // eV/K

template <typename F>
void parallel_for(long n, F fn) {
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 4096) nt = 1;
  std::vector<std::thread> th;
  const long chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; ++t) {
    const long b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    th.emplace_back([=] { for (long i = b; i < e; ++i) fn(i); });
  }
  for (auto& x : th) x.join();
}

}  // namespace

// remap locals into [lo, hi) (Domain::remap)
void eph_md::wrap() {
  for (int i = 0; i < nlocal; ++i)
    for (int d = 0; d < 3; ++d) {
      double& c = x[3 * (size_t)i + d];
      if (c < lo[d]) c += L[d];
      if (c >= hi[d]) c -= L[d];
      if (c < lo[d]) c = lo[d];
    }
}

// periodic images by dimension, x then y then z, each including the images
// made before it (so edge and corner images appear), as Comm::borders does
void eph_md::build_ghosts() {
  const double cut = std::max(cfg.cut_fix, cfg.cut_lj) + cfg.skin;
  x.resize(3 * (size_t)nlocal); v.resize(3 * (size_t)nlocal);
  type.resize(nlocal); mask.resize(nlocal); tag.resize(nlocal);
  owner.clear(); shift.clear();
  std::vector<int> src_owner(nlocal);
  std::iota(src_owner.begin(), src_owner.end(), 0);
  std::vector<double> src_shift(3 * (size_t)nlocal, 0.0);
  for (int d = 0; d < 3; ++d) {
    const int n = (int)src_owner.size();
    for (int k = 0; k < n; ++k) {
      const int o = src_owner[k];
      const double xd = x[3 * (size_t)o + d] + src_shift[3 * (size_t)k + d];
      for (int sgn = -1; sgn <= 1; sgn += 2) {
        const double img = xd + sgn * L[d];
        if (img >= lo[d] - cut && img < hi[d] + cut) {
          src_owner.push_back(o);
          for (int e = 0; e < 3; ++e) src_shift.push_back(src_shift[3 * (size_t)k + e]);
          src_shift[src_shift.size() - 3 + d] += sgn * L[d];
        }
      }
    }
  }
  nghost = (int)src_owner.size() - nlocal;
  owner.assign(src_owner.begin() + nlocal, src_owner.end());
  shift.assign(src_shift.begin() + 3 * (size_t)nlocal, src_shift.end());
  const size_t nall = (size_t)nlocal + nghost;
  x.resize(3 * nall); v.resize(3 * nall); type.resize(nall); mask.resize(nall); tag.resize(nall);
  for (int g = 0; g < nghost; ++g) {
    const int o = owner[g];
    const size_t gi = (size_t)nlocal + g;
    for (int e = 0; e < 3; ++e) {
      x[3 * gi + e] = x[3 * (size_t)o + e] + shift[3 * (size_t)g + e];
      v[3 * gi + e] = v[3 * (size_t)o + e];
    }
    type[gi] = type[o]; mask[gi] = mask[o]; tag[gi] = tag[o];
  }
}

// full lists at cut + skin over bins of half the larger cutoff; found in
// stencil bin order, then ascending index within a bin; the pair list keeps
// that order, the fix list is sorted by list_sort_key (distance, then index)
void eph_md::build_lists() {
  const double cf = cfg.cut_fix + cfg.skin, cl = cfg.cut_lj + cfg.skin;
  const double cmax = std::max(cf, cl);
  const double bs = 0.5 * cmax;
  const double glo[3] = {lo[0] - cmax, lo[1] - cmax, lo[2] - cmax};
  int nb[3];
  for (int d = 0; d < 3; ++d) nb[d] = std::max(1, (int)std::ceil((L[d] + 2 * cmax) / bs) + 1);
  const size_t nbins = (size_t)nb[0] * nb[1] * nb[2];
  const size_t nall = (size_t)nlocal + nghost;
  std::vector<int> bin_of(nall), bcount(nbins + 1, 0);
  auto bcoord = [&](size_t i, int d) {
    int b = (int)std::floor((x[3 * i + d] - glo[d]) / bs);
    return std::min(std::max(b, 0), nb[d] - 1);
  };
  for (size_t i = 0; i < nall; ++i) {
    const int b = bcoord(i, 0) + nb[0] * (bcoord(i, 1) + nb[1] * bcoord(i, 2));
    bin_of[i] = b;
    bcount[b + 1]++;
  }
  for (size_t b = 0; b < nbins; ++b) bcount[b + 1] += bcount[b];
  std::vector<int> bins(nall), fillp(bcount.begin(), bcount.end() - 1);
  for (size_t i = 0; i < nall; ++i) bins[fillp[bin_of[i]]++] = (int)i;
  const double cf2 = cf * cf, cl2 = cl * cl;
  const double sscale = list_sort_scale(cf2);
  std::vector<int> nf(nlocal), nl(nlocal);
  auto scan = [&](long i, bool write, int* of, int* ol) {
    const int bx = bcoord(i, 0), by = bcoord(i, 1), bz = bcoord(i, 2);
    int cf_ = 0, cl_ = 0;
    const double* xi = &x[3 * (size_t)i];
    for (int dz = -2; dz <= 2; ++dz) {
      const int z = bz + dz;
      if (z < 0 || z >= nb[2]) continue;
      for (int dy = -2; dy <= 2; ++dy) {
        const int y = by + dy;
        if (y < 0 || y >= nb[1]) continue;
        for (int dx = -2; dx <= 2; ++dx) {
          const int xx = bx + dx;
          if (xx < 0 || xx >= nb[0]) continue;
          const int b = xx + nb[0] * (y + nb[1] * z);
          for (int k = bcount[b]; k < bcount[b + 1]; ++k) {
            const int j = bins[k];
            if (j == i) continue;
            const double* xj = &x[3 * (size_t)j];
            const double ex = xj[0] - xi[0], ey = xj[1] - xi[1], ez = xj[2] - xi[2];
            const double r2 = ex * ex + ey * ey + ez * ez;
            if (r2 < cf2) { if (write) of[cf_] = j; ++cf_; }
            if (r2 < cl2) { if (write) ol[cl_] = j; ++cl_; }
          }
        }
      }
    }
    if (!write) {
      nf[i] = cf_;
      nl[i] = cl_;
      return;
    }
    std::vector<uint64_t> keys(cf_);
    const std::vector<int> found(of, of + cf_);
    for (int t = 0; t < cf_; ++t) {
      const double* xj = &x[3 * (size_t)found[t]];
      const double ex = xj[0] - xi[0], ey = xj[1] - xi[1], ez = xj[2] - xi[2];
      keys[t] = (static_cast<uint64_t>(list_sort_q(ex * ex + ey * ey + ez * ez, sscale)) << 32) |
                static_cast<uint32_t>(t);
    }
    std::sort(keys.begin(), keys.end());
    for (int t = 0; t < cf_; ++t) of[t] = found[static_cast<uint32_t>(keys[t])];
  };
  parallel_for(nlocal, [&](long i) { scan(i, false, nullptr, nullptr); });
  fix.off.assign(nlocal + 1, 0);
  lj.off.assign(nlocal + 1, 0);
  for (int i = 0; i < nlocal; ++i) {
    fix.off[i + 1] = fix.off[i] + nf[i];
    lj.off[i + 1] = lj.off[i] + nl[i];
  }
  fix.idx.resize(fix.off[nlocal]);
  lj.idx.resize(lj.off[nlocal]);
  parallel_for(nlocal, [&](long i) { scan(i, true, &fix.idx[fix.off[i]], &lj.idx[lj.off[i]]); });
}

int eph_md::upload(eph_context* ctx) {
  int rc;
  int mxf = 0, mx = 0;
  for (int i = 0; i < nlocal; ++i) {
    mxf = std::max(mxf, fix.off[i + 1] - fix.off[i]);
    mx = std::max(mx, lj.off[i + 1] - lj.off[i]);
  }
  // capacities for device rebuilds: ghosts and list widths with headroom
  ghost_cap = std::max(ghost_cap, nghost + nghost / 4 + 1024);
  // the device rebuild sorts fix lists of at most kListSortMax entries
  if (mxf > kListSortMax)
    return fail(EPH_ERR_UNSUPPORTED, "a fix list holds more than 512 entries (the device list sort's limit)");
  cap_f = std::min(std::max(cap_f, mxf + mxf / 4 + 16), kListSortMax);
  cap_l = std::max(cap_l, mx + mx / 4 + 16);
  if ((rc = ctx->reserve_atoms(nlocal, ghost_cap))) return fail(rc, ctx->err);
  EPH_HIP(this, ctx->nbr.reserve((size_t)cap_f * ctx->stride));
  EPH_HIP(this, ctx->act.reserve((size_t)cap_f * ctx->stride));
  EPH_HIP(this, ctx->cand.reserve((size_t)cap_f * ctx->stride));
  ctx->drop_candidates();
  EPH_HIP(this, ctx->ghost_owner.reserve(ghost_cap));
  EPH_HIP(this, ctx->ghost_shift.reserve(ghost_cap));
  if ((rc = ctx->set_atoms(nlocal, nghost, x.data(), v.data(), type.data(), mask.data(), tag.data())))
    return fail(rc, ctx->err);
  if ((rc = ctx->set_ghost_map(nghost, owner.data(), shift.data()))) return fail(rc, ctx->err);
  if ((rc = ctx->set_neighbors(nlocal, fix.off.data(), fix.idx.data(), ~0))) return fail(rc, ctx->err);
  // lj list in the context's ELL stride (indices checked like the fix list's)
  {
    const long nall = (long)nlocal + nghost;
    for (int j : lj.idx)
      if (j < 0 || j >= nall) return fail(EPH_ERR_STATE, "pair list index out of range");
  }
  mx = cap_l;
  lj_stride = ctx->stride;
  hipStream_t s = ctx->stream;
  EPH_HIP(this, lj_nbr.reserve((size_t)std::max(mx, 1) * lj_stride));
  EPH_HIP(this, lj_nn.reserve(lj_stride));
  EPH_HIP(this, lj_off.reserve(nlocal + 1));
  EPH_HIP(this, lj_idx.reserve(std::max<size_t>(lj.idx.size(), 1)));
  EPH_HIP(this, hipMemcpyAsync(lj_off.p, lj.off.data(), (nlocal + 1) * sizeof(int), hipMemcpyHostToDevice, s));
  if (!lj.idx.empty())
    EPH_HIP(this, hipMemcpyAsync(lj_idx.p, lj.idx.data(), lj.idx.size() * sizeof(int), hipMemcpyHostToDevice, s));
  EPH_HIP(this, ctx->list_bad.reserve(1));
  EPH_HIP(this, launch_fill_int(ctx->list_bad.p, 1, 0, s));
  EPH_HIP(this, launch_csr_to_ell(nlocal, lj_off.p, lj_idx.p, ~0, nlocal + nghost, lj_nbr.p, lj_nn.p,
                                  lj_stride, std::max(mx, 1), ctx->list_bad.p, s));
  EPH_HIP(this, xref.reserve(nlocal));
  EPH_HIP(this, hipMemcpyAsync(xref.p, ctx->pos.p, nlocal * sizeof(double4), hipMemcpyDeviceToDevice, s));
  EPH_HIP(this, partial.reserve(1024));
  // device rebuild workspace (single system: replica 0)
  const double cf = cfg.cut_fix + cfg.skin, cl = cfg.cut_lj + cfg.skin, cmax = std::max(cf, cl);
  const double bs = 0.5 * cmax;
  int nb[3];
  for (int d = 0; d < 3; ++d) nb[d] = std::max(1, (int)std::ceil((L[d] + 2 * cmax) / bs) + 1);
  nbins = nb[0] * nb[1] * nb[2];
  const size_t cap = (size_t)nlocal + ghost_cap;
  EPH_HIP(this, rep.reserve(cap));
  EPH_HIP(this, hipMemsetAsync(rep.p, 0, cap * sizeof(int), s));
  EPH_HIP(this, flags.reserve(8));
  EPH_HIP(this, hipMemsetAsync(flags.p, 0, 8 * sizeof(int), s));
  EPH_HIP(this, cnt.reserve(cap));
  EPH_HIP(this, start.reserve(cap + 1));
  EPH_HIP(this, bin_of.reserve(cap));
  EPH_HIP(this, items.reserve(cap));
  EPH_HIP(this, bcount.reserve(nbins + 1));
  EPH_HIP(this, bstart.reserve(nbins + 1));
  EPH_HIP(this, bfill.reserve(nbins + 1));
  EPH_HIP(this, scan_tmp.reserve((cap + nbins) / 4096 + 4));
  EPH_HIP(this, nbi.reserve(8));
  const double bx[9] = {lo[0], lo[1], lo[2], hi[0], hi[1], hi[2], L[0], L[1], L[2]};
  const double gl[3] = {lo[0] - cmax, lo[1] - cmax, lo[2] - cmax};
  EPH_HIP(this, box_d.reserve(12));
  EPH_HIP(this, hipMemcpyAsync(box_d.p, bx, sizeof bx, hipMemcpyHostToDevice, s));
  EPH_HIP(this, hipMemcpyAsync(box_d.p + 9, gl, sizeof gl, hipMemcpyHostToDevice, s));
  const int nbv[4] = {nb[0], nb[1], nb[2], 0};
  EPH_HIP(this, hipMemcpyAsync(nbi.p, nbv, sizeof nbv, hipMemcpyHostToDevice, s));
  if (!pinned) EPH_HIP(this, hipHostMalloc(reinterpret_cast<void**>(&pinned), 8 * sizeof(int)));
  NbBuild& w = nbw;
  w.nlocal = nlocal; w.ghost_cap = ghost_cap; w.nbins = nbins;
  w.at.pos = ctx->pos.p; w.at.vel = ctx->vel.p; w.at.type = ctx->type.p; w.at.mask = ctx->mask.p;
  w.at.pv = ctx->atoms_view().pv;
  w.at.tag = ctx->tag.p; w.at.rep = rep.p; w.at.owner = ctx->ghost_owner.p; w.at.shift = ctx->ghost_shift.p;
  w.box.lo = box_d.p; w.box.hi = box_d.p + 3; w.box.L = box_d.p + 6; w.box.glo = box_d.p + 9;
  w.box.nb = nbi.p; w.box.bin_base = nbi.p + 3; w.box.bs = bs; w.box.cg = cmax;
  w.lists.nbr_f = ctx->nbr.p; w.lists.nn_f = ctx->nnbr.p; w.lists.cap_f = cap_f; w.lists.cf2 = cf * cf;
  w.lists.sort_scale = list_sort_scale(w.lists.cf2);
  w.lists.nbr_l = lj_nbr.p; w.lists.nn_l = lj_nn.p; w.lists.cap_l = cap_l; w.lists.cl2 = cl * cl;
  w.lists.stride = ctx->stride;
  w.go = flags.p; w.overflow = flags.p + 1; w.nrebuild = flags.p + 2; w.nghost = flags.p + 4;
  w.cand_valid = ctx->cstate.p;
  w.cnt = cnt.p; w.start = start.p; w.bin_of = bin_of.p; w.bcount = bcount.p; w.bstart = bstart.p;
  w.bfill = bfill.p; w.items = items.p; w.scan_tmp = scan_tmp.p; w.xref = xref.p;
  EPH_HIP(this, hipMemcpyAsync(pinned, ctx->list_bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
  EPH_HIP(this, hipStreamSynchronize(s));
  if (pinned[0]) return fail(EPH_ERR_STATE, "pair list refused on the device (index or row width)");
  return EPH_OK;
}

// copy the device-built system back into the host arrays (for checking)
static int md_pull(eph_md* md, eph_context* ctx) {
  hipStream_t s = ctx->stream;
  const int nl = ctx->nlocal, ng = ctx->nghost;
  const size_t nall = (size_t)nl + ng;
  md->nghost = ng;
  md->x.resize(3 * nall); md->v.resize(3 * nall);
  int rc = eph_dev_get_atoms(ctx, md->x.data(), md->v.data(), nullptr);
  if (rc) return md->fail(rc, ctx->err);
  md->type.resize(nall); md->mask.resize(nall); md->tag.resize(nall);
  md->owner.resize(ng); md->shift.resize(3 * (size_t)ng);
  std::vector<double4> sh(ng);
  std::vector<int> nnf(nl), nnl(nl), ef((size_t)md->cap_f * ctx->stride), el((size_t)md->cap_l * ctx->stride);
  EPH_HIP(md, hipMemcpyAsync(md->type.data(), ctx->type.p, nall * sizeof(int), hipMemcpyDeviceToHost, s));
  EPH_HIP(md, hipMemcpyAsync(md->mask.data(), ctx->mask.p, nall * sizeof(int), hipMemcpyDeviceToHost, s));
  EPH_HIP(md, hipMemcpyAsync(md->tag.data(), ctx->tag.p, nall * sizeof(int), hipMemcpyDeviceToHost, s));
  if (ng) {
    EPH_HIP(md, hipMemcpyAsync(md->owner.data(), ctx->ghost_owner.p, ng * sizeof(int), hipMemcpyDeviceToHost, s));
    EPH_HIP(md, hipMemcpyAsync(sh.data(), ctx->ghost_shift.p, ng * sizeof(double4), hipMemcpyDeviceToHost, s));
  }
  EPH_HIP(md, hipMemcpyAsync(nnf.data(), ctx->nnbr.p, nl * sizeof(int), hipMemcpyDeviceToHost, s));
  EPH_HIP(md, hipMemcpyAsync(nnl.data(), md->lj_nn.p, nl * sizeof(int), hipMemcpyDeviceToHost, s));
  EPH_HIP(md, hipMemcpyAsync(ef.data(), ctx->nbr.p, ef.size() * sizeof(int), hipMemcpyDeviceToHost, s));
  EPH_HIP(md, hipMemcpyAsync(el.data(), md->lj_nbr.p, el.size() * sizeof(int), hipMemcpyDeviceToHost, s));
  EPH_HIP(md, hipStreamSynchronize(s));
  for (int g = 0; g < ng; ++g) {
    md->shift[3 * (size_t)g] = sh[g].x; md->shift[3 * (size_t)g + 1] = sh[g].y; md->shift[3 * (size_t)g + 2] = sh[g].z;
  }
  const int st = ctx->stride;
  md->fix.off.assign(nl + 1, 0);
  md->lj.off.assign(nl + 1, 0);
  for (int i = 0; i < nl; ++i) {
    md->fix.off[i + 1] = md->fix.off[i] + nnf[i];
    md->lj.off[i + 1] = md->lj.off[i] + nnl[i];
  }
  md->fix.idx.resize(md->fix.off[nl]);
  md->lj.idx.resize(md->lj.off[nl]);
  for (int i = 0; i < nl; ++i) {
    for (int k = 0; k < nnf[i]; ++k) md->fix.idx[md->fix.off[i] + k] = ef[(size_t)k * st + i];
    for (int k = 0; k < nnl[i]; ++k) md->lj.idx[md->lj.off[i] + k] = el[(size_t)k * st + i];
  }
  return EPH_OK;
}

extern "C" {

int eph_md_create_fcc(eph_md** out, const eph_md_config* cfg) {
  if (!out || !cfg) return EPH_ERR_ARGS;
  *out = nullptr;
  if (cfg->cells < 1 || !(cfg->a > 0) || cfg->ntypes < 1 || !(cfg->mass > 0))
    return EPH_ERR_ARGS;
  auto* m = new eph_md();
  m->cfg = *cfg;
  const int n = cfg->cells;
  for (int d = 0; d < 3; ++d) { m->lo[d] = 0.0; m->hi[d] = n * cfg->a; m->L[d] = m->hi[d] - m->lo[d]; }
  const double cut = std::max(cfg->cut_fix, cfg->cut_lj) + cfg->skin;
  if (m->L[0] <= cut) {
    m->err = "box must be wider than the ghost cutoff";
    *out = m;
    return EPH_ERR_ARGS;
  }
  const int N = 4 * n * n * n;
  m->nlocal = N;
  m->x.resize(3 * (size_t)N); m->v.resize(3 * (size_t)N);
  m->type.resize(N); m->mask.assign(N, 1); m->tag.resize(N);
  static const double basis[4][3] = {{0, 0, 0}, {0.5, 0.5, 0}, {0.5, 0, 0.5}, {0, 0.5, 0.5}};
  size_t k = 0;
  for (int iz = 0; iz < n; ++iz)
    for (int iy = 0; iy < n; ++iy)
      for (int ix = 0; ix < n; ++ix)
        for (int b = 0; b < 4; ++b, ++k) {
          m->x[3 * k + 0] = (ix + basis[b][0]) * cfg->a;
          m->x[3 * k + 1] = (iy + basis[b][1]) * cfg->a;
          m->x[3 * k + 2] = (iz + basis[b][2]) * cfg->a;
          m->tag[k] = (int)k + 1;
        }
  // species: equiatomic, shuffled by a seeded Fisher-Yates
  for (int i = 0; i < N; ++i) m->type[i] = 1 + i % cfg->ntypes;
  if (cfg->ntypes > 1) {
    for (int i = N - 1; i > 0; --i) {
      Philox4 r = philox4x32_10(Philox4{{(uint32_t)i, 0x53504543u, 0u, 0u}}, cfg->seed, 0x11u);
      const uint64_t u = ((uint64_t)r.v[0] << 32) | r.v[1];
      const int j = (int)(u % (uint64_t)(i + 1));
      std::swap(m->type[i], m->type[j]);
    }
  }
  // Maxwell velocities, zero momentum, rescaled to T with 3N - 3 dof
  double p[3] = {0, 0, 0};
  for (int i = 0; i < N; ++i) {
    double g[3];
    gaussian3(cfg->seed ^ 0x76656c6fu, (uint32_t)(i + 1), 0x7fffffffull, g);
    for (int d = 0; d < 3; ++d) { m->v[3 * (size_t)i + d] = g[d]; p[d] += g[d]; }
  }
  double ke = 0;
  for (int i = 0; i < N; ++i)
    for (int d = 0; d < 3; ++d) {
      double& c = m->v[3 * (size_t)i + d];
      c -= p[d] / N;
      ke += c * c;
    }
  if (cfg->T > 0 && N > 1) {
    const double t_now = ke * cfg->mass * kMvv2e / ((3.0 * N - 3.0) * kBoltz);
    const double sc = std::sqrt(cfg->T / t_now);
    for (auto& c : m->v) c *= sc;
  } else {
    std::fill(m->v.begin(), m->v.end(), 0.0);
  }
  // Gaussian lattice displacements
  if (cfg->disp_sigma > 0)
    for (int i = 0; i < N; ++i) {
      double g[3];
      gaussian3(cfg->seed ^ 0x64697370u, (uint32_t)(i + 1), 0x7ffffffeull, g);
      for (int d = 0; d < 3; ++d) m->x[3 * (size_t)i + d] += cfg->disp_sigma * g[d];
    }
  // primary knock-on atom: nearest to the box centre, seeded direction
  if (cfg->pka_eV > 0) {
    double best = 1e300;
    int ib = 0;
    for (int i = 0; i < N; ++i) {
      double r2 = 0;
      for (int d = 0; d < 3; ++d) {
        const double dd = m->x[3 * (size_t)i + d] - 0.5 * (m->lo[d] + m->hi[d]);
        r2 += dd * dd;
      }
      if (r2 < best) { best = r2; ib = i; }
    }
    double g[3];
    gaussian3(cfg->seed ^ 0x706b6121u, 1u, 0ull, g);
    const double nrm = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    const double speed = std::sqrt(2.0 * cfg->pka_eV / (cfg->mass * kMvv2e));
    for (int d = 0; d < 3; ++d) m->v[3 * (size_t)ib + d] = speed * g[d] / nrm;
  }
  // spatial order: atoms sorted by 2x2x2-cell blocks in Morton order so a
  // wavefront's atoms are neighbours in space
  {
    const double bsz = 2 * cfg->a;
    auto spread = [](uint32_t v) {
      uint64_t r = 0;
      for (int b = 0; b < 21; ++b) r |= (uint64_t)((v >> b) & 1u) << (3 * b);
      return r;
    };
    std::vector<uint64_t> key(N);
    for (int i = 0; i < N; ++i) {
      uint32_t c[3];
      for (int d = 0; d < 3; ++d) {
        double q = std::floor((m->x[3 * (size_t)i + d] - m->lo[d]) / bsz);
        q = std::max(0.0, q);
        c[d] = (uint32_t)q;
      }
      key[i] = spread(c[0]) | (spread(c[1]) << 1) | (spread(c[2]) << 2);
    }
    std::vector<int> perm(N);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return key[a] < key[b]; });
    std::vector<double> x2(3 * (size_t)N), v2(3 * (size_t)N);
    std::vector<int> t2(N), g2(N);
    for (int i = 0; i < N; ++i) {
      const int s = perm[i];
      for (int d = 0; d < 3; ++d) { x2[3 * (size_t)i + d] = m->x[3 * (size_t)s + d]; v2[3 * (size_t)i + d] = m->v[3 * (size_t)s + d]; }
      t2[i] = m->type[s];
      g2[i] = m->tag[s];
    }
    m->x.swap(x2); m->v.swap(v2); m->type.swap(t2); m->tag.swap(g2);
  }
  m->rebuild_all();
  *out = m;
  return EPH_OK;
}

void eph_md_destroy(eph_md* md) { delete md; }

const char* eph_md_last_error(const eph_md* md) { return md ? md->err.c_str() : ""; }

int eph_md_info(const eph_md* md, int* nlocal, int* nghost, long long* nfix, long long* nlj,
                double* boxlo, double* boxhi) {
  if (!md) return EPH_ERR_ARGS;
  if (nlocal) *nlocal = md->nlocal;
  if (nghost) *nghost = md->nghost;
  if (nfix) *nfix = (long long)md->fix.idx.size();
  if (nlj) *nlj = (long long)md->lj.idx.size();
  for (int d = 0; d < 3; ++d) {
    if (boxlo) boxlo[d] = md->lo[d];
    if (boxhi) boxhi[d] = md->hi[d];
  }
  return EPH_OK;
}

int eph_md_get(const eph_md* md, double* x, double* v, int* type, int* mask, int* tag,
               int* ghost_owner, double* ghost_shift, int* fix_off, int* fix_idx) {
  if (!md) return EPH_ERR_ARGS;
  if (x) std::copy(md->x.begin(), md->x.end(), x);
  if (v) std::copy(md->v.begin(), md->v.end(), v);
  if (type) std::copy(md->type.begin(), md->type.end(), type);
  if (mask) std::copy(md->mask.begin(), md->mask.end(), mask);
  if (tag) std::copy(md->tag.begin(), md->tag.end(), tag);
  if (ghost_owner) std::copy(md->owner.begin(), md->owner.end(), ghost_owner);
  if (ghost_shift) std::copy(md->shift.begin(), md->shift.end(), ghost_shift);
  if (fix_off) std::copy(md->fix.off.begin(), md->fix.off.end(), fix_off);
  if (fix_idx) std::copy(md->fix.idx.begin(), md->fix.idx.end(), fix_idx);
  return EPH_OK;
}

int eph_md_update_ghosts(const eph_md* md, double* x, double* v) {
  if (!md || !x) return EPH_ERR_ARGS;
  for (int g = 0; g < md->nghost; ++g) {
    const size_t gi = (size_t)md->nlocal + g, o = md->owner[g];
    for (int e = 0; e < 3; ++e) {
      x[3 * gi + e] = x[3 * o + e] + md->shift[3 * (size_t)g + e];
      if (v) v[3 * gi + e] = v[3 * o + e];
    }
  }
  return EPH_OK;
}

int eph_md_lj_host(const eph_md* md, const double* x, double* f) {
  if (!md || !x || !f) return EPH_ERR_ARGS;
  const double sg = md->cfg.lj_sigma;
  const double s6b = sg * sg * sg * sg * sg * sg;  // as the device pair kernel forms it
  const double lj1 = 48.0 * md->cfg.lj_eps * s6b * s6b, lj2 = 24.0 * md->cfg.lj_eps * s6b;
  const double rc2 = md->cfg.cut_lj * md->cfg.cut_lj;
  parallel_for(md->nlocal, [&](long i) {
    double fx = 0, fy = 0, fz = 0;
    const double* xi = &x[3 * i];
    for (int k = md->lj.off[i]; k < md->lj.off[i + 1]; ++k) {
      const double* xj = &x[3 * (size_t)md->lj.idx[k]];
      const double dx = xi[0] - xj[0], dy = xi[1] - xj[1], dz = xi[2] - xj[2];
      const double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 < rc2) {
        const double r2inv = 1.0 / r2;
        const double r6inv = r2inv * r2inv * r2inv;
        const double fpair = r6inv * (lj1 * r6inv - lj2) * r2inv;
        fx += dx * fpair; fy += dy * fpair; fz += dz * fpair;
      }
    }
    f[3 * i] = fx; f[3 * i + 1] = fy; f[3 * i + 2] = fz;
  });
  return EPH_OK;
}

int eph_md_needs_rebuild(const eph_md* md, const double* x) {
  if (!md || !x) return EPH_ERR_ARGS;
  const double lim = 0.25 * md->cfg.skin * md->cfg.skin;
  for (int i = 0; i < md->nlocal; ++i) {
    double r2 = 0;
    for (int d = 0; d < 3; ++d) {
      const double dd = x[3 * (size_t)i + d] - md->x_build[3 * (size_t)i + d];
      r2 += dd * dd;
    }
    if (r2 > lim) return 1;
  }
  return 0;
}

int eph_md_rebuild(eph_md* md, const double* x, const double* v) {
  if (!md || !x || !v) return EPH_ERR_ARGS;
  md->x.assign(x, x + 3 * (size_t)md->nlocal);
  md->v.assign(v, v + 3 * (size_t)md->nlocal);
  md->rebuild_all();
  return EPH_OK;
}

int eph_md_attach(eph_md* md, eph_context* ctx) {
  if (!md || !ctx) return EPH_ERR_ARGS;
  int rc = md->upload(ctx);
  if (rc) return rc;
  // Verlet::setup: pair forces of the initial configuration
  hipStream_t s = ctx->stream;
  EPH_HIP(md, launch_ghost_xv(ctx->atoms_view(), ctx->ghost_owner.p, ctx->ghost_shift.p, s));
  EPH_HIP(md, launch_lj(ctx->atoms_view(), md->lj_nbr.p, md->lj_nn.p, md->lj_stride, md->cfg.lj_eps,
                        md->cfg.lj_sigma, md->cfg.cut_lj * md->cfg.cut_lj, s));
  EPH_HIP(md, hipStreamSynchronize(s));
  return EPH_OK;
}

int eph_md_set_device_rebuild(eph_md* md, int on) {
  if (!md) return EPH_ERR_ARGS;
  md->device_rebuild = on != 0;
  return EPH_OK;
}

int eph_md_force_rebuild(eph_md* md) {
  if (!md) return EPH_ERR_ARGS;
  md->force_next = true;
  return EPH_OK;
}

int eph_md_pull(eph_md* md, eph_context* ctx) {
  if (!md || !ctx) return EPH_ERR_ARGS;
  return md_pull(md, ctx);
}

int eph_md_run(eph_md* md, eph_context* ctx, int nsteps, long long first_step, double* t,
               int* nrebuild) {
  if (!md || !ctx || nsteps < 0) return EPH_ERR_ARGS;
  hipStream_t s = ctx->stream;
  // five events per step, read after the last step: no step waits for the
  // previous one to finish, so each step's first kernels are queued behind it
  std::vector<hipEvent_t> evs(5 * static_cast<size_t>(nsteps));
  std::vector<char> host_rebuilt(nsteps, 0);
  for (auto& e : evs) EPH_HIP(md, hipEventCreate(&e));
  double t_fix = 0, t_total = 0, t_pair = 0, t_rebuild = 0;
  int nreb = 0;
  const double lim = 0.25 * md->cfg.skin * md->cfg.skin;
  const int nbp = std::min(1024, std::max(1, (md->nlocal + kBlock - 1) / kBlock));
  std::vector<double> part(nbp);
  int rc = EPH_OK;
  for (int st = 0; st < nsteps && rc == EPH_OK; ++st) {
    const long long step = first_step + st;
    ctx->step = step;
    hipEvent_t* ev = evs.data() + 5 * static_cast<size_t>(st);
    // a step starts where the previous one ended: the previous step's end
    // event serves as this step's start (one event less between two kernels)
    if (st == 0) EPH_HIP(md, hipEventRecord(ev[0], s));
    if ((rc = ctx->dev_initial_integrate())) break;
    EPH_HIP(md, hipEventRecord(ev[1], s));
    bool host_rebuild = false;
    if (md->device_rebuild) {
      // neighbour check and, when an atom moved more than skin/2, the device
      // rebuild of ghosts and both lists (and the density sweep's candidates,
      // at the margin in effect now)
      const WorkView cw = ctx->work_view();
      md->nbw.lists.cand = cw.rcand2 > 0.0 ? cw.cand : nullptr;
      md->nbw.lists.ncand = cw.ncand;
      md->nbw.lists.cand2 = cw.rcand2;
      md->nbw.cref = cw.cref;
      EPH_HIP(md, launch_nb_check(md->nlocal, ctx->pos.p, md->xref.p, lim, md->nbw.go, s));
      if (md->force_next) EPH_HIP(md, launch_nb_force(md->nbw.go, s));
      md->force_next = false;
      const hipError_t eb = launch_nb_build(md->nbw, s);
      // the candidate buffers belong to this step's launch only (the
      // context may reallocate them before the next)
      md->nbw.lists.cand = nullptr;
      md->nbw.lists.ncand = nullptr;
      md->nbw.cref = nullptr;
      EPH_HIP(md, eb);
      EPH_HIP(md, hipMemcpyAsync(md->pinned, md->flags.p, 8 * sizeof(int), hipMemcpyDeviceToHost, s));
      EPH_HIP(md, hipStreamSynchronize(s));
      if (md->pinned[0]) {
        if (md->pinned[1]) {
          host_rebuild = true;  // capacity exceeded: rebuild through the host
          EPH_HIP(md, launch_fill_int(md->flags.p + 1, 1, 0, s));
        } else {
          ctx->nghost = md->pinned[7];
          md->nghost = ctx->nghost;
          ++nreb;
        }
      }
    } else {
      EPH_HIP(md, launch_max_disp2(ctx->atoms_view(), md->xref.p, md->partial.p, nbp, s));
      EPH_HIP(md, hipMemcpyAsync(part.data(), md->partial.p, nbp * sizeof(double), hipMemcpyDeviceToHost, s));
      EPH_HIP(md, hipStreamSynchronize(s));
      host_rebuild = *std::max_element(part.begin(), part.end()) > lim || md->force_next;
      md->force_next = false;
    }
    if (host_rebuild) {
      const auto t0 = std::chrono::steady_clock::now();
      std::vector<double> xh(3 * (size_t)(ctx->nlocal + ctx->nghost)), vh(xh.size());
      if ((rc = eph_dev_get_atoms(ctx, xh.data(), vh.data(), nullptr))) break;
      md->x.assign(xh.begin(), xh.begin() + 3 * (size_t)md->nlocal);
      md->v.assign(vh.begin(), vh.begin() + 3 * (size_t)md->nlocal);
      md->rebuild_all();
      if ((rc = md->upload(ctx))) break;
      ++nreb;
      t_rebuild += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    EPH_HIP(md, hipEventRecord(ev[2], s));
    EPH_HIP(md, launch_ghost_xv(ctx->atoms_view(), ctx->ghost_owner.p, ctx->ghost_shift.p, s));
    EPH_HIP(md, launch_lj(ctx->atoms_view(), md->lj_nbr.p, md->lj_nn.p, md->lj_stride, md->cfg.lj_eps,
                          md->cfg.lj_sigma, md->cfg.cut_lj * md->cfg.cut_lj, s));
    EPH_HIP(md, hipEventRecord(ev[3], s));
    if ((rc = ctx->dev_post_force())) break;
    // nothing runs between the two hooks here: one pass where the deposit allows
    if ((rc = ctx->dev_final_integrate_end_of_step(step))) break;
    EPH_HIP(md, hipEventRecord(ev[4], s));
    host_rebuilt[st] = host_rebuild ? 1 : 0;
  }
  if (rc == EPH_OK && nsteps > 0) {
    EPH_HIP(md, hipEventSynchronize(evs[5 * static_cast<size_t>(nsteps) - 1]));
    for (int st = 0; st < nsteps; ++st) {
      const hipEvent_t* ev = evs.data() + 5 * static_cast<size_t>(st);
      const hipEvent_t start = st == 0 ? ev[0] : ev[-1];
      float a = 0, b = 0, c = 0, d = 0, e = 0;
      EPH_HIP(md, hipEventElapsedTime(&a, start, ev[1]));
      EPH_HIP(md, hipEventElapsedTime(&b, ev[3], ev[4]));
      EPH_HIP(md, hipEventElapsedTime(&c, ev[2], ev[3]));
      EPH_HIP(md, hipEventElapsedTime(&d, start, ev[4]));
      if (!host_rebuilt[st]) EPH_HIP(md, hipEventElapsedTime(&e, ev[1], ev[2]));
      t_fix += a + b;
      t_pair += c;
      t_total += d;
      t_rebuild += e;  // device neighbour check / rebuild
    }
  }
  for (auto& e : evs) (void)hipEventDestroy(e);
  if (rc != EPH_OK) {
    if (md->err.empty()) md->err = ctx->err;
    return rc;
  }
  if (t) { t[0] = t_fix; t[1] = t_total; t[2] = t_pair; t[3] = t_rebuild; }
  if (nrebuild) *nrebuild = nreb;
  return ctx->check_flags();
}

}  // extern "C"
