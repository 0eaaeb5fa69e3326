// eph_adjoint.cpp -- trajectory gradients of the device-resident run
// (eph_adjoint.h): forward pass with checkpoints, segment recording, reverse
// pass over the adjoint kernels of eph_adjoint.hip, and the Jacobian of the
// Akima table build for the chain to the beta knots.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "eph_adjoint.hpp"
#include "eph_harness.hpp"

extern "C" {
#include "eph_adjoint.h"
}

using namespace eph;

namespace {

thread_local std::string g_adj_error;

// Forward-mode value with gradient with respect to the n knot values.
struct Dual {
  double v = 0;
  std::vector<double> d;
};

Dual dconst(double v, size_t n) { Dual r; r.v = v; r.d.assign(n, 0.0); return r; }
Dual operator+(const Dual& a, const Dual& b) { Dual r; r.v = a.v + b.v; r.d.resize(a.d.size()); for (size_t k = 0; k < a.d.size(); ++k) r.d[k] = a.d[k] + b.d[k]; return r; }
Dual operator-(const Dual& a, const Dual& b) { Dual r; r.v = a.v - b.v; r.d.resize(a.d.size()); for (size_t k = 0; k < a.d.size(); ++k) r.d[k] = a.d[k] - b.d[k]; return r; }
Dual operator*(const Dual& a, const Dual& b) { Dual r; r.v = a.v * b.v; r.d.resize(a.d.size()); for (size_t k = 0; k < a.d.size(); ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k]; return r; }
Dual operator/(const Dual& a, const Dual& b) { Dual r; r.v = a.v / b.v; r.d.resize(a.d.size()); for (size_t k = 0; k < a.d.size(); ++k) r.d[k] = (a.d[k] * b.v - a.v * b.d[k]) / (b.v * b.v); return r; }
Dual operator*(double s, const Dual& a) { Dual r; r.v = s * a.v; r.d.resize(a.d.size()); for (size_t k = 0; k < a.d.size(); ++k) r.d[k] = s * a.d[k]; return r; }
Dual operator/(const Dual& a, double s) { Dual r; r.v = a.v / s; r.d.resize(a.d.size()); for (size_t k = 0; k < a.d.size(); ++k) r.d[k] = a.d[k] / s; return r; }
Dual dabs(const Dual& a) { Dual r = a; if (a.v < 0) { r.v = -a.v; for (auto& x : r.d) x = -x; } else if (a.v == 0) { for (auto& x : r.d) x = 0; } return r; }

// The table of build_spline (eph_host.cpp) with derivatives: the same
// secant extension, slope choice and absolute-x cubic solve.
bool spline_jacobian(double dx, const std::vector<double>& y, std::vector<double>& jac) {
  const size_t n = y.size();
  if (!(dx > 0) || n < 4) return false;
  std::vector<Dual> Y(n);
  for (size_t i = 0; i < n; ++i) { Y[i] = dconst(y[i], n); Y[i].d[i] = 1.0; }
  std::vector<Dual> s(n + 3, dconst(0.0, n));
  Dual* sec = s.data() + 2;
  for (size_t i = 0; i + 1 < n; ++i) sec[i] = (Y[i + 1] - Y[i]) / dx;
  sec[-1] = 2.0 * sec[0] - sec[1];
  sec[-2] = 2.0 * sec[-1] - sec[0];
  sec[n - 1] = 2.0 * sec[n - 2];
  sec[n] = 2.0 * sec[n - 1];
  std::vector<Dual> t(n);
  for (size_t i = 0; i < n; ++i) {
    const Dual& d_2 = sec[(long)i - 2];
    const Dual& d_1 = sec[(long)i - 1];
    const Dual& d0 = sec[i];
    const Dual wr = dabs(sec[i + 1] - sec[i]);
    const Dual wl = dabs(sec[(long)i - 1] - sec[(long)i - 2]);
    const bool top = (i == n - 1);
    const bool d0_eq_d1 = top ? false : (d0.v == sec[i + 1].v);
    if (d_2.v == d_1.v && !d0_eq_d1) t[i] = d_1;
    else if (d0_eq_d1 && d_2.v == d_1.v) t[i] = d0;
    else if (d_1.v == d0.v) t[i] = d0;
    else t[i] = (d_1 * wr + d0 * wl) / (wr + wl);
  }
  jac.assign(n * 4 * n, 0.0);
  const double h3 = dx * dx * dx;
  for (size_t i = 0; i + 1 < n; ++i) {
    const double xl = i * dx, xl2 = i * dx * xl, xl3 = i * dx * xl2;
    const double xr = (i + 1) * dx, xr2 = (i + 1) * dx * xr, xr3 = (i + 1) * dx * xr2;
    const Dual& tl = t[i];
    const Dual& tr = t[i + 1];
    const Dual d = ((-1.0 * tl) * dconst(xl, n) - tr * dconst(xl, n) + tl * dconst(xr, n) +
                    tr * dconst(xr, n) + 2.0 * Y[i] - 2.0 * Y[i + 1]) / h3;
    const Dual c = ((-1.0 * tl) + tr + (3.0 * xl2) * d - (3.0 * xr2) * d) / 2.0 / dx;
    const Dual b = (xl2 * c + xl3 * d - xr2 * c - xr3 * d - Y[i] + Y[i + 1]) / dx;
    const Dual a = Y[i] - xl * b - xl2 * c - xl3 * d;
    const Dual* co[4] = {&a, &b, &c, &d};
    for (int q = 0; q < 4; ++q)
      for (size_t m = 0; m < n; ++m) jac[(i * 4 + q) * n + m] = co[q]->d[m];
  }
  jac[((n - 1) * 4 + 0) * n + (n - 1)] = 1.0;  // the top knot's constant piece
  return true;
}

}  // namespace

struct eph_adjoint {
  std::string err;
  eph_md* md = nullptr;
  eph_context* ctx = nullptr;
  int nsteps = 0, every = 1, nl = 0, G = 0, n_elem = 0, n_beta = 0;
  bool grid_term = false;
  std::vector<double> T_target;
  DevArray<double> Ttar;

  // a full restart point of the harness + fix
  struct State {
    DevArray<double4> pos, vel, shift, xref;
    DevArray<double> f, T, scal;
    DevArray<int> owner, nbr, nnbr, lj, ljn;
    int nghost = 0;
  };
  State start;
  std::vector<State> ckpt;  // state after step k * every
  std::vector<std::vector<double>> beta_file, beta_cur;
  std::vector<double> v_start;  // start local velocities [nlocal][3]
  std::vector<double> v_cur;

  // one recorded step
  struct Rec {
    DevArray<ForceRec> rec;
    DevArray<PosVel> pv;
    DevArray<double4> vn;
    DevArray<double> fE, fR, TB, DT;
    DevArray<int> act, nact, lj, ljn, owner;
    int nghost = 0;
  };
  std::vector<Rec> seg;

  // adjoint accumulators
  DevArray<double4> xb, vhb, wb, vnb, fEb, fRb, ftb;
  DevArray<double> ab, rb, Tb, Tb2, dTb, Gb, Tsub, kb;
  bool have_forward = false;

  int fail(int code, const std::string& m) { err = m; return code; }
  int fail_hip(hipError_t e, const char* what) {
    err = std::string("HIP error ") + hipGetErrorString(e) + " in " + what;
    return e == hipErrorOutOfMemory ? EPH_ERR_OOM : EPH_ERR_HIP;
  }
  size_t nall_cap() const { return (size_t)nl + md->ghost_cap; }
  int save(State& s);
  int restore(const State& s);
  int set_tables(const std::vector<std::vector<double>>& beta);
  int step(long long n, Rec* r);
  int backward_step(Rec& r, double* g_beta_unused);
};

#define HCPY(dst, src, n) EPH_HIP(this, hipMemcpyAsync((dst), (src), (n), hipMemcpyDeviceToDevice, s))

int eph_adjoint::save(State& st) {
  hipStream_t s = ctx->stream;
  const size_t na = nall_cap(), gc = md->ghost_cap, sf = (size_t)md->cap_f * ctx->stride,
               sl = (size_t)md->cap_l * ctx->stride;
  EPH_HIP(this, st.pos.reserve(na)); EPH_HIP(this, st.vel.reserve(na));
  EPH_HIP(this, st.shift.reserve(gc)); EPH_HIP(this, st.owner.reserve(gc));
  EPH_HIP(this, st.xref.reserve(nl)); EPH_HIP(this, st.f.reserve(3 * (size_t)ctx->fstride));
  EPH_HIP(this, st.T.reserve(G)); EPH_HIP(this, st.scal.reserve(4));
  EPH_HIP(this, st.nbr.reserve(sf)); EPH_HIP(this, st.nnbr.reserve(ctx->stride));
  EPH_HIP(this, st.lj.reserve(sl)); EPH_HIP(this, st.ljn.reserve(ctx->stride));
  HCPY(st.pos.p, ctx->pos.p, na * sizeof(double4));
  HCPY(st.vel.p, ctx->vel.p, na * sizeof(double4));
  HCPY(st.shift.p, ctx->ghost_shift.p, gc * sizeof(double4));
  HCPY(st.owner.p, ctx->ghost_owner.p, gc * sizeof(int));
  HCPY(st.xref.p, md->xref.p, nl * sizeof(double4));
  HCPY(st.f.p, ctx->f.p, 3 * (size_t)ctx->fstride * sizeof(double));
  HCPY(st.T.p, ctx->grid->T.p, G * sizeof(double));
  HCPY(st.scal.p, ctx->scal.p, 2 * sizeof(double));
  HCPY(st.nbr.p, ctx->nbr.p, sf * sizeof(int));
  HCPY(st.nnbr.p, ctx->nnbr.p, ctx->stride * sizeof(int));
  HCPY(st.lj.p, md->lj_nbr.p, sl * sizeof(int));
  HCPY(st.ljn.p, md->lj_nn.p, ctx->stride * sizeof(int));
  st.nghost = ctx->nghost;
  EPH_HIP(this, hipStreamSynchronize(s));
  return EPH_OK;
}

int eph_adjoint::restore(const State& st) {
  hipStream_t s = ctx->stream;
  const size_t na = nall_cap(), gc = md->ghost_cap, sf = (size_t)md->cap_f * ctx->stride,
               sl = (size_t)md->cap_l * ctx->stride;
  HCPY(ctx->pos.p, st.pos.p, na * sizeof(double4));
  HCPY(ctx->vel.p, st.vel.p, na * sizeof(double4));
  HCPY(ctx->ghost_shift.p, st.shift.p, gc * sizeof(double4));
  HCPY(ctx->ghost_owner.p, st.owner.p, gc * sizeof(int));
  HCPY(md->xref.p, st.xref.p, nl * sizeof(double4));
  HCPY(ctx->f.p, st.f.p, 3 * (size_t)ctx->fstride * sizeof(double));
  HCPY(ctx->grid->T.p, st.T.p, G * sizeof(double));
  HCPY(ctx->scal.p, st.scal.p, 2 * sizeof(double));
  HCPY(ctx->nbr.p, st.nbr.p, sf * sizeof(int));
  HCPY(ctx->nnbr.p, st.nnbr.p, ctx->stride * sizeof(int));
  ctx->drop_candidates();
  ctx->pv_fresh = false;
  HCPY(md->lj_nbr.p, st.lj.p, sl * sizeof(int));
  HCPY(md->lj_nn.p, st.ljn.p, ctx->stride * sizeof(int));
  EPH_HIP(this, launch_fill_double(ctx->grid->dT.p, G, 0.0, s));
  EPH_HIP(this, launch_fill_int(md->flags.p, 8, 0, s));
  ctx->nghost = st.nghost;
  md->nghost = st.nghost;
  EPH_HIP(this, hipStreamSynchronize(s));
  return EPH_OK;
}

// beta knots -> beta and alpha (= table of sqrt(beta) knots) tables, eph_beta.h:113-124
int eph_adjoint::set_tables(const std::vector<std::vector<double>>& beta) {
  BetaModel& m = ctx->beta;
  hipStream_t s = ctx->stream;
  for (int e = 0; e < n_elem; ++e) {
    if (beta[e] == m.beta_y[e]) continue;
    for (double b : beta[e])
      if (!(b >= 0.0)) return fail(EPH_ERR_ARGS, "beta knots must be non-negative");
    std::string err2;
    std::vector<double> ak(n_beta);
    for (int j = 0; j < n_beta; ++j) ak[j] = std::sqrt(beta[e][j]);
    if (!build_spline(m.drho, beta[e], m.beta[e], err2) || !build_spline(m.drho, ak, m.alpha[e], err2))
      return fail(EPH_ERR_ARGS, err2);
    m.beta_y[e] = beta[e];
    EPH_HIP(this, hipMemcpyAsync(ctx->t_beta.p + (size_t)e * n_beta, m.beta[e].knots.data(),
                                 n_beta * sizeof(Cubic), hipMemcpyHostToDevice, s));
    EPH_HIP(this, hipMemcpyAsync(ctx->t_alpha.p + (size_t)e * n_beta, m.alpha[e].knots.data(),
                                 n_beta * sizeof(Cubic), hipMemcpyHostToDevice, s));
  }
  EPH_HIP(this, hipStreamSynchronize(s));
  return EPH_OK;
}

// one Verlet step (the order of eph_md_run with device rebuilds); with r, the
// quantities the adjoint step needs are kept
int eph_adjoint::step(long long n, Rec* r) {
  hipStream_t s = ctx->stream;
  int rc;
  ctx->step = n;
  if ((rc = ctx->dev_initial_integrate())) return fail(rc, ctx->err);
  const double lim = 0.25 * md->cfg.skin * md->cfg.skin;
  EPH_HIP(this, launch_nb_check(nl, ctx->pos.p, md->xref.p, lim, md->nbw.go, s));
  EPH_HIP(this, launch_nb_build(md->nbw, s));
  EPH_HIP(this, hipMemcpyAsync(md->pinned, md->flags.p, 8 * sizeof(int), hipMemcpyDeviceToHost, s));
  EPH_HIP(this, hipStreamSynchronize(s));
  if (md->pinned[0]) {
    if (md->pinned[1]) return fail(EPH_ERR_STATE, "neighbour capacity exceeded during the trajectory");
    ctx->nghost = md->pinned[7];
    md->nghost = ctx->nghost;
  }
  EPH_HIP(this, launch_ghost_xv(ctx->atoms_view(), ctx->ghost_owner.p, ctx->ghost_shift.p, s));
  EPH_HIP(this, launch_lj(ctx->atoms_view(), md->lj_nbr.p, md->lj_nn.p, md->lj_stride, md->cfg.lj_eps,
                          md->cfg.lj_sigma, md->cfg.cut_lj * md->cfg.cut_lj, s));
  if ((rc = ctx->dev_post_force())) return fail(rc, ctx->err);
  if ((rc = ctx->dev_final_integrate())) return fail(rc, ctx->err);
  if (r) HCPY(r->TB.p, ctx->grid->T.p, G * sizeof(double));
  if ((rc = ctx->dev_deposit(false, nullptr))) return fail(rc, ctx->err);
  if (r) HCPY(r->DT.p, ctx->grid->dT.p, G * sizeof(double));
  if (ctx->flags & EPH_FDM) {
    if ((rc = ctx->grid->solve(nullptr))) return fail(rc, ctx->grid->err);
  }
  if (r) {
    const size_t na = (size_t)nl + ctx->nghost, st = ctx->stride;
    HCPY(r->rec.p, ctx->rec.p, na * sizeof(ForceRec));
    HCPY(r->pv.p, ctx->pv.p, na * sizeof(PosVel));
    HCPY(r->vn.p, ctx->vel.p, nl * sizeof(double4));
    HCPY(r->fE.p, ctx->fE.p, 3 * (size_t)ctx->fstride * sizeof(double));
    HCPY(r->fR.p, ctx->fR.p, 3 * (size_t)ctx->fstride * sizeof(double));
    HCPY(r->act.p, ctx->act.p, (size_t)md->cap_f * st * sizeof(int));
    HCPY(r->nact.p, ctx->nact.p, st * sizeof(int));
    HCPY(r->lj.p, md->lj_nbr.p, (size_t)md->cap_l * st * sizeof(int));
    HCPY(r->ljn.p, md->lj_nn.p, st * sizeof(int));
    if (ctx->nghost) HCPY(r->owner.p, ctx->ghost_owner.p, ctx->nghost * sizeof(int));
    r->nghost = ctx->nghost;
  }
  return EPH_OK;
}

// the reverse of one recorded step; on entry Tb = dL/dT_e after its solve,
// vnb, ftb, xb = dL/d(v', f', x') from the later steps
int eph_adjoint::backward_step(Rec& r, double*) {
  hipStream_t s = ctx->stream;
  const size_t na = nall_cap();
  AdjBuffers b;
  b.xb = xb.p; b.vhb = vhb.p; b.wb = wb.p; b.vnb = vnb.p; b.fEb = fEb.p; b.fRb = fRb.p;
  b.ftb = ftb.p; b.ab = ab.p; b.rb = rb.p; b.Gb = Gb.p;
  EPH_HIP(this, launch_fill_double(reinterpret_cast<double*>(wb.p), 4 * (long)na, 0.0, s));
  EPH_HIP(this, launch_fill_double(reinterpret_cast<double*>(fEb.p), 4 * (long)nl, 0.0, s));
  EPH_HIP(this, launch_fill_double(reinterpret_cast<double*>(fRb.p), 4 * (long)nl, 0.0, s));
  EPH_HIP(this, launch_fill_double(ab.p, (long)na, 0.0, s));
  EPH_HIP(this, launch_fill_double(rb.p, (long)na, 0.0, s));
  EPH_HIP(this, launch_fill_double(dTb.p, G, 0.0, s));
  // FDM: Tb (after) -> Tb2 (before), dTb
  GridView gv = ctx->grid->v;
  gv.dT = r.DT.p;
  double* cur = Tb.p;
  if (ctx->flags & EPH_FDM) {
    eph_grid* grid = ctx->grid;
    const bool tdyn = grid->any_tdyn;
    double rr = 0, inner = 0;
    unsigned S = 0;
    if (tdyn) {
      // this step's solve: C_e(T), kappa(T) at its starting T_e, then the
      // stability extrema and the sub-step count from them
      GridView g0 = gv;
      g0.T = r.TB.p;
      EPH_HIP(this, launch_grid_tdyn_update(g0, s));
      EPH_HIP(this, launch_grid_minmax(g0, grid->minmax.p, nullptr, s));
      double mm[3];
      EPH_HIP(this, hipMemcpyAsync(mm, grid->minmax.p, sizeof mm, hipMemcpyDeviceToHost, s));
      EPH_HIP(this, hipStreamSynchronize(s));
      grid->cmin = mm[0]; grid->rmin = mm[1]; grid->kmax = mm[2];
      EPH_HIP(this, kb.reserve(G));
      EPH_HIP(this, launch_fill_double(kb.p, G, 0.0, s));
    } else {
      grid->refresh_minmax();
    }
    grid->substeps(&rr, &S, &inner);
    EPH_HIP(this, Tsub.reserve((size_t)(S + 1) * G));
    HCPY(Tsub.p, r.TB.p, G * sizeof(double));
    for (unsigned k = 0; k < S; ++k)
      EPH_HIP(this, launch_fdm_substep(gv, Tsub.p + (size_t)k * G, Tsub.p + (size_t)(k + 1) * G,
                                       inner, tdyn, s));
    double* other = Tb2.p;
    for (int k = (int)S - 1; k >= 0; --k) {
      EPH_HIP(this, launch_fill_double(other, G, 0.0, s));
      EPH_HIP(this, launch_adj_fdm(gv, Tsub.p + (size_t)k * G, cur, other, dTb.p, tdyn ? kb.p : nullptr,
                                   inner, s));
      std::swap(cur, other);
    }
    if (tdyn) EPH_HIP(this, launch_adj_kappa_T(gv, r.TB.p, kb.p, cur, s));
  }
  b.Tb = cur;  // dL/dT_e before this step's solve; the force sweep adds its part
  AdjStepView sv;
  sv.rec = r.rec.p; sv.pv = r.pv.p; sv.vn = r.vn.p; sv.fE = r.fE.p; sv.fR = r.fR.p;
  sv.fstride = ctx->fstride; sv.act = r.act.p; sv.nact = r.nact.p; sv.lj = r.lj.p; sv.ljn = r.ljn.p;
  sv.owner = r.owner.p; sv.nghost = r.nghost;
  AtomsView a = ctx->atoms_view();
  a.nghost = r.nghost;
  GridView gt = gv;
  gt.T = r.TB.p;
  AdjStepFlags f;
  f.fric = (ctx->flags & EPH_FRICTION) != 0;
  f.rnd = (ctx->flags & EPH_RANDOM) != 0;
  f.addE = f.fric && !(ctx->flags & EPH_NOFRICTION);
  f.addR = f.rnd && !(ctx->flags & EPH_NORANDOM);
  f.noint = (ctx->flags & EPH_NOINT) != 0;
  f.model4 = ctx->model == 4;
  const double sg = md->cfg.lj_sigma, s6 = sg * sg * sg * sg * sg * sg;
  f.lj1 = 48.0 * md->cfg.lj_eps * s6 * s6;
  f.lj2 = 24.0 * md->cfg.lj_eps * s6;
  f.ljrc2 = md->cfg.cut_lj * md->cfg.cut_lj;
  EPH_HIP(this, launch_adj_step(sv, a, ctx->tabs, gt, ctx->step_params(), f,
                                (ctx->flags & EPH_FDM) ? dTb.p : nullptr, b, s));
  if (cur != Tb.p) HCPY(Tb.p, cur, G * sizeof(double));
  return EPH_OK;
}

extern "C" {

int eph_spline_jacobian(double dx, const double* y, int n, double* jac) {
  if (!y || !jac || n < 4 || !(dx > 0)) return EPH_ERR_ARGS;
  std::vector<double> yy(y, y + n), J;
  if (!spline_jacobian(dx, yy, J)) return EPH_ERR_ARGS;
  std::memcpy(jac, J.data(), J.size() * sizeof(double));
  return EPH_OK;
}

int eph_adjoint_create(eph_adjoint** out, eph_md* md, eph_context* ctx, int nsteps, int every,
                       const double* T_target) {
  if (!out) return EPH_ERR_ARGS;
  *out = nullptr;
  if (!md || !ctx || nsteps < 1 || every < 1) {
    g_adj_error = "bad arguments";
    return EPH_ERR_ARGS;
  }
  if (!ctx->have_atoms || ctx->nlocal != md->nlocal || !md->device_rebuild) {
    g_adj_error = "the harness must be attached to the fix (device rebuilds on)";
    return EPH_ERR_STATE;
  }
  if (ctx->model != 4 || ctx->colour >= 2) {
    g_adj_error = "only the PRL model of the eph and eph/coloured styles is differentiated";
    return EPH_ERR_UNSUPPORTED;
  }
  auto* a = new eph_adjoint();
  *out = a;
  a->md = md; a->ctx = ctx; a->nsteps = nsteps; a->every = every;
  a->nl = ctx->nlocal; a->G = (int)ctx->grid->h.ntotal();
  a->n_elem = ctx->tabs.n_elem; a->n_beta = ctx->tabs.n_beta;
  a->beta_file = ctx->beta.beta_y;
  a->beta_cur = a->beta_file;
  hipStream_t s = ctx->stream;
  if (T_target) {
    a->grid_term = true;
    a->T_target.assign(T_target, T_target + a->G);
  }
  int rc = a->save(a->start);
  if (rc) return rc;
  // start velocities of the locals
  std::vector<double4> v4(a->nl);
  EPH_HIP(a, hipMemcpyAsync(v4.data(), ctx->vel.p, a->nl * sizeof(double4), hipMemcpyDeviceToHost, s));
  EPH_HIP(a, hipStreamSynchronize(s));
  a->v_start.resize(3 * (size_t)a->nl);
  for (int i = 0; i < a->nl; ++i) {
    a->v_start[3 * i] = v4[i].x; a->v_start[3 * i + 1] = v4[i].y; a->v_start[3 * i + 2] = v4[i].z;
  }
  a->v_cur = a->v_start;
  const size_t na = a->nall_cap();
  EPH_HIP(a, a->xb.reserve(na)); EPH_HIP(a, a->vhb.reserve(na)); EPH_HIP(a, a->wb.reserve(na));
  EPH_HIP(a, a->vnb.reserve(a->nl)); EPH_HIP(a, a->fEb.reserve(a->nl)); EPH_HIP(a, a->fRb.reserve(a->nl));
  EPH_HIP(a, a->ftb.reserve(a->nl)); EPH_HIP(a, a->ab.reserve(na)); EPH_HIP(a, a->rb.reserve(na));
  EPH_HIP(a, a->Tb.reserve(a->G)); EPH_HIP(a, a->Tb2.reserve(a->G)); EPH_HIP(a, a->dTb.reserve(a->G));
  EPH_HIP(a, a->Gb.reserve(4 * (size_t)a->n_elem * a->n_beta));
  const int nck = (nsteps + every - 1) / every;
  a->ckpt.resize(nck);
  return EPH_OK;
}

void eph_adjoint_destroy(eph_adjoint* a) { delete a; }

const char* eph_adjoint_last_error(const eph_adjoint* a) { return a ? a->err.c_str() : g_adj_error.c_str(); }

int eph_adjoint_info(const eph_adjoint* a, int* n_elem, int* n_beta, int* nlocal, int* ncells) {
  if (!a) return EPH_ERR_ARGS;
  if (n_elem) *n_elem = a->n_elem;
  if (n_beta) *n_beta = a->n_beta;
  if (nlocal) *nlocal = a->nl;
  if (ncells) *ncells = a->G;
  return EPH_OK;
}

int eph_adjoint_get_beta(const eph_adjoint* a, double* beta) {
  if (!a || !beta) return EPH_ERR_ARGS;
  for (int e = 0; e < a->n_elem; ++e)
    std::copy(a->beta_file[e].begin(), a->beta_file[e].end(), beta + (size_t)e * a->n_beta);
  return EPH_OK;
}

int eph_adjoint_get_v0(const eph_adjoint* a, double* v0) {
  if (!a || !v0) return EPH_ERR_ARGS;
  std::copy(a->v_start.begin(), a->v_start.end(), v0);
  return EPH_OK;
}

int eph_adjoint_forward(eph_adjoint* a, const double* beta, const double* v0, double* loss,
                        double* parts) {
  if (!a) return EPH_ERR_ARGS;
  eph_context* ctx = a->ctx;
  hipStream_t s = ctx->stream;
  a->have_forward = false;
  std::vector<std::vector<double>> bk = a->beta_file;
  if (beta)
    for (int e = 0; e < a->n_elem; ++e) bk[e].assign(beta + (size_t)e * a->n_beta, beta + (size_t)(e + 1) * a->n_beta);
  int rc = a->set_tables(bk);
  if (rc) return rc;
  a->beta_cur = bk;
  if ((rc = a->restore(a->start))) return rc;
  a->v_cur = v0 ? std::vector<double>(v0, v0 + 3 * (size_t)a->nl) : a->v_start;
  if (v0) {
    std::vector<double4> v4(a->nl);
    for (int i = 0; i < a->nl; ++i) v4[i] = make_double4(v0[3 * i], v0[3 * i + 1], v0[3 * i + 2], 0.0);
    EPH_HIP(a, hipMemcpyAsync(ctx->vel.p, v4.data(), a->nl * sizeof(double4), hipMemcpyHostToDevice, s));
    EPH_HIP(a, hipStreamSynchronize(s));
    ctx->pv_fresh = false;
  }
  for (int n = 1; n <= a->nsteps; ++n) {
    if ((n - 1) % a->every == 0 && (rc = a->save(a->ckpt[(n - 1) / a->every]))) return rc;
    if ((rc = a->step(n, nullptr))) return rc;
  }
  double E[2] = {0, 0};
  std::vector<double> T(a->G);
  EPH_HIP(a, hipMemcpyAsync(E, ctx->scal.p, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  EPH_HIP(a, hipMemcpyAsync(T.data(), ctx->grid->T.p, a->G * sizeof(double), hipMemcpyDeviceToHost, s));
  EPH_HIP(a, hipStreamSynchronize(s));
  double sq = 0.0;
  if (a->grid_term)
    for (int c = 0; c < a->G; ++c) sq += (T[c] - a->T_target[c]) * (T[c] - a->T_target[c]);
  if (!std::isfinite(E[1]) || !std::isfinite(sq)) return a->fail(EPH_ERR_NONFINITE, "non-finite trajectory");
  if (loss) *loss = E[1] + sq;
  if (parts) { parts[0] = E[1]; parts[1] = sq; }
  a->have_forward = true;
  return EPH_OK;
}

int eph_adjoint_backward(eph_adjoint* a, double* g_beta, double* g_v0) {
  if (!a) return EPH_ERR_ARGS;
  if (!a->have_forward) return a->fail(EPH_ERR_STATE, "backward needs a forward call first");
  eph_context* ctx = a->ctx;
  hipStream_t s = ctx->stream;
  const int G = a->G, nl = a->nl;
  // dL/dT_e at the end
  {
    std::vector<double> T(G), tb(G, 0.0);
    EPH_HIP(a, hipMemcpyAsync(T.data(), ctx->grid->T.p, G * sizeof(double), hipMemcpyDeviceToHost, s));
    EPH_HIP(a, hipStreamSynchronize(s));
    if (a->grid_term)
      for (int c = 0; c < G; ++c) tb[c] = 2.0 * (T[c] - a->T_target[c]);
    EPH_HIP(a, hipMemcpyAsync(a->Tb.p, tb.data(), G * sizeof(double), hipMemcpyHostToDevice, s));
  }
  const size_t na = a->nall_cap();
  EPH_HIP(a, launch_fill_double(reinterpret_cast<double*>(a->xb.p), 4 * (long)na, 0.0, s));
  EPH_HIP(a, launch_fill_double(reinterpret_cast<double*>(a->vhb.p), 4 * (long)na, 0.0, s));
  EPH_HIP(a, launch_fill_double(reinterpret_cast<double*>(a->vnb.p), 4 * (long)nl, 0.0, s));
  EPH_HIP(a, launch_fill_double(reinterpret_cast<double*>(a->ftb.p), 4 * (long)nl, 0.0, s));
  EPH_HIP(a, launch_fill_double(a->Gb.p, 4 * (long)a->n_elem * a->n_beta, 0.0, s));
  // record buffers for one segment
  const size_t st = ctx->stride, gc = a->md->ghost_cap;
  a->seg.resize(a->every);
  for (auto& r : a->seg) {
    EPH_HIP(a, r.rec.reserve(na)); EPH_HIP(a, r.pv.reserve(na)); EPH_HIP(a, r.vn.reserve(nl));
    EPH_HIP(a, r.fE.reserve(3 * (size_t)ctx->fstride)); EPH_HIP(a, r.fR.reserve(3 * (size_t)ctx->fstride));
    EPH_HIP(a, r.TB.reserve(G)); EPH_HIP(a, r.DT.reserve(G));
    EPH_HIP(a, r.act.reserve((size_t)a->md->cap_f * st)); EPH_HIP(a, r.nact.reserve(st));
    EPH_HIP(a, r.lj.reserve((size_t)a->md->cap_l * st)); EPH_HIP(a, r.ljn.reserve(st));
    EPH_HIP(a, r.owner.reserve(gc));
  }
  int rc;
  const int nseg = (int)a->ckpt.size();
  for (int sgi = nseg - 1; sgi >= 0; --sgi) {
    const int first = sgi * a->every + 1, last = std::min(a->nsteps, (sgi + 1) * a->every);
    if ((rc = a->restore(a->ckpt[sgi]))) return rc;
    for (int n = first; n <= last; ++n)
      if ((rc = a->step(n, &a->seg[n - first]))) return rc;
    for (int n = last; n >= first; --n) {
      ctx->step = n;
      if ((rc = a->backward_step(a->seg[n - first], nullptr))) return rc;
    }
  }
  // dL/dv0 = the v' adjoint before step 1
  std::vector<double4> vb(nl);
  std::vector<double> Gh(4 * (size_t)a->n_elem * a->n_beta);
  EPH_HIP(a, hipMemcpyAsync(vb.data(), a->vnb.p, nl * sizeof(double4), hipMemcpyDeviceToHost, s));
  EPH_HIP(a, hipMemcpyAsync(Gh.data(), a->Gb.p, Gh.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  EPH_HIP(a, hipStreamSynchronize(s));
  if (g_v0)
    for (int i = 0; i < nl; ++i) {
      g_v0[3 * i] = vb[i].x; g_v0[3 * i + 1] = vb[i].y; g_v0[3 * i + 2] = vb[i].z;
    }
  if (g_beta) {
    // alpha table coefficients -> alpha knots -> beta knots (alpha_k = sqrt(beta_k))
    const int nb = a->n_beta;
    for (int e = 0; e < a->n_elem; ++e) {
      std::vector<double> ak(nb), J;
      for (int j = 0; j < nb; ++j) ak[j] = std::sqrt(a->beta_cur[e][j]);
      if (!spline_jacobian(ctx->beta.drho, ak, J)) return a->fail(EPH_ERR_ARGS, "spline Jacobian");
      for (int m = 0; m < nb; ++m) {
        double ybar = 0.0;
        for (int k = 0; k < nb; ++k)
          for (int q = 0; q < 4; ++q) ybar += Gh[4 * ((size_t)e * nb + k) + q] * J[((size_t)k * 4 + q) * nb + m];
        g_beta[(size_t)e * nb + m] = ak[m] > 0.0 ? ybar / (2.0 * ak[m]) : 0.0;
      }
    }
  }
  return EPH_OK;
}

}  // extern "C"
