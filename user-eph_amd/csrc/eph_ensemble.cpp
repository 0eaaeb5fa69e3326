// eph_ensemble.cpp -- batched ensembles of fix eph systems (eph_ensemble.h).
//
// One device system holds every replica: locals of replica 0, 1, ... then
// the ghosts the device rebuild writes (pass-major, which keeps each
// replica's images and list entries in the order of a separate run).  The
// fix phases use the chunk-mapped kernels of eph_kernels.hip; the harness
// phases (integrate, ghost update, pair, rebuild) are replica-agnostic.
#include <algorithm>
#include <cmath>

#include "eph_harness.hpp"

extern "C" {
#include "eph_ensemble.h"
}

using namespace eph;

namespace {
thread_local std::string g_ens_error;
}

struct eph_ensemble {
  int device = 0;
  std::string err;
  struct Replica {
    eph_md* md = nullptr;
    eph_context* fix = nullptr;
    int atom_off = 0, nlocal = 0;
    int cell_off = 0, ncells = 0;
    int chunk0 = 0, nchunks = 0;
    unsigned substeps = 0;
    double inner_dt = 0;
  };
  std::vector<Replica> reps;
  std::vector<std::string> args0;
  int ntypes = 0;
  double dt = 0, boltz = 0, ftm2v = 0;
  std::vector<double> masses;
  eph_context* sys = nullptr;  // the concatenated system (atoms, tables, fix parameters)
  bool ready = false;
  int nl = 0, G = 0, nchunk = 0, max_sub = 0;
  int ghost_cap = 0, cap_f = 0, cap_l = 0, nbins = 0;
  long long step0 = 1, next_step = 1;
  uint64_t log_first = 0;  // host source of the device copy (pinned by lifetime)
  eph_md_config cfg0{};
  // harness
  DevArray<int> rep, flags, cnt, start, bin_of, bcount, bstart, bfill, items, scan_tmp, nbi,
      lj_nbr, lj_nn, rlog;
  DevArray<double> box_d;
  DevArray<double4> xref;
  NbBuild nbw;
  // fix batching
  DevArray<Chunk> chunks;
  DevArray<uint32_t> seeds;
  DevArray<uint64_t> step_dev;
  DevArray<GridView> gviews;
  DevArray<int> cell_rep, cell_off, chunk0, nchunks, sub_n, nonfinite;
  DevArray<double> inner_dt, E_local, Ee;
  DevArray<double> T, Tn, dT, S, rho, C, kappa;
  DevArray<int8_t> flag, tdyn;
  // temperature-dependent C_e, kappa: the sub-step counts are found on the
  // device each step and the largest one is read back before the stencil runs
  bool any_tdyn = false;
  DevArray<int8_t> rep_tdyn;
  DevArray<double> grid_dt;
  DevArray<int> grid_steps, max_sub_dev;
  int* max_sub_host = nullptr;  // pinned
  DepositBuffers dep;
  GridView gall;  // the concatenated grid as one flat array (deposit, clears)
  // recorded step
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_nlog = -1;
  int var_density = 0, var_force = 0;  // kernel forms chosen by autotune()

  int fail(int code, const std::string& m) { err = m; return code; }
  int fail_hip(hipError_t e, const char* what) {
    err = std::string("HIP error ") + hipGetErrorString(e) + " in " + what;
    return e == hipErrorOutOfMemory ? EPH_ERR_OOM : EPH_ERR_HIP;
  }
  ~eph_ensemble() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto& r : reps) {
      eph_destroy(r.fix);
      eph_md_destroy(r.md);
    }
    if (max_sub_host) (void)hipHostFree(max_sub_host);
    eph_destroy(sys);
  }
  int setup(long long first);
  int enqueue_step(int nlog);
  int enqueue_fdm(int nsub);
  int fdm_after_counts();
  int record(int nlog);
  int autotune(double* ms);
};

int eph_ensemble::setup(long long first) {
  const int R = static_cast<int>(reps.size());
  if (R == 0) return fail(EPH_ERR_STATE, "no replicas");
  // the concatenated system's fix: replica 0's command (tables, flags, dt)
  {
    std::vector<const char*> av;
    for (auto& s : args0) av.push_back(s.c_str());
    const eph_md* m0 = reps[0].md;
    int rc = eph_create(&sys, (int)av.size(), av.data(), ntypes, m0->lo, m0->hi, dt, boltz, ftm2v,
                        device);
    if (rc != EPH_OK) return fail(rc, eph_last_error(sys));
    if ((rc = eph_set_masses(sys, masses.data(), ntypes))) return fail(rc, sys->err);
    if ((rc = eph_init(sys, 3, 0, 0, nullptr))) return fail(rc, sys->err);
  }
  hipStream_t s = sys->stream;
  // replica layout
  nl = 0; G = 0; nchunk = 0; max_sub = 0;
  int ng_est = 0, mxf = 0, mxl = 0;
  for (int r = 0; r < R; ++r) {
    Replica& p = reps[r];
    eph_md* m = p.md;
    p.atom_off = nl; p.nlocal = m->nlocal; nl += m->nlocal;
    eph_grid* g = p.fix->grid;
    any_tdyn = any_tdyn || g->any_tdyn;
    p.cell_off = G; p.ncells = (int)g->h.ntotal(); G += p.ncells;
    p.chunk0 = nchunk; p.nchunks = (m->nlocal + kBlock - 1) / kBlock; nchunk += p.nchunks;
    g->refresh_minmax();
    double rr = 0;
    g->substeps(&rr, &p.substeps, &p.inner_dt);
    if (sys->flags & EPH_FDM) max_sub = std::max(max_sub, (int)p.substeps);
    ng_est += m->nghost;
    for (int i = 0; i < m->nlocal; ++i) {
      mxf = std::max(mxf, m->fix.off[i + 1] - m->fix.off[i]);
      mxl = std::max(mxl, m->lj.off[i + 1] - m->lj.off[i]);
    }
  }
  ghost_cap = ng_est + ng_est / 4 + 1024 * R;
  if (mxf > kListSortMax)
    return fail(EPH_ERR_UNSUPPORTED, "a fix list holds more than 512 entries (the device list sort's limit)");
  cap_f = std::min(mxf + mxf / 4 + 16, kListSortMax);
  cap_l = mxl + mxl / 4 + 16;
  // concatenated locals
  std::vector<double> x(3 * (size_t)nl), v(3 * (size_t)nl);
  std::vector<int> ty(nl), mk(nl), tg(nl), rp((size_t)nl + ghost_cap, 0);
  for (int r = 0; r < R; ++r) {
    const Replica& p = reps[r];
    const eph_md* m = p.md;
    std::copy(m->x.begin(), m->x.begin() + 3 * (size_t)m->nlocal, x.begin() + 3 * (size_t)p.atom_off);
    std::copy(m->v.begin(), m->v.begin() + 3 * (size_t)m->nlocal, v.begin() + 3 * (size_t)p.atom_off);
    std::copy(m->type.begin(), m->type.begin() + m->nlocal, ty.begin() + p.atom_off);
    std::copy(m->mask.begin(), m->mask.begin() + m->nlocal, mk.begin() + p.atom_off);
    std::copy(m->tag.begin(), m->tag.begin() + m->nlocal, tg.begin() + p.atom_off);
    std::fill(rp.begin() + p.atom_off, rp.begin() + p.atom_off + m->nlocal, r);
  }
  int rc;
  if ((rc = sys->reserve_atoms(nl, ghost_cap))) return fail(rc, sys->err);
  if ((rc = sys->set_atoms(nl, 0, x.data(), v.data(), ty.data(), mk.data(), tg.data())))
    return fail(rc, sys->err);
  const int st = sys->stride;
  EPH_HIP(this, sys->nbr.reserve((size_t)cap_f * st));
  EPH_HIP(this, sys->act.reserve((size_t)cap_f * st));
  EPH_HIP(this, sys->cand.reserve((size_t)cap_f * st));  // the density sweep's candidate lists
  EPH_HIP(this, sys->ghost_owner.reserve(ghost_cap));
  EPH_HIP(this, sys->ghost_shift.reserve(ghost_cap));
  EPH_HIP(this, lj_nbr.reserve((size_t)cap_l * st));
  EPH_HIP(this, lj_nn.reserve(st));
  EPH_HIP(this, xref.reserve(nl));
  const size_t cap = (size_t)nl + ghost_cap;
  EPH_HIP(this, rep.reserve(cap));
  EPH_HIP(this, hipMemcpyAsync(rep.p, rp.data(), cap * sizeof(int), hipMemcpyHostToDevice, s));
  // per-replica boxes and bin grids
  const double cf = cfg0.cut_fix + cfg0.skin, cl = cfg0.cut_lj + cfg0.skin, cmax = std::max(cf, cl);
  const double bs = 0.5 * cmax;
  std::vector<double> bx(12 * (size_t)R);
  std::vector<int> nbv(4 * (size_t)R);
  nbins = 0;
  for (int r = 0; r < R; ++r) {
    const eph_md* m = reps[r].md;
    for (int d = 0; d < 3; ++d) {
      bx[3 * r + d] = m->lo[d];
      bx[3 * R + 3 * r + d] = m->hi[d];
      bx[6 * R + 3 * r + d] = m->L[d];
      bx[9 * R + 3 * r + d] = m->lo[d] - cmax;
      nbv[3 * r + d] = std::max(1, (int)std::ceil((m->L[d] + 2 * cmax) / bs) + 1);
    }
    nbv[3 * R + r] = nbins;
    nbins += nbv[3 * r] * nbv[3 * r + 1] * nbv[3 * r + 2];
  }
  EPH_HIP(this, box_d.reserve(bx.size()));
  EPH_HIP(this, nbi.reserve(nbv.size()));
  EPH_HIP(this, hipMemcpyAsync(box_d.p, bx.data(), bx.size() * sizeof(double), hipMemcpyHostToDevice, s));
  EPH_HIP(this, hipMemcpyAsync(nbi.p, nbv.data(), nbv.size() * sizeof(int), hipMemcpyHostToDevice, s));
  EPH_HIP(this, flags.reserve(8));
  EPH_HIP(this, cnt.reserve(cap));
  EPH_HIP(this, start.reserve(cap + 1));
  EPH_HIP(this, bin_of.reserve(cap));
  EPH_HIP(this, items.reserve(cap));
  EPH_HIP(this, bcount.reserve(nbins + 1));
  EPH_HIP(this, bstart.reserve(nbins + 1));
  EPH_HIP(this, bfill.reserve(nbins + 1));
  EPH_HIP(this, scan_tmp.reserve((cap + nbins) / 4096 + 4));
  NbBuild& w = nbw;
  w.nlocal = nl; w.ghost_cap = ghost_cap; w.nbins = nbins;
  w.at.pos = sys->pos.p; w.at.vel = sys->vel.p; w.at.type = sys->type.p; w.at.mask = sys->mask.p;
  w.at.tag = sys->tag.p; w.at.rep = rep.p; w.at.owner = sys->ghost_owner.p; w.at.shift = sys->ghost_shift.p;
  w.box.lo = box_d.p; w.box.hi = box_d.p + 3 * R; w.box.L = box_d.p + 6 * R; w.box.glo = box_d.p + 9 * R;
  w.box.nb = nbi.p; w.box.bin_base = nbi.p + 3 * R; w.box.bs = bs; w.box.cg = cmax;
  w.lists.nbr_f = sys->nbr.p; w.lists.nn_f = sys->nnbr.p; w.lists.cap_f = cap_f; w.lists.cf2 = cf * cf;
  w.lists.sort_scale = list_sort_scale(w.lists.cf2);
  w.lists.nbr_l = lj_nbr.p; w.lists.nn_l = lj_nn.p; w.lists.cap_l = cap_l; w.lists.cl2 = cl * cl;
  w.lists.stride = st;
  w.go = flags.p; w.overflow = flags.p + 1; w.nrebuild = flags.p + 2; w.nghost = flags.p + 4;
  w.cnt = cnt.p; w.start = start.p; w.bin_of = bin_of.p; w.bcount = bcount.p; w.bstart = bstart.p;
  w.bfill = bfill.p; w.items = items.p; w.scan_tmp = scan_tmp.p; w.xref = xref.p;
  // each rebuild also writes the density sweep's candidate lists, valid at
  // the rebuild's positions (the buffers keep their size for the ensemble's life)
  {
    sys->drop_candidates();
    const WorkView cw = sys->work_view();
    w.lists.cand = cw.rcand2 > 0.0 ? cw.cand : nullptr;
    w.lists.ncand = cw.ncand;
    w.lists.cand2 = cw.rcand2;
    w.cref = cw.cref;
    w.cand_valid = sys->cstate.p;
  }
  // initial ghosts and lists: a forced device build
  EPH_HIP(this, launch_fill_int(flags.p, 8, 0, s));
  EPH_HIP(this, launch_nb_force(w.go, s));
  EPH_HIP(this, launch_nb_build(w, s));
  int fl[8];
  EPH_HIP(this, hipMemcpyAsync(fl, flags.p, sizeof fl, hipMemcpyDeviceToHost, s));
  EPH_HIP(this, hipStreamSynchronize(s));
  if (fl[1]) return fail(EPH_ERR_STATE, "ensemble ghost or list capacity exceeded at set-up");
  sys->nghost = fl[7];
  sys->have_list = true;
  // Verlet::setup: pair forces of the initial configuration
  EPH_HIP(this, launch_lj(sys->atoms_view(), lj_nbr.p, lj_nn.p, st, cfg0.lj_eps, cfg0.lj_sigma,
                          cfg0.cut_lj * cfg0.cut_lj, s));
  // concatenated T_e grids and per-replica views
  EPH_HIP(this, T.reserve(G)); EPH_HIP(this, Tn.reserve(G)); EPH_HIP(this, dT.reserve(G));
  EPH_HIP(this, S.reserve(G)); EPH_HIP(this, rho.reserve(G)); EPH_HIP(this, C.reserve(G));
  EPH_HIP(this, kappa.reserve(G)); EPH_HIP(this, flag.reserve(G)); EPH_HIP(this, tdyn.reserve(G));
  std::vector<GridView> gv(R);
  std::vector<int> crep(G), coff(R), c0(R), nc(R), subn(R);
  std::vector<double> idt(R), gdt(R);
  std::vector<int> gsteps(R);
  std::vector<int8_t> rtd(R);
  std::vector<uint32_t> sd(R);
  std::vector<Chunk> ch(nchunk);
  for (int r = 0; r < R; ++r) {
    const Replica& p = reps[r];
    eph_grid* g = p.fix->grid;
    EPH_HIP(this, hipStreamSynchronize(g->stream));
    const size_t o = p.cell_off, n = p.ncells, bd = n * sizeof(double);
    EPH_HIP(this, hipMemcpyAsync(T.p + o, g->T.p, bd, hipMemcpyDeviceToDevice, s));
    EPH_HIP(this, hipMemcpyAsync(Tn.p + o, g->T.p, bd, hipMemcpyDeviceToDevice, s));
    EPH_HIP(this, hipMemcpyAsync(dT.p + o, g->dT.p, bd, hipMemcpyDeviceToDevice, s));
    EPH_HIP(this, hipMemcpyAsync(S.p + o, g->S.p, bd, hipMemcpyDeviceToDevice, s));
    EPH_HIP(this, hipMemcpyAsync(rho.p + o, g->rho.p, bd, hipMemcpyDeviceToDevice, s));
    EPH_HIP(this, hipMemcpyAsync(C.p + o, g->C.p, bd, hipMemcpyDeviceToDevice, s));
    EPH_HIP(this, hipMemcpyAsync(kappa.p + o, g->kappa.p, bd, hipMemcpyDeviceToDevice, s));
    EPH_HIP(this, hipMemcpyAsync(flag.p + o, g->flag.p, n, hipMemcpyDeviceToDevice, s));
    EPH_HIP(this, hipMemcpyAsync(tdyn.p + o, g->tdyn.p, n, hipMemcpyDeviceToDevice, s));
    GridView v = g->v;
    v.T = T.p + o; v.Tn = Tn.p + o; v.dT = dT.p + o; v.S = S.p + o; v.rho = rho.p + o;
    v.C = C.p + o; v.kappa = kappa.p + o; v.flag = flag.p + o; v.tdyn = tdyn.p + o;
    gv[r] = v;
    std::fill(crep.begin() + o, crep.begin() + o + n, r);
    coff[r] = p.cell_off; c0[r] = p.chunk0; nc[r] = p.nchunks;
    subn[r] = (int)p.substeps; idt[r] = p.inner_dt;
    gdt[r] = g->h.dt; gsteps[r] = g->h.steps; rtd[r] = g->any_tdyn ? 1 : 0;
    sd[r] = static_cast<uint32_t>(p.fix->seed);
    for (int b = 0; b < p.nchunks; ++b) {
      Chunk c;
      c.rep = r; c.start = p.atom_off + b * kBlock;
      c.count = std::min(kBlock, p.nlocal - b * kBlock); c.pad = 0;
      ch[p.chunk0 + b] = c;
    }
  }
  gall = GridView();
  gall.nx = G; gall.ny = 1; gall.nz = 1; gall.T = T.p; gall.Tn = Tn.p; gall.dT = dT.p;
  EPH_HIP(this, gviews.reserve(R)); EPH_HIP(this, cell_rep.reserve(G)); EPH_HIP(this, cell_off.reserve(R));
  EPH_HIP(this, chunk0.reserve(R)); EPH_HIP(this, nchunks.reserve(R)); EPH_HIP(this, sub_n.reserve(R));
  EPH_HIP(this, inner_dt.reserve(R)); EPH_HIP(this, seeds.reserve(R)); EPH_HIP(this, chunks.reserve(nchunk));
  EPH_HIP(this, E_local.reserve(R)); EPH_HIP(this, Ee.reserve(R)); EPH_HIP(this, step_dev.reserve(2));
  EPH_HIP(this, nonfinite.reserve(1));
  EPH_HIP(this, rep_tdyn.reserve(R)); EPH_HIP(this, grid_dt.reserve(R));
  EPH_HIP(this, grid_steps.reserve(R)); EPH_HIP(this, max_sub_dev.reserve(1));
  if (!max_sub_host) EPH_HIP(this, hipHostMalloc(reinterpret_cast<void**>(&max_sub_host), sizeof(int)));
  *max_sub_host = max_sub;
  EPH_HIP(this, hipMemcpyAsync(rep_tdyn.p, rtd.data(), R * sizeof(int8_t), hipMemcpyHostToDevice, s));
  EPH_HIP(this, hipMemcpyAsync(grid_dt.p, gdt.data(), R * sizeof(double), hipMemcpyHostToDevice, s));
  EPH_HIP(this, hipMemcpyAsync(grid_steps.p, gsteps.data(), R * sizeof(int), hipMemcpyHostToDevice, s));
  EPH_HIP(this, hipMemcpyAsync(gviews.p, gv.data(), R * sizeof(GridView), hipMemcpyHostToDevice, s));
  EPH_HIP(this, hipMemcpyAsync(cell_rep.p, crep.data(), G * sizeof(int), hipMemcpyHostToDevice, s));
  EPH_HIP(this, hipMemcpyAsync(cell_off.p, coff.data(), R * sizeof(int), hipMemcpyHostToDevice, s));
  EPH_HIP(this, hipMemcpyAsync(chunk0.p, c0.data(), R * sizeof(int), hipMemcpyHostToDevice, s));
  EPH_HIP(this, hipMemcpyAsync(nchunks.p, nc.data(), R * sizeof(int), hipMemcpyHostToDevice, s));
  EPH_HIP(this, hipMemcpyAsync(sub_n.p, subn.data(), R * sizeof(int), hipMemcpyHostToDevice, s));
  EPH_HIP(this, hipMemcpyAsync(inner_dt.p, idt.data(), R * sizeof(double), hipMemcpyHostToDevice, s));
  EPH_HIP(this, hipMemcpyAsync(seeds.p, sd.data(), R * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  EPH_HIP(this, hipMemcpyAsync(chunks.p, ch.data(), nchunk * sizeof(Chunk), hipMemcpyHostToDevice, s));
  const uint64_t s0 = static_cast<uint64_t>(first - 1);
  EPH_HIP(this, hipMemcpyAsync(step_dev.p, &s0, sizeof s0, hipMemcpyHostToDevice, s));
  EPH_HIP(this, launch_fill_double(E_local.p, R, 0.0, s));
  EPH_HIP(this, launch_fill_double(Ee.p, R, 0.0, s));
  EPH_HIP(this, launch_fill_int(nonfinite.p, 1, 0, s));
  EPH_HIP(this, dep.reserve(nl, G, nchunk));
  EPH_HIP(this, hipStreamSynchronize(s));
  step0 = first;
  next_step = first;
  ready = true;
  return EPH_OK;
}

// one Verlet step of every replica (the step order of eph_md_run)
int eph_ensemble::enqueue_step(int nlog) {
  hipStream_t s = sys->stream;
  const int st = sys->stride;
  const int* ng = flags.p + 7;
  int rc;
  EPH_HIP(this, launch_step_advance(step_dev.p, s));
  if ((rc = sys->dev_initial_integrate())) return fail(rc, sys->err);
  const double lim = 0.25 * cfg0.skin * cfg0.skin;
  EPH_HIP(this, launch_nb_check(nl, sys->pos.p, xref.p, lim, nbw.go, s));
  EPH_HIP(this, launch_nb_build(nbw, s));
  if (nlog > 0) EPH_HIP(this, launch_step_log(nbw.go, step_dev.p, step_dev.p + 1, rlog.p, nlog, s));
  AtomsView a = sys->atoms_view();
  a.nghost = ghost_cap;  // the ghost kernels below read the count on the device
  const WorkView w = sys->work_view();
  EPH_HIP(this, launch_ghost_xv_dev(a, sys->ghost_owner.p, sys->ghost_shift.p, ng, ghost_cap, s));
  EPH_HIP(this, launch_lj(a, lj_nbr.p, lj_nn.p, st, cfg0.lj_eps, cfg0.lj_sigma,
                          cfg0.cut_lj * cfg0.cut_lj, s));
  // post_force (fix_eph.cpp:841-907)
  StepParams p = sys->step_params();
  p.xi_mode = (sys->flags & EPH_RANDOM) ? 2 : 0;
  EPH_HIP(this, launch_pack_pv_dev(a, w, ng, ghost_cap, s));
  EPH_HIP(this, launch_density_w_batch(a, sys->tabs, p, w, chunks.p, nchunk, seeds.p, step_dev.p, s,
                                       var_density, ng, rep.p));
  EPH_HIP(this, launch_ghost_rec_dev(a, w, ng, ghost_cap, s));
  if (sys->model == 4)
    EPH_HIP(this, launch_prl_forces_batch(a, sys->tabs, gviews.p, p, w, chunks.p, nchunk, s,
                                          var_force));
  if ((rc = sys->dev_final_integrate())) return fail(rc, sys->err);
  // end_of_step (fix_eph.cpp:350-429) and EPH_FDM::solve per replica
  const bool fric = sys->flags & EPH_FRICTION, rnd = sys->flags & EPH_RANDOM;
  const int R = static_cast<int>(reps.size());
  if (fric || rnd) {
    DepositView d = dep.view(true);
    EPH_HIP(this, launch_fill_int(d.count, G + 1, 0, s));
    EPH_HIP(this, launch_deposit_atoms_batch(a, gviews.p, cell_off.p, p, w, d, chunks.p, nchunk, s));
    DepositView dc = d;
    if (!fric) dc.v1 = nullptr;
    if (!rnd) dc.v2 = nullptr;
    EPH_HIP(this, launch_deposit_cells(gall, dc, nl, dep.qbuf.p, s));
    EPH_HIP(this, launch_energy_reduce_batch(d.partial, chunk0.p, nchunks.p, R, E_local.p, Ee.p, s));
  } else {
    EPH_HIP(this, launch_fill_double(E_local.p, R, 0.0, s));
  }
  if (sys->flags & EPH_FDM) {
    if (!any_tdyn) return enqueue_fdm(max_sub);
    // C_e(T), kappa(T), then each replica's sub-step count (eph_fdm.h:280-314);
    // the stencil launches follow once the host has read the largest count
    EPH_HIP(this, launch_grid_tdyn_update_batch(gviews.p, cell_rep.p, cell_off.p, G, s));
    EPH_HIP(this, launch_fill_int(max_sub_dev.p, 1, 0, s));
    EPH_HIP(this, launch_fdm_substeps_batch(gviews.p, rep_tdyn.p, grid_dt.p, grid_steps.p, sub_n.p,
                                            inner_dt.p, max_sub_dev.p, R, s));
    EPH_HIP(this, hipMemcpyAsync(max_sub_host, max_sub_dev.p, sizeof(int), hipMemcpyDeviceToHost, s));
  }
  return EPH_OK;
}

// the stencil sub-steps of every replica: nsub launches, replicas with fewer
// sub-steps copy in the extra ones; then dT_e is cleared for the next step
int eph_ensemble::enqueue_fdm(int nsub) {
  hipStream_t s = sys->stream;
  for (int n = 0; n < nsub; ++n) {
    const double* in = (n % 2 == 0) ? T.p : Tn.p;
    double* out = (n % 2 == 0) ? Tn.p : T.p;
    EPH_HIP(this, launch_fdm_substep_batch(gviews.p, cell_rep.p, cell_off.p, in, out, inner_dt.p,
                                           sub_n.p, n, G, any_tdyn, s));
  }
  if (nsub % 2 == 1) EPH_HIP(this, launch_copy_double(Tn.p, T.p, G, s));
  EPH_HIP(this, launch_fill_double(dT.p, G, 0.0, s));
  EPH_HIP(this, launch_nonfinite(T.p, G, nonfinite.p, s));
  return EPH_OK;
}

// temperature-dependent grids: wait for the step's largest sub-step count
// and run the stencil
int eph_ensemble::fdm_after_counts() {
  if (!any_tdyn || !(sys->flags & EPH_FDM)) return EPH_OK;
  EPH_HIP(this, hipStreamSynchronize(sys->stream));
  const int n = *max_sub_host;
  if (n < 1 || n > (1 << 22))
    return fail(EPH_ERR_NONFINITE, "T_e sub-step count out of range (" + std::to_string(n) +
                                       "): C_e, rho_e or kappa of a grid is not usable");
  return enqueue_fdm(n);
}

// Time every form of the two neighbour sweeps on the current system and keep
// the fastest of each: one chunk per 256-thread block (tables through the
// caches), held to 4, 3 or 2 blocks per CU, or persistent blocks of 1024, 512
// or 256 threads with the tables in LDS.  The density sweep only rewrites its
// own outputs; the force sweep is timed with the force application switched
// off, so f is untouched.  Either way the next step recomputes everything
// these launches wrote.
int eph_ensemble::autotune(double* ms) {
  hipStream_t s = sys->stream;
  AtomsView a = sys->atoms_view();
  a.nghost = ghost_cap;
  const WorkView w = sys->work_view();
  const int* ng = flags.p + 7;
  StepParams p = sys->step_params();
  p.xi_mode = (sys->flags & EPH_RANDOM) ? 2 : 0;
  EPH_HIP(this, launch_pack_pv_dev(a, w, ng, ghost_cap, s));
  EPH_HIP(this, launch_density_w_batch(a, sys->tabs, p, w, chunks.p, nchunk, seeds.p, step_dev.p, s, 0, ng, rep.p));
  EPH_HIP(this, launch_ghost_rec_dev(a, w, ng, ghost_cap, s));
  StepParams pf = p;
  pf.flags |= EPH_NOFRICTION | EPH_NORANDOM;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t he = hipEventCreate(&e0);
  if (he == hipSuccess) he = hipEventCreate(&e1);
  constexpr int nf = EPH_ENSEMBLE_FORMS;
  double t[2 * nf];
  std::fill(t, t + 2 * nf, -1.0);  // forms that cannot run here stay at -1
  const bool lds_ok = batch_lds_possible(sys->tabs);
  for (int phase = 0; phase < 2 && he == hipSuccess; ++phase) {
    if (phase == 1 && sys->model != 4) break;
    for (int v = 0; v < nf && he == hipSuccess; ++v) {
      if (v >= kLdsForm1024 && v <= kLdsForm256 && !lds_ok) continue;
      double acc = 0.0;
      for (int trial = 0; trial < 4 && he == hipSuccess; ++trial) {  // first launch warms up
        he = hipEventRecord(e0, s);
        if (he == hipSuccess)
          he = phase == 0 ? launch_density_w_batch(a, sys->tabs, p, w, chunks.p, nchunk, seeds.p,
                                                   step_dev.p, s, v, ng, rep.p)
                          : launch_prl_forces_batch(a, sys->tabs, gviews.p, pf, w, chunks.p, nchunk,
                                                    s, v);
        if (he == hipSuccess) he = hipEventRecord(e1, s);
        if (he == hipSuccess) he = hipEventSynchronize(e1);
        float m = 0;
        if (he == hipSuccess) he = hipEventElapsedTime(&m, e0, e1);
        if (trial > 0) acc += m / 3.0;
      }
      t[phase * nf + v] = acc;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (he != hipSuccess) return fail_hip(he, "autotune");
  auto best = [&](const double* tt) {
    int b = 0;
    for (int v = 1; v < nf; ++v)
      if (tt[v] >= 0.0 && tt[v] < tt[b]) b = v;
    return b;
  };
  var_density = best(t);
  var_force = sys->model == 4 ? best(t + nf) : 0;
  if (ms) std::copy(t, t + 2 * nf, ms);
  // a recorded step holds the old forms
  if (exec) { (void)hipGraphExecDestroy(exec); exec = nullptr; }
  if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
  graph_nlog = -1;
  return EPH_OK;
}

int eph_ensemble::record(int nlog) {
  if (exec && graph_nlog == nlog) return EPH_OK;
  if (exec) { (void)hipGraphExecDestroy(exec); exec = nullptr; }
  if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
  hipStream_t s = sys->stream;
  EPH_HIP(this, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int rc = enqueue_step(nlog);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(s, &g);
  if (rc != EPH_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail_hip(e, "hipStreamEndCapture");
  graph = g;
  EPH_HIP(this, hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  graph_nlog = nlog;
  return EPH_OK;
}

extern "C" {

int eph_ensemble_create(eph_ensemble** out, int device) {
  if (!out) return EPH_ERR_ARGS;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    g_ens_error = "no HIP device";
    return EPH_ERR_NO_DEVICE;
  }
  if (device < 0 || device >= ndev) {
    g_ens_error = "device index out of range";
    return EPH_ERR_ARGS;
  }
  auto* e = new eph_ensemble();
  e->device = device;
  *out = e;
  return EPH_OK;
}

void eph_ensemble_destroy(eph_ensemble* e) { delete e; }

const char* eph_ensemble_last_error(const eph_ensemble* e) {
  return e ? e->err.c_str() : g_ens_error.c_str();
}

int eph_ensemble_add(eph_ensemble* e, const eph_md_config* cfg, int narg, const char* const* arg,
                     int ntypes, double dt, double boltz, double ftm2v, const double* masses) {
  if (!e || !cfg || !arg || narg < 1 || ntypes < 1 || !masses) return EPH_ERR_ARGS;
  if (e->ready) return e->fail(EPH_ERR_STATE, "replicas are added before set-up");
  if (!e->reps.empty()) {
    const eph_md_config& c = e->cfg0;
    if (cfg->cut_fix != c.cut_fix || cfg->cut_lj != c.cut_lj || cfg->skin != c.skin ||
        cfg->lj_eps != c.lj_eps || cfg->lj_sigma != c.lj_sigma || ntypes != e->ntypes || dt != e->dt)
      return e->fail(EPH_ERR_ARGS, "replicas must share cutoffs, skin, pair, types and dt");
    for (int t = 1; t <= ntypes; ++t)
      if (masses[t] != e->masses[t]) return e->fail(EPH_ERR_ARGS, "replicas must share masses");
  }
  eph_ensemble::Replica r;
  int rc = eph_md_create_fcc(&r.md, cfg);
  if (rc != EPH_OK) {
    const std::string m = r.md ? r.md->err : "harness creation failed";
    eph_md_destroy(r.md);
    return e->fail(rc, m);
  }
  rc = eph_create(&r.fix, narg, arg, ntypes, r.md->lo, r.md->hi, dt, boltz, ftm2v, e->device);
  if (rc != EPH_OK) {
    e->err = eph_last_error(r.fix);
    eph_destroy(r.fix);
    eph_md_destroy(r.md);
    return rc;
  }
  if (r.fix->colour >= 2 || r.fix->model != 4) {
    e->err = "ensemble replicas run the PRL model of the eph or eph/coloured styles";
    eph_destroy(r.fix);
    eph_md_destroy(r.md);
    return EPH_ERR_UNSUPPORTED;
  }
  if ((rc = eph_set_masses(r.fix, masses, ntypes))) {
    e->err = r.fix->err;
    eph_destroy(r.fix);
    eph_md_destroy(r.md);
    return rc;
  }
  if (e->reps.empty()) {
    e->cfg0 = *cfg;
    e->args0.assign(arg, arg + narg);
    e->ntypes = ntypes;
    e->dt = dt; e->boltz = boltz; e->ftm2v = ftm2v;
    e->masses.assign(masses, masses + ntypes + 1);
  } else {
    const eph_context* f0 = e->reps[0].fix;
    bool same = narg == (int)e->args0.size();
    for (int k = 16; same && k < narg; ++k) same = e->args0[k] == arg[k];  // .beta file, elements
    if (r.fix->flags != f0->flags || r.fix->model != f0->model || !same) {
      eph_destroy(r.fix);
      eph_md_destroy(r.md);
      return e->fail(EPH_ERR_ARGS, "replicas must share the .beta file, flags and model");
    }
  }
  e->reps.push_back(r);
  return EPH_OK;
}

int eph_ensemble_setup(eph_ensemble* e, long long first_step) {
  if (!e) return EPH_ERR_ARGS;
  if (e->ready) return e->fail(EPH_ERR_STATE, "already set up");
  return e->setup(first_step);
}

int eph_ensemble_run(eph_ensemble* e, int nsteps, int use_graph, double* ms, int* rebuilt) {
  if (!e || nsteps < 0) return EPH_ERR_ARGS;
  if (!e->ready) return e->fail(EPH_ERR_STATE, "ensemble not set up");
  hipStream_t s = e->sys->stream;
  const int nlog = rebuilt ? nsteps : 0;
  if (nlog > 0) {
    EPH_HIP(e, e->rlog.reserve(nlog));
    EPH_HIP(e, launch_fill_int(e->rlog.p, nlog, 0, s));
  }
  // the log indexes steps from the first step of this call
  e->step0 = e->next_step;
  e->log_first = static_cast<uint64_t>(e->step0);
  EPH_HIP(e, hipMemcpyAsync(e->step_dev.p + 1, &e->log_first, sizeof(uint64_t), hipMemcpyHostToDevice, s));
  int rc;
  if (use_graph && (rc = e->record(nlog))) return rc;
  hipEvent_t a, b;
  EPH_HIP(e, hipEventCreate(&a));
  EPH_HIP(e, hipEventCreate(&b));
  EPH_HIP(e, hipEventRecord(a, s));
  for (int k = 0; k < nsteps; ++k) {
    if (use_graph) {
      EPH_HIP(e, hipGraphLaunch(e->exec, s));
    } else if ((rc = e->enqueue_step(nlog))) {
      return rc;
    }
    if ((rc = e->fdm_after_counts())) return rc;
  }
  EPH_HIP(e, hipEventRecord(b, s));
  EPH_HIP(e, hipEventSynchronize(b));
  float t = 0;
  EPH_HIP(e, hipEventElapsedTime(&t, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  e->next_step += nsteps;
  e->sys->step = e->next_step - 1;
  if (ms) *ms = t;
  int fl[8], nf = 0;
  EPH_HIP(e, hipMemcpyAsync(fl, e->flags.p, sizeof fl, hipMemcpyDeviceToHost, s));
  EPH_HIP(e, hipMemcpyAsync(&nf, e->nonfinite.p, sizeof nf, hipMemcpyDeviceToHost, s));
  if (nlog > 0) EPH_HIP(e, hipMemcpyAsync(rebuilt, e->rlog.p, nlog * sizeof(int), hipMemcpyDeviceToHost, s));
  EPH_HIP(e, hipStreamSynchronize(s));
  e->sys->nghost = fl[7];
  if (fl[1]) return e->fail(EPH_ERR_STATE, "ensemble ghost or list capacity exceeded during the run");
  if (nf & 2) return e->fail(EPH_ERR_NONFINITE, "T_e grid became non-finite (FDM instability)");
  return EPH_OK;
}

int eph_ensemble_autotune(eph_ensemble* e, int* variants, double* ms) {
  if (!e) return EPH_ERR_ARGS;
  if (!e->ready) return e->fail(EPH_ERR_STATE, "ensemble not set up");
  int rc = e->autotune(ms);
  if (rc) return rc;
  if (variants) { variants[0] = e->var_density; variants[1] = e->var_force; }
  return EPH_OK;
}

int eph_ensemble_set_variants(eph_ensemble* e, int density, int force) {
  if (!e || density < 0 || density >= EPH_ENSEMBLE_FORMS || force < 0 || force >= EPH_ENSEMBLE_FORMS)
    return EPH_ERR_ARGS;
  e->var_density = density;
  e->var_force = force;
  if (e->exec) { (void)hipGraphExecDestroy(e->exec); e->exec = nullptr; }
  if (e->graph) { (void)hipGraphDestroy(e->graph); e->graph = nullptr; }
  e->graph_nlog = -1;
  return EPH_OK;
}

int eph_ensemble_size(const eph_ensemble* e, int* nrep, long long* natoms, long long* ncells) {
  if (!e) return EPH_ERR_ARGS;
  long long na = 0, nc = 0;
  for (const auto& r : e->reps) {
    na += r.md->nlocal;
    nc += (long long)r.fix->grid->h.ntotal();
  }
  if (nrep) *nrep = static_cast<int>(e->reps.size());
  if (natoms) *natoms = na;
  if (ncells) *ncells = nc;
  return EPH_OK;
}

int eph_ensemble_info(const eph_ensemble* e, int r, int* nlocal, int* ncells, int* substeps) {
  if (!e || r < 0 || r >= (int)e->reps.size()) return EPH_ERR_ARGS;
  const auto& p = e->reps[r];
  if (nlocal) *nlocal = p.md->nlocal;
  if (ncells) *ncells = (int)p.fix->grid->h.ntotal();
  if (substeps) {
    int n = 0;
    if (e->ready) {  // the count of the last step (set up on the host, or found on the device)
      hipStream_t s = e->sys->stream;
      if (hipMemcpyAsync(&n, e->sub_n.p + r, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess)
        return EPH_ERR_HIP;
    } else {
      double rr = 0, idt = 0;
      unsigned u = 0;
      p.fix->grid->refresh_minmax();
      p.fix->grid->substeps(&rr, &u, &idt);
      n = (int)u;
    }
    *substeps = (e->reps[0].fix->flags & EPH_FDM) ? n : 0;
  }
  return EPH_OK;
}

int eph_ensemble_get(eph_ensemble* e, int r, int which, double* out) {
  if (!e || !out || r < 0 || r >= (int)e->reps.size()) return EPH_ERR_ARGS;
  if (!e->ready) return e->fail(EPH_ERR_STATE, "ensemble not set up");
  const auto& p = e->reps[r];
  eph_context* c = e->sys;
  hipStream_t s = c->stream;
  const int n = p.nlocal;
  switch (which) {
    case 0:
    case 1:
      EPH_HIP(e, launch_d4_into_rows((which == 0 ? c->pos.p : c->vel.p) + p.atom_off, n, c->stage.p, s));
      EPH_HIP(e, hipMemcpyAsync(out, c->stage.p, 3 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s));
      break;
    case 2:
      EPH_HIP(e, launch_soa_into_rows(c->f.p + p.atom_off, n, c->fstride, c->stage.p, s));
      EPH_HIP(e, hipMemcpyAsync(out, c->stage.p, 3 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s));
      break;
    case 3:
      EPH_HIP(e, hipMemcpyAsync(out, e->T.p + p.cell_off, p.ncells * sizeof(double), hipMemcpyDeviceToHost, s));
      break;
    case 4:
      EPH_HIP(e, hipMemcpyAsync(out, e->E_local.p + r, sizeof(double), hipMemcpyDeviceToHost, s));
      EPH_HIP(e, hipMemcpyAsync(out + 1, e->Ee.p + r, sizeof(double), hipMemcpyDeviceToHost, s));
      break;
    case 5: {  // candidate-list statistics of the last batch density sweep
      int st[8] = {};
      if (c->cstate.p) EPH_HIP(e, hipMemcpyAsync(st, c->cstate.p, sizeof st, hipMemcpyDeviceToHost, s));
      EPH_HIP(e, hipStreamSynchronize(s));
      long long waves = 0;  // chunks are whole blocks of one replica, so waves per replica
      for (const auto& q : e->reps) waves += (q.nlocal + 63) / 64;
      out[0] = st[3]; out[1] = static_cast<double>(waves); out[2] = st[4]; out[3] = st[5];
      return EPH_OK;
    }
    default:
      return e->fail(EPH_ERR_ARGS, "unknown quantity");
  }
  EPH_HIP(e, hipStreamSynchronize(s));
  return EPH_OK;
}

}  // extern "C"
