// eph_kernels.hip -- gfx950 kernels of the fix eph hot path.
//
// Arithmetic contract: every kernel evaluates the reference's FP64
// expressions in the reference's operation order, one thread per atom (or
// per grid cell) summing its neighbours in list order, with FMA contraction
// disabled at build time.  Results are therefore bit-identical to the serial
// CPU FixEPH on the same inputs; no atomics touch floating-point sums.  In
// the cooperative forms a wavefront fetches its atoms' neighbour records
// together through LDS; the sums themselves stay one thread per atom.
#include "eph_devfn.hpp"
#include "eph_rng.hpp"

namespace eph {


// ---------------------------------------------------------------- comm
__global__ __launch_bounds__(kBlock) void k_ghost_xv(AtomsView a, const int* __restrict__ owner,
                                                     const double4* __restrict__ shift) {
  const int g = blockIdx.x * kBlock + threadIdx.x;
  if (g >= a.nghost) return;
  const int o = owner[g];
  const double4 p = a.pos[o];
  const double4 s = shift[g];
  const double4 x = make_double4(p.x + s.x, p.y + s.y, p.z + s.z, p.w);
  const double4 v = a.vel[o];
  a.pos[a.nlocal + g] = x;
  a.vel[a.nlocal + g] = v;
  if (a.pv) {
    PosVel q;
    q.p = x;
    q.v = v;
    a.pv[a.nlocal + g] = q;
  }
}

// CSR -> transposed ELL; an index outside [0, nall) raises *bad and is stored
// as 0, so that the list never addresses outside the atom arrays
__global__ __launch_bounds__(kBlock) void k_csr_to_ell(int nlocal, const int* __restrict__ off,
                                                       const int* __restrict__ idx, int neighmask,
                                                       int nall, int* __restrict__ ell,
                                                       int* __restrict__ nn, int stride, int width,
                                                       int* bad) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= nlocal) return;
  const int b = off[i];
  int e = off[i + 1];
  // a row longer than the ELL width (or with decreasing offsets) is refused
  // and left empty: the device copy of the offsets is never trusted to stay
  // inside the buffer
  if (e - b > width || e < b) {
    atomicOr(bad, 2);
    e = b;
  }
  for (int k = b; k < e; ++k) {
    int j = idx[k] & neighmask;
    if (j < 0 || j >= nall) {
      atomicOr(bad, 1);
      j = 0;
    }
    ell[(k - b) * stride + i] = j;
  }
  nn[i] = e - b;
}

// ------------------------------------------------- cooperative gathers
// A wave fetches the k-th neighbours of its 64 atoms together: the lanes
// copying one record take consecutive 16-B parts of it, straight into LDS
// (global_load_lds), so a gather instruction touches one cache line per
// record instead of one per lane; each lane then reads its own record from
// LDS.  Every copy fills a 1 KB block (the force sweep's add 128 B of
// padding); inside a record the parts are rotated per record so that the
// per-lane reads are free of bank conflicts.  The whole wave runs these loops
// (to its longest list), lanes past their own list doing only copies.
static __device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int kCoopStage = 384;  // 16-B units of LDS per wave and force-sweep entry (64 x 96 B)
constexpr int kCoopF = 1;        // force-sweep entries per iteration
// density sweep: entries per iteration of the rho loop (128 x 16 B of LDS
// each) and of the w loop (256 x 16 B each); their copies need no padding,
// since lanes reading from different copies are never served together
constexpr int kCoopRho = 4, kCoopW = 2;
constexpr int kCoopStageW = (128 * kCoopRho > 256 * kCoopW) ? 128 * kCoopRho : 256 * kCoopW;

static __device__ __forceinline__ int wave_max(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

using lds_void = __attribute__((address_space(3))) void*;

// 32-B position records, 2 lanes per record: in copy m (0, 1) lane 2q + s
// fetches slot s of the record of lane 32 m + q, which holds part s ^ g(q)
static __device__ __forceinline__ int pos_rot(int q) { return ((q >> 2) ^ (q >> 3)) & 1; }
static __device__ __forceinline__ void coop_fetch_pos(const double2* __restrict__ pos2, int j, int lane,
                                                      double2* stage) {
  const int q = lane >> 1, part = (lane & 1) ^ pos_rot(lane >> 1);
  int r0 = __shfl(j, q), r1 = __shfl(j, 32 + q);
  asm volatile("" : "+v"(r0), "+v"(r1));
  __builtin_amdgcn_global_load_lds(pos2 + static_cast<size_t>(r0) * 2 + part, (lds_void)(stage + 0), 16, 0, 0);
  __builtin_amdgcn_global_load_lds(pos2 + static_cast<size_t>(r1) * 2 + part, (lds_void)(stage + 64), 16, 0, 0);
}
static __device__ __forceinline__ double4 coop_read_pos(const double2* stage, int lane) {
  const int q = lane & 31, g = pos_rot(q);
  const double2* row = stage + 64 * (lane >> 5) + 2 * q;
  const double2 a = row[g], b = row[g ^ 1];
  return make_double4(a.x, a.y, b.x, b.y);
}

// 64-B (position, velocity) records, 4 lanes per record: in copy m (0..3)
// lane 4q + s fetches slot s of the record of lane 16 m + q, which holds part
// (s - h(q)) mod 4
static __device__ __forceinline__ int pv_rot(int q) { return ((q >> 1) + (q >> 3)) & 3; }
static __device__ __forceinline__ void coop_fetch_pv(const double2* __restrict__ pv2, int j, int lane,
                                                     double2* stage) {
  const int q = lane >> 2, part = ((lane & 3) - pv_rot(lane >> 2)) & 3;
  int r0 = __shfl(j, q), r1 = __shfl(j, 16 + q), r2 = __shfl(j, 32 + q), r3 = __shfl(j, 48 + q);
  asm volatile("" : "+v"(r0), "+v"(r1), "+v"(r2), "+v"(r3));
  const double2* g = pv2 + part;
  __builtin_amdgcn_global_load_lds(g + static_cast<size_t>(r0) * 4, (lds_void)(stage + 0), 16, 0, 0);
  __builtin_amdgcn_global_load_lds(g + static_cast<size_t>(r1) * 4, (lds_void)(stage + 64), 16, 0, 0);
  __builtin_amdgcn_global_load_lds(g + static_cast<size_t>(r2) * 4, (lds_void)(stage + 128), 16, 0, 0);
  __builtin_amdgcn_global_load_lds(g + static_cast<size_t>(r3) * 4, (lds_void)(stage + 192), 16, 0, 0);
}
static __device__ __forceinline__ void coop_read_pv(const double2* stage, int lane, double4* pj,
                                                    double4* vj) {
  const int q = lane & 15, h = pv_rot(q);
  const double2* row = stage + 64 * (lane >> 4) + 4 * q;
  const double2 a = row[h], b = row[(1 + h) & 3], c = row[(2 + h) & 3], d = row[(3 + h) & 3];
  *pj = make_double4(a.x, a.y, b.x, b.y);
  *vj = make_double4(c.x, c.y, d.x, d.y);
}

// ------------------------------------------------- density + w (stage 1)
// FixEPH::calculate_environment (fix_eph.cpp:431-466) followed by stage 1
// of force_prl (fix_eph.cpp:702-741) and the noise draw of
// fix_eph.cpp:854-862.  The density sweep also compacts the entries with
// r^2 < rc^2 into an active list, which the w sweep and the force sweep
// re-read in the same order (they skip exactly the other entries).
// `rtab` holds the rho(r^2) tables, in global memory or staged in LDS.
// alpha(rho_i) (eph_beta.h:164-198 via fix_eph.cpp:711); rho above the
// table's range gives 0 and raises the warning flag (eph_beta.h:189-194)
template <bool MULTI>
__device__ __forceinline__ double density_alpha(const bool grp, const double rho, const double4& pi,
                                                const TablesView& t, const WorkView& w) {
  double alpha = 0.0;
  if (grp && rho > 0.0) {
    if (rho > t.rho_cut) {
      atomicOr(w.flags_out, 1);  // eph_beta.h:189-194: alpha = 0 with a warning
    } else {
      const int ei = MULTI ? static_cast<int>(pi.w) : 0;
      alpha = cubic_at(t.alpha + ei * t.n_beta, t.n_beta, t.inv_drho, rho);
    }
  }
  return alpha;
}

// one active neighbour's term of w_i (fix_eph.cpp:719-737)
template <bool MULTI>
__device__ __forceinline__ void density_w_pair(const double4& pi, const double4& vi, const double4& pj,
                                               const double4& vj, const double rho, const double alpha,
                                               const TablesView& t, const double4* __restrict__ rtab,
                                               double* wx, double* wy, double* wz) {
  const double ex = pj.x - pi.x, ey = pj.y - pi.y, ez = pj.z - pi.z;
  const double r2 = ex * ex + ey * ey + ez * ez;
  const int ej = MULTI ? static_cast<int>(pj.w) : 0;
  const double vr = rho_r2_at(rtab, t, ej, r2);
  const double pre = alpha * vr / (rho * r2);
  const double ev1 = ex * vi.x + ey * vi.y + ez * vi.z;
  const double var1 = pre * ev1;
  const double ev2 = ex * vj.x + ey * vj.y + ez * vj.z;
  const double var2 = pre * ev2;
  const double dvar = var1 - var2;
  *wx += dvar * ex;
  *wy += dvar * ey;
  *wz += dvar * ez;
}

// the noise xi_i (drawn, injected or none), the v1 filter, and atom i's record
template <int XI>
__device__ __forceinline__ void density_finish(const int i, const bool grp, const double4& pi,
                                               const double rho, const double alpha, const double wx,
                                               const double wy, const double wz, const AtomsView& a,
                                               const StepParams& p, const WorkView& w) {
  double x0 = 0.0, x1 = 0.0, x2 = 0.0;
  if (XI == 1 && grp) {
    x0 = p.xi_in[3 * i + 0];
    x1 = p.xi_in[3 * i + 1];
    x2 = p.xi_in[3 * i + 2];
  } else if (XI == 2 && grp) {
    double g3[3];
    gaussian3(p.seed, static_cast<uint32_t>(a.tag[i]), p.step, g3);
    x0 = g3[0]; x1 = g3[1]; x2 = g3[2];
  }
  if (XI != 0 && p.colour == 3) {
    // eph/coloured/exp/v1 (fix_eph_coloured_exp_v1.cpp:645-652): every local
    // atom's noise is filtered, zi <- zi (1 - zeta) + zeta xi, and the random
    // sums use zi
    const double om = 1. - p.zeta;
    double4 z = w.cB[i];
    z.x = z.x * om + p.zeta * x0;
    z.y = z.y * om + p.zeta * x1;
    z.z = z.z * om + p.zeta * x2;
    w.cB[i] = z;
    x0 = z.x; x1 = z.y; x2 = z.z;
  }
  ForceRec* r = w.rec + i;
  r->p = pi;
  r->rw = make_double4(wx, wy, wz, rho);
  r->xa = make_double4(x0, x1, x2, alpha);
}

// Per-thread form (used by the LDS-table kernels): each thread gathers its own
// neighbours, one iteration ahead of their use.
template <int XI, bool FRIC, bool MULTI>
__device__ __forceinline__ void density_w_atom(const int i, const AtomsView& a, const TablesView& t,
                                               const double4* __restrict__ rtab,
                                               const StepParams& p, const WorkView& w) {
  const int stride = a.stride;
  const double4 pi = a.pos[i];
  const bool grp = (a.mask[i] & a.groupbit) != 0;
  double rho = 0.0;
  int na = 0;
  if (grp) {
    // each neighbour's position is loaded one iteration ahead of its use
    const int nn = w.nnbr[i];
    int jn = nn > 0 ? w.nbr[i] : 0;
    double4 pn = nn > 0 ? a.pos[jn] : pi;
    for (int k = 0; k < nn; ++k) {
      const int j = jn;
      const double4 pj = pn;
      if (k + 1 < nn) {
        jn = w.nbr[(k + 1) * stride + i];
        pn = a.pos[jn];
      }
      const double ex = pj.x - pi.x, ey = pj.y - pi.y, ez = pj.z - pi.z;
      const double r2 = ex * ex + ey * ey + ez * ez;
      if (r2 < t.rc2) {
        const int ej = MULTI ? static_cast<int>(pj.w) : 0;
        rho += rho_r2_at(rtab, t, ej, r2);
        w.act[na * stride + i] = j;
        ++na;
      }
    }
  }
  w.nact[i] = na;
  const double alpha = density_alpha<MULTI>(grp, rho, pi, t, w);
  double wx = 0.0, wy = 0.0, wz = 0.0;
  if (FRIC && grp && rho > 0.0) {
    const double4 vi = w.vw[i];
    int jn = na > 0 ? w.act[i] : 0;
    double4 pn = pi, vn = vi;
    if (na > 0) {
      pn = w.pv[jn].p;
      vn = w.pv[jn].v;
    }
    for (int k = 0; k < na; ++k) {
      const double4 pj = pn, vj = vn;
      if (k + 1 < na) {
        jn = w.act[(k + 1) * stride + i];
        pn = w.pv[jn].p;
        vn = w.pv[jn].v;
      }
      density_w_pair<MULTI>(pi, vi, pj, vj, rho, alpha, t, rtab, &wx, &wy, &wz);
    }
  }
  density_finish<XI>(i, grp, pi, rho, alpha, wx, wy, wz, a, p, w);
}

// Is an atom listed as moved within w.reach of any of this wave's atoms?  The
// moved positions are tested against the wave's bounding box grown by reach,
// a superset of the spheres around its atoms; lanes share the list.  With
// rep >= 0 (a batch of replicas, whose lists never pair atoms of two
// replicas) only the moved atoms of replica rep count: the check writes each
// moved atom's replica into .w.
__device__ __forceinline__ bool wave_near_moved(const bool valid, const double4& pi,
                                                const WorkView& w, const int nm, const int rep) {
  const double inf = __builtin_huge_val();
  double lx = valid ? pi.x : inf, ly = valid ? pi.y : inf, lz = valid ? pi.z : inf;
  double hx = valid ? pi.x : -inf, hy = valid ? pi.y : -inf, hz = valid ? pi.z : -inf;
  for (int o = 32; o > 0; o >>= 1) {
    lx = fmin(lx, __shfl_xor(lx, o)); ly = fmin(ly, __shfl_xor(ly, o)); lz = fmin(lz, __shfl_xor(lz, o));
    hx = fmax(hx, __shfl_xor(hx, o)); hy = fmax(hy, __shfl_xor(hy, o)); hz = fmax(hz, __shfl_xor(hz, o));
  }
  const double r = w.reach;
  lx -= r; ly -= r; lz -= r;
  hx += r; hy += r; hz += r;
  const int lane = threadIdx.x & 63;
  bool hit = false;
  const int n = nm < w.moved_cap ? nm : w.moved_cap;
  for (int b = 0; b < n; b += 64) {
    const int m = b + lane;
    if (m < n) {
      const double4 q = w.moved[m];
      hit = hit || ((rep < 0 || q.w == static_cast<double>(rep)) && q.x >= lx && q.x <= hx &&
                    q.y >= ly && q.y <= hy && q.z >= lz && q.z <= hz);
    }
  }
  return __any(hit) != 0;
}

// Cooperative form of density_w_atom (same sums in the same order): the
// positions of the rho loop and the (position, velocity) records of the w loop
// come through LDS, see coop_fetch_pos / coop_fetch_pv.  The records of entry
// k + 1 are in flight while entry k is computed.  `valid` = lane i is a local
// atom; stage: kCoopStageW x 16 B of LDS owned by the calling wave; rep: the
// wave's replica in a batch, -1 otherwise.
template <int XI, bool FRIC, bool MULTI>
__device__ __forceinline__ void density_w_atom_coop(const int i_lane, const bool valid, const AtomsView& a,
                                                    const TablesView& t, const double4* __restrict__ rtab,
                                                    const StepParams& p, const WorkView& w,
                                                    double2* stage, const int rep = -1) {
  const int i = valid ? i_lane : 0;  // lanes past the end read atom 0 and store nothing
  const int stride = a.stride;
  const int lane = threadIdx.x & 63;
  const double4 pi = a.pos[i];
  const bool grp = valid && (a.mask[i] & a.groupbit) != 0;
  // the rho loop walks the candidate list when it is valid, the moved list
  // holds every moved atom and none is near this wave; it scans the full
  // list otherwise, and takes the candidates on the way when they are being
  // retaken for all atoms (the reference positions are the current ones)
  const bool keep = w.rcand2 > 0.0;
  bool use_cand = false, take = false;
  if (keep) {
    const int nm = w.cstate[2];
    take = w.cstate[0] == 0;
    use_cand = !take && nm <= w.moved_cap && !(nm > 0 && wave_near_moved(valid, pi, w, nm, rep));
    if (!use_cand && lane == 0) atomicAdd(w.cstate + 1, 1);
  }
  const int* __restrict__ list = use_cand ? w.cand : w.nbr;
  const int nn = grp ? (use_cand ? w.ncand[i] : w.nnbr[i]) : 0;
  const int nnmax = wave_max(nn);
  double rho = 0.0;
  int na = 0, nc = 0;
  {
    const double2* __restrict__ pos2 = reinterpret_cast<const double2*>(a.pos);
    // kCoopRho entries per iteration: the positions of entry k + e land at
    // stage + 128 e, and those of the next kCoopRho entries are in flight while
    // these are summed (in list order)
    constexpr int U = kCoopRho;
    int jj[U], nx[U];
#pragma unroll
    for (int e = 0; e < U; ++e) jj[e] = e < nn ? list[e * stride + i] : 0;
#pragma unroll
    for (int e = 0; e < U; ++e) nx[e] = U + e < nn ? list[(U + e) * stride + i] : 0;
    if (nnmax > 0) {
#pragma unroll
      for (int e = 0; e < U; ++e) coop_fetch_pos(pos2, jj[e], lane, stage + 128 * e);
    }
    for (int k = 0; k < nnmax; k += U) {
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      wave_lds_sync();
      double4 pq[U];
#pragma unroll
      for (int e = 0; e < U; ++e) pq[e] = coop_read_pos(stage + 128 * e, lane);
      asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
      wave_lds_sync();
      int jc[U];
#pragma unroll
      for (int e = 0; e < U; ++e) jc[e] = jj[e];
      if (k + U < nnmax) {
#pragma unroll
        for (int e = 0; e < U; ++e) coop_fetch_pos(pos2, nx[e], lane, stage + 128 * e);
#pragma unroll
        for (int e = 0; e < U; ++e) {
          jj[e] = nx[e];
          nx[e] = k + 2 * U + e < nn ? list[(k + 2 * U + e) * stride + i] : 0;
        }
      }
#pragma unroll
      for (int e = 0; e < U; ++e) {
        if (k + e < nn) {
          const double ex = pq[e].x - pi.x, ey = pq[e].y - pi.y, ez = pq[e].z - pi.z;
          const double r2 = ex * ex + ey * ey + ez * ez;
          if (r2 < t.rc2) {
            const int ej = MULTI ? static_cast<int>(pq[e].w) : 0;
            rho += rho_r2_at(rtab, t, ej, r2);
            w.act[na * stride + i] = jc[e];
            ++na;
          }
          if (take && r2 < w.rcand2) {
            w.cand[nc * stride + i] = jc[e];
            ++nc;
          }
        }
      }
    }
  }
  if (valid) w.nact[i] = na;
  if (take && valid) w.ncand[i] = nc;
  const double alpha = density_alpha<MULTI>(grp, rho, pi, t, w);
  double wx = 0.0, wy = 0.0, wz = 0.0;
  if (FRIC) {
    const int nw = (grp && rho > 0.0) ? na : 0;
    const int nwmax = wave_max(nw);
    const double4 vi = w.vw[i];
    const double2* __restrict__ pv2 = reinterpret_cast<const double2*>(w.pv);
    // kCoopW entries per iteration, (position, velocity) of entry k + e at
    // stage + 256 e
    constexpr int U = kCoopW;
    int jj[U];
#pragma unroll
    for (int e = 0; e < U; ++e) jj[e] = U + e < nw ? w.act[(U + e) * stride + i] : 0;
    if (nwmax > 0) {
#pragma unroll
      for (int e = 0; e < U; ++e) coop_fetch_pv(pv2, e < nw ? w.act[e * stride + i] : 0, lane, stage + 256 * e);
    }
    for (int k = 0; k < nwmax; k += U) {
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      wave_lds_sync();
      double4 pq[U], vq[U];
#pragma unroll
      for (int e = 0; e < U; ++e) coop_read_pv(stage + 256 * e, lane, &pq[e], &vq[e]);
      asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
      wave_lds_sync();
      if (k + U < nwmax) {
#pragma unroll
        for (int e = 0; e < U; ++e) coop_fetch_pv(pv2, jj[e], lane, stage + 256 * e);
#pragma unroll
        for (int e = 0; e < U; ++e) jj[e] = k + 2 * U + e < nw ? w.act[(k + 2 * U + e) * stride + i] : 0;
      }
#pragma unroll
      for (int e = 0; e < U; ++e)
        if (k + e < nw) density_w_pair<MULTI>(pi, vi, pq[e], vq[e], rho, alpha, t, rtab, &wx, &wy, &wz);
    }
  }
  if (valid) density_finish<XI>(i, grp, pi, rho, alpha, wx, wy, wz, a, p, w);
}

// Before the sweep, over all nall = nlocal + nghost atoms: on a step that
// retakes the candidates, the reference positions are the current ones;
// otherwise every atom 0.45 delta or farther from its reference position (a
// non-finite distance counts) puts its current position on the moved list.
__global__ __launch_bounds__(kBlock) void k_cand_check(int nall, const double4* __restrict__ pos,
                                                       WorkView w) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (w.cstate[0] == 0) {
    if (i < nall) w.cref[i] = pos[i];
    return;
  }
  if (i >= nall) return;
  const double4 x = pos[i], r = w.cref[i];
  const double dx = x.x - r.x, dy = x.y - r.y, dz = x.z - r.z;
  if (!(dx * dx + dy * dy + dz * dz < w.half_delta2)) {
    const int k = atomicAdd(w.cstate + 2, 1);
    if (k < w.moved_cap) w.moved[k] = x;
  }
}

// the same check for a batch of replicas, with the ghost count read on the
// device (the ensemble's recorded steps, whose ghost count changes between
// replays; the grid covers nlocal + the ghost capacity); a moved atom's entry
// carries its replica rep[i] in .w
__global__ __launch_bounds__(kBlock) void k_cand_check_dev(int nlocal, const int* __restrict__ ng,
                                                           const double4* __restrict__ pos,
                                                           const int* __restrict__ rep, WorkView w) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const int nall = nlocal + *ng;
  if (w.cstate[0] == 0) {
    if (i < nall) w.cref[i] = pos[i];
    return;
  }
  if (i >= nall) return;
  const double4 x = pos[i], r = w.cref[i];
  const double dx = x.x - r.x, dy = x.y - r.y, dz = x.z - r.z;
  if (!(dx * dx + dy * dy + dz * dz < w.half_delta2)) {
    const int k = atomicAdd(w.cstate + 2, 1);
    if (k < w.moved_cap) w.moved[k] = make_double4(x.x, x.y, x.z, static_cast<double>(rep[i]));
  }
}

// After the sweep: a sweep that retook the candidates leaves them valid.  A
// moved list that overflowed sent every wave to its full list; the reference
// positions then no longer bound how far the atoms are from where they were
// when the candidates were taken, so the next check takes new reference
// positions and the next sweep retakes the candidates.  The moved list starts
// empty at the next check.
__global__ void k_cand_flags(int* cstate, int moved_cap) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int retake = cstate[0] == 0;
  cstate[3] = cstate[1];
  cstate[4] = cstate[2];
  cstate[5] = retake;
  cstate[0] = retake || cstate[2] <= moved_cap ? 1 : 0;
  cstate[1] = 0;
  cstate[2] = 0;
}

// 256 atoms per block, tables through the caches, cooperative gathers
template <int XI, bool FRIC, bool MULTI>
__global__ __launch_bounds__(kBlock) void k_density_w(AtomsView a, TablesView t, StepParams p,
                                                      WorkView w) {
  // (kBlock / 64) * kCoopStageW units; a larger request only lowers the
  // blocks per CU
  extern __shared__ double2 s_stage[];
  const int base = xcd_block() * kBlock;
  if (base + static_cast<int>(threadIdx.x & ~63u) >= a.nlocal) return;  // whole wave past the end
  const int i = base + threadIdx.x;
  density_w_atom_coop<XI, FRIC, MULTI>(i, i < a.nlocal, a, t, t.rho_r2, p, w,
                                       s_stage + (threadIdx.x >> 6) * kCoopStageW);
}

// Persistent form: each block stages the rho(r^2) tables in LDS once and then
// walks tiles of kBigBlock atoms; the tiles of XCD x form one contiguous range
// shared by the blocks resident on x.
__device__ __forceinline__ void xcd_tiles(int ntiles, int* first, int* step) {
  constexpr int kXcd = 8;
  const int nb = gridDim.x, b = blockIdx.x;
  const int x = b % kXcd, l = b / kXcd;
  const int nbx = nb / kXcd + (x < nb % kXcd ? 1 : 0);  // blocks on this XCD
  const int lo = static_cast<int>((static_cast<long>(ntiles) * x) / kXcd);
  *first = lo + l;
  *step = nbx;
  (void)lo;
}

__device__ __forceinline__ int xcd_tiles_end(int ntiles) {
  constexpr int kXcd = 8;
  const int x = blockIdx.x % kXcd;
  return static_cast<int>((static_cast<long>(ntiles) * (x + 1)) / kXcd);
}

// BLK threads per block: one block per CU either way (the tables take most
// of the LDS), so BLK sets the waves per SIMD (1024: 4, 512: 2, 256: 1).
template <int XI, bool FRIC, bool MULTI, int BLK>
__global__ __launch_bounds__(BLK) void k_density_w_lds(AtomsView a, TablesView t,
                                                       StepParams p, WorkView w) {
  extern __shared__ double4 s_tab[];
  const int ntab = t.n_elem * t.n_rho;
  for (int k = threadIdx.x; k < ntab; k += BLK) s_tab[k] = t.rho_r2[k];
  __syncthreads();
  const int ntiles = (a.nlocal + BLK - 1) / BLK;
  int tile, step;
  xcd_tiles(ntiles, &tile, &step);
  const int end = xcd_tiles_end(ntiles);
  for (; tile < end; tile += step) {
    const int i = tile * BLK + threadIdx.x;
    if (i < a.nlocal) density_w_atom<XI, FRIC, MULTI>(i, a, t, s_tab, p, w);
  }
}

// forward communication of rho, w (fix_eph.cpp:743-744, 870-871) and xi
// (:863-864) on one process: ghost <- owner, next to the ghost's own position
__device__ __forceinline__ void ghost_rec(const int g, const AtomsView& a, const WorkView& w) {
  const int o = w.ghost_owner[g];
  ForceRec* r = w.rec + (a.nlocal + g);
  r->p = a.pos[a.nlocal + g];
  r->rw = w.rec[o].rw;
  r->xa = w.rec[o].xa;
}

// the halo values of the local atoms, [nlocal][7] = (xi, rho, w): what the
// XI, RHO and WI forward exchanges of fix_eph.cpp:951-985 send
__global__ __launch_bounds__(kBlock) void k_halo_pack(const ForceRec* __restrict__ rec, int n,
                                                      double* __restrict__ out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double4 rw = rec[i].rw, xa = rec[i].xa;
  double* o = out + 7 * static_cast<size_t>(i);
  o[0] = xa.x; o[1] = xa.y; o[2] = xa.z; o[3] = rw.w; o[4] = rw.x; o[5] = rw.y; o[6] = rw.z;
}

// a ghost's record from its exchanged (xi, rho, w), [nghost][7]; alpha from
// rho as for a local atom (the same function, so the same bits as its owner's)
template <bool MULTI>
__global__ __launch_bounds__(kBlock) void k_ghost_halo(AtomsView a, TablesView t, WorkView w,
                                                       const double* __restrict__ in) {
  const int g = blockIdx.x * kBlock + threadIdx.x;
  if (g >= a.nghost) return;
  const int i = a.nlocal + g;
  const double* h = in + 7 * static_cast<size_t>(g);
  const double4 pi = a.pos[i];
  const bool grp = (a.mask[i] & a.groupbit) != 0;
  const double rho = h[3];
  const double alpha = density_alpha<MULTI>(grp, rho, pi, t, w);
  ForceRec* r = w.rec + i;
  r->p = pi;
  r->rw = make_double4(h[4], h[5], h[6], rho);
  r->xa = make_double4(h[0], h[1], h[2], alpha);
}

__global__ __launch_bounds__(kBlock) void k_ghost_rec(AtomsView a, WorkView w) {
  const int g = blockIdx.x * kBlock + threadIdx.x;
  if (g < a.nghost) ghost_rec(g, a, w);
}

// positions and velocities of all atoms, packed for the w sweep
__global__ __launch_bounds__(kBlock) void k_pack_pv(AtomsView a, WorkView w, int n) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  PosVel q;
  q.p = a.pos[i];
  q.v = w.vw[i];
  w.pv[i] = q;
}

// ------------------------------------------- force_prl stages 2 and 3 fused
// fix_eph.cpp:748-787 (friction) and :791-836 (random), then the force
// application of :892-906.  Both sums share the neighbour gather; each
// accumulates in list order exactly as its reference loop does.

// The terms of one neighbour j (rho_j > 0) of atom i, added to the friction
// (fe) and random (fr) sums: one definition for every form of the sweep, so
// that all of them round alike.
template <bool FRIC, bool RAND, bool MULTI>
__device__ __forceinline__ void prl_pair_terms(const double4& pi, const double4& rwi, const double4& xai,
                                               const double4& pj, const double4& rwj, const double4& xaj,
                                               const double r2, const double4& cji, const double4& cij,
                                               double* te, double* tr) {
  const double rho_i = rwi.w, alpha_i = xai.w;
  const double rho_j = rwj.w, alpha_j = xaj.w;
  const double ex = pj.x - pi.x, ey = pj.y - pi.y, ez = pj.z - pi.z;
  const double vr_ji = cji.x + r2 * (cji.y + r2 * (cji.z + r2 * cji.w));
  const double vr_ij = MULTI ? cij.x + r2 * (cij.y + r2 * (cij.z + r2 * cij.w)) : vr_ji;
  if (FRIC) {
    const double ev1 = ex * rwi.x + ey * rwi.y + ez * rwi.z;
    const double var1 = alpha_i * vr_ji * ev1 / (rho_i * r2);
    const double ev2 = ex * rwj.x + ey * rwj.y + ez * rwj.z;
    const double var2 = alpha_j * vr_ij * ev2 / (rho_j * r2);
    const double dvar = var1 - var2;
    te[0] = dvar * ex;
    te[1] = dvar * ey;
    te[2] = dvar * ez;
  }
  if (RAND) {
    const double ex1 = ex * xai.x + ey * xai.y + ez * xai.z;
    const double var1 = alpha_i * vr_ji * ex1 / (rho_i * r2);
    const double ex2 = ex * xaj.x + ey * xaj.y + ez * xaj.z;
    const double var2 = alpha_j * vr_ij * ex2 / (rho_j * r2);
    const double dvar = var1 - var2;
    tr[0] = dvar * ex;
    tr[1] = dvar * ey;
    tr[2] = dvar * ez;
  }
}

// the pair's terms taken into the sums: f_EPH -= dvar e, f_RNG += dvar e
template <bool FRIC, bool RAND, bool MULTI>
__device__ __forceinline__ void prl_pair_coef(const double4& pi, const double4& rwi, const double4& xai,
                                              const double4& pj, const double4& rwj, const double4& xaj,
                                              const double r2, const double4& cji, const double4& cij,
                                              double* fe, double* fr) {
  double te[3], tr[3];
  prl_pair_terms<FRIC, RAND, MULTI>(pi, rwi, xai, pj, rwj, xaj, r2, cji, cij, te, tr);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (FRIC) fe[d] -= te[d];
    if (RAND) fr[d] += tr[d];
  }
}

// the same with the coefficients looked up here (rho_r2_at)
template <bool FRIC, bool RAND, bool MULTI>
__device__ __forceinline__ void prl_pair(const double4& pi, const double4& rwi, const double4& xai,
                                         const int ei, const double4& pj, const double4& rwj,
                                         const double4& xaj, const TablesView& t,
                                         const double4* __restrict__ rtab, double* fe, double* fr) {
  const double ex = pj.x - pi.x, ey = pj.y - pi.y, ez = pj.z - pi.z;
  const double r2 = ex * ex + ey * ey + ez * ez;
  int kn = static_cast<int>(r2 * t.inv_dr2);
  kn = kn < 0 ? 0 : (kn < t.n_rho - 1 ? kn : t.n_rho - 1);
  const int ej = MULTI ? static_cast<int>(pj.w) : 0;
  const double4 cji = rtab[kn * t.n_elem + ej];
  const double4 cij = MULTI ? rtab[kn * t.n_elem + ei] : cji;
  prl_pair_coef<FRIC, RAND, MULTI>(pi, rwi, xai, pj, rwj, xaj, r2, cji, cij, fe, fr);
}

// After the sums: the sqrt(T_e) factor of the random force, the coloured/exp
// filters, the stored f_EPH / f_RNG and the force update.  `on` = atom i is in
// the group with rho_i > 0 (otherwise its sums are zero and stay so).
template <bool FRIC, bool RAND>
__device__ __forceinline__ void prl_finish(const int i, const bool on, const double4& pi,
                                           double* fe, double* fr, const AtomsView& a,
                                           const GridView& g, const StepParams& p,
                                           const WorkView& w) {
  double fex = fe[0], fey = fe[1], fez = fe[2], frx = fr[0], fry = fr[1], frz = fr[2];
  if (on) {
    if (RAND) {
      const double Te = g.T[cell_of(g, pi.x, pi.y, pi.z)];
      const double var = p.eta * sqrt(Te);
      frx *= var;
      fry *= var;
      frz *= var;
    }
    if (p.colour == 2) {
      // eph/coloured/exp (fix_eph_coloured_exp.cpp:562-568, 619-625): the
      // forces applied and deposited are the filtered f_dis, f_sto
      const double om = 1. - p.zeta;
      if (FRIC) {
        double4 d = w.cA[i];
        d.x = d.x * om + p.zeta * fex;
        d.y = d.y * om + p.zeta * fey;
        d.z = d.z * om + p.zeta * fez;
        w.cA[i] = d;
        fex = d.x; fey = d.y; fez = d.z;
      }
      if (RAND) {
        double4 d = w.cB[i];
        d.x = d.x * om + p.zeta * frx;
        d.y = d.y * om + p.zeta * fry;
        d.z = d.z * om + p.zeta * frz;
        w.cB[i] = d;
        frx = d.x; fry = d.y; frz = d.z;
      }
    }
  }
  const int fs = a.fstride;
  if (FRIC) {
    w.fE[i] = fex; w.fE[fs + i] = fey; w.fE[2 * fs + i] = fez;
  }
  if (RAND) {
    w.fR[i] = frx; w.fR[fs + i] = fry; w.fR[2 * fs + i] = frz;
  }
  const bool addE = FRIC && !(p.flags & 0x10);
  const bool addR = RAND && !(p.flags & 0x20);
  if (addE || addR) {
    double fx = a.f[i], fy = a.f[fs + i], fz = a.f[2 * fs + i];
    if (addE) { fx = fx + fex; fy = fy + fey; fz = fz + fez; }
    if (addR) { fx = fx + frx; fy = fy + fry; fz = fz + frz; }
    a.f[i] = fx; a.f[fs + i] = fy; a.f[2 * fs + i] = fz;
  }
}

// One thread per atom, each gathering its own neighbours' records.  Each
// record is loaded one iteration ahead of its use; with DEEP its index is
// loaded two ahead, so the record load never waits for an index load issued
// in the same iteration (more registers: the LDS-table forms, whose occupancy
// is set by LDS, use it).
template <bool FRIC, bool RAND, bool MULTI, bool DEEP = false>
__device__ __forceinline__ void prl_forces_atom(const int i, const AtomsView& a, const TablesView& t,
                                                const double4* __restrict__ rtab,
                                                const GridView& g, const StepParams& p,
                                                const WorkView& w) {
  const int stride = a.stride;
  const bool grp = (a.mask[i] & a.groupbit) != 0;
  const double4 rwi = w.rec[i].rw;
  const double rho_i = rwi.w;
  double fe[3] = {0.0, 0.0, 0.0}, fr[3] = {0.0, 0.0, 0.0};
  const bool on = grp && rho_i > 0.0;
  double4 pi = rwi;
  if (on) {
    pi = w.rec[i].p;
    const double4 xai = w.rec[i].xa;
    const int ei = MULTI ? static_cast<int>(pi.w) : 0;
    const int na = w.nact[i];
    int jn = na > 0 ? w.act[i] : 0;
    double4 nrw = rwi, np = pi, nxa = xai;
    if (na > 0) {
      nrw = w.rec[jn].rw; np = w.rec[jn].p; nxa = w.rec[jn].xa;
    }
    if (DEEP) jn = na > 1 ? w.act[stride + i] : 0;  // index of entry k + 1
    for (int k = 0; k < na; ++k) {
      const double4 rwj = nrw, pj = np, xaj = nxa;
      if (k + 1 < na) {
        if (DEEP) {
          const int jc = jn;
          if (k + 2 < na) jn = w.act[(k + 2) * stride + i];
          nrw = w.rec[jc].rw; np = w.rec[jc].p; nxa = w.rec[jc].xa;
        } else {
          jn = w.act[(k + 1) * stride + i];
          nrw = w.rec[jn].rw; np = w.rec[jn].p; nxa = w.rec[jn].xa;
        }
      }
      if (!(rwj.w > 0.0)) continue;
      prl_pair<FRIC, RAND, MULTI>(pi, rwi, xai, ei, pj, rwj, xaj, t, rtab, fe, fr);
    }
  }
  prl_finish<FRIC, RAND>(i, on, pi, fe, fr, a, g, p, w);
}

// Cooperative form: the wave fetches its 64 atoms' k-th neighbour records
// together, straight into LDS (global_load_lds, 16 B per lane), and each lane
// then reads its own record from there.  Only parts 0-5 of each 128-B record
// (p, (w, rho), (xi, alpha): 96 B) are copied, packed: 6 copies of 64 lanes
// for the wave's 64 records, every lane busy.  In copy m lane l takes slot
// g = 64 m + l of the stage, which receives part g mod 6 of the record of
// lane g / 6, so record r lands at stage + 6 r (16-B units).  A gather
// instruction touches at most 11 cache lines instead of 64.  Every lane of the
// wave runs the loop to the wave's longest active list (lanes past their own
// do only the copies), so the whole wave must call this; `valid` = lane i is
// a local atom.  stage: kCoopStage x 16 B of LDS owned by the calling wave.
static __device__ __forceinline__ void coop_fetch6(const double2* __restrict__ recs, int j, int lane,
                                                   double2* stage) {
  using lds_ptr = __attribute__((address_space(3))) void*;
  int r[6];  // the record indices, all read before the first copy is issued
#pragma unroll
  for (int m = 0; m < 6; ++m) r[m] = __shfl(j, (64 * m + lane) / 6);
  asm volatile("" : "+v"(r[0]), "+v"(r[1]), "+v"(r[2]), "+v"(r[3]), "+v"(r[4]), "+v"(r[5]));
#pragma unroll
  for (int m = 0; m < 6; ++m)
    __builtin_amdgcn_global_load_lds(recs + static_cast<size_t>(r[m]) * 8 + (64 * m + lane) % 6,
                                     (lds_ptr)(stage + 64 * m), 16, 0, 0);
}

template <bool FRIC, bool RAND, bool MULTI>
__device__ __forceinline__ void prl_forces_atom_coop(const int i_lane, const bool valid,
                                                     const AtomsView& a, const TablesView& t,
                                                     const double4* __restrict__ rtab,
                                                     const GridView& g, const StepParams& p,
                                                     const WorkView& w, double2* stage) {
  const int i = valid ? i_lane : 0;  // lanes past the end read atom 0 and store nothing
  const int stride = a.stride;
  const int lane = threadIdx.x & 63;
  const bool grp = valid && (a.mask[i] & a.groupbit) != 0;
  const double4 rwi = w.rec[i].rw;
  const bool on = grp && rwi.w > 0.0;
  const double4 pi = w.rec[i].p;
  const double4 xai = w.rec[i].xa;
  const int ei = MULTI ? static_cast<int>(pi.w) : 0;
  const int na = on ? w.nact[i] : 0;
  int nmax = na;
  for (int o = 32; o > 0; o >>= 1) nmax = max(nmax, __shfl_xor(nmax, o));
  double fe[3] = {0.0, 0.0, 0.0}, fr[3] = {0.0, 0.0, 0.0};
  const double2* __restrict__ recs = reinterpret_cast<const double2*>(w.rec);
  const double2* mine6 = stage + 6 * lane;
  static_assert(kCoopF == 1, "one entry per iteration");
  // Two-stage pipeline.  Iteration k waits for everything in flight (the
  // coefficients of entry k, the records of entry k + 1), reads entry k + 1's
  // records from LDS, sends out its coefficient loads and the copies of entry
  // k + 2, and only then computes pair k, whose inputs are all in registers:
  // the pair's arithmetic runs while the next loads are in flight, and no
  // wait depends on the order in which loads of different kinds return.
  // Loads past the wave's longest list fetch record 0 and knot 0, which
  // nothing uses.
  constexpr unsigned kWaitVm0 = 0x0F70;  // s_waitcnt vmcnt(0), expcnt and lgkmcnt untouched
  auto row = [&](int r) { return r < nmax ? r : (nmax > 0 ? nmax - 1 : 0); };
  auto knot = [&](double r2, bool act) {
    int kn = static_cast<int>(r2 * t.inv_dr2);
    kn = kn < 0 ? 0 : (kn < t.n_rho - 1 ? kn : t.n_rho - 1);
    return act ? kn : 0;
  };
  auto read_entry = [&](double4& pj, double4& rwj, double4& xaj) {
    double2 sv[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) sv[c] = mine6[c];
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
    pj = make_double4(sv[0].x, sv[0].y, sv[1].x, sv[1].y);
    rwj = make_double4(sv[2].x, sv[2].y, sv[3].x, sv[3].y);
    xaj = make_double4(sv[4].x, sv[4].y, sv[5].x, sv[5].y);
  };
  auto dist2 = [&](const double4& pj) {
    const double ex = pj.x - pi.x, ey = pj.y - pi.y, ez = pj.z - pi.z;
    return ex * ex + ey * ey + ez * ez;
  };
  // the state of one entry: its records, coefficients, r^2 and whether it counts
  struct Entry {
    double4 pj, rwj, xaj, cji, cij;
    double r2;
    bool act;
  };
  // read entry k's records from the stage and send out its coefficient loads
  auto take = [&](Entry& e, int k) {
    read_entry(e.pj, e.rwj, e.xaj);
    wave_lds_sync();
    e.r2 = dist2(e.pj);
    e.act = k < na && e.rwj.w > 0.0;
    const int kn = knot(e.r2, e.act);
    e.cji = rtab[kn * t.n_elem + (MULTI ? static_cast<int>(e.pj.w) : 0)];
    e.cij = MULTI ? rtab[kn * t.n_elem + ei] : e.cji;
  };
  // pair k's terms into the sums (only an active lane's)
  auto add = [&](const Entry& e) {
    double te[3], tr[3];
    prl_pair_terms<FRIC, RAND, MULTI>(pi, rwi, xai, e.pj, e.rwj, e.xaj, e.r2, e.cji, e.cij, te, tr);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (FRIC) fe[d] = e.act ? fe[d] - te[d] : fe[d];
      if (RAND) fr[d] = e.act ? fr[d] + tr[d] : fr[d];
    }
  };
  int jraw = 0;
  // one iteration: entry k + 1 taken, entry k + 2's copies and entry k + 3's
  // index sent, then pair k (cur) computed while they are in flight
  auto iter = [&](int k, Entry& cur, Entry& nxt) {
    __builtin_amdgcn_s_waitcnt(kWaitVm0);
    wave_lds_sync();
    take(nxt, k + 1);
    coop_fetch6(recs, k + 2 < na ? jraw : 0, lane, stage);
    jraw = w.act[row(k + 3) * stride + i];
    // the loads above go out first; pair k's arithmetic (whose inputs are
    // all in registers, its loads having completed at the top) comes after
    // them, so it runs while they are in flight
    asm volatile("" : "+v"(cur.pj.x), "+v"(cur.pj.y), "+v"(cur.pj.z), "+v"(cur.r2) :: "memory");
    add(cur);
    // ... and ends before the next iteration's wait
    asm volatile("" : "+v"(fe[0]), "+v"(fe[1]), "+v"(fe[2]), "+v"(fr[0]), "+v"(fr[1]), "+v"(fr[2]));
  };
  Entry e0, e1;
  if (nmax > 0) {
    coop_fetch6(recs, na > 0 ? w.act[i] : 0, lane, stage);
    jraw = w.act[row(1) * stride + i];
    __builtin_amdgcn_s_waitcnt(kWaitVm0);
    wave_lds_sync();
    take(e0, 0);
    coop_fetch6(recs, 1 < na ? jraw : 0, lane, stage);
    jraw = w.act[row(2) * stride + i];
  }
  // two entries per trip, the two Entry slots taking turns (no copies of
  // registers whose loads are still in flight)
  int k = 0;
  for (; k + 1 < nmax; k += 2) {
    iter(k, e0, e1);
    iter(k + 1, e1, e0);
  }
  if (k < nmax) iter(k, e0, e1);
  // the last copies land before the wave may leave
  __builtin_amdgcn_s_waitcnt(kWaitVm0);
  if (valid) prl_finish<FRIC, RAND>(i, on, pi, fe, fr, a, g, p, w);
}

// 256 atoms per block, tables through the caches, cooperative gathers
template <bool FRIC, bool RAND, bool MULTI>
__global__ __launch_bounds__(kBlock) void k_prl_forces(AtomsView a, TablesView t, GridView g,
                                                       StepParams p, WorkView w) {
  // (kBlock / 64) * kCoopStage * kCoopF units; a larger request only lowers
  // the blocks per CU
  extern __shared__ double2 s_stage[];
  const int base = xcd_block() * kBlock;
  if (base + static_cast<int>(threadIdx.x & ~63u) >= a.nlocal) return;  // whole wave past the end
  const int i = base + threadIdx.x;
  prl_forces_atom_coop<FRIC, RAND, MULTI>(i, i < a.nlocal, a, t, t.rho_r2, g, p, w,
                                          s_stage + (threadIdx.x >> 6) * kCoopStage * kCoopF);
}

template <bool FRIC, bool RAND, bool MULTI, int BLK>
__global__ __launch_bounds__(BLK) void k_prl_forces_lds(AtomsView a, TablesView t, GridView g,
                                                        StepParams p, WorkView w) {
  extern __shared__ double4 s_tab[];
  const int ntab = t.n_elem * t.n_rho;
  for (int k = threadIdx.x; k < ntab; k += BLK) s_tab[k] = t.rho_r2[k];
  __syncthreads();
  const int ntiles = (a.nlocal + BLK - 1) / BLK;
  int tile, step;
  xcd_tiles(ntiles, &tile, &step);
  const int end = xcd_tiles_end(ntiles);
  for (; tile < end; tile += step) {
    const int i = tile * BLK + threadIdx.x;
    if (i < a.nlocal) prl_forces_atom<FRIC, RAND, MULTI, true>(i, a, t, s_tab, g, p, w);
  }
}

// ---------------------------------------------- fix eph models 1 and 2
// beta(rho) and alpha(rho) with the reference's cutoff rule: zero (and the
// warning flag) above rho_cutoff (eph_beta.h:171-198)
__device__ __forceinline__ double beta_at(const TablesView& t, int e, double rho, int* flag) {
  if (rho > t.rho_cut) {
    atomicOr(flag, 1);
    return 0.;
  }
  return cubic_at(t.beta + e * t.n_beta, t.n_beta, t.inv_drho, rho);
}
__device__ __forceinline__ double alpha_at(const TablesView& t, int e, double rho, int* flag) {
  if (rho > t.rho_cut) {
    atomicOr(flag, 1);
    return 0.;
  }
  return cubic_at(t.alpha + e * t.n_beta, t.n_beta, t.inv_drho, rho);
}

// the random force of models 1 and 2 (fix_eph.cpp:489-502, 554-567):
// f_RNG = eta alpha(rho_i) sqrt(T_e) xi_i
__device__ __forceinline__ void random_per_atom(const int i, const TablesView& t, const GridView& g,
                                                const StepParams& p, const WorkView& w,
                                                const double4& pi, double rho, double* fr) {
  const int ei = static_cast<int>(pi.w);
  const double Te = g.T[cell_of(g, pi.x, pi.y, pi.z)];
  const double var = p.eta * alpha_at(t, ei, rho, w.flags_out) * sqrt(Te);
  const double4 xa = w.rec[i].xa;
  fr[0] = var * xa.x;
  fr[1] = var * xa.y;
  fr[2] = var * xa.z;
}

// f_EPH, f_RNG out; f += them unless NOFRICTION / NORANDOM (fix_eph.cpp:892-906)
__device__ __forceinline__ void store_forces(const int i, const AtomsView& a, const StepParams& p,
                                             const WorkView& w, bool fric, bool rnd,
                                             const double* fe, const double* fr) {
  const int fs = a.fstride;
  if (fric) { w.fE[i] = fe[0]; w.fE[fs + i] = fe[1]; w.fE[2 * fs + i] = fe[2]; }
  if (rnd) { w.fR[i] = fr[0]; w.fR[fs + i] = fr[1]; w.fR[2 * fs + i] = fr[2]; }
  const bool addE = fric && !(p.flags & 0x10);
  const bool addR = rnd && !(p.flags & 0x20);
  if (addE || addR) {
    double fx = a.f[i], fy = a.f[fs + i], fz = a.f[2 * fs + i];
    if (addE) { fx = fx + fe[0]; fy = fy + fe[1]; fz = fz + fe[2]; }
    if (addR) { fx = fx + fr[0]; fy = fy + fr[1]; fz = fz + fr[2]; }
    a.f[i] = fx; a.f[fs + i] = fy; a.f[2 * fs + i] = fz;
  }
}

// model 1, FixEPH::force_ttm (fix_eph.cpp:468-503): f_EPH = -beta(rho_i) v_i
template <bool FRIC, bool RAND>
__global__ __launch_bounds__(kBlock) void k_ttm_forces(AtomsView a, TablesView t, GridView g,
                                                       StepParams p, WorkView w) {
  const int i = xcd_block() * kBlock + threadIdx.x;
  if (i >= a.nlocal) return;
  double fe[3] = {0.0, 0.0, 0.0}, fr[3] = {0.0, 0.0, 0.0};
  if ((a.mask[i] & a.groupbit) != 0) {
    const double4 pi = a.pos[i];
    const double rho = w.rec[i].rw.w;
    if (FRIC) {
      const double var = -beta_at(t, static_cast<int>(pi.w), rho, w.flags_out);
      const double4 v = a.vel[i];
      fe[0] = var * v.x; fe[1] = var * v.y; fe[2] = var * v.z;
    }
    if (RAND) random_per_atom(i, t, g, p, w, pi, rho, fr);
  }
  store_forces(i, a, p, w, FRIC, RAND, fe, fr);
}

// model 2, FixEPH::force_prb (fix_eph.cpp:505-568): f_EPH = beta(rho_i)
// (v_i - sum_j rho_{jtype-1}(r_ij) v_j / rho_i) over the entries with
// r^2 < r_c^2 in list order (the density sweep's active list), rho(r) of the
// unmapped element jtype - 1 at r = sqrt(r^2)
template <bool FRIC, bool RAND>
__global__ __launch_bounds__(kBlock) void k_prb_forces(AtomsView a, TablesView t, GridView g,
                                                       StepParams p, WorkView w) {
  const int i = xcd_block() * kBlock + threadIdx.x;
  if (i >= a.nlocal) return;
  double fe[3] = {0.0, 0.0, 0.0}, fr[3] = {0.0, 0.0, 0.0};
  if ((a.mask[i] & a.groupbit) != 0) {
    const double4 pi = a.pos[i];
    const double rho = w.rec[i].rw.w;
    if (FRIC && rho > 0.0) {
      const double4 vi = a.vel[i];
      fe[0] = vi.x; fe[1] = vi.y; fe[2] = vi.z;
      const int na = w.nact[i];
      for (int k = 0; k < na; ++k) {
        const int j = w.act[k * a.stride + i];
        const double4 pj = a.pos[j];
        const double ex = pj.x - pi.x, ey = pj.y - pi.y, ez = pj.z - pi.z;
        const double r2 = ex * ex + ey * ey + ez * ez;
        const int ej = a.type[j] - 1;
        const double var = cubic_at(t.rho_r + ej * t.n_rho, t.n_rho, t.inv_dr, sqrt(r2)) / rho;
        const double4 vj = a.vel[j];
        fe[0] -= var * vj.x;
        fe[1] -= var * vj.y;
        fe[2] -= var * vj.z;
      }
      const double b = beta_at(t, static_cast<int>(pi.w), rho, w.flags_out);
      fe[0] *= b; fe[1] *= b; fe[2] *= b;
    }
    if (RAND) random_per_atom(i, t, g, p, w, pi, rho, fr);
  }
  store_forces(i, a, p, w, FRIC, RAND, fe, fr);
}

// eph/coloured/exp/v1 (fix_eph_coloured_exp_v1.cpp:477-487): every local
// atom's velocity filtered, then the ghosts take their owners' values
__global__ __launch_bounds__(kBlock) void k_colour_v_local(AtomsView a, StepParams p, WorkView w) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.nlocal) return;
  const double om = 1. - p.zeta;
  const double4 v = a.vel[i];
  double4 z = w.cA[i];
  z.x = z.x * om + p.zeta * v.x;
  z.y = z.y * om + p.zeta * v.y;
  z.z = z.z * om + p.zeta * v.z;
  w.cA[i] = z;
}

__global__ __launch_bounds__(kBlock) void k_colour_v_ghost(AtomsView a, WorkView w) {
  const int g = blockIdx.x * kBlock + threadIdx.x;
  if (g >= a.nghost) return;
  w.cA[a.nlocal + g] = w.cA[w.ghost_owner[g]];
}

// --------------------------------------------------------- integrator
// FixEPH::initial_integrate / final_integrate (fix_eph.cpp:305-348)
template <bool INITIAL>
__global__ __launch_bounds__(kBlock) void k_integrate(AtomsView a, StepParams p) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.nlocal) return;
  if (!(a.mask[i] & a.groupbit)) return;
  const int fs = a.fstride;
  const double dtfm = p.dtf / p.mass[a.type[i]];
  double4 v = a.vel[i];
  v.x += dtfm * a.f[i];
  v.y += dtfm * a.f[fs + i];
  v.z += dtfm * a.f[2 * fs + i];
  a.vel[i] = v;
  if (INITIAL) {
    double4 x = a.pos[i];
    x.x += p.dtv * v.x;
    x.y += p.dtv * v.y;
    x.z += p.dtv * v.z;
    a.pos[i] = x;
    if (a.pv) {
      PosVel q;
      q.p = x;
      q.v = v;
      a.pv[i] = q;
    }
  }
  // the final kick leaves the records at (x', v_half) of this step's
  // post_force (what the trajectory adjoint reads back); the next initial
  // kick rewrites them before any sweep reads them
}

__global__ void k_fill_int(int* p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_double(double* p, long n, double v) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------- deposit
// FixEPH::end_of_step (fix_eph.cpp:364-390): per-atom energies of the
// friction and random forces at end-of-step v, their cell, and block sums
// of E_local.
// one atom of the deposit; `coff` is added to its cell (replica offset)
template <bool FRIC, bool RAND>
__device__ __forceinline__ void deposit_atom(const int i, const AtomsView& a, const GridView& g,
                                             const StepParams& p, const WorkView& w,
                                             const DepositView& d, const int coff, double* eF,
                                             double* eR) {
  int c = -1;
  double v1 = 0.0, v2 = 0.0;
  if (a.mask[i] & a.groupbit) {
    const double4 x = a.pos[i];
    const double4 v = a.vel[i];
    c = cell_of(g, x.x, x.y, x.z) + coff;
    const int fs = a.fstride;
    const double prescale = g.dV * p.dt;
    if (FRIC) {
      double dE = 0.0;
      dE -= w.fE[i] * v.x * p.dt;
      dE -= w.fE[fs + i] * v.y * p.dt;
      dE -= w.fE[2 * fs + i] * v.z * p.dt;
      *eF = dE;
      v1 = dE / prescale;
    }
    if (RAND) {
      double dE = 0.0;
      dE -= w.fR[i] * v.x * p.dt;
      dE -= w.fR[fs + i] * v.y * p.dt;
      dE -= w.fR[2 * fs + i] * v.z * p.dt;
      *eR = dE;
      v2 = dE / prescale;
    }
    atomicAdd(&d.count[c], 1);
  }
  d.cell[i] = c;
  d.v1[i] = v1;
  d.v2[i] = v2;
}

template <bool FRIC, bool RAND>
__global__ __launch_bounds__(kBlock) void k_deposit_atoms(AtomsView a, GridView g, StepParams p,
                                                          WorkView w, DepositView d) {
  __shared__ double sh[kBlock];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double eF = 0.0, eR = 0.0;
  if (i < a.nlocal) deposit_atom<FRIC, RAND>(i, a, g, p, w, d, 0, &eF, &eR);
  const double sF = block_sum(eF, sh);
  const double sR = block_sum(eR, sh);
  if (threadIdx.x == 0) {
    d.partial[2 * blockIdx.x] = sF;
    d.partial[2 * blockIdx.x + 1] = sR;
  }
}

// EPH_FDM::insert_energy for a list of points (eph_fdm.h:172-179)
__global__ __launch_bounds__(kBlock) void k_deposit_points(int n, const double* x, const double* y,
                                                           const double* z, const double* E,
                                                           double dt, GridView g, DepositView d) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int c = cell_of(g, x[i], y[i], z[i]);
  d.cell[i] = c;
  d.v1[i] = E[i] / (g.dV * dt);
  atomicAdd(&d.count[c], 1);
}

// exclusive scan of count[0..n) into start[0..n] (three phases)
constexpr int kScanBlock = 1024;
constexpr int kScanItems = 4;  // per thread
__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(const int* in, int n, int* block_sums) {
  __shared__ int sh[kScanBlock];
  const int base = blockIdx.x * kScanBlock * kScanItems;
  int s = 0;
  for (int k = 0; k < kScanItems; ++k) {
    const int idx = base + k * kScanBlock + threadIdx.x;
    if (idx < n) s += in[idx];
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = s;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_blocks(int* block_sums, int nb) {
  // single block: exclusive scan of nb block sums, chunked by kScanBlock
  __shared__ int sh[kScanBlock];
  int carry = 0;
  for (int base = 0; base < nb; base += kScanBlock) {
    const int idx = base + threadIdx.x;
    const int v = idx < nb ? block_sums[idx] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
      const int add = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
      __syncthreads();
      sh[threadIdx.x] += add;
      __syncthreads();
    }
    const int incl = sh[threadIdx.x];
    const int total = sh[kScanBlock - 1];
    if (idx < nb) block_sums[idx] = carry + incl - v;
    __syncthreads();
    carry += total;
  }
}

__global__ __launch_bounds__(kScanBlock) void k_scan_apply(const int* in, int n, const int* block_off,
                                                           int* out) {
  __shared__ int sh[kScanBlock];
  const int base = blockIdx.x * kScanBlock * kScanItems;
  int carry = block_off[blockIdx.x];
  for (int k = 0; k < kScanItems; ++k) {
    const int idx = base + k * kScanBlock + threadIdx.x;
    const int v = idx < n ? in[idx] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
      const int add = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
      __syncthreads();
      sh[threadIdx.x] += add;
      __syncthreads();
    }
    if (idx < n) out[idx] = carry + sh[threadIdx.x] - v;
    if (idx == n - 1) out[n] = carry + sh[threadIdx.x];
    const int total = sh[kScanBlock - 1];
    __syncthreads();
    carry += total;
  }
}

__global__ __launch_bounds__(kBlock) void k_scatter(int n, DepositView d) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int c = d.cell[i];
  if (c < 0) return;
  const int slot = d.start[c] + atomicAdd(&d.fill[c], 1);
  d.items[slot] = i;
}

// Per cell: dT_e[c] += v1 of its atoms in ascending atom order, then += v2 in
// ascending order -- the serial order of the two loops of fix_eph.cpp:364-390.
// Cells with up to kMaxSortInCell items sort them by insertion in place;
// larger cells are queued for a workgroup (medium) or an ordered scan over
// all items (huge).  d.count is reused as the queue: count[G] holds the
// number of queued cells and the queue follows in fill[].
__global__ __launch_bounds__(kBlock) void k_cell_sum_small(GridView g, DepositView d, int ncell,
                                                           int* queue, int* qn) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= ncell) return;
  const int s = d.start[c], n = d.start[c + 1] - s;
  if (n == 0) return;
  if (n > kMaxSortInCell) {
    const int q = atomicAdd(qn, 1);
    queue[q] = c;
    return;
  }
  int* it = d.items + s;
  for (int k = 1; k < n; ++k) {
    const int v = it[k];
    int m = k - 1;
    while (m >= 0 && it[m] > v) {
      it[m + 1] = it[m];
      --m;
    }
    it[m + 1] = v;
  }
  double acc = g.dT[c];
  if (d.v1)
    for (int k = 0; k < n; ++k) acc += d.v1[it[k]];
  if (d.v2)
    for (int k = 0; k < n; ++k) acc += d.v2[it[k]];
  g.dT[c] = acc;
}

constexpr int kMedBlock = 1024;
constexpr int kMedMax = 8192;
// one workgroup per queued cell: bitonic sort of up to kMedMax items in LDS,
// then the ordered sums; larger cells fall through to k_cell_sum_huge
__global__ __launch_bounds__(kMedBlock) void k_cell_sum_medium(GridView g, DepositView d,
                                                               const int* queue, const int* qn,
                                                               int* huge, int* hn) {
  __shared__ int key[kMedMax];
  const int nq = *qn;
  for (int q = blockIdx.x; q < nq; q += gridDim.x) {
    const int c = queue[q];
    const int s = d.start[c], n = d.start[c + 1] - s;
    if (n > kMedMax) {
      if (threadIdx.x == 0) huge[atomicAdd(hn, 1)] = c;
      continue;
    }
    int m = 1;
    while (m < n) m <<= 1;
    for (int k = threadIdx.x; k < m; k += kMedBlock) key[k] = k < n ? d.items[s + k] : 0x7fffffff;
    __syncthreads();
    for (int size = 2; size <= m; size <<= 1) {
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        for (int k = threadIdx.x; k < m; k += kMedBlock) {
          const int l = k ^ stride;
          if (l > k) {
            const bool up = (k & size) == 0;
            const int a = key[k], b = key[l];
            if ((a > b) == up) {
              key[k] = b;
              key[l] = a;
            }
          }
        }
        __syncthreads();
      }
    }
    if (threadIdx.x == 0) {
      double acc = g.dT[c];
      if (d.v1)
        for (int k = 0; k < n; ++k) acc += d.v1[key[k]];
      if (d.v2)
        for (int k = 0; k < n; ++k) acc += d.v2[key[k]];
      g.dT[c] = acc;
    }
    __syncthreads();
  }
}

// cells beyond kMedMax items: stream all items in ascending order; thread 0
// adds the values of this cell one by one (serial order by construction)
__global__ __launch_bounds__(kBlock) void k_cell_sum_huge(GridView g, DepositView d, int n_items,
                                                          const int* huge, const int* hn) {
  __shared__ double val[kBlock];
  __shared__ int hit[kBlock];
  const int nh = *hn;
  for (int q = blockIdx.x; q < nh; q += gridDim.x) {
    const int c = huge[q];
    double acc = g.dT[c];  // meaningful in thread 0 only
    for (int pass = 0; pass < 2; ++pass) {
      const double* v = pass == 0 ? d.v1 : d.v2;
      if (!v) continue;
      for (int base = 0; base < n_items; base += kBlock) {
        const int i = base + threadIdx.x;
        const bool in = i < n_items && d.cell[i] == c;
        hit[threadIdx.x] = in;
        val[threadIdx.x] = in ? v[i] : 0.0;
        __syncthreads();
        if (threadIdx.x == 0)
          for (int k = 0; k < kBlock; ++k)
            if (hit[k]) acc += val[k];
        __syncthreads();
      }
    }
    if (threadIdx.x == 0) g.dT[c] = acc;
    __syncthreads();
  }
}

// E_local = sum of friction then random energies; Ee += E_local
__device__ __forceinline__ void energy_reduce(const double* partial, int nb, double* E_local,
                                              double* Ee, double* sh) {
  double sF = 0.0, sR = 0.0;
  for (int b = threadIdx.x; b < nb; b += 1024) {
    sF += partial[2 * b];
    sR += partial[2 * b + 1];
  }
  sF = block_sum(sF, sh);
  sR = block_sum(sR, sh);
  if (threadIdx.x == 0) {
    const double E = sF + sR;
    *E_local = E;
    *Ee += E;
  }
}

__global__ __launch_bounds__(1024) void k_energy_reduce(const double* partial, int nb,
                                                        double* E_local, double* Ee) {
  __shared__ double sh[1024];
  energy_reduce(partial, nb, E_local, Ee, sh);
}

// ----------------------------------------------- bucketed deposit
// The same deposit for grids with a few atoms per cell (fcc at one cell per
// unit cell holds 4 on average, at most 14 while every atom stays within
// half a cell of its site).  Each cell collects its atoms in a bucket of
// kBucket slots as they are counted; one thread per cell then sorts its
// bucket and adds the values in ascending atom order, friction then random,
// the serial order of fix_eph.cpp:364-390.  A cell that receives more atoms
// than its bucket holds is queued and summed by k_deposit_finish from an
// ordered scan of every atom's cell.  Three launches in all.
// one atom of the bucketed deposit at end-of-step velocity v: its energies
// into the bucket of its cell and the values the cell sums read
template <bool FRIC, bool RAND>
__device__ __forceinline__ void deposit_bucket_atom(const int i, const double4& v, const AtomsView& a,
                                                    const GridView& g, const StepParams& p,
                                                    const WorkView& w, const DepositView& d,
                                                    const BucketView& b, double* eF, double* eR) {
  const double4 x = a.pos[i];
  const int c = cell_of(g, x.x, x.y, x.z);
  const int fs = a.fstride;
  const double prescale = g.dV * p.dt;
  double v1 = 0.0, v2 = 0.0;
  if (FRIC) {
    double dE = 0.0;
    dE -= w.fE[i] * v.x * p.dt;
    dE -= w.fE[fs + i] * v.y * p.dt;
    dE -= w.fE[2 * fs + i] * v.z * p.dt;
    *eF = dE;
    v1 = dE / prescale;
  }
  if (RAND) {
    double dE = 0.0;
    dE -= w.fR[i] * v.x * p.dt;
    dE -= w.fR[fs + i] * v.y * p.dt;
    dE -= w.fR[2 * fs + i] * v.z * p.dt;
    *eR = dE;
    v2 = dE / prescale;
  }
  const int slot = atomicAdd(&b.count[c], 1);
  if (slot < kBucket) b.items[static_cast<size_t>(c) * kBucket + slot] = i;
  if (FRIC) d.v1[i] = v1;
  if (RAND) d.v2[i] = v2;
}

// the per-atom pass; with KICK it first applies the final velocity-Verlet
// kick (k_integrate<false>'s expressions), so that one pass serves
// final_integrate and the deposit of end_of_step when nothing runs between
template <bool FRIC, bool RAND, bool KICK>
__global__ __launch_bounds__(kBlock) void k_deposit_bucket(AtomsView a, GridView g, StepParams p,
                                                           WorkView w, DepositView d,
                                                           BucketView b) {
  __shared__ double sh[kBlock];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i == 0) *b.nqueue = 0;  // the queue of this step starts empty
  double eF = 0.0, eR = 0.0;
  if (i < a.nlocal && (a.mask[i] & a.groupbit)) {
    double4 v = a.vel[i];
    if (KICK) {
      const int fs = a.fstride;
      const double dtfm = p.dtf / p.mass[a.type[i]];
      v.x += dtfm * a.f[i];
      v.y += dtfm * a.f[fs + i];
      v.z += dtfm * a.f[2 * fs + i];
      a.vel[i] = v;
    }
    deposit_bucket_atom<FRIC, RAND>(i, v, a, g, p, w, d, b, &eF, &eR);
  }
  const double sF = block_sum(eF, sh);
  const double sR = block_sum(eR, sh);
  if (threadIdx.x == 0) {
    d.partial[2 * blockIdx.x] = sF;
    d.partial[2 * blockIdx.x + 1] = sR;
  }
}

// compare-exchange of a fixed-size register array (indices known after
// unrolling, so the array stays in registers)
static __device__ __forceinline__ void cswap(int* k, int a, int b) {
  const int x = k[a], y = k[b];
  k[a] = min(x, y);
  k[b] = max(x, y);
}

// one thread per cell: bitonic sort of the bucket (empty slots as INT_MAX),
// then the ordered sums into dT_e; the bucket count is left at zero
__global__ __launch_bounds__(kBlock) void k_deposit_bucket_cells(GridView g, DepositView d,
                                                                 BucketView b, int ncell) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= ncell) return;
  const int n = b.count[c];
  if (n == 0) return;
  b.count[c] = 0;
  if (n > kBucket) {
    b.queue[atomicAdd(b.nqueue, 1)] = c;
    return;
  }
  int k[kBucket];
  const int* src = b.items + static_cast<size_t>(c) * kBucket;
#pragma unroll
  for (int e = 0; e < kBucket; ++e) k[e] = e < n ? src[e] : 0x7fffffff;
#pragma unroll
  for (int size = 2; size <= kBucket; size <<= 1)
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1)
#pragma unroll
      for (int e = 0; e < kBucket; ++e) {
        const int l = e ^ stride;
        if (l > e) {
          if ((e & size) == 0) cswap(k, e, l);
          else cswap(k, l, e);
        }
      }
  double acc = g.dT[c];
  if (d.v1) {
#pragma unroll
    for (int e = 0; e < kBucket; ++e)
      if (e < n) acc += d.v1[k[e]];
  }
  if (d.v2) {
#pragma unroll
    for (int e = 0; e < kBucket; ++e)
      if (e < n) acc += d.v2[k[e]];
  }
  g.dT[c] = acc;
}

// block 0: E_local and Ee from the block partial sums (as k_energy_reduce);
// blocks 1..: the queued cells, each by an ordered scan over all atoms:
// every wave ballots the atoms of the cell among its 64, and thread 0 adds
// their values in ascending order
__global__ __launch_bounds__(1024) void k_deposit_finish(AtomsView a, GridView g, DepositView d,
                                                         BucketView b, int nblocks, double* E_local,
                                                         double* Ee) {
  const int n_items = a.nlocal;
  __shared__ double sh[1024];
  __shared__ unsigned long long hits[16];
  if (blockIdx.x == 0) {
    energy_reduce(d.partial, nblocks, E_local, Ee, sh);
    return;
  }
  const int nq = *b.nqueue;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int q = blockIdx.x - 1; q < nq; q += gridDim.x - 1) {
    const int c = b.queue[q];
    double acc = g.dT[c];  // meaningful in thread 0 only
    for (int pass = 0; pass < 2; ++pass) {
      const double* v = pass == 0 ? d.v1 : d.v2;
      if (!v) continue;
      for (int base = 0; base < n_items; base += 1024) {
        const int i = base + threadIdx.x;
        bool in = false;
        if (i < n_items && (a.mask[i] & a.groupbit)) {
          const double4 x = a.pos[i];
          in = cell_of(g, x.x, x.y, x.z) == c;
        }
        const unsigned long long m = __ballot(in);
        if (lane == 0) hits[wv] = m;
        __syncthreads();
        if (threadIdx.x == 0)
          for (int w2 = 0; w2 < 16; ++w2)
            for (unsigned long long bits = hits[w2]; bits; bits &= bits - 1)
              acc += v[base + 64 * w2 + __builtin_ctzll(bits)];
        __syncthreads();
      }
    }
    if (threadIdx.x == 0) g.dT[c] = acc;
  }
}


// ------------------------------------------------------------- grid
// C_e(T), kappa(T) of temperature-dependent cells (eph_fdm.h:280-287)
__global__ __launch_bounds__(kBlock) void k_grid_tdyn_update(GridView g, int n) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= n || !g.tdyn[c]) return;
  const double T = g.T[c];
  g.C[c] = cubic_at(g.C_T, g.n_CT, g.inv_dT_C, T);
  g.kappa[c] = cubic_at(g.kappa_T, g.n_kT, g.inv_dT_k, T);
}

// smallest C_e and rho_e, largest kappa over non-constant cells, seeded with
// cell 0 (eph_fdm.h:289-300); min/max are exact, so the order is free
__global__ __launch_bounds__(1024) void k_grid_minmax(GridView g, int n, double* out3) {
  __shared__ double s0[1024], s1[1024], s2[1024];
  double cmin = g.C[0], rmin = g.rho[0], kmax = g.kappa[0];
  for (int c = 1 + threadIdx.x; c < n; c += 1024) {
    if (g.flag[c] != 0) {
      cmin = g.C[c] < cmin ? g.C[c] : cmin;
      rmin = g.rho[c] < rmin ? g.rho[c] : rmin;
      kmax = g.kappa[c] > kmax ? g.kappa[c] : kmax;
    }
  }
  s0[threadIdx.x] = cmin; s1[threadIdx.x] = rmin; s2[threadIdx.x] = kmax;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      s0[threadIdx.x] = s0[threadIdx.x + s] < s0[threadIdx.x] ? s0[threadIdx.x + s] : s0[threadIdx.x];
      s1[threadIdx.x] = s1[threadIdx.x + s] < s1[threadIdx.x] ? s1[threadIdx.x + s] : s1[threadIdx.x];
      s2[threadIdx.x] = s2[threadIdx.x + s] > s2[threadIdx.x] ? s2[threadIdx.x + s] : s2[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { out3[0] = s0[0]; out3[1] = s1[0]; out3[2] = s2[0]; }
}

// one dimension's two terms of the 7-point variable-kappa stencil
// (eph_fdm.h:337-339, 350-352, 363-365), in the reference's order; p and q
// are the cell's lower and upper neighbours after the wall rule
static __device__ __forceinline__ double fdm_dim(double ddT, const double kr, const double Tr,
                                                 const double kp, const double kq, const double Tp,
                                                 const double Tq, const double d) {
  ddT += (kq - kp) * (Tq - Tp) / d / d / 4.0;
  ddT += kr * ((Tq + Tp - 2.0 * Tr) / d / d);
  return ddT;
}

// the update of a cell from its stencil sum, then the clamp (eph_fdm.h:376-395)
template <bool TDYN>
static __device__ __forceinline__ double fdm_update(const GridView& g, const int r, const int fr,
                                                    double T, const double ddT,
                                                    const double inner_dt) {
  if (fr == kCellDynamic) {
    if (TDYN && g.tdyn[r] == 1) {
      double E = linear_fwd(g, T);
      E += (ddT + g.dT[r] + g.S[r]) / g.rho[r] * inner_dt;
      T = linear_rev(g, E);
    } else {
      const double prescaler = g.rho[r] * g.C[r];
      T += (ddT + g.dT[r] + g.S[r]) / prescaler * inner_dt;
    }
  }
  if (T < 0.0) T = 0.0;
  return T;
}

// one explicit sub-step of EPH_FDM::solve (eph_fdm.h:315-395): the 7-point
// variable-kappa stencil on Tin, the update and the clamp, into Tout
template <bool TDYN>
__device__ __forceinline__ void fdm_cell(const GridView& g, const double* __restrict__ Tin,
                                         double* __restrict__ Tout, const int r,
                                         const double inner_dt) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int i = r % nx, j = (r / nx) % ny, k = r / (nx * ny);
  const int fr = g.flag[r];
  double ddT = 0.0;
  if (fr != kCellZeroDerivative) {
    const double Tr = Tin[r], kr = g.kappa[r];
    int p, q;
    // x
    p = i > 0 ? r - 1 : r + (nx - 1);
    q = i < nx - 1 ? r + 1 : r - (nx - 1);
    if (g.flag[q] == kCellZeroDerivative) q = r;
    else if (g.flag[p] == kCellZeroDerivative) p = r;
    ddT = fdm_dim(ddT, kr, Tr, g.kappa[p], g.kappa[q], Tin[p], Tin[q], g.dx);
    // y
    p = j > 0 ? r - nx : r + (ny - 1) * nx;
    q = j < ny - 1 ? r + nx : r - (ny - 1) * nx;
    if (g.flag[q] == kCellZeroDerivative) q = r;
    else if (g.flag[p] == kCellZeroDerivative) p = r;
    ddT = fdm_dim(ddT, kr, Tr, g.kappa[p], g.kappa[q], Tin[p], Tin[q], g.dy);
    // z
    const int pl = nx * ny;
    p = k > 0 ? r - pl : r + (nz - 1) * pl;
    q = k < nz - 1 ? r + pl : r - (nz - 1) * pl;
    if (g.flag[q] == kCellZeroDerivative) q = r;
    else if (g.flag[p] == kCellZeroDerivative) p = r;
    ddT = fdm_dim(ddT, kr, Tr, g.kappa[p], g.kappa[q], Tin[p], Tin[q], g.dz);
  }
  Tout[r] = fdm_update<TDYN>(g, r, fr, Tin[r], ddT, inner_dt);
}

template <bool TDYN>
__global__ __launch_bounds__(kBlock) void k_fdm_substep(GridView g, const double* __restrict__ Tin,
                                                        double* __restrict__ Tout, double inner_dt) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= g.nx * g.ny * g.nz) return;
  fdm_cell<TDYN>(g, Tin, Tout, r, inner_dt);
}

__global__ __launch_bounds__(kBlock) void k_grid_lookup(GridView g, int n, const double* x,
                                                        const double* y, const double* z, int* idx,
                                                        double* T) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int c = cell_of(g, x[i], y[i], z[i]);
  if (idx) idx[i] = c;
  if (T) T[i] = g.T[c];
}

// deterministic two-level sum (block partials in fixed order)
__global__ __launch_bounds__(kBlock) void k_sum_partial(const double* v, int n, double* partial) {
  __shared__ double sh[kBlock];
  double s = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) s += v[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void k_sum_final(const double* partial, int nb, double* out) {
  __shared__ double sh[kBlock];
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += kBlock) s += partial[b];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) *out = s;
}

__global__ __launch_bounds__(kBlock) void k_nonfinite(const double* v, int n, int* flag) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n && !isfinite(v[i])) atomicOr(flag, 2);
}

// per-atom array of fix_eph.cpp:406-428
__global__ __launch_bounds__(kBlock) void k_peratom(AtomsView a, TablesView t, WorkView w,
                                                    double* out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.nlocal) return;
  double* o = out + 8 * static_cast<long>(i);
  if (a.mask[i] & a.groupbit) {
    const double rho = w.rec[i].rw.w;
    const int e = static_cast<int>(a.pos[i].w);
    o[0] = rho;
    o[1] = rho > t.rho_cut ? 0.0 : cubic_at(t.beta + e * t.n_beta, t.n_beta, t.inv_drho, rho);
    const int fs = a.fstride;
    o[2] = w.fE ? w.fE[i] : 0.0;
    o[3] = w.fE ? w.fE[fs + i] : 0.0;
    o[4] = w.fE ? w.fE[2 * fs + i] : 0.0;
    o[5] = w.fR ? w.fR[i] : 0.0;
    o[6] = w.fR ? w.fR[fs + i] : 0.0;
    o[7] = w.fR ? w.fR[2 * fs + i] : 0.0;
  } else {
    for (int k = 0; k < 8; ++k) o[k] = 0.0;
  }
}

// ---------------------------------------------------------- harness
// pair lj/cut, full list, Newton off: f_i = sum_j fpair (x_i - x_j), summed
// in list order.  The neighbours' positions come through LDS as in the
// density sweep's rho loop (coop_fetch_pos, kLjU entries per iteration, the
// next entries' copies in flight while these are summed); the whole wave runs
// to its longest list.
constexpr int kLjU = 4;
__global__ __launch_bounds__(kBlock) void k_lj(AtomsView a, const int* __restrict__ nbr,
                                               const int* __restrict__ nn, int stride, double lj1,
                                               double lj2, double rc2) {
  __shared__ double2 s_lj[(kBlock / 64) * kLjU * 128];
  const int base = xcd_block() * kBlock;
  if (base + static_cast<int>(threadIdx.x & ~63u) >= a.nlocal) return;  // whole wave past the end
  const int i_lane = base + threadIdx.x;
  const bool valid = i_lane < a.nlocal;
  const int i = valid ? i_lane : 0;
  const int lane = threadIdx.x & 63;
  double2* stage = s_lj + (threadIdx.x >> 6) * kLjU * 128;
  const double2* __restrict__ pos2 = reinterpret_cast<const double2*>(a.pos);
  const double4 pi = a.pos[i];
  double fx = 0.0, fy = 0.0, fz = 0.0;
  const int n = valid ? nn[i] : 0;
  const int nmax = wave_max(n);
  constexpr int U = kLjU;
  int jj[U], nx[U];
#pragma unroll
  for (int e = 0; e < U; ++e) jj[e] = e < n ? nbr[e * stride + i] : 0;
#pragma unroll
  for (int e = 0; e < U; ++e) nx[e] = U + e < n ? nbr[(U + e) * stride + i] : 0;
  if (nmax > 0) {
#pragma unroll
    for (int e = 0; e < U; ++e) coop_fetch_pos(pos2, jj[e], lane, stage + 128 * e);
  }
  for (int k = 0; k < nmax; k += U) {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    wave_lds_sync();
    double4 pq[U];
#pragma unroll
    for (int e = 0; e < U; ++e) pq[e] = coop_read_pos(stage + 128 * e, lane);
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    wave_lds_sync();
    if (k + U < nmax) {
#pragma unroll
      for (int e = 0; e < U; ++e) coop_fetch_pos(pos2, nx[e], lane, stage + 128 * e);
#pragma unroll
      for (int e = 0; e < U; ++e) nx[e] = k + 2 * U + e < n ? nbr[(k + 2 * U + e) * stride + i] : 0;
    }
#pragma unroll
    for (int e = 0; e < U; ++e) {
      if (k + e < n) {
        const double dx = pi.x - pq[e].x, dy = pi.y - pq[e].y, dz = pi.z - pq[e].z;
        const double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 < rc2) {
          const double r2inv = 1.0 / r2;
          const double r6inv = r2inv * r2inv * r2inv;
          const double fpair = r6inv * (lj1 * r6inv - lj2) * r2inv;
          fx += dx * fpair;
          fy += dy * fpair;
          fz += dz * fpair;
        }
      }
    }
  }
  if (!valid) return;
  const int fs = a.fstride;
  a.f[i] = fx;
  a.f[fs + i] = fy;
  a.f[2 * fs + i] = fz;
}

__global__ __launch_bounds__(kBlock) void k_max_disp2(AtomsView a, const double4* ref,
                                                      double* partial) {
  __shared__ double sh[kBlock];
  double m = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.nlocal; i += gridDim.x * kBlock) {
    const double4 p = a.pos[i], q = ref[i];
    const double dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
    const double d2 = dx * dx + dy * dy + dz * dz;
    m = d2 > m ? d2 : m;
  }
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] = sh[threadIdx.x + s] > sh[threadIdx.x] ? sh[threadIdx.x + s] : sh[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// ============================================================ batched replicas
// Several independent systems share one launch per phase.  A chunk is up to
// kBlock consecutive atoms of one replica and block b handles chunk b, so the
// replica's noise seed and T_e grid are uniform in a block and its energy
// partials line up with the blocks of a single-system launch: every replica
// gets the same bits it would get alone.
// the step number is read from device memory so a recorded step replays
template <int XI, bool FRIC, bool MULTI>
__global__ __launch_bounds__(kBlock) void k_density_w_batch(AtomsView a, TablesView t, StepParams p,
                                                            WorkView w,
                                                            const Chunk* __restrict__ chunks,
                                                            const uint32_t* __restrict__ seeds,
                                                            const uint64_t* __restrict__ step) {
  // the staging area, sized at launch (a larger request holds fewer blocks per CU)
  extern __shared__ double2 s_bstage[];
  double2* s_stage = s_bstage;
  const Chunk c = chunks[blockIdx.x];
  if (static_cast<int>(threadIdx.x & ~63u) >= c.count) return;  // whole wave past the chunk
  StepParams pr = p;
  pr.seed = seeds[c.rep];
  pr.step = *step;
  density_w_atom_coop<XI, FRIC, MULTI>(c.start + threadIdx.x, static_cast<int>(threadIdx.x) < c.count,
                                       a, t, t.rho_r2, pr, w, s_stage + (threadIdx.x >> 6) * kCoopStageW,
                                       c.rep);
}

// start of a recorded step: step += 1
__global__ void k_step_advance(uint64_t* step) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1;
}

// log[step - first] = *flag (which steps rebuilt)
__global__ void k_step_log(const int* __restrict__ flag, const uint64_t* __restrict__ step,
                           const uint64_t* __restrict__ first, int* log, int nlog) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const long k = static_cast<long>(*step - *first);
  if (k >= 0 && k < nlog) log[k] = *flag;
}

template <bool FRIC, bool RAND, bool MULTI>
__global__ __launch_bounds__(kBlock) void k_prl_forces_batch(AtomsView a, TablesView t,
                                                             const GridView* __restrict__ grids,
                                                             StepParams p, WorkView w,
                                                             const Chunk* __restrict__ chunks) {
  extern __shared__ double2 s_bstage[];
  double2* s_stage = s_bstage;
  const Chunk c = chunks[blockIdx.x];
  if (static_cast<int>(threadIdx.x & ~63u) >= c.count) return;  // whole wave past the chunk
  prl_forces_atom_coop<FRIC, RAND, MULTI>(c.start + threadIdx.x, static_cast<int>(threadIdx.x) < c.count,
                                          a, t, t.rho_r2, grids[c.rep], p, w,
                                          s_stage + (threadIdx.x >> 6) * kCoopStage * kCoopF);
}

// Persistent LDS-table forms: each block of BS threads (1024, 512 or 256)
// stages the rho(r^2) tables once and walks tiles of BS / 256 consecutive
// chunks (one chunk per 256 threads); the tiles of XCD x form one contiguous
// range.  Same per-atom arithmetic as the chunk-per-block forms.
template <int XI, bool FRIC, bool MULTI, int BS>
__global__ __launch_bounds__(BS) void k_density_w_batch_lds(AtomsView a, TablesView t,
                                                                   StepParams p, WorkView w,
                                                                   const Chunk* __restrict__ chunks,
                                                                   int nchunks,
                                                                   const uint32_t* __restrict__ seeds,
                                                                   const uint64_t* __restrict__ step) {
  constexpr int kChunksPerTile = BS / kBlock;
  extern __shared__ double4 s_tab[];
  const int ntab = t.n_elem * t.n_rho;
  for (int k = threadIdx.x; k < ntab; k += BS) s_tab[k] = t.rho_r2[k];
  __syncthreads();
  const int ntiles = (nchunks + kChunksPerTile - 1) / kChunksPerTile;
  int tile, stp;
  xcd_tiles(ntiles, &tile, &stp);
  const int end = xcd_tiles_end(ntiles);
  StepParams pr = p;
  pr.step = *step;
  const int lane = static_cast<int>(threadIdx.x) % kBlock;
  for (; tile < end; tile += stp) {
    const int ci = tile * kChunksPerTile + static_cast<int>(threadIdx.x) / kBlock;
    if (ci >= nchunks) continue;
    const Chunk c = chunks[ci];
    if (lane >= c.count) continue;
    pr.seed = seeds[c.rep];
    density_w_atom<XI, FRIC, MULTI>(c.start + lane, a, t, s_tab, pr, w);
  }
}

template <bool FRIC, bool RAND, bool MULTI, int BS>
__global__ __launch_bounds__(BS) void k_prl_forces_batch_lds(AtomsView a, TablesView t,
                                                                    const GridView* __restrict__ grids,
                                                                    StepParams p, WorkView w,
                                                                    const Chunk* __restrict__ chunks,
                                                                    int nchunks) {
  constexpr int kChunksPerTile = BS / kBlock;
  extern __shared__ double4 s_tab[];
  const int ntab = t.n_elem * t.n_rho;
  for (int k = threadIdx.x; k < ntab; k += BS) s_tab[k] = t.rho_r2[k];
  __syncthreads();
  const int ntiles = (nchunks + kChunksPerTile - 1) / kChunksPerTile;
  int tile, stp;
  xcd_tiles(ntiles, &tile, &stp);
  const int end = xcd_tiles_end(ntiles);
  const int lane = static_cast<int>(threadIdx.x) % kBlock;
  for (; tile < end; tile += stp) {
    const int ci = tile * kChunksPerTile + static_cast<int>(threadIdx.x) / kBlock;
    if (ci >= nchunks) continue;
    const Chunk c = chunks[ci];
    if (lane >= c.count) continue;
    prl_forces_atom<FRIC, RAND, MULTI, true>(c.start + lane, a, t, s_tab, grids[c.rep], p, w);
  }
}

template <bool FRIC, bool RAND>
__global__ __launch_bounds__(kBlock) void k_deposit_atoms_batch(AtomsView a,
                                                                const GridView* __restrict__ grids,
                                                                const int* __restrict__ cell_off,
                                                                StepParams p, WorkView w,
                                                                DepositView d,
                                                                const Chunk* __restrict__ chunks) {
  __shared__ double sh[kBlock];
  const Chunk c = chunks[blockIdx.x];
  double eF = 0.0, eR = 0.0;
  if (static_cast<int>(threadIdx.x) < c.count)
    deposit_atom<FRIC, RAND>(c.start + threadIdx.x, a, grids[c.rep], p, w, d, cell_off[c.rep], &eF,
                             &eR);
  const double sF = block_sum(eF, sh);
  const double sR = block_sum(eR, sh);
  if (threadIdx.x == 0) {
    d.partial[2 * blockIdx.x] = sF;
    d.partial[2 * blockIdx.x + 1] = sR;
  }
}

__global__ __launch_bounds__(1024) void k_energy_reduce_batch(const double* partial,
                                                              const int* __restrict__ chunk0,
                                                              const int* __restrict__ nchunks,
                                                              double* E_local, double* Ee) {
  __shared__ double sh[1024];
  const int r = blockIdx.x;
  energy_reduce(partial + 2 * chunk0[r], nchunks[r], E_local + r, Ee + r, sh);
}

// one sub-step over all replicas' cells; a replica whose sub-step count is
// below n copies its cells, so the two buffers stay in step for everyone
template <bool TDYN>
__global__ __launch_bounds__(kBlock) void k_fdm_substep_batch(const GridView* __restrict__ grids,
                                                              const int* __restrict__ cell_rep,
                                                              const int* __restrict__ cell_off,
                                                              const double* __restrict__ Tin,
                                                              double* __restrict__ Tout,
                                                              const double* __restrict__ inner_dt,
                                                              const int* __restrict__ steps, int n,
                                                              int G) {
  const int gc = blockIdx.x * kBlock + threadIdx.x;
  if (gc >= G) return;
  const int r = cell_rep[gc];
  const int off = cell_off[r];
  if (n >= steps[r]) {
    Tout[gc] = Tin[gc];
    return;
  }
  fdm_cell<TDYN>(grids[r], Tin + off, Tout + off, gc - off, inner_dt[r]);
}

// C_e(T), kappa(T) of the temperature-dependent cells of every replica
// (eph_fdm.h:280-287 per replica grid, with that grid's tables)
__global__ __launch_bounds__(kBlock) void k_grid_tdyn_update_batch(const GridView* __restrict__ grids,
                                                                   const int* __restrict__ cell_rep,
                                                                   const int* __restrict__ cell_off,
                                                                   int G) {
  const int gc = blockIdx.x * kBlock + threadIdx.x;
  if (gc >= G) return;
  const int r = cell_rep[gc];
  const GridView& g = grids[r];
  const int c = gc - cell_off[r];
  if (!g.tdyn[c]) return;
  const double T = g.T[c];
  g.C[c] = cubic_at(g.C_T, g.n_CT, g.inv_dT_C, T);
  g.kappa[c] = cubic_at(g.kappa_T, g.n_kT, g.inv_dT_k, T);
}

// Sub-step count and inner dt of each replica with temperature-dependent
// cells, from its stability extrema (smallest C_e and rho_e, largest kappa
// over non-constant cells seeded with cell 0, eph_fdm.h:289-314), in the
// arithmetic of the single-grid solve; the other replicas keep the counts set
// up for them.  Block r handles replica r; *max_sub (cleared before) receives
// the largest count of the step.
__global__ __launch_bounds__(256) void k_fdm_substeps_batch(const GridView* __restrict__ grids,
                                                            const int8_t* __restrict__ rep_tdyn,
                                                            const double* __restrict__ grid_dt,
                                                            const int* __restrict__ grid_steps,
                                                            int* __restrict__ sub_n,
                                                            double* __restrict__ inner_dt,
                                                            int* __restrict__ max_sub) {
  __shared__ double s0[256], s1[256], s2[256];
  const int r = blockIdx.x;
  if (rep_tdyn[r]) {
    const GridView& g = grids[r];
    const int n = g.nx * g.ny * g.nz;
    double cmin = g.C[0], rmin = g.rho[0], kmax = g.kappa[0];
    for (int c = 1 + threadIdx.x; c < n; c += 256) {
      if (g.flag[c] != 0) {
        cmin = g.C[c] < cmin ? g.C[c] : cmin;
        rmin = g.rho[c] < rmin ? g.rho[c] : rmin;
        kmax = g.kappa[c] > kmax ? g.kappa[c] : kmax;
      }
    }
    s0[threadIdx.x] = cmin; s1[threadIdx.x] = rmin; s2[threadIdx.x] = kmax;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if (threadIdx.x < h) {
        s0[threadIdx.x] = s0[threadIdx.x + h] < s0[threadIdx.x] ? s0[threadIdx.x + h] : s0[threadIdx.x];
        s1[threadIdx.x] = s1[threadIdx.x + h] < s1[threadIdx.x] ? s1[threadIdx.x + h] : s1[threadIdx.x];
        s2[threadIdx.x] = s2[threadIdx.x + h] > s2[threadIdx.x] ? s2[threadIdx.x + h] : s2[threadIdx.x];
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      const double dt = grid_dt[r];
      double idt = dt / static_cast<double>(grid_steps[r]);
      const double dtdxdydz = idt * (1.0 / g.dx / g.dx + 1.0 / g.dy / g.dy + 1.0 / g.dz / g.dz);
      const double q = dtdxdydz / s0[0] / s1[0] * s2[0];
      unsigned ns = static_cast<unsigned>(grid_steps[r]);
      if (q > 0.4) {
        idt = 0.4 * idt / q;
        const unsigned k = static_cast<unsigned>(dt / idt);
        ns = k > 1u ? k : 1u;
        idt = dt / ns;
      }
      sub_n[r] = static_cast<int>(ns);
      inner_dt[r] = idt;
    }
  }
  if (threadIdx.x == 0) atomicMax(max_sub, sub_n[r]);
}

__global__ __launch_bounds__(kBlock) void k_copy_double(const double* __restrict__ src,
                                                        double* __restrict__ dst, int n) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// ghost updates with the ghost count read on the device (graph-capturable)
__global__ __launch_bounds__(kBlock) void k_ghost_xv_dev(AtomsView a, const int* __restrict__ owner,
                                                         const double4* __restrict__ shift,
                                                         const int* __restrict__ ng) {
  const int g = blockIdx.x * kBlock + threadIdx.x;
  if (g >= *ng) return;
  const int o = owner[g];
  const double4 p = a.pos[o];
  const double4 s = shift[g];
  a.pos[a.nlocal + g] = make_double4(p.x + s.x, p.y + s.y, p.z + s.z, p.w);
  a.vel[a.nlocal + g] = a.vel[o];
}

__global__ __launch_bounds__(kBlock) void k_ghost_rec_dev(AtomsView a, WorkView w,
                                                          const int* __restrict__ ng) {
  const int g = blockIdx.x * kBlock + threadIdx.x;
  if (g < *ng) ghost_rec(g, a, w);
}

__global__ __launch_bounds__(kBlock) void k_pack_pv_dev(AtomsView a, WorkView w,
                                                        const int* __restrict__ ng) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.nlocal + *ng) return;
  PosVel q;
  q.p = a.pos[i];
  q.v = w.vw[i];
  w.pv[i] = q;
}

// ---------------------------------------------------- host layout staging
// LAMMPS [n][3] rows <-> device double4 (keeping .w) and [3][stride] arrays
__global__ __launch_bounds__(kBlock) void k_rows_into_d4(const double* __restrict__ rows, int n,
                                                         double4* __restrict__ out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double4 o = out[i];
  o.x = rows[3 * i]; o.y = rows[3 * i + 1]; o.z = rows[3 * i + 2];
  out[i] = o;
}

// rows -> (x, y, z, w) with w = element of the atom's type (species map
// given, types already checked to lie in 1..ntypes) or 0
__global__ __launch_bounds__(kBlock) void k_rows_into_d4_w(const double* __restrict__ rows, int n,
                                                           const int* __restrict__ type,
                                                           const int* __restrict__ species,
                                                           double4* __restrict__ out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double w = species ? static_cast<double>(species[type[i] - 1]) : 0.0;
  out[i] = make_double4(rows[3 * i], rows[3 * i + 1], rows[3 * i + 2], w);
}

__global__ __launch_bounds__(kBlock) void k_d4_into_rows(const double4* __restrict__ in, int n,
                                                         int stride4, double* __restrict__ rows) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double4 o = in[static_cast<long>(i) * stride4];
  rows[3 * i] = o.x; rows[3 * i + 1] = o.y; rows[3 * i + 2] = o.z;
}

__global__ __launch_bounds__(kBlock) void k_rows_into_soa(const double* __restrict__ rows, int n,
                                                          double* __restrict__ soa, int stride) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  soa[i] = rows[3 * i]; soa[stride + i] = rows[3 * i + 1]; soa[2 * stride + i] = rows[3 * i + 2];
}

__global__ __launch_bounds__(kBlock) void k_soa_into_rows(const double* __restrict__ soa, int n,
                                                          int stride, double* __restrict__ rows) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  rows[3 * i] = soa[i]; rows[3 * i + 1] = soa[stride + i]; rows[3 * i + 2] = soa[2 * stride + i];
}

__global__ __launch_bounds__(kBlock) void k_d4_component(const double4* __restrict__ in, int n,
                                                         int stride4, int comp,
                                                         double* __restrict__ out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double4 o = in[static_cast<long>(i) * stride4];
  out[i] = comp == 0 ? o.x : comp == 1 ? o.y : comp == 2 ? o.z : o.w;
}

// ================================================================ wrappers
static inline int nblk(long n, int b = kBlock) { return static_cast<int>((n + b - 1) / b); }

hipError_t launch_ghost_xv(const AtomsView& a, const int* owner, const double4* shift, hipStream_t s) {
  if (a.nghost == 0) return hipSuccess;
  k_ghost_xv<<<nblk(a.nghost), kBlock, 0, s>>>(a, owner, shift);
  return hipGetLastError();
}

hipError_t launch_csr_to_ell(int nlocal, const int* off, const int* idx, int neighmask, int nall,
                             int* ell, int* nn, int stride, int width, int* bad, hipStream_t s) {
  if (nlocal == 0) return hipSuccess;
  k_csr_to_ell<<<nblk(nlocal), kBlock, 0, s>>>(nlocal, off, idx, neighmask, nall, ell, nn, stride,
                                               width, bad);
  return hipGetLastError();
}

// Grid of a persistent LDS-table launch: all blocks that fit at once, a
// multiple of the 8 XCDs, no more than the tiles need.
template <typename K>
static int lds_grid(K kernel, size_t lds_bytes, int ntiles, int block = kBigBlock) {
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                            hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds_bytes));
  int occ = 0, dev = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, block, lds_bytes) != hipSuccess ||
      occ < 1)
    occ = 1;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 8)
    cus = 8;
  int g = occ * cus;
  const int need = ((ntiles + 7) / 8) * 8;
  g = g < need ? g : need;
  g = (g / 8) * 8;
  return g < 8 ? 8 : g;
}

// The tables go to LDS only when the device reports room for them in one
// workgroup (gfx950: 160 KB); otherwise the sweeps read them from global memory.
static bool lds_tables_fit(const TablesView& t) {
  const size_t bytes = static_cast<size_t>(t.n_elem) * t.n_rho * sizeof(double4);
  if (bytes > kLdsTableMax) return false;
  int dev = 0, lds_max = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    return false;
  return bytes <= static_cast<size_t>(lds_max);
}

// the default choice when no form is set: LDS tables for large systems
static bool use_lds_tables(const AtomsView& a, const TablesView& t) {
  return a.nlocal >= kLdsMinAtoms && lds_tables_fit(t);
}

bool lds_tables_possible(const TablesView& t) { return lds_tables_fit(t); }

// The LDS-table form of a sweep (kLdsForm1024 .. kLdsForm256), or 0 for the
// form that reads the tables through the caches.  -1 picks by system size,
// with `dflt` for large systems.
static int resolve_form(const AtomsView& a, const TablesView& t, int form, int dflt) {
  if (form == 0 || form >= kChunkOcc4) return form;  // no tables in LDS
  if (!lds_tables_fit(t)) return 0;
  if (form < 0) return use_lds_tables(a, t) ? dflt : 0;
  return form;
}

// LDS request of a cache-table launch: the staging its waves use, or more
// so that at most 4, 3 or 2 blocks share a CU (160 KB of LDS per CU)
static size_t chunk_lds(int form, size_t need) {
  const size_t want = form == kChunkOcc4 ? 38912 : form == kChunkOcc3 ? 49152
                    : form == kChunkOcc2 ? 65536 : 0;
  return want > need ? want : need;
}
constexpr size_t kDensityStageBytes = (kBlock / 64) * kCoopStageW * sizeof(double2);
constexpr size_t kForceStageBytes = (kBlock / 64) * kCoopStage * kCoopF * sizeof(double2);

template <int XI, bool F, bool M>
static void density_lds(int blk, size_t lds, const AtomsView& a, const TablesView& t,
                        const StepParams& p, const WorkView& w, hipStream_t s) {
#define EPH_DWL(B)                                                                               \
  k_density_w_lds<XI, F, M, B><<<lds_grid(k_density_w_lds<XI, F, M, B>, lds,                    \
                                          (a.nlocal + B - 1) / B, B), B, lds, s>>>(a, t, p, w)
  if (blk == 256) EPH_DWL(256);
  else if (blk == 512) EPH_DWL(512);
  else EPH_DWL(1024);
#undef EPH_DWL
}

static int lds_block(int form) { return form == kLdsForm256 ? 256 : form == kLdsForm512 ? 512 : 1024; }

hipError_t launch_density_w(const AtomsView& a, const TablesView& t, const StepParams& p,
                            const WorkView& w, hipStream_t s, int form, hipEvent_t* ev) {
  if (a.nlocal == 0) return hipSuccess;
  const bool fric = p.flags & 0x01;
  const bool multi = t.n_elem > 1;
  const int f = resolve_form(a, t, form, 0);
  if (f >= kLdsForm1024 && f <= kLdsForm256) {
    const size_t lds = static_cast<size_t>(t.n_elem) * t.n_rho * sizeof(double4);
    const int blk = lds_block(f);
    if (ev) (void)hipEventRecord(ev[0], s);
#define EPH_DWL_M(XI, F) if (multi) density_lds<XI, F, true>(blk, lds, a, t, p, w, s); \
                         else density_lds<XI, F, false>(blk, lds, a, t, p, w, s)
#define EPH_DWL_F(XI) if (fric) { EPH_DWL_M(XI, true); } else { EPH_DWL_M(XI, false); }
    switch (p.xi_mode) {
      case 1: EPH_DWL_F(1); break;
      case 2: EPH_DWL_F(2); break;
      default: EPH_DWL_F(0); break;
    }
#undef EPH_DWL_F
#undef EPH_DWL_M
    if (ev) (void)hipEventRecord(ev[1], s);
    return hipGetLastError();
  }
  const int g = nblk(a.nlocal);
  const bool keep = w.rcand2 > 0.0 && w.cstate;
  WorkView wc = w;
  if (!keep) wc.rcand2 = 0.0;
  if (keep) k_cand_check<<<nblk(a.nlocal + a.nghost), kBlock, 0, s>>>(a.nlocal + a.nghost, a.pos, wc);
  if (ev) (void)hipEventRecord(ev[0], s);  // the sweep alone, without the candidate check
  const size_t dl = chunk_lds(f, kDensityStageBytes);
#define EPH_DW(XI, F, M) k_density_w<XI, F, M><<<g, kBlock, dl, s>>>(a, t, p, wc)
#define EPH_DW_M(XI, F) if (multi) EPH_DW(XI, F, true); else EPH_DW(XI, F, false)
#define EPH_DW_F(XI) if (fric) { EPH_DW_M(XI, true); } else { EPH_DW_M(XI, false); }
  switch (p.xi_mode) {
    case 1: EPH_DW_F(1); break;
    case 2: EPH_DW_F(2); break;
    default: EPH_DW_F(0); break;
  }
#undef EPH_DW_F
#undef EPH_DW_M
#undef EPH_DW
  if (ev) (void)hipEventRecord(ev[1], s);
  if (keep) k_cand_flags<<<1, 64, 0, s>>>(w.cstate, w.moved_cap);
  return hipGetLastError();
}

hipError_t launch_pack_pv(const AtomsView& a, const WorkView& w, hipStream_t s) {
  const int n = a.nlocal + a.nghost;
  if (n == 0) return hipSuccess;
  k_pack_pv<<<nblk(n), kBlock, 0, s>>>(a, w, n);
  return hipGetLastError();
}

hipError_t launch_halo_pack(const ForceRec* rec, int n, double* out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_halo_pack<<<nblk(n), kBlock, 0, s>>>(rec, n, out);
  return hipGetLastError();
}

hipError_t launch_ghost_halo(const AtomsView& a, const TablesView& t, const WorkView& w,
                             const double* in, hipStream_t s) {
  if (a.nghost == 0) return hipSuccess;
  if (t.n_elem > 1) k_ghost_halo<true><<<nblk(a.nghost), kBlock, 0, s>>>(a, t, w, in);
  else k_ghost_halo<false><<<nblk(a.nghost), kBlock, 0, s>>>(a, t, w, in);
  return hipGetLastError();
}

hipError_t launch_ghost_rec(const AtomsView& a, const WorkView& w, hipStream_t s) {
  if (a.nghost == 0) return hipSuccess;
  k_ghost_rec<<<nblk(a.nghost), kBlock, 0, s>>>(a, w);
  return hipGetLastError();
}

template <bool F, bool R, bool M>
static void forces_lds(int blk, size_t lds, const AtomsView& a, const TablesView& t,
                       const GridView& g, const StepParams& p, const WorkView& w, hipStream_t s) {
#define EPH_PFL(B)                                                                              \
  k_prl_forces_lds<F, R, M, B><<<lds_grid(k_prl_forces_lds<F, R, M, B>, lds,                   \
                                          (a.nlocal + B - 1) / B, B), B, lds, s>>>(a, t, g, p, w)
  if (blk == 256) EPH_PFL(256);
  else if (blk == 512) EPH_PFL(512);
  else EPH_PFL(1024);
#undef EPH_PFL
}

hipError_t launch_prl_forces(const AtomsView& a, const TablesView& t, const GridView& g,
                             const StepParams& p, const WorkView& w, hipStream_t s, int form) {
  if (a.nlocal == 0) return hipSuccess;
  const bool fric = p.flags & 0x01, rnd = p.flags & 0x02, multi = t.n_elem > 1;
  if (!fric && !rnd) return hipSuccess;
  const int f = resolve_form(a, t, form, 0);
  if (f >= kLdsForm1024 && f <= kLdsForm256) {
    const size_t lds = static_cast<size_t>(t.n_elem) * t.n_rho * sizeof(double4);
    const int blk = lds_block(f);
#define EPH_PFL_M(F, R) if (multi) forces_lds<F, R, true>(blk, lds, a, t, g, p, w, s); \
                        else forces_lds<F, R, false>(blk, lds, a, t, g, p, w, s)
    if (fric && rnd) { EPH_PFL_M(true, true); }
    else if (fric) { EPH_PFL_M(true, false); }
    else { EPH_PFL_M(false, true); }
#undef EPH_PFL_M
    return hipGetLastError();
  }
  const int gr = nblk(a.nlocal);
  const size_t fl = chunk_lds(f, kForceStageBytes);
#define EPH_PF(F, R, M) k_prl_forces<F, R, M><<<gr, kBlock, fl, s>>>(a, t, g, p, w)
#define EPH_PF_M(F, R) if (multi) EPH_PF(F, R, true); else EPH_PF(F, R, false)
  if (fric && rnd) { EPH_PF_M(true, true); }
  else if (fric) { EPH_PF_M(true, false); }
  else { EPH_PF_M(false, true); }
#undef EPH_PF_M
#undef EPH_PF
  return hipGetLastError();
}

hipError_t launch_ttm_forces(const AtomsView& a, const TablesView& t, const GridView& g,
                             const StepParams& p, const WorkView& w, hipStream_t s) {
  if (a.nlocal == 0) return hipSuccess;
  const bool fric = p.flags & 0x01, rnd = p.flags & 0x02;
  const int gr = nblk(a.nlocal);
  if (fric && rnd) k_ttm_forces<true, true><<<gr, kBlock, 0, s>>>(a, t, g, p, w);
  else if (fric) k_ttm_forces<true, false><<<gr, kBlock, 0, s>>>(a, t, g, p, w);
  else if (rnd) k_ttm_forces<false, true><<<gr, kBlock, 0, s>>>(a, t, g, p, w);
  return hipGetLastError();
}

hipError_t launch_prb_forces(const AtomsView& a, const TablesView& t, const GridView& g,
                             const StepParams& p, const WorkView& w, hipStream_t s) {
  if (a.nlocal == 0) return hipSuccess;
  const bool fric = p.flags & 0x01, rnd = p.flags & 0x02;
  const int gr = nblk(a.nlocal);
  if (fric && rnd) k_prb_forces<true, true><<<gr, kBlock, 0, s>>>(a, t, g, p, w);
  else if (fric) k_prb_forces<true, false><<<gr, kBlock, 0, s>>>(a, t, g, p, w);
  else if (rnd) k_prb_forces<false, true><<<gr, kBlock, 0, s>>>(a, t, g, p, w);
  return hipGetLastError();
}

hipError_t launch_colour_v(const AtomsView& a, const StepParams& p, const WorkView& w, hipStream_t s) {
  if (a.nlocal > 0) k_colour_v_local<<<nblk(a.nlocal), kBlock, 0, s>>>(a, p, w);
  if (a.nghost > 0) k_colour_v_ghost<<<nblk(a.nghost), kBlock, 0, s>>>(a, w);
  return hipGetLastError();
}

hipError_t launch_integrate(const AtomsView& a, const StepParams& p, bool initial, hipStream_t s) {
  if (a.nlocal == 0) return hipSuccess;
  if (initial) k_integrate<true><<<nblk(a.nlocal), kBlock, 0, s>>>(a, p);
  else k_integrate<false><<<nblk(a.nlocal), kBlock, 0, s>>>(a, p);
  return hipGetLastError();
}

hipError_t launch_fill_int(int* p, int n, int v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_fill_int<<<nblk(n), kBlock, 0, s>>>(p, n, v);
  return hipGetLastError();
}

hipError_t launch_fill_double(double* p, long n, double v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_fill_double<<<nblk(n), kBlock, 0, s>>>(p, n, v);
  return hipGetLastError();
}

hipError_t launch_deposit_atoms(const AtomsView& a, const GridView& g, const StepParams& p,
                                const WorkView& w, const DepositView& d, int nblocks, hipStream_t s) {
  const bool fric = p.flags & 0x01, rnd = p.flags & 0x02;
  if (nblocks == 0) return hipSuccess;
  if (fric && rnd) k_deposit_atoms<true, true><<<nblocks, kBlock, 0, s>>>(a, g, p, w, d);
  else if (fric) k_deposit_atoms<true, false><<<nblocks, kBlock, 0, s>>>(a, g, p, w, d);
  else if (rnd) k_deposit_atoms<false, true><<<nblocks, kBlock, 0, s>>>(a, g, p, w, d);
  else k_deposit_atoms<false, false><<<nblocks, kBlock, 0, s>>>(a, g, p, w, d);
  return hipGetLastError();
}

hipError_t launch_deposit_points(int n, const double* x, const double* y, const double* z,
                                 const double* E, double dt, const GridView& g, const DepositView& d,
                                 hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_deposit_points<<<nblk(n), kBlock, 0, s>>>(n, x, y, z, E, dt, g, d);
  return hipGetLastError();
}

// count[] -> start[] -> items[] -> ordered per-cell sums into g.dT
hipError_t launch_deposit_cells(const GridView& g, const DepositView& d, int n_items, int* qbuf,
                                hipStream_t s) {
  const int G = g.nx * g.ny * g.nz;
  const int per = kScanBlock * kScanItems;
  const int nb = (G + per - 1) / per;
  k_scan_reduce<<<nb, kScanBlock, 0, s>>>(d.count, G, d.scan_tmp);
  k_scan_blocks<<<1, kScanBlock, 0, s>>>(d.scan_tmp, nb);
  k_scan_apply<<<nb, kScanBlock, 0, s>>>(d.count, G, d.scan_tmp, d.start);
  // qbuf layout: [0] medium count, [1] huge count, [2 ..] medium queue,
  // [2 + G ..] huge queue
  int* qn = qbuf;
  int* hn = qbuf + 1;
  k_fill_int<<<1, 64, 0, s>>>(qbuf, 2, 0);
  k_fill_int<<<nblk(G), kBlock, 0, s>>>(d.fill, G, 0);
  if (n_items > 0) k_scatter<<<nblk(n_items), kBlock, 0, s>>>(n_items, d);
  int* queue = qbuf + 2;
  int* huge = qbuf + 2 + G;
  k_cell_sum_small<<<nblk(G), kBlock, 0, s>>>(g, d, G, queue, qn);
  k_cell_sum_medium<<<64, kMedBlock, 0, s>>>(g, d, queue, qn, huge, hn);
  k_cell_sum_huge<<<16, kBlock, 0, s>>>(g, d, n_items, huge, hn);
  return hipGetLastError();
}

hipError_t launch_deposit_bucket(const AtomsView& a, const GridView& g, const StepParams& p,
                                 const WorkView& w, const DepositView& d, const BucketView& b,
                                 bool fric, bool rand, bool kick, double* E_local, double* Ee,
                                 hipStream_t s) {
  const int G = g.nx * g.ny * g.nz;
  const int nb = nblk(a.nlocal);
  if (a.nlocal > 0) {
#define EPH_DB(F, R) if (kick) k_deposit_bucket<F, R, true><<<nb, kBlock, 0, s>>>(a, g, p, w, d, b); \
                     else k_deposit_bucket<F, R, false><<<nb, kBlock, 0, s>>>(a, g, p, w, d, b)
    if (fric && rand) { EPH_DB(true, true); }
    else if (fric) { EPH_DB(true, false); }
    else { EPH_DB(false, true); }
#undef EPH_DB
  } else {
    k_fill_int<<<1, 64, 0, s>>>(b.nqueue, 1, 0);
  }
  DepositView dc = d;
  if (!fric) dc.v1 = nullptr;
  if (!rand) dc.v2 = nullptr;
  k_deposit_bucket_cells<<<nblk(G), kBlock, 0, s>>>(g, dc, b, G);
  k_deposit_finish<<<1 + 16, 1024, 0, s>>>(a, g, dc, b, a.nlocal > 0 ? nb : 0, E_local, Ee);
  return hipGetLastError();
}

hipError_t launch_energy_reduce(const DepositView& d, int nblocks, double* E_local, double* Ee,
                                hipStream_t s) {
  k_energy_reduce<<<1, 1024, 0, s>>>(d.partial, nblocks, E_local, Ee);
  return hipGetLastError();
}

hipError_t launch_grid_tdyn_update(const GridView& g, hipStream_t s) {
  const int n = g.nx * g.ny * g.nz;
  k_grid_tdyn_update<<<nblk(n), kBlock, 0, s>>>(g, n);
  return hipGetLastError();
}

hipError_t launch_grid_minmax(const GridView& g, double* out3, double*, hipStream_t s) {
  k_grid_minmax<<<1, 1024, 0, s>>>(g, g.nx * g.ny * g.nz, out3);
  return hipGetLastError();
}

hipError_t launch_fdm_substep(const GridView& g, const double* Tin, double* Tout, double inner_dt,
                              bool any_tdyn, hipStream_t s) {
  const int n = g.nx * g.ny * g.nz;
  if (any_tdyn) k_fdm_substep<true><<<nblk(n), kBlock, 0, s>>>(g, Tin, Tout, inner_dt);
  else k_fdm_substep<false><<<nblk(n), kBlock, 0, s>>>(g, Tin, Tout, inner_dt);
  return hipGetLastError();
}

hipError_t launch_grid_lookup(const GridView& g, int n, const double* x, const double* y,
                              const double* z, int* idx, double* T, hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_grid_lookup<<<nblk(n), kBlock, 0, s>>>(g, n, x, y, z, idx, T);
  return hipGetLastError();
}

hipError_t launch_sum(const double* v, int n, double* out, double* scratch, hipStream_t s) {
  const int nb = 256;
  k_sum_partial<<<nb, kBlock, 0, s>>>(v, n, scratch);
  k_sum_final<<<1, kBlock, 0, s>>>(scratch, nb, out);
  return hipGetLastError();
}

hipError_t launch_nonfinite(const double* v, int n, int* flag, hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_nonfinite<<<nblk(n), kBlock, 0, s>>>(v, n, flag);
  return hipGetLastError();
}

hipError_t launch_peratom(const AtomsView& a, const TablesView& t, const WorkView& w, double* out,
                          hipStream_t s) {
  if (a.nlocal == 0) return hipSuccess;
  k_peratom<<<nblk(a.nlocal), kBlock, 0, s>>>(a, t, w, out);
  return hipGetLastError();
}

hipError_t launch_lj(const AtomsView& a, const int* nbr, const int* nn, int stride, double eps,
                     double sigma, double rc2, hipStream_t s) {
  if (a.nlocal == 0) return hipSuccess;
  const double s6 = sigma * sigma * sigma * sigma * sigma * sigma;
  const double lj1 = 48.0 * eps * s6 * s6, lj2 = 24.0 * eps * s6;
  k_lj<<<nblk(a.nlocal), kBlock, 0, s>>>(a, nbr, nn, stride, lj1, lj2, rc2);
  return hipGetLastError();
}

hipError_t launch_max_disp2(const AtomsView& a, const double4* ref, double* partial, int nblocks,
                            hipStream_t s) {
  k_max_disp2<<<nblocks, kBlock, 0, s>>>(a, ref, partial);
  return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_copy_probe(double4* __restrict__ dst,
                                                       const double4* __restrict__ src, long n) {
  for (long i = static_cast<long>(blockIdx.x) * kBlock + threadIdx.x; i < n;
       i += static_cast<long>(gridDim.x) * kBlock)
    dst[i] = src[i];
}

hipError_t launch_copy_probe(double4* dst, const double4* src, long n, hipStream_t s) {
  k_copy_probe<<<4096, kBlock, 0, s>>>(dst, src, n);
  return hipGetLastError();
}

hipError_t launch_rows_into_d4(const double* rows, int n, double4* out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_rows_into_d4<<<nblk(n), kBlock, 0, s>>>(rows, n, out);
  return hipGetLastError();
}
hipError_t launch_rows_into_d4_w(const double* rows, int n, const int* type, const int* species,
                                 double4* out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_rows_into_d4_w<<<nblk(n), kBlock, 0, s>>>(rows, n, type, species, out);
  return hipGetLastError();
}
hipError_t launch_d4_into_rows(const double4* in, int n, double* rows, hipStream_t s, int stride4) {
  if (n == 0) return hipSuccess;
  k_d4_into_rows<<<nblk(n), kBlock, 0, s>>>(in, n, stride4, rows);
  return hipGetLastError();
}
hipError_t launch_rows_into_soa(const double* rows, int n, double* soa, int stride, hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_rows_into_soa<<<nblk(n), kBlock, 0, s>>>(rows, n, soa, stride);
  return hipGetLastError();
}
hipError_t launch_soa_into_rows(const double* soa, int n, int stride, double* rows, hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_soa_into_rows<<<nblk(n), kBlock, 0, s>>>(soa, n, stride, rows);
  return hipGetLastError();
}
hipError_t launch_d4_component(const double4* in, int n, int comp, double* out, hipStream_t s,
                               int stride4) {
  if (n == 0) return hipSuccess;
  k_d4_component<<<nblk(n), kBlock, 0, s>>>(in, n, stride4, comp, out);
  return hipGetLastError();
}


// ---------------------------------------------------------- batched wrappers
bool batch_lds_possible(const TablesView& t) {
  const size_t bytes = static_cast<size_t>(t.n_elem) * t.n_rho * sizeof(double4);
  int dev = 0, lds_max = 0;
  if (bytes > kLdsTableMax || hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    return false;
  return bytes <= static_cast<size_t>(lds_max);
}

// the form's block size of a persistent LDS-table launch (forms 1-3), else 0
static int batch_lds_block(int form) {
  return form == kLdsForm1024 ? 1024 : form == kLdsForm512 ? 512 : form == kLdsForm256 ? 256 : 0;
}

hipError_t launch_density_w_batch(const AtomsView& a, const TablesView& t, const StepParams& p,
                                  const WorkView& w_in, const Chunk* chunks, int nchunks,
                                  const uint32_t* seeds, const uint64_t* step, hipStream_t s,
                                  int variant, const int* ng, const int* rep) {
  WorkView w = w_in;
  if (nchunks == 0) return hipSuccess;
  const bool fric = p.flags & 0x01, multi = t.n_elem > 1;
  const int bs = batch_lds_block(variant);
  if (bs > 0 && batch_lds_possible(t)) {
    w.rcand2 = 0.0;  // the LDS-table forms walk the full lists
    const size_t lds = static_cast<size_t>(t.n_elem) * t.n_rho * sizeof(double4);
    const int ntiles = (nchunks + bs / kBlock - 1) / (bs / kBlock);
#define EPH_DWBL(XI, F, M, BS)                                                                   \
  k_density_w_batch_lds<XI, F, M, BS><<<lds_grid(k_density_w_batch_lds<XI, F, M, BS>, lds, ntiles, BS), \
                                        BS, lds, s>>>(a, t, p, w, chunks, nchunks, seeds, step)
#define EPH_DWBL_B(XI, F, M)                 \
  if (bs == 1024) EPH_DWBL(XI, F, M, 1024);  \
  else if (bs == 512) EPH_DWBL(XI, F, M, 512); \
  else EPH_DWBL(XI, F, M, 256)
#define EPH_DWBL_M(XI, F) if (multi) { EPH_DWBL_B(XI, F, true); } else { EPH_DWBL_B(XI, F, false); }
#define EPH_DWBL_F(XI) if (fric) { EPH_DWBL_M(XI, true); } else { EPH_DWBL_M(XI, false); }
    switch (p.xi_mode) {
      case 1: EPH_DWBL_F(1); break;
      case 2: EPH_DWBL_F(2); break;
      default: EPH_DWBL_F(0); break;
    }
#undef EPH_DWBL_F
#undef EPH_DWBL_M
#undef EPH_DWBL_B
#undef EPH_DWBL
    return hipGetLastError();
  }
  // candidate lists over all replicas' atoms, one moved list for all (a wave
  // looks only at its own replica's moved atoms); they need the device ghost
  // count and the replica of every atom
  const bool keep = w.rcand2 > 0.0 && w.cstate && ng && rep;
  if (!keep) w.rcand2 = 0.0;
  if (keep) k_cand_check_dev<<<nblk(a.nlocal + a.nghost), kBlock, 0, s>>>(a.nlocal, ng, a.pos, rep, w);
  const size_t dl = chunk_lds(variant, kDensityStageBytes);
#define EPH_DWB(XI, F, M) k_density_w_batch<XI, F, M><<<nchunks, kBlock, dl, s>>>(a, t, p, w, chunks, seeds, step)
#define EPH_DWB_M(XI, F) if (multi) EPH_DWB(XI, F, true); else EPH_DWB(XI, F, false)
#define EPH_DWB_F(XI) if (fric) { EPH_DWB_M(XI, true); } else { EPH_DWB_M(XI, false); }
  switch (p.xi_mode) {
    case 1: EPH_DWB_F(1); break;
    case 2: EPH_DWB_F(2); break;
    default: EPH_DWB_F(0); break;
  }
#undef EPH_DWB_F
#undef EPH_DWB_M
#undef EPH_DWB
  if (keep) k_cand_flags<<<1, 64, 0, s>>>(w.cstate, w.moved_cap);
  return hipGetLastError();
}

hipError_t launch_prl_forces_batch(const AtomsView& a, const TablesView& t, const GridView* grids,
                                   const StepParams& p, const WorkView& w, const Chunk* chunks,
                                   int nchunks, hipStream_t s, int variant) {
  const bool fric = p.flags & 0x01, rnd = p.flags & 0x02, multi = t.n_elem > 1;
  if (nchunks == 0 || (!fric && !rnd)) return hipSuccess;
  const int bs = batch_lds_block(variant);
  if (bs > 0 && batch_lds_possible(t)) {
    const size_t lds = static_cast<size_t>(t.n_elem) * t.n_rho * sizeof(double4);
    const int ntiles = (nchunks + bs / kBlock - 1) / (bs / kBlock);
#define EPH_PFBL(F, R, M, BS)                                                                     \
  k_prl_forces_batch_lds<F, R, M, BS><<<lds_grid(k_prl_forces_batch_lds<F, R, M, BS>, lds, ntiles, BS), \
                                        BS, lds, s>>>(a, t, grids, p, w, chunks, nchunks)
#define EPH_PFBL_B(F, R, M)                  \
  if (bs == 1024) EPH_PFBL(F, R, M, 1024);   \
  else if (bs == 512) EPH_PFBL(F, R, M, 512); \
  else EPH_PFBL(F, R, M, 256)
#define EPH_PFBL_M(F, R) if (multi) { EPH_PFBL_B(F, R, true); } else { EPH_PFBL_B(F, R, false); }
    if (fric && rnd) { EPH_PFBL_M(true, true); }
    else if (fric) { EPH_PFBL_M(true, false); }
    else { EPH_PFBL_M(false, true); }
#undef EPH_PFBL_M
#undef EPH_PFBL_B
#undef EPH_PFBL
    return hipGetLastError();
  }
  const size_t fl = chunk_lds(variant, kForceStageBytes);
#define EPH_PFB(F, R, M) k_prl_forces_batch<F, R, M><<<nchunks, kBlock, fl, s>>>(a, t, grids, p, w, chunks)
#define EPH_PFB_M(F, R) if (multi) EPH_PFB(F, R, true); else EPH_PFB(F, R, false)
  if (fric && rnd) { EPH_PFB_M(true, true); }
  else if (fric) { EPH_PFB_M(true, false); }
  else { EPH_PFB_M(false, true); }
#undef EPH_PFB_M
#undef EPH_PFB
  return hipGetLastError();
}

hipError_t launch_deposit_atoms_batch(const AtomsView& a, const GridView* grids, const int* cell_off,
                                      const StepParams& p, const WorkView& w, const DepositView& d,
                                      const Chunk* chunks, int nchunks, hipStream_t s) {
  const bool fric = p.flags & 0x01, rnd = p.flags & 0x02;
  if (nchunks == 0) return hipSuccess;
  if (fric && rnd) k_deposit_atoms_batch<true, true><<<nchunks, kBlock, 0, s>>>(a, grids, cell_off, p, w, d, chunks);
  else if (fric) k_deposit_atoms_batch<true, false><<<nchunks, kBlock, 0, s>>>(a, grids, cell_off, p, w, d, chunks);
  else if (rnd) k_deposit_atoms_batch<false, true><<<nchunks, kBlock, 0, s>>>(a, grids, cell_off, p, w, d, chunks);
  else k_deposit_atoms_batch<false, false><<<nchunks, kBlock, 0, s>>>(a, grids, cell_off, p, w, d, chunks);
  return hipGetLastError();
}

hipError_t launch_energy_reduce_batch(const double* partial, const int* chunk0, const int* nchunks,
                                      int nrep, double* E_local, double* Ee, hipStream_t s) {
  if (nrep == 0) return hipSuccess;
  k_energy_reduce_batch<<<nrep, 1024, 0, s>>>(partial, chunk0, nchunks, E_local, Ee);
  return hipGetLastError();
}

hipError_t launch_fdm_substep_batch(const GridView* grids, const int* cell_rep, const int* cell_off,
                                    const double* Tin, double* Tout, const double* inner_dt,
                                    const int* steps, int n, int G, bool tdyn, hipStream_t s) {
  if (G == 0) return hipSuccess;
  if (tdyn)
    k_fdm_substep_batch<true><<<nblk(G), kBlock, 0, s>>>(grids, cell_rep, cell_off, Tin, Tout, inner_dt,
                                                         steps, n, G);
  else
    k_fdm_substep_batch<false><<<nblk(G), kBlock, 0, s>>>(grids, cell_rep, cell_off, Tin, Tout, inner_dt,
                                                          steps, n, G);
  return hipGetLastError();
}

hipError_t launch_grid_tdyn_update_batch(const GridView* grids, const int* cell_rep, const int* cell_off,
                                         int G, hipStream_t s) {
  if (G == 0) return hipSuccess;
  k_grid_tdyn_update_batch<<<nblk(G), kBlock, 0, s>>>(grids, cell_rep, cell_off, G);
  return hipGetLastError();
}

hipError_t launch_fdm_substeps_batch(const GridView* grids, const int8_t* rep_tdyn, const double* grid_dt,
                                     const int* grid_steps, int* sub_n, double* inner_dt, int* max_sub,
                                     int R, hipStream_t s) {
  if (R == 0) return hipSuccess;
  k_fdm_substeps_batch<<<R, 256, 0, s>>>(grids, rep_tdyn, grid_dt, grid_steps, sub_n, inner_dt, max_sub);
  return hipGetLastError();
}

hipError_t launch_copy_double(const double* src, double* dst, int n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_copy_double<<<nblk(n), kBlock, 0, s>>>(src, dst, n);
  return hipGetLastError();
}

hipError_t launch_ghost_xv_dev(const AtomsView& a, const int* owner, const double4* shift,
                               const int* ng, int cap, hipStream_t s) {
  if (cap == 0) return hipSuccess;
  k_ghost_xv_dev<<<nblk(cap), kBlock, 0, s>>>(a, owner, shift, ng);
  return hipGetLastError();
}

hipError_t launch_pack_pv_dev(const AtomsView& a, const WorkView& w, const int* ng, int cap,
                              hipStream_t s) {
  const long n = static_cast<long>(a.nlocal) + cap;
  if (n == 0) return hipSuccess;
  k_pack_pv_dev<<<nblk(n), kBlock, 0, s>>>(a, w, ng);
  return hipGetLastError();
}

hipError_t launch_ghost_rec_dev(const AtomsView& a, const WorkView& w, const int* ng, int cap,
                                hipStream_t s) {
  if (cap == 0) return hipSuccess;
  k_ghost_rec_dev<<<nblk(cap), kBlock, 0, s>>>(a, w, ng);
  return hipGetLastError();
}

hipError_t launch_step_advance(uint64_t* step, hipStream_t s) {
  k_step_advance<<<1, 64, 0, s>>>(step);
  return hipGetLastError();
}

hipError_t launch_step_log(const int* flag, const uint64_t* step, const uint64_t* first, int* log,
                           int nlog, hipStream_t s) {
  k_step_log<<<1, 64, 0, s>>>(flag, step, first, log, nlog);
  return hipGetLastError();
}

}  // namespace eph
