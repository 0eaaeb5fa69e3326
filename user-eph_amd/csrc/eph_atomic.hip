// eph_atomic.hip -- gfx950 kernels of the fix eph/atomic style.
//
// One thread per local atom, neighbours in list order, no FMA contraction and
// no floating-point atomics: every per-atom quantity is bit-identical to the
// serial FixEPHAtomic on the same inputs (fix_eph_atomic.cpp).
#include "eph_atomic.hpp"
#include "eph_devfn.hpp"
#include "eph_rng.hpp"

namespace eph {

// EPH_Linear::operator() (eph_linear.h:42-49) on one table; the top knot,
// whose slope the reference reads past its table, returns the knot value
__device__ __forceinline__ double lin_fwd(const double* y, const double* dy, int n, double dx,
                                          double x) {
  if (!(x >= 0.0)) return 0.0;
  const double q = x / dx;
  if (q >= static_cast<double>(n)) return 0.0;
  const int idx = static_cast<int>(q);
  const double delta = x - idx * dx;
  if (idx >= n - 1) return y[idx];
  return y[idx] + dy[idx] * delta;
}

// EPH_Linear::reverse_lookup (eph_linear.h:52-61): upper_bound, step back
__device__ __forceinline__ double lin_rev(const double* y, const double* dy, int n, double dx,
                                          double v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (v < y[mid]) hi = mid; else lo = mid + 1;
  }
  if (lo >= n || lo == 0) return 0.0;
  const int idx = lo - 1;
  return idx * dx + 1. / dy[idx] * (v - y[idx]);
}

__device__ __forceinline__ double T_of_E(const AtomicView& k, int kel, double E) {
  return lin_rev(k.ET_y + kel * k.nT, k.ET_dy + kel * k.nT, k.nT, k.dT, E);
}

// forward communication of E (state EI, fix_eph_atomic.cpp:778-779, 690-691)
// and the per-atom temperatures (and conductivities) every atom's neighbours
// read; a ghost reads its owner's E, which this kernel does not write
__global__ __launch_bounds__(kBlock) void k_atomic_sync(AtomsView a, WorkView w, AtomicView k,
                                                        bool with_K) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.nlocal + a.nghost) return;
  double E;
  if (i < a.nlocal) {
    E = k.E[i];
  } else {
    E = k.E[w.ghost_owner[i - a.nlocal]];
    k.E[i] = E;
  }
  const int kel = k.kmap[a.type[i] - 1];
  const double T = T_of_E(k, kel, E);
  k.T[i] = T;
  if (with_K) k.K[i] = lin_fwd(k.K_y + kel * k.nT, k.K_dy + kel * k.nT, k.nT, k.dT, T);
}

// calculate_environment (fix_eph_atomic.cpp:450-490): rho_i and rho_a_i over
// group neighbours, each with its own cutoff; the group neighbours inside the
// .beta cutoff form the active list of the PRL sums (which skip exactly the
// others).  Then alpha_i, the w sum of force_prl (:504-545) and xi_i.
template <int XI, bool FRIC>
__global__ __launch_bounds__(kBlock) void k_atomic_env(AtomsView a, TablesView t, StepParams p,
                                                       WorkView w, AtomicView k) {
  const int i = xcd_block() * kBlock + threadIdx.x;
  if (i >= a.nlocal) return;
  const int stride = a.stride;
  const double4 pi = a.pos[i];
  const bool grp = (a.mask[i] & a.groupbit) != 0;
  double rho = 0.0, rho_a = 0.0;
  int na = 0;
  if (grp) {
    const int nn = w.nnbr[i];
    for (int kk = 0; kk < nn; ++kk) {
      const int j = w.nbr[kk * stride + i];
      if (!(a.mask[j] & a.groupbit)) continue;
      const double4 pj = a.pos[j];
      const double ex = pj.x - pi.x, ey = pj.y - pi.y, ez = pj.z - pi.z;
      const double r2 = ex * ex + ey * ey + ez * ez;
      if (r2 < t.rc2) {
        rho += rho_r2_at(t.rho_r2, t, static_cast<int>(pj.w), r2);
        w.act[na * stride + i] = j;
        ++na;
      }
      if (r2 < k.k_rc2) {
        const int kel = k.kmap[a.type[j] - 1];
        rho_a += cubic_at(k.krho + kel * k.k_nr, k.k_nr, k.k_inv_dr2, r2);
      }
    }
  }
  w.nact[i] = na;
  k.rho_a[i] = rho_a;
  double alpha = 0.0;
  if (grp && rho > 0.0) {
    if (rho > t.rho_cut) {
      atomicOr(w.flags_out, 1);  // eph_beta.h:186-198: alpha = 0 with a warning
    } else {
      alpha = cubic_at(t.alpha + static_cast<int>(pi.w) * t.n_beta, t.n_beta, t.inv_drho, rho);
    }
  }
  double wx = 0.0, wy = 0.0, wz = 0.0;
  if (FRIC && grp && rho > 0.0) {
    const double4 vi = a.vel[i];
    for (int kk = 0; kk < na; ++kk) {
      const int j = w.act[kk * stride + i];
      const PosVel q = w.pv[j];
      const double ex = q.p.x - pi.x, ey = q.p.y - pi.y, ez = q.p.z - pi.z;
      const double r2 = ex * ex + ey * ey + ez * ez;
      const double vr = rho_r2_at(t.rho_r2, t, static_cast<int>(q.p.w), r2);
      const double pre = alpha * vr / (rho * r2);
      const double var1 = pre * (ex * vi.x + ey * vi.y + ez * vi.z);
      const double var2 = pre * (ex * q.v.x + ey * q.v.y + ez * q.v.z);
      const double dvar = var1 - var2;
      wx += dvar * ex;
      wy += dvar * ey;
      wz += dvar * ez;
    }
  }
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
  ForceRec* r = w.rec + i;
  r->p = pi;
  r->rw = make_double4(wx, wy, wz, rho);
  r->xa = make_double4(x0, x1, x2, alpha);
}

// forward communication of rho, rho_a (state RHO), w (WI) and xi (XI)
__global__ __launch_bounds__(kBlock) void k_atomic_ghost(AtomsView a, WorkView w, AtomicView k) {
  const int g = blockIdx.x * kBlock + threadIdx.x;
  if (g >= a.nghost) return;
  const int o = w.ghost_owner[g];
  ForceRec* r = w.rec + (a.nlocal + g);
  r->p = a.pos[a.nlocal + g];
  r->rw = w.rec[o].rw;
  r->xa = w.rec[o].xa;
  k.rho_a[a.nlocal + g] = k.rho_a[o];
}

// force_prl (fix_eph_atomic.cpp:551-657): the friction pass, then the random
// pass with sqrt(T) of both atoms inside each pair term and eta applied per
// pair; each pass adds half the pair power to the atom's deposit, in list
// order, friction terms first.  Forces are applied to group atoms
// (:806-822).
template <bool FRIC, bool RAND>
__global__ __launch_bounds__(kBlock) void k_atomic_forces(AtomsView a, TablesView t, StepParams p,
                                                          WorkView w, AtomicView k) {
  const int i = xcd_block() * kBlock + threadIdx.x;
  if (i >= a.nlocal) return;
  const int stride = a.stride;
  const bool grp = (a.mask[i] & a.groupbit) != 0;
  const double4 rwi = w.rec[i].rw;
  const double rho_i = rwi.w;
  const double l_dt = p.dt;
  double fex = 0.0, fey = 0.0, fez = 0.0, frx = 0.0, fry = 0.0, frz = 0.0, dE = 0.0;
  if (grp && rho_i > 0.0) {
    const double4 pi = w.rec[i].p;
    const double4 xai = w.rec[i].xa;
    const double4 vi = a.vel[i];
    const double alpha_i = xai.w;
    const int ei = static_cast<int>(pi.w);
    const int na = w.nact[i];
    if (FRIC) {
      const bool dep = !(p.flags & 0x10);
      for (int kk = 0; kk < na; ++kk) {
        const int j = w.act[kk * stride + i];
        const double4 rwj = w.rec[j].rw;
        const double rho_j = rwj.w;
        if (!(rho_j > 0.0)) continue;
        const double4 pj = w.rec[j].p;
        const double alpha_j = w.rec[j].xa.w;
        const double ex = pj.x - pi.x, ey = pj.y - pi.y, ez = pj.z - pi.z;
        const double r2 = ex * ex + ey * ey + ez * ez;
        const int ej = static_cast<int>(pj.w);
        const double vr_ji = rho_r2_at(t.rho_r2, t, ej, r2);
        const double var1 = alpha_i * vr_ji * (ex * rwi.x + ey * rwi.y + ez * rwi.z) / (rho_i * r2);
        const double vr_ij = rho_r2_at(t.rho_r2, t, ei, r2);
        const double var2 = alpha_j * vr_ij * (ex * rwj.x + ey * rwj.y + ez * rwj.z) / (rho_j * r2);
        const double dvar = var1 - var2;
        const double f0 = dvar * ex, f1 = dvar * ey, f2 = dvar * ez;
        fex -= f0; fey -= f1; fez -= f2;
        if (dep) {
          const double4 vj = a.vel[j];
          dE += 0.5 * f0 * (vi.x - vj.x) * l_dt;
          dE += 0.5 * f1 * (vi.y - vj.y) * l_dt;
          dE += 0.5 * f2 * (vi.z - vj.z) * l_dt;
        }
      }
    }
    if (RAND) {
      const bool dep = !(p.flags & 0x20);
      const double v_Ti = sqrt(k.T[i]);
      for (int kk = 0; kk < na; ++kk) {
        const int j = w.act[kk * stride + i];
        const double4 rwj = w.rec[j].rw;
        const double rho_j = rwj.w;
        if (!(rho_j > 0.0)) continue;
        const double4 pj = w.rec[j].p;
        const double4 xaj = w.rec[j].xa;
        const double alpha_j = xaj.w;
        const double v_Tj = sqrt(k.T[j]);
        const double ex = pj.x - pi.x, ey = pj.y - pi.y, ez = pj.z - pi.z;
        const double r2 = ex * ex + ey * ey + ez * ez;
        const int ej = static_cast<int>(pj.w);
        const double vr_ji = rho_r2_at(t.rho_r2, t, ej, r2);
        const double var1 =
            v_Ti * alpha_i * vr_ji * (ex * xai.x + ey * xai.y + ez * xai.z) / (rho_i * r2);
        const double vr_ij = rho_r2_at(t.rho_r2, t, ei, r2);
        const double var2 =
            v_Tj * alpha_j * vr_ij * (ex * xaj.x + ey * xaj.y + ez * xaj.z) / (rho_j * r2);
        const double dvar = p.eta * (var1 - var2);
        const double f0 = dvar * ex, f1 = dvar * ey, f2 = dvar * ez;
        frx += f0; fry += f1; frz += f2;
        if (dep) {
          const double4 vj = a.vel[j];
          dE -= 0.5 * f0 * (vi.x - vj.x) * l_dt;
          dE -= 0.5 * f1 * (vi.y - vj.y) * l_dt;
          dE -= 0.5 * f2 * (vi.z - vj.z) * l_dt;
        }
      }
    }
  }
  const int fs = a.fstride;
  if (FRIC) { w.fE[i] = fex; w.fE[fs + i] = fey; w.fE[2 * fs + i] = fez; }
  if (RAND) { w.fR[i] = frx; w.fR[fs + i] = fry; w.fR[2 * fs + i] = frz; }
  k.dE[i] = dE;
  const bool addE = FRIC && grp && !(p.flags & 0x10);
  const bool addR = RAND && grp && !(p.flags & 0x20);
  if (addE || addR) {
    double fx = a.f[i], fy = a.f[fs + i], fz = a.f[2 * fs + i];
    if (addE) { fx = fx + fex; fy = fy + fey; fz = fz + fez; }
    if (addR) { fx = fx + frx; fy = fy + fry; fz = fz + frz; }
    a.f[i] = fx; a.f[fs + i] = fy; a.f[2 * fs + i] = fz;
  }
}

// heat_solve (fix_eph_atomic.cpp:660-762), one inner loop in three launches:
// the deposit share and the clamp, (after the sync of E, T, K) the diffusion
// over group neighbours inside the .kappa cutoff, and the copy back
__global__ __launch_bounds__(kBlock) void k_heat_add(AtomsView a, AtomicView k, double scaling) {
  const int j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= a.nlocal || !(a.mask[j] & a.groupbit)) return;
  double E = k.E[j] + k.dE[j] * scaling;
  if (E < 0.0) E = 0.0;
  k.E[j] = E;
}

__global__ __launch_bounds__(kBlock) void k_heat_diffuse(AtomsView a, WorkView w, AtomicView k,
                                                         double dt) {
  const int j = xcd_block() * kBlock + threadIdx.x;
  if (j >= a.nlocal) return;
  const double E0 = k.E[j];
  double E1 = E0;
  if (a.mask[j] & a.groupbit) {
    const int stride = a.stride;
    const double4 pj = a.pos[j];
    const int kj = k.kmap[a.type[j] - 1];
    double l_dE_j = 0.;
    const double l_T_j = k.T[j];
    const double l_K_j = k.K[j];
    const double rho_j = k.rho_a[j];
    const double rho_j_inv = 1. / k.rho_a[j];
    const int nn = w.nnbr[j];
    for (int kk = 0; kk < nn; ++kk) {
      const int q = w.nbr[kk * stride + j];
      if (!(a.mask[q] & a.groupbit)) continue;
      const double rho_k = k.rho_a[q];
      const double rho_k_inv = 1. / k.rho_a[q];
      const double l_K = 0.5 * (l_K_j + k.K[q]);
      const double v_dT = k.T[q] - l_T_j;
      const double4 pq = a.pos[q];
      const double ex = pq.x - pj.x, ey = pq.y - pj.y, ez = pq.z - pj.z;
      const double r2 = ex * ex + ey * ey + ez * ez;
      if (r2 >= k.k_rc2) continue;
      const int kq = k.kmap[a.type[q] - 1];
      const double v_rho_j = cubic_at(k.krho + kj * k.k_nr, k.k_nr, k.k_inv_dr2, r2);
      const double v_rho_k = cubic_at(k.krho + kq * k.k_nr, k.k_nr, k.k_inv_dr2, r2);
      if (rho_j > 0.) l_dE_j += l_K * v_rho_k * rho_j_inv * v_dT;
      if (rho_k > 0.) l_dE_j += l_K * v_rho_j * rho_k_inv * v_dT;
    }
    E1 = E0 + 0.5 * l_dE_j * dt;
    if (E1 < 0) E1 = 0.0;
  }
  k.E1[j] = E1;
}

__global__ __launch_bounds__(kBlock) void k_heat_copy(AtomsView a, AtomicView k) {
  const int j = blockIdx.x * kBlock + threadIdx.x;
  if (j < a.nlocal) k.E[j] = k.E1[j];
}

__global__ __launch_bounds__(kBlock) void k_group_values(AtomsView a, const double* v, double* out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < a.nlocal) out[i] = (a.mask[i] & a.groupbit) ? v[i] : 0.0;
}

// populate_array (fix_eph_atomic.cpp:416-448): rho, beta(rho), f_EPH, f_RNG,
// rho_a, E, dE, T per local atom, zeros outside the group
__global__ __launch_bounds__(kBlock) void k_peratom_atomic(AtomsView a, TablesView t, WorkView w,
                                                           AtomicView k, double* out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.nlocal) return;
  double* o = out + 12 * static_cast<size_t>(i);
  if (!(a.mask[i] & a.groupbit)) {
    for (int c = 0; c < 12; ++c) o[c] = 0.0;
    return;
  }
  const double rho = w.rec[i].rw.w;
  const int e = static_cast<int>(a.pos[i].w);
  const int fs = a.fstride;
  o[0] = rho;
  o[1] = rho > t.rho_cut ? 0.0 : cubic_at(t.beta + e * t.n_beta, t.n_beta, t.inv_drho, rho);
  o[2] = w.fE[i]; o[3] = w.fE[fs + i]; o[4] = w.fE[2 * fs + i];
  o[5] = w.fR[i]; o[6] = w.fR[fs + i]; o[7] = w.fR[2 * fs + i];
  o[8] = k.rho_a[i]; o[9] = k.E[i]; o[10] = k.dE[i]; o[11] = k.T[i];
}

static inline int nblk(long n) { return static_cast<int>((n + kBlock - 1) / kBlock); }

hipError_t launch_atomic_sync(const AtomsView& a, const WorkView& w, const AtomicView& k,
                              bool with_K, hipStream_t s) {
  const int n = a.nlocal + a.nghost;
  if (n) k_atomic_sync<<<nblk(n), kBlock, 0, s>>>(a, w, k, with_K);
  return hipGetLastError();
}

hipError_t launch_atomic_env(const AtomsView& a, const TablesView& t, const StepParams& p,
                             const WorkView& w, const AtomicView& k, hipStream_t s) {
  if (a.nlocal == 0) return hipSuccess;
  const int g = nblk(a.nlocal);
  const bool fric = p.flags & 0x01;
#define EPH_AE(XI) \
  if (fric) k_atomic_env<XI, true><<<g, kBlock, 0, s>>>(a, t, p, w, k); \
  else k_atomic_env<XI, false><<<g, kBlock, 0, s>>>(a, t, p, w, k)
  switch (p.xi_mode) {
    case 1: EPH_AE(1); break;
    case 2: EPH_AE(2); break;
    default: EPH_AE(0); break;
  }
#undef EPH_AE
  return hipGetLastError();
}

hipError_t launch_atomic_ghost(const AtomsView& a, const WorkView& w, const AtomicView& k,
                               hipStream_t s) {
  if (a.nghost) k_atomic_ghost<<<nblk(a.nghost), kBlock, 0, s>>>(a, w, k);
  return hipGetLastError();
}

hipError_t launch_atomic_forces(const AtomsView& a, const TablesView& t, const StepParams& p,
                                const WorkView& w, const AtomicView& k, hipStream_t s) {
  if (a.nlocal == 0) return hipSuccess;
  const int g = nblk(a.nlocal);
  const bool fric = p.flags & 0x01, rnd = p.flags & 0x02;
  if (fric && rnd) k_atomic_forces<true, true><<<g, kBlock, 0, s>>>(a, t, p, w, k);
  else if (fric) k_atomic_forces<true, false><<<g, kBlock, 0, s>>>(a, t, p, w, k);
  else if (rnd) k_atomic_forces<false, true><<<g, kBlock, 0, s>>>(a, t, p, w, k);
  else k_atomic_forces<false, false><<<g, kBlock, 0, s>>>(a, t, p, w, k);
  return hipGetLastError();
}

hipError_t launch_heat_add(const AtomsView& a, const AtomicView& k, double scaling, hipStream_t s) {
  if (a.nlocal) k_heat_add<<<nblk(a.nlocal), kBlock, 0, s>>>(a, k, scaling);
  return hipGetLastError();
}

hipError_t launch_heat_diffuse(const AtomsView& a, const WorkView& w, const AtomicView& k,
                               double dt, hipStream_t s) {
  if (a.nlocal) k_heat_diffuse<<<nblk(a.nlocal), kBlock, 0, s>>>(a, w, k, dt);
  return hipGetLastError();
}

hipError_t launch_heat_copy(const AtomsView& a, const AtomicView& k, hipStream_t s) {
  if (a.nlocal) k_heat_copy<<<nblk(a.nlocal), kBlock, 0, s>>>(a, k);
  return hipGetLastError();
}

hipError_t launch_group_values(const AtomsView& a, const double* v, double* out, hipStream_t s) {
  if (a.nlocal) k_group_values<<<nblk(a.nlocal), kBlock, 0, s>>>(a, v, out);
  return hipGetLastError();
}

hipError_t launch_peratom_atomic(const AtomsView& a, const TablesView& t, const WorkView& w,
                                 const AtomicView& k, double* out, hipStream_t s) {
  if (a.nlocal) k_peratom_atomic<<<nblk(a.nlocal), kBlock, 0, s>>>(a, t, w, k, out);
  return hipGetLastError();
}

}  // namespace eph
