// eph_adjoint.hip -- reverse-mode (adjoint) kernels of one MD step with the
// fix: vector-Jacobian products of the FDM stencil, the deposit, the Verlet
// kicks, pair lj/cut, the force sweep (stages 2 + 3 of force_prl), the w sweep
// (stage 1), alpha(rho) and the density sweep.  Forward expressions are those
// of eph_kernels.hip (fix_eph.cpp:431-466, 702-836, 364-390; eph_fdm.h:315-395);
// each pair term is differentiated by hand.  Contributions to a neighbour's
// adjoint are scattered with FP64 atomics, so gradients are deterministic up to
// the order of those additions.
#include "eph_adjoint.hpp"
#include "eph_devfn.hpp"

namespace eph {

namespace {

// value and d/dr^2 of rho(r^2) of element e at x = r^2 (the interval and cubic
// of rho_r2_at, interleaved tables)
__device__ __forceinline__ void rho_r2_vd(const TablesView& t, int e, double x, double* v, double* dv) {
  int i = static_cast<int>(x * t.inv_dr2);
  i = i < 0 ? 0 : (i < t.n_rho - 1 ? i : t.n_rho - 1);
  const double4 c = t.rho_r2[i * t.n_elem + e];
  *v = c.x + x * (c.y + x * (c.z + x * c.w));
  *dv = c.y + x * (2.0 * c.z + 3.0 * x * c.w);
}

__device__ __forceinline__ void atomic_add3(double4* p, double x, double y, double z) {
  atomicAdd(&p->x, x);
  atomicAdd(&p->y, y);
  atomicAdd(&p->z, z);
}

inline int nblk(long n) { return static_cast<int>((n + kBlock - 1) / kBlock); }

}  // namespace

// ------------------------------------------------------------------ FDM
// one sub-step of k_fdm_substep in reverse:
// Tb_in += (dT_out/dT_in)^T Tb_out, dTb += (dT_out/d dT)^T Tb_out, and with
// temperature-dependent cells (TDYN) kb += (dT_out/d kappa)^T Tb_out.  A
// temperature-dependent cell updates in energy space, T_out =
// E^-1(E(T_in) + (ddT + dT + S) / rho dt_in), whose derivatives are the
// slopes of the two piecewise-linear lookups; the others update
// T_out = T_in + (ddT + dT + S) / (rho C) dt_in.
template <bool TDYN>
__global__ __launch_bounds__(kBlock) void k_adj_fdm(GridView g, const double* __restrict__ Tin,
                                                    const double* __restrict__ Tb_out,
                                                    double* Tb_in, double* dTb, double* kb,
                                                    double inner_dt, int G) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= G) return;
  const double tb = Tb_out[r];
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int i = r % nx, j = (r / nx) % ny, k = r / (nx * ny);
  const int fr = g.flag[r];
  // forward value before the clamp
  double ddT = 0.0;
  int pp[3], qq[3];
  const double dd[3] = {g.dx, g.dy, g.dz};
  if (fr != kCellZeroDerivative) {
    int p, q;
    p = i > 0 ? r - 1 : r + (nx - 1);
    q = i < nx - 1 ? r + 1 : r - (nx - 1);
    if (g.flag[q] == kCellZeroDerivative) q = r;
    else if (g.flag[p] == kCellZeroDerivative) p = r;
    pp[0] = p; qq[0] = q;
    p = j > 0 ? r - nx : r + (ny - 1) * nx;
    q = j < ny - 1 ? r + nx : r - (ny - 1) * nx;
    if (g.flag[q] == kCellZeroDerivative) q = r;
    else if (g.flag[p] == kCellZeroDerivative) p = r;
    pp[1] = p; qq[1] = q;
    const int pl = nx * ny;
    p = k > 0 ? r - pl : r + (nz - 1) * pl;
    q = k < nz - 1 ? r + pl : r - (nz - 1) * pl;
    if (g.flag[q] == kCellZeroDerivative) q = r;
    else if (g.flag[p] == kCellZeroDerivative) p = r;
    pp[2] = p; qq[2] = q;
    const double Tr = Tin[r], kr = g.kappa[r];
    for (int d = 0; d < 3; ++d) {
      const int P = pp[d], Q = qq[d];
      ddT += (g.kappa[Q] - g.kappa[P]) * (Tin[Q] - Tin[P]) / dd[d] / dd[d] / 4.0;
      ddT += kr * ((Tin[Q] + Tin[P] - 2.0 * Tr) / dd[d] / dd[d]);
    }
  }
  double T = Tin[r];
  double kfac = 0.0;  // dT_out / d ddT (= dT_out / d dT)
  double sfac = 1.0;  // dT_out / dT_in through the cell's own update
  if (fr == kCellDynamic) {
    if (TDYN && g.tdyn[r] == 1) {
      const double E = linear_fwd(g, T) + (ddT + g.dT[r] + g.S[r]) / g.rho[r] * inner_dt;
      const double inv = linear_rev_slope(g, E);
      sfac = inv * linear_fwd_slope(g, T);
      kfac = inv * inner_dt / g.rho[r];
      T = linear_rev(g, E);
    } else {
      kfac = inner_dt / (g.rho[r] * g.C[r]);
      T += (ddT + g.dT[r] + g.S[r]) / (g.rho[r] * g.C[r]) * inner_dt;
    }
  }
  if (T < 0.0 || tb == 0.0) return;  // clamped to 0 K: no dependence
  atomicAdd(&Tb_in[r], tb * sfac);
  if (fr != kCellDynamic) return;
  atomicAdd(&dTb[r], tb * kfac);
  const double kr = g.kappa[r];
  for (int d = 0; d < 3; ++d) {
    const int P = pp[d], Q = qq[d];
    const double h2 = dd[d] * dd[d];
    const double cq = (g.kappa[Q] - g.kappa[P]) / h2 / 4.0 + kr / h2;
    const double cp = -(g.kappa[Q] - g.kappa[P]) / h2 / 4.0 + kr / h2;
    atomicAdd(&Tb_in[Q], tb * kfac * cq);
    atomicAdd(&Tb_in[P], tb * kfac * cp);
    atomicAdd(&Tb_in[r], tb * kfac * (-2.0 * kr / h2));
    if (TDYN) {
      // ddT's dependence on kappa at q, p and the cell itself
      const double dq = (Tin[Q] - Tin[P]) / h2 / 4.0;
      atomicAdd(&kb[Q], tb * kfac * dq);
      atomicAdd(&kb[P], -tb * kfac * dq);
      atomicAdd(&kb[r], tb * kfac * ((Tin[Q] + Tin[P] - 2.0 * Tin[r]) / h2));
    }
  }
}

// kappa(T) of the temperature-dependent cells is evaluated once per solve at
// the solve's starting T_e (eph_fdm.h:280-287): its gradient, gathered over
// the sub-steps in kb, goes to that T_e through the spline's slope
__global__ __launch_bounds__(kBlock) void k_adj_kappa_T(GridView g, const double* __restrict__ T0,
                                                        const double* __restrict__ kb, double* Tb,
                                                        int G) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= G || !g.tdyn[c] || kb[c] == 0.0) return;
  Tb[c] += kb[c] * cubic_slope(g.kappa_T, g.n_kT, g.inv_dT_k, T0[c]);
}

// ------------------------------------------------------------- deposit
// dE_i = -(f_i . v'_i) dt for the friction and random parts; each feeds the
// tally (weight Eb) and dT_e of its cell (weight dTb / (dV dt))
__global__ __launch_bounds__(kBlock) void k_adj_deposit(AdjStepView s, GridView g, AtomsView a,
                                                        double dt, int fric, int rnd, double Eb,
                                                        const double* __restrict__ dTb,
                                                        AdjBuffers b) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.nlocal || !(a.mask[i] & a.groupbit)) return;
  const double4 x = s.rec[i].p;
  const double4 v = s.vn[i];
  const int c = cell_of(g, x.x, x.y, x.z);
  const double w = Eb + (dTb ? dTb[c] / (g.dV * dt) : 0.0);
  const int fs = s.fstride;
  double4 vb = b.vnb[i];
  if (fric) {
    const double fx = s.fE[i], fy = s.fE[fs + i], fz = s.fE[2 * fs + i];
    double4 fb = b.fEb[i];
    fb.x -= w * dt * v.x; fb.y -= w * dt * v.y; fb.z -= w * dt * v.z;
    b.fEb[i] = fb;
    vb.x -= w * dt * fx; vb.y -= w * dt * fy; vb.z -= w * dt * fz;
  }
  if (rnd) {
    const double fx = s.fR[i], fy = s.fR[fs + i], fz = s.fR[2 * fs + i];
    double4 fb = b.fRb[i];
    fb.x -= w * dt * v.x; fb.y -= w * dt * v.y; fb.z -= w * dt * v.z;
    b.fRb[i] = fb;
    vb.x -= w * dt * fx; vb.y -= w * dt * fy; vb.z -= w * dt * fz;
  }
  b.vnb[i] = vb;
}

// ------------------------------------------------------ final integrate
// v' = vh + (dtf / m) f  (group atoms; others keep vh)
__global__ __launch_bounds__(kBlock) void k_adj_final(AtomsView a, StepParams p, int noint,
                                                      AdjBuffers b) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.nlocal) return;
  const double4 vn = b.vnb[i];
  double4 vh = b.vhb[i];
  vh.x += vn.x; vh.y += vn.y; vh.z += vn.z;
  b.vhb[i] = vh;
  if (!noint && (a.mask[i] & a.groupbit)) {
    const double c = p.dtf / p.mass[a.type[i]];
    double4 ft = b.ftb[i];
    ft.x += c * vn.x; ft.y += c * vn.y; ft.z += c * vn.z;
    b.ftb[i] = ft;
  }
}

// -------------------------------------------------------------- lj/cut
__global__ __launch_bounds__(kBlock) void k_adj_lj(AdjStepView s, AtomsView a, double lj1,
                                                   double lj2, double rc2, AdjBuffers b) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.nlocal) return;
  const double4 u = b.ftb[i];
  if (u.x == 0.0 && u.y == 0.0 && u.z == 0.0) return;
  const double4 pi = s.pv[i].p;
  const int n = s.ljn[i];
  double bx = 0.0, by = 0.0, bz = 0.0;
  for (int k = 0; k < n; ++k) {
    const int j = s.lj[k * a.stride + i];
    const double4 pj = s.pv[j].p;
    const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
    const double r2 = dx * dx + dy * dy + dz * dz;
    if (!(r2 < rc2)) continue;
    // F(r^2) = lj1 r^-14 - lj2 r^-8 (the fpair of k_lj), dF/d(r^2)
    const double r2i = 1.0 / r2;
    const double r4i = r2i * r2i, r6i = r4i * r2i, r8i = r4i * r4i;
    const double F = r6i * (lj1 * r6i - lj2) * r2i;
    const double dF = -7.0 * lj1 * r8i * r8i + 4.0 * lj2 * r8i * r2i;
    const double ud = u.x * dx + u.y * dy + u.z * dz;
    const double gx = F * u.x + 2.0 * dF * ud * dx;
    const double gy = F * u.y + 2.0 * dF * ud * dy;
    const double gz = F * u.z + 2.0 * dF * ud * dz;
    bx += gx; by += gy; bz += gz;
    atomic_add3(b.xb + j, -gx, -gy, -gz);
  }
  atomic_add3(b.xb + i, bx, by, bz);
}

// ------------------------------------------------- force sweep (2 + 3)
// f_E,i = -sum_j [B_ij (e.w_i) - B_ji (e.w_j)] e,
// f_R,i = eta sqrt(T_i) sum_j [B_ij (e.xi_i) - B_ji (e.xi_j)] e,
// B_ij = alpha_i rho_{t_j}(r^2) / (rho_i r^2), e = x_j - x_i.
__global__ __launch_bounds__(kBlock) void k_adj_force(AdjStepView s, AtomsView a, TablesView t,
                                                      GridView g, StepParams p, int fric, int rnd,
                                                      int addE, int addR, AdjBuffers b) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.nlocal || !(a.mask[i] & a.groupbit)) return;
  const ForceRec ri = s.rec[i];
  const double rho_i = ri.rw.w;
  if (!(rho_i > 0.0)) return;
  const double4 pi = ri.p, wi = ri.rw, xi = ri.xa;
  const double alpha_i = xi.w;
  const double4 ft = b.ftb[i];
  double4 uE = fric ? b.fEb[i] : make_double4(0, 0, 0, 0);
  double4 uR = rnd ? b.fRb[i] : make_double4(0, 0, 0, 0);
  if (addE && fric) { uE.x += ft.x; uE.y += ft.y; uE.z += ft.z; }
  if (addR && rnd) { uR.x += ft.x; uR.y += ft.y; uR.z += ft.z; }
  // random part: f_R = eta sqrt(T) S -> S_bar = eta sqrt(T) uR, T_bar += uR.f_R / (2 T)
  double sx = 0.0, sy = 0.0, sz = 0.0;
  if (rnd) {
    const int c = cell_of(g, pi.x, pi.y, pi.z);
    const double Te = g.T[c];
    const double f = p.eta * sqrt(Te);
    const int fs = s.fstride;
    if (Te > 0.0) {
      const double fr = uR.x * s.fR[i] + uR.y * s.fR[fs + i] + uR.z * s.fR[2 * fs + i];
      atomicAdd(b.Tb + c, fr / (2.0 * Te));
    }
    sx = f * uR.x; sy = f * uR.y; sz = f * uR.z;
  }
  const bool multi = t.n_elem > 1;
  const int ei = multi ? static_cast<int>(pi.w) : 0;
  double wbx = 0, wby = 0, wbz = 0, abi = 0, rbi = 0, xbx = 0, xby = 0, xbz = 0;
  const int na = s.nact[i];
  for (int k = 0; k < na; ++k) {
    const int j = s.act[k * a.stride + i];
    const ForceRec rj = s.rec[j];
    const double rho_j = rj.rw.w;
    if (!(rho_j > 0.0)) continue;
    const double4 pj = rj.p, wj = rj.rw, xj = rj.xa;
    const double alpha_j = xj.w;
    const double ex = pj.x - pi.x, ey = pj.y - pi.y, ez = pj.z - pi.z;
    const double r2 = ex * ex + ey * ey + ez * ez;
    const int ej = multi ? static_cast<int>(pj.w) : 0;
    double vji, dvji, vij, dvij;
    rho_r2_vd(t, ej, r2, &vji, &dvji);
    rho_r2_vd(t, ei, r2, &vij, &dvij);
    const double B1 = alpha_i * vji / (rho_i * r2), B2 = alpha_j * vij / (rho_j * r2);
    double B1b = 0, B2b = 0, gx = 0, gy = 0, gz = 0;  // gx..: d/de
    if (fric) {
      const double a1 = ex * wi.x + ey * wi.y + ez * wi.z;
      const double a2 = ex * wj.x + ey * wj.y + ez * wj.z;
      const double sc = -(B1 * a1 - B2 * a2);
      const double sb = uE.x * ex + uE.y * ey + uE.z * ez;
      gx += sc * uE.x; gy += sc * uE.y; gz += sc * uE.z;
      B1b -= sb * a1;
      B2b += sb * a2;
      const double a1b = -sb * B1, a2b = sb * B2;
      wbx += a1b * ex; wby += a1b * ey; wbz += a1b * ez;
      atomic_add3(b.wb + j, a2b * ex, a2b * ey, a2b * ez);
      gx += a1b * wi.x + a2b * wj.x;
      gy += a1b * wi.y + a2b * wj.y;
      gz += a1b * wi.z + a2b * wj.z;
    }
    if (rnd) {
      const double b1 = ex * xi.x + ey * xi.y + ez * xi.z;
      const double b2 = ex * xj.x + ey * xj.y + ez * xj.z;
      const double sc = B1 * b1 - B2 * b2;
      const double sb = sx * ex + sy * ey + sz * ez;
      gx += sc * sx; gy += sc * sy; gz += sc * sz;
      B1b += sb * b1;
      B2b -= sb * b2;
      const double b1b = sb * B1, b2b = -sb * B2;
      gx += b1b * xi.x + b2b * xj.x;
      gy += b1b * xi.y + b2b * xj.y;
      gz += b1b * xi.z + b2b * xj.z;
    }
    // B1 = alpha_i v_ji / (rho_i r2), B2 = alpha_j v_ij / (rho_j r2)
    abi += B1b * vji / (rho_i * r2);
    rbi -= B1b * B1 / rho_i;
    const double r2b = B1b * alpha_i * (dvji * r2 - vji) / (rho_i * r2 * r2) +
                       B2b * alpha_j * (dvij * r2 - vij) / (rho_j * r2 * r2);
    atomicAdd(b.ab + j, B2b * vij / (rho_j * r2));
    atomicAdd(b.rb + j, -B2b * B2 / rho_j);
    gx += 2.0 * r2b * ex; gy += 2.0 * r2b * ey; gz += 2.0 * r2b * ez;
    xbx -= gx; xby -= gy; xbz -= gz;
    atomic_add3(b.xb + j, gx, gy, gz);
  }
  if (fric) atomic_add3(b.wb + i, wbx, wby, wbz);
  atomicAdd(b.ab + i, abi);
  atomicAdd(b.rb + i, rbi);
  atomic_add3(b.xb + i, xbx, xby, xbz);
}

// --------------------------------------------------- w sweep (stage 1)
// w_i = sum_j B_ij (e.vh_i - e.vh_j) e over the active list (no rho_j test)
__global__ __launch_bounds__(kBlock) void k_adj_w(AdjStepView s, AtomsView a, TablesView t,
                                                  AdjBuffers b) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.nlocal || !(a.mask[i] & a.groupbit)) return;
  const ForceRec ri = s.rec[i];
  const double rho_i = ri.rw.w;
  if (!(rho_i > 0.0)) return;
  const double4 W = b.wb[i];
  if (W.x == 0.0 && W.y == 0.0 && W.z == 0.0) return;
  const double4 pi = s.pv[i].p, vi = s.pv[i].v;
  const double alpha_i = ri.xa.w;
  const bool multi = t.n_elem > 1;
  double vbx = 0, vby = 0, vbz = 0, abi = 0, rbi = 0, xbx = 0, xby = 0, xbz = 0;
  const int na = s.nact[i];
  for (int k = 0; k < na; ++k) {
    const int j = s.act[k * a.stride + i];
    const double4 pj = s.pv[j].p, vj = s.pv[j].v;
    const double ex = pj.x - pi.x, ey = pj.y - pi.y, ez = pj.z - pi.z;
    const double r2 = ex * ex + ey * ey + ez * ez;
    const int ej = multi ? static_cast<int>(pj.w) : 0;
    double vr, dvr;
    rho_r2_vd(t, ej, r2, &vr, &dvr);
    const double B = alpha_i * vr / (rho_i * r2);
    const double c1 = ex * vi.x + ey * vi.y + ez * vi.z;
    const double c2 = ex * vj.x + ey * vj.y + ez * vj.z;
    const double sc = B * (c1 - c2);
    const double sb = W.x * ex + W.y * ey + W.z * ez;
    double gx = sc * W.x, gy = sc * W.y, gz = sc * W.z;
    const double Bb = sb * (c1 - c2);
    const double c1b = sb * B, c2b = -sb * B;
    vbx += c1b * ex; vby += c1b * ey; vbz += c1b * ez;
    atomic_add3(b.vhb + j, c2b * ex, c2b * ey, c2b * ez);
    gx += c1b * vi.x + c2b * vj.x;
    gy += c1b * vi.y + c2b * vj.y;
    gz += c1b * vi.z + c2b * vj.z;
    abi += Bb * vr / (rho_i * r2);
    rbi -= Bb * B / rho_i;
    const double r2b = Bb * alpha_i * (dvr * r2 - vr) / (rho_i * r2 * r2);
    gx += 2.0 * r2b * ex; gy += 2.0 * r2b * ey; gz += 2.0 * r2b * ez;
    xbx -= gx; xby -= gy; xbz -= gz;
    atomic_add3(b.xb + j, gx, gy, gz);
  }
  atomic_add3(b.vhb + i, vbx, vby, vbz);
  atomicAdd(b.ab + i, abi);
  atomicAdd(b.rb + i, rbi);
  atomic_add3(b.xb + i, xbx, xby, xbz);
}

// ------------------------------------------------------------ alpha(rho)
// alpha_i = a + rho (b + rho (c + rho d)) on the interval of rho_i: rho_b +=
// alpha_b alpha', table gradient Gb[e][k] += alpha_b (1, rho, rho^2, rho^3).
// The atoms of a wave mostly sit on one or two knots (rho is nearly uniform in
// a crystal), so the table gradient is summed over the wave per distinct knot
// first and added with one atomic per knot and wave; all lanes take part.
__global__ __launch_bounds__(kBlock) void k_adj_alpha(AdjStepView s, AtomsView a, TablesView t,
                                                      AdjBuffers b) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  bool valid = i < a.nlocal && (a.mask[i] & a.groupbit);
  double rho = 0.0, ab = 0.0;
  if (valid) {
    rho = s.rec[i].rw.w;
    ab = b.ab[i];
    valid = rho > 0.0 && !(rho > t.rho_cut) && ab != 0.0;
  }
  int key = -1;
  if (valid) {
    const int e = t.n_elem > 1 ? static_cast<int>(s.rec[i].p.w) : 0;
    int k = static_cast<int>(rho * t.inv_drho);
    k = k < 0 ? 0 : (k < t.n_beta - 1 ? k : t.n_beta - 1);
    key = e * t.n_beta + k;
    const double4 c = t.alpha[key];
    b.rb[i] += ab * (c.y + rho * (2.0 * c.z + 3.0 * rho * c.w));
  }
  const double g0 = valid ? ab : 0.0, g1 = valid ? ab * rho : 0.0;
  const double g2 = valid ? ab * rho * rho : 0.0, g3 = valid ? ab * rho * rho * rho : 0.0;
  const int lane = static_cast<int>(threadIdx.x) & (warpSize - 1);
  unsigned long long left = __ballot(valid);
  while (left) {
    const int lead = __ffsll(static_cast<long long>(left)) - 1;
    const int k0 = __shfl(key, lead);
    const bool mine = valid && key == k0;
    double s0 = mine ? g0 : 0.0, s1 = mine ? g1 : 0.0, s2 = mine ? g2 : 0.0, s3 = mine ? g3 : 0.0;
    for (int off = warpSize / 2; off > 0; off >>= 1) {
      s0 += __shfl_xor(s0, off);
      s1 += __shfl_xor(s1, off);
      s2 += __shfl_xor(s2, off);
      s3 += __shfl_xor(s3, off);
    }
    if (lane == lead) {
      double* G = b.Gb + 4 * static_cast<long>(k0);
      atomicAdd(G + 0, s0);
      atomicAdd(G + 1, s1);
      atomicAdd(G + 2, s2);
      atomicAdd(G + 3, s3);
    }
    left &= ~__ballot(mine);
  }
}

// ------------------------------------------------------- density sweep
// rho_i = sum_j rho_{t_j}(r^2) over the active list
__global__ __launch_bounds__(kBlock) void k_adj_rho(AdjStepView s, AtomsView a, TablesView t,
                                                    AdjBuffers b) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.nlocal || !(a.mask[i] & a.groupbit)) return;
  const double rb = b.rb[i];
  if (rb == 0.0) return;
  const double4 pi = s.rec[i].p;
  const bool multi = t.n_elem > 1;
  double xbx = 0, xby = 0, xbz = 0;
  const int na = s.nact[i];
  for (int k = 0; k < na; ++k) {
    const int j = s.act[k * a.stride + i];
    const double4 pj = s.rec[j].p;
    const double ex = pj.x - pi.x, ey = pj.y - pi.y, ez = pj.z - pi.z;
    const double r2 = ex * ex + ey * ey + ez * ez;
    const int ej = multi ? static_cast<int>(pj.w) : 0;
    double vr, dvr;
    rho_r2_vd(t, ej, r2, &vr, &dvr);
    const double r2b = rb * dvr;
    const double gx = 2.0 * r2b * ex, gy = 2.0 * r2b * ey, gz = 2.0 * r2b * ez;
    xbx -= gx; xby -= gy; xbz -= gz;
    atomic_add3(b.xb + j, gx, gy, gz);
  }
  atomic_add3(b.xb + i, xbx, xby, xbz);
}

// ------------------------------------------------------------ ghosts
// ghosts are copies of their owners (x + shift, v, w, alpha, rho): move each
// ghost's adjoint onto the owner and clear it
__global__ __launch_bounds__(kBlock) void k_adj_fold(AdjStepView s, AtomsView a, int what,
                                                     AdjBuffers b) {
  const int g = blockIdx.x * kBlock + threadIdx.x;
  if (g >= s.nghost) return;
  const int o = s.owner[g];
  const int h = a.nlocal + g;
  if (what & 1) {
    const double4 v = b.xb[h];
    atomic_add3(b.xb + o, v.x, v.y, v.z);
    b.xb[h] = make_double4(0, 0, 0, 0);
  }
  if (what & 2) {
    const double4 v = b.vhb[h];
    atomic_add3(b.vhb + o, v.x, v.y, v.z);
    b.vhb[h] = make_double4(0, 0, 0, 0);
  }
  if (what & 4) {
    const double4 v = b.wb[h];
    atomic_add3(b.wb + o, v.x, v.y, v.z);
    b.wb[h] = make_double4(0, 0, 0, 0);
  }
  if (what & 8) {
    atomicAdd(b.ab + o, b.ab[h]);
    atomicAdd(b.rb + o, b.rb[h]);
    b.ab[h] = 0.0;
    b.rb[h] = 0.0;
  }
}

// ---------------------------------------------------- initial integrate
// vh = v + (dtf / m) f_prev, x' = x + dtv vh  (group atoms)
__global__ __launch_bounds__(kBlock) void k_adj_initial(AtomsView a, StepParams p, int noint,
                                                        AdjBuffers b) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.nlocal) return;
  const double4 xb = b.xb[i];
  double4 vh = b.vhb[i];
  double4 ft = make_double4(0, 0, 0, 0);
  if (!noint && (a.mask[i] & a.groupbit)) {
    vh.x += p.dtv * xb.x; vh.y += p.dtv * xb.y; vh.z += p.dtv * xb.z;
    const double c = p.dtf / p.mass[a.type[i]];
    ft = make_double4(c * vh.x, c * vh.y, c * vh.z, 0.0);
  }
  // the previous step's v' and total force
  b.vnb[i] = make_double4(vh.x, vh.y, vh.z, 0.0);
  b.ftb[i] = ft;
  b.vhb[i] = make_double4(0, 0, 0, 0);
}

// ------------------------------------------------------------ wrappers
hipError_t launch_adj_fdm(const GridView& g, const double* Tin, const double* Tb_out, double* Tb_in,
                          double* dTb, double* kb, double inner_dt, hipStream_t s) {
  const int G = g.nx * g.ny * g.nz;
  if (kb) k_adj_fdm<true><<<nblk(G), kBlock, 0, s>>>(g, Tin, Tb_out, Tb_in, dTb, kb, inner_dt, G);
  else k_adj_fdm<false><<<nblk(G), kBlock, 0, s>>>(g, Tin, Tb_out, Tb_in, dTb, nullptr, inner_dt, G);
  return hipGetLastError();
}

hipError_t launch_adj_kappa_T(const GridView& g, const double* T0, const double* kb, double* Tb,
                              hipStream_t s) {
  const int G = g.nx * g.ny * g.nz;
  k_adj_kappa_T<<<nblk(G), kBlock, 0, s>>>(g, T0, kb, Tb, G);
  return hipGetLastError();
}

hipError_t launch_adj_step(const AdjStepView& sv, const AtomsView& a, const TablesView& t,
                           const GridView& g, const StepParams& p, const AdjStepFlags& f,
                           const double* dTb, const AdjBuffers& b, hipStream_t s) {
  const int nl = a.nlocal;
  if (nl == 0) return hipSuccess;
  const int nb = nblk(nl), ngb = nblk(sv.nghost > 0 ? sv.nghost : 1);
  if (f.fric || f.rnd)
    k_adj_deposit<<<nb, kBlock, 0, s>>>(sv, g, a, p.dt, f.fric, f.rnd, 1.0, dTb, b);
  k_adj_final<<<nb, kBlock, 0, s>>>(a, p, f.noint, b);
  k_adj_lj<<<nb, kBlock, 0, s>>>(sv, a, f.lj1, f.lj2, f.ljrc2, b);
  if (f.model4 && (f.fric || f.rnd))
    k_adj_force<<<nb, kBlock, 0, s>>>(sv, a, t, g, p, f.fric, f.rnd, f.addE, f.addR, b);
  k_adj_fold<<<ngb, kBlock, 0, s>>>(sv, a, 1 | 4 | 8, b);
  if (f.fric) k_adj_w<<<nb, kBlock, 0, s>>>(sv, a, t, b);
  k_adj_fold<<<ngb, kBlock, 0, s>>>(sv, a, 1 | 2 | 8, b);
  k_adj_alpha<<<nb, kBlock, 0, s>>>(sv, a, t, b);
  k_adj_rho<<<nb, kBlock, 0, s>>>(sv, a, t, b);
  k_adj_fold<<<ngb, kBlock, 0, s>>>(sv, a, 1, b);
  k_adj_initial<<<nb, kBlock, 0, s>>>(a, p, f.noint, b);
  return hipGetLastError();
}

}  // namespace eph
