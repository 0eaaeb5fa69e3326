// eph_devfn.hpp -- small device helpers shared by the kernel files.
#pragma once

#include "eph_device.hpp"

namespace eph {

// Blocks are dealt round-robin to the 8 XCDs (block b -> XCD b % 8).  Remap so
// that each XCD owns one contiguous range of atoms: neighbour gathers of a
// range then stay in that XCD's L2.
static __device__ __forceinline__ int xcd_block() {
  constexpr int kXcd = 8;
  const int nb = gridDim.x, b = blockIdx.x;
  const int per = nb / kXcd, rem = nb % kXcd;
  const int x = b % kXcd, local = b / kXcd;
  return x * per + (x < rem ? x : rem) + local;
}

static __device__ __forceinline__ double cubic_at(const double4* __restrict__ tab, int n, double inv,
                                           double x) {
  int i = static_cast<int>(x * inv);
  i = i < 0 ? 0 : (i < n - 1 ? i : n - 1);
  const double4 c = tab[i];
  return c.x + x * (c.y + x * (c.z + x * c.w));
}

// rho(r^2) of element e (EPH_Beta::get_rho_r_sq, eph_beta.h:164-170) from the
// interleaved tables of TablesView::rho_r2 (or their copy in LDS): the cubic of
// cubic_at, with the element's knot k at k * n_elem + e
static __device__ __forceinline__ double rho_r2_at(const double4* __restrict__ tab, const TablesView& t,
                                                   int e, double r2) {
  int i = static_cast<int>(r2 * t.inv_dr2);
  i = i < 0 ? 0 : (i < t.n_rho - 1 ? i : t.n_rho - 1);
  const double4 c = tab[i * t.n_elem + e];
  return c.x + r2 * (c.y + r2 * (c.z + r2 * c.w));
}

// EPH_FDM::get_index (eph_fdm.h:494-509): floor, then periodic fold
static __device__ __forceinline__ int cell_of(const GridView& g, double x, double y, double z) {
  int lx = static_cast<int>(floor((x - g.x0) / g.dx));
  int px = static_cast<int>(floor(static_cast<double>(lx) / g.nx));
  lx -= px * g.nx;
  int ly = static_cast<int>(floor((y - g.y0) / g.dy));
  int py = static_cast<int>(floor(static_cast<double>(ly) / g.ny));
  ly -= py * g.ny;
  int lz = static_cast<int>(floor((z - g.z0) / g.dz));
  int pz = static_cast<int>(floor(static_cast<double>(lz) / g.nz));
  lz -= pz * g.nz;
  const int c = lx + ly * g.nx + lz * g.nx * g.ny;
  // finite coordinates always land in range; non-finite ones (a run that has
  // blown up) must not index outside the grid
  return (c >= 0 && c < g.nx * g.ny * g.nz) ? c : 0;
}

template <typename T>
static __device__ __forceinline__ T block_sum(T v, T* sh) {
  // fixed-order tree: deterministic for a given block size
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] = sh[t] + sh[t + s];
    __syncthreads();
  }
  T r = sh[0];
  __syncthreads();
  return r;
}


// E_e(T) of a temperature-dependent cell: EPH_Linear::operator()
// (eph_linear.h:42-49); the top knot, whose slope the reference reads past
// its table, returns the knot value
static __device__ __forceinline__ double linear_fwd(const GridView& g, double x) {
  if (!(x >= 0.0)) return 0.0;
  const double q = x / g.E_dx;
  if (q >= static_cast<double>(g.n_E)) return 0.0;
  const int idx = static_cast<int>(q);
  const double delta = x - idx * g.E_dx;
  if (idx >= g.n_E - 1) return g.E_y[idx];
  return g.E_y[idx] + g.E_dy[idx] * delta;
}

// the slope of linear_fwd at x (0 where it is constant)
static __device__ __forceinline__ double linear_fwd_slope(const GridView& g, double x) {
  if (!(x >= 0.0)) return 0.0;
  const double q = x / g.E_dx;
  if (q >= static_cast<double>(g.n_E)) return 0.0;
  const int idx = static_cast<int>(q);
  return idx >= g.n_E - 1 ? 0.0 : g.E_dy[idx];
}

// the segment of E_e holding y (EPH_Linear::reverse_lookup's upper_bound,
// then one step back), or -1 where the lookup returns 0
static __device__ __forceinline__ int linear_rev_segment(const GridView& g, double y) {
  int lo = 0, hi = g.n_E;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (y < g.E_y[mid]) hi = mid; else lo = mid + 1;
  }
  return (lo >= g.n_E || lo == 0) ? -1 : lo - 1;
}

// T of energy y: EPH_Linear::reverse_lookup (eph_linear.h:52-61)
static __device__ __forceinline__ double linear_rev(const GridView& g, double y) {
  const int idx = linear_rev_segment(g, y);
  if (idx < 0) return 0.0;
  return idx * g.E_dx + 1. / g.E_dy[idx] * (y - g.E_y[idx]);
}

// the slope of linear_rev at y
static __device__ __forceinline__ double linear_rev_slope(const GridView& g, double y) {
  const int idx = linear_rev_segment(g, y);
  return idx < 0 ? 0.0 : 1. / g.E_dy[idx];
}

// d/dx of cubic_at (the same interval, absolute x)
static __device__ __forceinline__ double cubic_slope(const double4* __restrict__ tab, int n, double inv,
                                                     double x) {
  int i = static_cast<int>(x * inv);
  i = i < 0 ? 0 : (i < n - 1 ? i : n - 1);
  const double4 c = tab[i];
  return c.y + x * (2.0 * c.z + x * 3.0 * c.w);
}

}  // namespace eph
