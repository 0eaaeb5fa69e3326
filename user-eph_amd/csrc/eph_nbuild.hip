// eph_nbuild.hip -- device rebuild of ghosts and neighbour lists for the MD
// harness (the LAMMPS stand-in), bit-identical to the host build in eph_md.cpp:
//   * locals remapped into the box,
//   * periodic images made dimension by dimension (x, y, z), each pass
//     imaging every atom present before it, images in source order and
//     lower image first,
//   * atoms binned on bins of half the larger list cutoff, each bin holding
//     its atoms in ascending index,
//   * full lists found in stencil order (z outer, then y, then x, offsets
//     -2..2) and ascending index within a bin; the pair list keeps that
//     order, the fix list is then sorted by list_sort_key (distance, index).
// Several independent systems ("replicas") can be built at once: atom i of
// replica rep[i] only images and neighbours atoms of its own replica.
// Every kernel first reads the device flag `go` and returns when it is 0, so
// the whole sequence can be recorded once and replayed every step.
#include "eph_nbuild.hpp"

namespace eph {

namespace {

__device__ __forceinline__ int bin_coord(double x, double glo, double bs, int nb) {
  int b = static_cast<int>(floor((x - glo) / bs));
  return b < 0 ? 0 : (b > nb - 1 ? nb - 1 : b);
}

__device__ __forceinline__ double comp(const double4& p, int d) {
  return d == 0 ? p.x : (d == 1 ? p.y : p.z);
}

}  // namespace

// max squared displacement since the last build > lim -> go = 1
__global__ __launch_bounds__(kBlock) void k_nb_check(int nlocal, const double4* __restrict__ pos,
                                                     const double4* __restrict__ ref, double lim,
                                                     int* go) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= nlocal) return;
  const double4 p = pos[i], q = ref[i];
  const double dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
  if (dx * dx + dy * dy + dz * dz > lim) atomicOr(go, 1);
}

// Domain::remap of the locals
__global__ __launch_bounds__(kBlock) void k_nb_wrap(int nlocal, double4* pos, PosVel* pv,
                                                    const int* rep, NbBoxes b, const int* go) {
  if (!*go) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= nlocal) return;
  const int r = rep[i];
  double4 p = pos[i];
  double c[3] = {p.x, p.y, p.z};
  for (int d = 0; d < 3; ++d) {
    const double lo = b.lo[3 * r + d], hi = b.hi[3 * r + d], L = b.L[3 * r + d];
    if (c[d] < lo) c[d] += L;
    if (c[d] >= hi) c[d] -= L;
    if (c[d] < lo) c[d] = lo;
  }
  if (c[0] == p.x && c[1] == p.y && c[2] == p.z) return;
  p.x = c[0]; p.y = c[1]; p.z = c[2];
  pos[i] = p;
  if (pv) pv[i].p = p;
}

// images of source s in dimension d: count (0..2)
__global__ __launch_bounds__(kBlock) void k_nb_ghost_count(int d, int cap, int nlocal,
                                                           const int* nghost_in,
                                                           const double4* __restrict__ pos,
                                                           const int* __restrict__ rep, NbBoxes b,
                                                           int* cnt, const int* go) {
  if (!*go) return;
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= cap) return;
  const int nsrc = nlocal + *nghost_in;
  int c = 0;
  if (s < nsrc) {
    const int r = rep[s];
    const double x = comp(pos[s], d);
    const double lo = b.lo[3 * r + d] - b.cg, hi = b.hi[3 * r + d] + b.cg, L = b.L[3 * r + d];
    for (int sgn = -1; sgn <= 1; sgn += 2) {
      const double img = x + sgn * L;
      if (img >= lo && img < hi) ++c;
    }
  }
  cnt[s] = c;
}

// write the images of pass d after the ghosts already present
__global__ __launch_bounds__(kBlock) void k_nb_ghost_write(int d, int cap, int nlocal,
                                                           const int* nghost_in, int* nghost_out,
                                                           const int* __restrict__ cnt,
                                                           const int* __restrict__ start,
                                                           NbAtoms at, NbBoxes b, int ghost_cap,
                                                           int* overflow, const int* go) {
  if (!*go) return;
  const int s = blockIdx.x * kBlock + threadIdx.x;
  const int ng0 = *nghost_in;
  const int nsrc = nlocal + ng0;
  if (s >= nsrc || s >= cap) return;
  if (s == nsrc - 1) {
    const int total = ng0 + start[s] + cnt[s];
    *nghost_out = total <= ghost_cap ? total : ghost_cap;
    if (total > ghost_cap) atomicOr(overflow, 1);
  }
  if (cnt[s] == 0) return;
  const int r = at.rep[s];
  const double x = comp(at.pos[s], d);
  const double lo = b.lo[3 * r + d] - b.cg, hi = b.hi[3 * r + d] + b.cg, L = b.L[3 * r + d];
  const int o = s < nlocal ? s : at.owner[s - nlocal];
  const double4 sh0 = s < nlocal ? make_double4(0.0, 0.0, 0.0, 0.0) : at.shift[s - nlocal];
  int c = 0;
  for (int sgn = -1; sgn <= 1; sgn += 2) {
    const double img = x + sgn * L;
    if (!(img >= lo && img < hi)) continue;
    const int g = ng0 + start[s] + c;
    ++c;
    if (g >= ghost_cap) continue;
    double4 sh = sh0;
    if (d == 0) sh.x += sgn * L; else if (d == 1) sh.y += sgn * L; else sh.z += sgn * L;
    at.owner[g] = o;
    at.shift[g] = sh;
    const double4 po = at.pos[o];
    const int gi = nlocal + g;
    at.pos[gi] = make_double4(po.x + sh.x, po.y + sh.y, po.z + sh.z, po.w);
    at.vel[gi] = at.vel[o];
    at.type[gi] = at.type[o];
    at.mask[gi] = at.mask[o];
    at.tag[gi] = at.tag[o];
    at.rep[gi] = at.rep[o];
  }
}

__global__ __launch_bounds__(kBlock) void k_nb_bin(int cap, int nlocal, const int* nghost,
                                                    const double4* __restrict__ pos,
                                                    const int* __restrict__ rep, NbBoxes b,
                                                    int* bin_of, int* bcount, const int* go) {
  if (!*go) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= cap || i >= nlocal + *nghost) return;
  const int r = rep[i];
  const double4 p = pos[i];
  const int* nb = b.nb + 3 * r;
  const double* glo = b.glo + 3 * r;
  const int bx = bin_coord(p.x, glo[0], b.bs, nb[0]);
  const int by = bin_coord(p.y, glo[1], b.bs, nb[1]);
  const int bz = bin_coord(p.z, glo[2], b.bs, nb[2]);
  const int bin = b.bin_base[r] + bx + nb[0] * (by + nb[1] * bz);
  bin_of[i] = bin;
  atomicAdd(&bcount[bin], 1);
}

__global__ __launch_bounds__(kBlock) void k_nb_bin_scatter(int cap, int nlocal, const int* nghost,
                                                           const int* __restrict__ bin_of,
                                                           const int* __restrict__ bstart,
                                                           int* bfill, int* items, const int* go) {
  if (!*go) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= cap || i >= nlocal + *nghost) return;
  const int bin = bin_of[i];
  items[bstart[bin] + atomicAdd(&bfill[bin], 1)] = i;
}

// atoms of each bin in ascending index (insertion sort; bins hold few atoms)
__global__ __launch_bounds__(kBlock) void k_nb_bin_sort(int nbins, const int* __restrict__ bstart,
                                                        int* items, const int* go) {
  if (!*go) return;
  const int bin = blockIdx.x * kBlock + threadIdx.x;
  if (bin >= nbins) return;
  int* it = items + bstart[bin];
  const int n = bstart[bin + 1] - bstart[bin];
  for (int k = 1; k < n; ++k) {
    const int v = it[k];
    int m = k - 1;
    while (m >= 0 && it[m] > v) {
      it[m + 1] = it[m];
      --m;
    }
    it[m + 1] = v;
  }
}

// both full lists of every local atom, transposed ELL
__global__ __launch_bounds__(kBlock) void k_nb_lists(int nlocal, const double4* __restrict__ pos,
                                                     const int* __restrict__ rep, NbBoxes b,
                                                     const int* __restrict__ bstart,
                                                     const int* __restrict__ items, NbLists l,
                                                     int* overflow, const int* go) {
  if (!*go) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= nlocal) return;
  const int r = rep[i];
  const double4 pi = pos[i];
  const int* nb = b.nb + 3 * r;
  const double* glo = b.glo + 3 * r;
  const int bx = bin_coord(pi.x, glo[0], b.bs, nb[0]);
  const int by = bin_coord(pi.y, glo[1], b.bs, nb[1]);
  const int bz = bin_coord(pi.z, glo[2], b.bs, nb[2]);
  // Squared distance from the atom to each bin slab of the 5 x 5 x 5 stencil,
  // per dimension.  A bin whose box lies farther than the larger cutoff
  // holds no atom of either list and is skipped, which leaves the lists (and
  // their order) unchanged.  Edge bins also hold the atoms clamped into them,
  // so they are never skipped; the margin covers the rounding of the slab
  // bounds.
  double sx[5], sy[5], sz[5];
  {
    const double pc[3] = {pi.x, pi.y, pi.z};
    const int bc[3] = {bx, by, bz};
    double* sd[3] = {sx, sy, sz};
    for (int d = 0; d < 3; ++d) {
      for (int o = 0; o < 5; ++o) {
        const int c = bc[d] + o - 2;
        const double lo = glo[d] + c * b.bs, hi = lo + b.bs;
        const double g = (c <= 0 || c >= nb[d] - 1) ? 0.0 : (pc[d] < lo ? lo - pc[d] : (pc[d] > hi ? pc[d] - hi : 0.0));
        sd[d][o] = g * g;
      }
    }
  }
  const double reach2 = b.cg * b.cg * (1.0 + 1e-9);
  int nf = 0, nl = 0, nc = 0;
  for (int dz = -2; dz <= 2; ++dz) {
    const int z = bz + dz;
    if (z < 0 || z >= nb[2]) continue;
    for (int dy = -2; dy <= 2; ++dy) {
      const int y = by + dy;
      if (y < 0 || y >= nb[1]) continue;
      for (int dx = -2; dx <= 2; ++dx) {
        const int x = bx + dx;
        if (x < 0 || x >= nb[0]) continue;
        if (sx[dx + 2] + sy[dy + 2] + sz[dz + 2] > reach2) continue;
        const int bin = b.bin_base[r] + x + nb[0] * (y + nb[1] * z);
        for (int k = bstart[bin]; k < bstart[bin + 1]; ++k) {
          const int j = items[k];
          if (j == i) continue;
          const double4 pj = pos[j];
          const double ex = pj.x - pi.x, ey = pj.y - pi.y, ez = pj.z - pi.z;
          const double r2 = ex * ex + ey * ey + ez * ez;
          if (r2 < l.cf2) {
            if (nf < l.cap_f) l.nbr_f[nf * l.stride + i] = j;
            ++nf;
            if (l.cand && r2 < l.cand2) {
              if (nc < l.cap_f) l.cand[nc * l.stride + i] = j;
              ++nc;
            }
          }
          if (r2 < l.cl2) {
            if (nl < l.cap_l) l.nbr_l[nl * l.stride + i] = j;
            ++nl;
          }
        }
      }
    }
  }
  if (nf > l.cap_f || nl > l.cap_l) atomicOr(overflow, 2);
  l.nn_f[i] = nf < l.cap_f ? nf : l.cap_f;
  l.nn_l[i] = nl < l.cap_l ? nl : l.cap_l;
  if (l.cand) l.ncand[i] = nc < l.cap_f ? nc : l.cap_f;
}

// Each atom's fix list into (q(r^2), found position) order, and its
// candidates (the entries with r^2 < cand2, in list order) rewritten to
// match: one wave per atom, a bitonic sort of N = 64 E 32-bit keys q x 2^9 + t
// (element n = lane E + s; padding keys ~0 sort last; nn_f <= cap_f <= N).
// The found position t in a key gives the entry's index, read back before the
// sorted list is written over the found one.  A fixed grid of waves walks the
// atoms, so a step without a rebuild costs one flag read per block.
template <int E>
__device__ __forceinline__ void nb_sort_atom(const int i, const int lane, const double4* __restrict__ pos,
                                             const NbLists& l) {
  const int nf = l.nn_f[i];
  const double4 pi = pos[i];
  uint32_t key[E];
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const int n = lane * E + s;
    key[s] = ~0u;
    if (n < nf) {
      const double4 pj = pos[l.nbr_f[n * l.stride + i]];
      const double ex = pj.x - pi.x, ey = pj.y - pi.y, ez = pj.z - pi.z;
      key[s] = (list_sort_q(ex * ex + ey * ey + ez * ez, l.sort_scale) << 9) | static_cast<uint32_t>(n);
    }
  }
  constexpr int N = 64 * E;
#pragma unroll
  for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
    for (int d = k >> 1; d > 0; d >>= 1) {
      if (d < E) {  // both elements of a pair in this lane
#pragma unroll
        for (int s = 0; s < E; ++s) {
          const int ps = s ^ d;
          if (ps > s) {
            const bool up = ((lane * E + s) & k) == 0;
            const uint32_t lo = min(key[s], key[ps]), hi = max(key[s], key[ps]);
            key[s] = up ? lo : hi;
            key[ps] = up ? hi : lo;
          }
        }
      } else {  // the partner element sits in lane ^ (d / E), same slot
#pragma unroll
        for (int s = 0; s < E; ++s) {
          const int n = lane * E + s;
          const uint32_t mine = key[s];
          const uint32_t other = static_cast<uint32_t>(__shfl_xor(static_cast<int>(mine), d / E));
          const bool keep_min = ((n & d) == 0) == ((n & k) == 0);
          key[s] = keep_min ? min(mine, other) : max(mine, other);
        }
      }
    }
  }
  // the indices of the sorted entries, all read before any is written
  int jv[E];
#pragma unroll
  for (int s = 0; s < E; ++s)
    jv[s] = lane * E + s < nf ? l.nbr_f[static_cast<int>(key[s] & 511u) * l.stride + i] : 0;
  __builtin_amdgcn_wave_barrier();
  int mine = 0;  // candidates among this lane's entries
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const int n = lane * E + s;
    bool c = false;
    if (n < nf) {
      l.nbr_f[n * l.stride + i] = jv[s];
      if (l.cand) {
        const double4 pj = pos[jv[s]];
        const double ex = pj.x - pi.x, ey = pj.y - pi.y, ez = pj.z - pi.z;
        c = ex * ex + ey * ey + ez * ez < l.cand2;
      }
    }
    mine |= (c ? 1 : 0) << s;
  }
  if (!l.cand) return;
  // exclusive prefix of the candidate counts over the lanes
  const int cnt = __builtin_popcount(mine);
  int incl = cnt;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int v = __shfl_up(incl, o);
    if (lane >= o) incl += v;
  }
  int at = incl - cnt;
#pragma unroll
  for (int s = 0; s < E; ++s) {
    if ((mine >> s) & 1) {
      if (at < l.cap_f) l.cand[at * l.stride + i] = jv[s];
      ++at;
    }
  }
  if (lane == 63) l.ncand[i] = incl < l.cap_f ? incl : l.cap_f;
}

// The sort network of nb_sort_atom on one wave's E keys per lane
template <int E>
__device__ __forceinline__ void bitonic_wave(uint32_t (&key)[E], const int lane) {
  constexpr int N = 64 * E;
#pragma unroll
  for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
    for (int d = k >> 1; d > 0; d >>= 1) {
      if (d < E) {
#pragma unroll
        for (int s = 0; s < E; ++s) {
          const int ps = s ^ d;
          if (ps > s) {
            const bool up = ((lane * E + s) & k) == 0;
            const uint32_t lo = min(key[s], key[ps]), hi = max(key[s], key[ps]);
            key[s] = up ? lo : hi;
            key[ps] = up ? hi : lo;
          }
        }
      } else {
#pragma unroll
        for (int s = 0; s < E; ++s) {
          const int n = lane * E + s;
          const uint32_t mine = key[s];
          const uint32_t other = static_cast<uint32_t>(__shfl_xor(static_cast<int>(mine), d / E));
          const bool keep_min = ((n & d) == 0) == ((n & k) == 0);
          key[s] = keep_min ? min(mine, other) : max(mine, other);
        }
      }
    }
  }
}

// The same sort for lists of at most 256 entries, 64 atoms per block: the
// list rows of the block's atoms are read and written whole (coalesced)
// through an LDS tile, and each of the 4 waves sorts 16 of the atoms from it.
// Rows are padded by one word so that a row-wise and a column-wise pass both
// avoid bank conflicts.
template <int E>
__global__ __launch_bounds__(kBlock) void k_nb_sort_tile(int nlocal, const double4* __restrict__ pos,
                                                         NbLists l, const int* go) {
  constexpr int R = 64 * E + 1;
  __shared__ int s_j[64 * R];        // the block's lists: found order, then sorted
  __shared__ int s_c[64 * R];        // their candidates
  __shared__ unsigned char s_cf[64 * 64 * E];  // candidate flag per found position
  __shared__ int s_nf[64], s_nc[64];
  if (!*go) return;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int i0 = blockIdx.x * 64; i0 < nlocal; i0 += gridDim.x * 64) {
    if (threadIdx.x < 64) s_nf[lane] = i0 + lane < nlocal ? l.nn_f[i0 + lane] : 0;
    __syncthreads();
    int mx = s_nf[lane];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o));
    for (int k = wv; k < mx; k += 4)
      if (k < s_nf[lane]) s_j[lane * R + k] = l.nbr_f[k * l.stride + i0 + lane];
    __syncthreads();
    for (int a = wv; a < 64; a += 4) {
      const int nf = s_nf[a];  // the same for the whole wave
      if (nf == 0) {
        if (lane == 0) s_nc[a] = 0;
        continue;
      }
      const int i = i0 + a;
      const double4 pi = pos[i];
      uint32_t key[E];
#pragma unroll
      for (int s = 0; s < E; ++s) {
        const int n = lane * E + s;
        key[s] = ~0u;
        if (n < nf) {
          const double4 pj = pos[s_j[a * R + n]];
          const double ex = pj.x - pi.x, ey = pj.y - pi.y, ez = pj.z - pi.z;
          const double r2 = ex * ex + ey * ey + ez * ez;
          key[s] = (list_sort_q(r2, l.sort_scale) << 9) | static_cast<uint32_t>(n);
          s_cf[a * 64 * E + n] = r2 < l.cand2 ? 1 : 0;
        }
      }
      bitonic_wave<E>(key, lane);
      int jv[E];
      int mine = 0;
#pragma unroll
      for (int s = 0; s < E; ++s) {
        const int t = static_cast<int>(key[s] & 511u);
        const bool in = lane * E + s < nf;
        jv[s] = in ? s_j[a * R + t] : 0;
        mine |= (in && s_cf[a * 64 * E + t] ? 1 : 0) << s;
      }
      __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): every read of row a is back
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int s = 0; s < E; ++s)
        if (lane * E + s < nf) s_j[a * R + lane * E + s] = jv[s];
      const int cnt = __builtin_popcount(mine);
      int incl = cnt;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o);
        if (lane >= o) incl += v;
      }
      int at = incl - cnt;
#pragma unroll
      for (int s = 0; s < E; ++s)
        if ((mine >> s) & 1) s_c[a * R + at++] = jv[s];
      if (lane == 63) s_nc[a] = incl;
    }
    __syncthreads();
    for (int k = wv; k < mx; k += 4) {
      if (k < s_nf[lane]) l.nbr_f[k * l.stride + i0 + lane] = s_j[lane * R + k];
      if (l.cand && k < s_nc[lane]) l.cand[k * l.stride + i0 + lane] = s_c[lane * R + k];
    }
    if (l.cand && threadIdx.x < 64 && i0 + lane < nlocal) l.ncand[i0 + lane] = s_nc[lane];
    __syncthreads();
  }
}

constexpr int kSortBlocks = 2048;
template <int E>
__global__ __launch_bounds__(kBlock) void k_nb_sort(int nlocal, const double4* __restrict__ pos, NbLists l,
                                                    const int* go) {
  if (!*go) return;
  const int lane = threadIdx.x & 63;
  const int nw = gridDim.x * (kBlock / 64);
  for (int i = blockIdx.x * (kBlock / 64) + static_cast<int>(threadIdx.x >> 6); i < nlocal; i += nw)
    nb_sort_atom<E>(i, lane, pos, l);  // i is the same for the whole wave
}

__global__ __launch_bounds__(kBlock) void k_nb_finish(int nlocal, const int* ng,
                                                      const double4* __restrict__ pos,
                                                      double4* ref, int* go, int* nreb,
                                                      int* cand_valid, double4* cref,
                                                      bool cand_written) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (!*go) return;
  if (i < nlocal) ref[i] = pos[i];
  // the candidates of the new lists, if written, are valid at the positions
  // of this rebuild (locals and the ghosts just made); otherwise the next
  // density sweep retakes them
  if (cand_written && i < nlocal + *ng) cref[i] = pos[i];
  if (i == 0 && cand_valid) *cand_valid = cand_written ? 1 : 0;
  // the last block clears the flag only after every block has read it: the
  // flag is cleared by the next step's k_nb_reset instead
  if (i == 0) atomicAdd(nreb, 1);
}

__global__ void k_nb_reset(int* go) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *go = 0;
}

__global__ void k_nb_set(int* p, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

__global__ __launch_bounds__(kBlock) void k_nb_zero(int* p, int n, const int* go) {
  if (!*go) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ------------------------------------------------------------------ scan
constexpr int kSB = 1024, kSI = 4;
__global__ __launch_bounds__(kSB) void k_nb_scan_reduce(const int* in, int n, int* bsum, const int* go) {
  if (!*go) return;
  __shared__ int sh[kSB];
  const int base = blockIdx.x * kSB * kSI;
  int s = 0;
  for (int k = 0; k < kSI; ++k) {
    const int idx = base + k * kSB + threadIdx.x;
    if (idx < n) s += in[idx];
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int st = kSB / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(kSB) void k_nb_scan_blocks(int* bsum, int nb, const int* go) {
  if (!*go) return;
  __shared__ int sh[kSB];
  int carry = 0;
  for (int base = 0; base < nb; base += kSB) {
    const int idx = base + threadIdx.x;
    const int v = idx < nb ? bsum[idx] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kSB; off <<= 1) {
      const int add = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
      __syncthreads();
      sh[threadIdx.x] += add;
      __syncthreads();
    }
    if (idx < nb) bsum[idx] = carry + sh[threadIdx.x] - v;
    const int total = sh[kSB - 1];
    __syncthreads();
    carry += total;
  }
}

__global__ __launch_bounds__(kSB) void k_nb_scan_apply(const int* in, int n, const int* boff,
                                                       int* out, const int* go) {
  if (!*go) return;
  __shared__ int sh[kSB];
  const int base = blockIdx.x * kSB * kSI;
  int carry = boff[blockIdx.x];
  for (int k = 0; k < kSI; ++k) {
    const int idx = base + k * kSB + threadIdx.x;
    const int v = idx < n ? in[idx] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kSB; off <<= 1) {
      const int add = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
      __syncthreads();
      sh[threadIdx.x] += add;
      __syncthreads();
    }
    if (idx < n) out[idx] = carry + sh[threadIdx.x] - v;
    if (idx == n - 1) out[n] = carry + sh[threadIdx.x];
    const int total = sh[kSB - 1];
    __syncthreads();
    carry += total;
  }
}

static inline int nbk(long n, int b = kBlock) { return static_cast<int>((n + b - 1) / b); }

static hipError_t scan(const int* in, int n, int* out, int* tmp, const int* go, hipStream_t s) {
  const int per = kSB * kSI;
  const int nb = (n + per - 1) / per;
  k_nb_scan_reduce<<<nb, kSB, 0, s>>>(in, n, tmp, go);
  k_nb_scan_blocks<<<1, kSB, 0, s>>>(tmp, nb, go);
  k_nb_scan_apply<<<nb, kSB, 0, s>>>(in, n, tmp, out, go);
  return hipGetLastError();
}

hipError_t launch_nb_check(int nlocal, const double4* pos, const double4* ref, double lim, int* go,
                           hipStream_t s) {
  k_nb_reset<<<1, 64, 0, s>>>(go);
  if (nlocal > 0) k_nb_check<<<nbk(nlocal), kBlock, 0, s>>>(nlocal, pos, ref, lim, go);
  return hipGetLastError();
}

hipError_t launch_nb_force(int* go, hipStream_t s) {
  k_nb_set<<<1, 64, 0, s>>>(go, 1);
  return hipGetLastError();
}

// the rebuild proper, guarded by *go
hipError_t launch_nb_build(const NbBuild& w, hipStream_t s) {
  const int cap = w.nlocal + w.ghost_cap;
  const int* go = w.go;
  k_nb_wrap<<<nbk(w.nlocal), kBlock, 0, s>>>(w.nlocal, w.at.pos, w.at.pv, w.at.rep, w.box, go);
  // ghost passes: nghost[0] = 0, nghost[d + 1] = after pass d
  k_nb_zero<<<1, 64, 0, s>>>(w.nghost, 4, go);
  for (int d = 0; d < 3; ++d) {
    k_nb_ghost_count<<<nbk(cap), kBlock, 0, s>>>(d, cap, w.nlocal, w.nghost + d, w.at.pos, w.at.rep,
                                                 w.box, w.cnt, go);
    (void)scan(w.cnt, cap, w.start, w.scan_tmp, go, s);
    k_nb_ghost_write<<<nbk(cap), kBlock, 0, s>>>(d, cap, w.nlocal, w.nghost + d, w.nghost + d + 1,
                                                 w.cnt, w.start, w.at, w.box, w.ghost_cap,
                                                 w.overflow, go);
  }
  const int* ng = w.nghost + 3;
  k_nb_zero<<<nbk(w.nbins + 1), kBlock, 0, s>>>(w.bcount, w.nbins + 1, go);
  k_nb_zero<<<nbk(w.nbins + 1), kBlock, 0, s>>>(w.bfill, w.nbins + 1, go);
  k_nb_bin<<<nbk(cap), kBlock, 0, s>>>(cap, w.nlocal, ng, w.at.pos, w.at.rep, w.box, w.bin_of,
                                       w.bcount, go);
  (void)scan(w.bcount, w.nbins, w.bstart, w.scan_tmp, go, s);
  k_nb_bin_scatter<<<nbk(cap), kBlock, 0, s>>>(cap, w.nlocal, ng, w.bin_of, w.bstart, w.bfill,
                                               w.items, go);
  k_nb_bin_sort<<<nbk(w.nbins), kBlock, 0, s>>>(w.nbins, w.bstart, w.items, go);
  NbLists found = w.lists;
  found.cand = nullptr;  // the candidates are written in the sorted order by k_nb_sort
  k_nb_lists<<<nbk(w.nlocal), kBlock, 0, s>>>(w.nlocal, w.at.pos, w.at.rep, w.box, w.bstart,
                                              w.items, found, w.overflow, go);
  const int ek = w.lists.cap_f <= 128 ? 2 : (w.lists.cap_f <= 256 ? 4 : (w.lists.cap_f <= kListSortMax ? 8 : 0));
  if (ek == 0) return hipErrorInvalidValue;  // more fix-list entries per atom than the sort takes
  const int sb = std::min(kSortBlocks, nbk(w.nlocal, kBlock / 64));
  const int tb = std::min(kSortBlocks, nbk(w.nlocal, 64));
  if (ek == 2) k_nb_sort_tile<2><<<tb, kBlock, 0, s>>>(w.nlocal, w.at.pos, w.lists, go);
  else if (ek == 4) k_nb_sort_tile<4><<<tb, kBlock, 0, s>>>(w.nlocal, w.at.pos, w.lists, go);
  else k_nb_sort<8><<<sb, kBlock, 0, s>>>(w.nlocal, w.at.pos, w.lists, go);
  const bool cw = w.lists.cand && w.cref && w.cand_valid;
  k_nb_finish<<<nbk(cw ? cap : w.nlocal), kBlock, 0, s>>>(w.nlocal, ng, w.at.pos, w.xref, w.go,
                                                          w.nrebuild, w.cand_valid, w.cref, cw);
  return hipGetLastError();
}

}  // namespace eph
