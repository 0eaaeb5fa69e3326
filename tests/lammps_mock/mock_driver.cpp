// mock_driver.cpp -- runs the fix eph/amd adaptor inside the LAMMPS-shaped test
// harness (TEST INFRASTRUCTURE ONLY).
//
//   mock_driver <in.bin> <out.bin> <fix words...>
//
// in.bin (native endianness): int32 nlocal, nghost, ntypes, nsteps, nnz;
// float64 dt, boltz, ftm2v, boxlo[3], boxhi[3], mass[ntypes + 1], x[nall][3],
// v[nall][3]; int32 type[nall], mask[nall], tag[nall], owner[nghost];
// float64 shift[nghost][3]; int32 off[nlocal + 1], idx[nnz]; int32 sort_step
// and, when sort_step > 0, int32 perm[nlocal] (new slot s takes old atom
// perm[s]), owner[nghost], off[nlocal + 1], nnz, idx[nnz] after the sort.
// Each step follows LAMMPS's Verlet loop for a fix without a pair style:
// initial_integrate, [on sort_step: pre_exchange, local atoms re-ordered
// through a scratch slot with the fixes' copy_arrays, ghosts and list
// rebuilt], ghost update (x_owner + shift, v_owner), force clear, post_force,
// final_integrate, end_of_step.
// out.bin: float64 x[nlocal][3], v[nlocal][3], f[nlocal][3], Ee, mean T_e,
// array[nlocal][ncols].
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "atom.h"
#include "comm.h"
#include "domain.h"
#include "error.h"
#include "fix_eph_amd.h"
#include "force.h"
#include "memory.h"
#include "neigh_list.h"
#include "neighbor.h"
#include "update.h"

using namespace LAMMPS_NS;

namespace {
template <typename T>
void get(FILE* f, T* p, size_t n) {
  if (n && std::fread(p, sizeof(T), n, f) != n) throw std::runtime_error("short input file");
}
template <typename T>
void put(FILE* f, const T* p, size_t n) {
  if (n && std::fwrite(p, sizeof(T), n, f) != n) throw std::runtime_error("cannot write output");
}
}  // namespace

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: mock_driver <in.bin> <out.bin> <fix words...>\n");
    return 2;
  }
  try {
    LAMMPS lmp;
    Atom atom;
    Comm comm;
    Domain domain;
    Error error;
    Force force;
    Memory memory;
    Neighbor neighbor;
    Update update;
    lmp.atom = &atom; lmp.comm = &comm; lmp.domain = &domain; lmp.error = &error;
    lmp.force = &force; lmp.memory = &memory; lmp.neighbor = &neighbor; lmp.update = &update;

    FILE* in = std::fopen(argv[1], "rb");
    if (!in) throw std::runtime_error(std::string("cannot open ") + argv[1]);
    int hdr[5];
    get(in, hdr, 5);
    const int nl = hdr[0], ng = hdr[1], nt = hdr[2], nsteps = hdr[3], nnz = hdr[4];
    const int nall = nl + ng;
    double c[3];
    get(in, c, 3);
    update.dt = c[0]; force.boltz = c[1]; force.ftm2v = c[2];
    get(in, domain.boxlo, 3);
    get(in, domain.boxhi, 3);
    std::vector<double> mass(nt + 1);
    get(in, mass.data(), mass.size());
    memory.grow(atom.x, nall + 1, 3, "x");  // + the scratch slot of a sort
    memory.grow(atom.v, nall + 1, 3, "v");
    memory.grow(atom.f, nall + 1, 3, "f");
    get(in, &atom.x[0][0], 3 * static_cast<size_t>(nall));
    get(in, &atom.v[0][0], 3 * static_cast<size_t>(nall));
    std::vector<int> type(nall + 1), mask(nall + 1), tag(nall + 1), owner(ng), off(nl + 1), idx(nnz);
    std::vector<double> shift(3 * static_cast<size_t>(ng));
    get(in, type.data(), nall); get(in, mask.data(), nall); get(in, tag.data(), nall);
    get(in, owner.data(), ng); get(in, shift.data(), shift.size());
    get(in, off.data(), off.size()); get(in, idx.data(), idx.size());
    int sort_step = 0;
    get(in, &sort_step, 1);
    std::vector<int> perm, owner2, off2, idx2;
    if (sort_step > 0) {
      perm.resize(nl); owner2.resize(ng); off2.resize(nl + 1);
      get(in, perm.data(), nl); get(in, owner2.data(), ng); get(in, off2.data(), off2.size());
      int nnz2 = 0;
      get(in, &nnz2, 1);
      idx2.resize(nnz2);
      get(in, idx2.data(), idx2.size());
    }
    std::fclose(in);
    atom.type = type.data(); atom.mask = mask.data(); atom.tag = tag.data();
    atom.mass = mass.data();
    atom.nlocal = nl; atom.nghost = ng; atom.ntypes = nt; atom.nmax = nall + 1; atom.natoms = nl;
    for (int i = 0; i < nl; ++i) atom.tag_map[tag[i]] = i;

    NeighList list;
    std::vector<int> ilist(nl), numneigh(nl);
    std::vector<int*> firstneigh(nl);
    for (int i = 0; i < nl; ++i) {
      ilist[i] = i;
      numneigh[i] = off[i + 1] - off[i];
      firstneigh[i] = idx.data() + off[i];
    }
    list.inum = nl; list.ilist = ilist.data(); list.numneigh = numneigh.data();
    list.firstneigh = firstneigh.data();
    neighbor.lastcall = 0;  // the list was built at setup
    neighbor.ago = 0;

    std::vector<char*> words(argv + 3, argv + argc);
    FixEPHAMD fix(&lmp, static_cast<int>(words.size()), words.data());
    fix.init();
    if (neighbor.request.cutoff <= 0.0) throw std::runtime_error("no list cutoff requested");
    fix.init_list(0, &list);
    const bool wants_exchange = (fix.setmask() & FixConst::PRE_EXCHANGE) != 0;
    // one atom's row from slot a to slot b, with the fix's per-atom rows
    auto move = [&](int a, int b) {
      for (int d = 0; d < 3; ++d) {
        atom.x[b][d] = atom.x[a][d];
        atom.v[b][d] = atom.v[a][d];
        atom.f[b][d] = atom.f[a][d];
      }
      type[b] = type[a]; mask[b] = mask[a]; tag[b] = tag[a];
      fix.copy_arrays(a, b, 0);
    };
    for (int step = 1; step <= nsteps; ++step) {
      update.ntimestep = step;
      fix.initial_integrate(0);
      if (step == sort_step) {
        if (wants_exchange) fix.pre_exchange();
        // the locals take the order perm, cycle by cycle through slot nall
        std::vector<char> done(nl, 0);
        for (int s0 = 0; s0 < nl; ++s0) {
          if (done[s0] || perm[s0] == s0) { done[s0] = 1; continue; }
          move(s0, nall);
          int s = s0;
          for (;;) {
            done[s] = 1;
            const int src = perm[s];
            if (src == s0) { move(nall, s); break; }
            move(src, s);
            s = src;
          }
        }
        atom.tag_map.clear();
        for (int i = 0; i < nl; ++i) atom.tag_map[tag[i]] = i;
        owner = owner2;
        off = off2;
        idx = idx2;
        for (int i = 0; i < nl; ++i) {
          numneigh[i] = off[i + 1] - off[i];
          firstneigh[i] = idx.data() + off[i];
        }
        neighbor.lastcall = step;
        neighbor.ago = 0;
      }
      for (int g = 0; g < ng; ++g)
        for (int d = 0; d < 3; ++d) {
          atom.x[nl + g][d] = atom.x[owner[g]][d] + shift[3 * g + d];
          atom.v[nl + g][d] = atom.v[owner[g]][d];
        }
      for (int i = 0; i < nall; ++i) atom.f[i][0] = atom.f[i][1] = atom.f[i][2] = 0.0;
      fix.post_force(0);
      fix.final_integrate();
      fix.end_of_step();
      ++neighbor.ago;
    }
    const double Ee = fix.compute_vector(0), Te = fix.compute_vector(1);
    FILE* out = std::fopen(argv[2], "wb");
    if (!out) throw std::runtime_error(std::string("cannot write ") + argv[2]);
    put(out, &atom.x[0][0], 3 * static_cast<size_t>(nl));
    put(out, &atom.v[0][0], 3 * static_cast<size_t>(nl));
    put(out, &atom.f[0][0], 3 * static_cast<size_t>(nl));
    put(out, &Ee, 1);
    put(out, &Te, 1);
    // fix's per-atom array: size_peratom_cols columns (read through Fix's public field)
    const Fix& base = fix;
    put(out, &base.array_atom[0][0], static_cast<size_t>(nl) * base.size_peratom_cols);
    std::fclose(out);
    memory.destroy(atom.x);
    memory.destroy(atom.v);
    memory.destroy(atom.f);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "mock_driver: %s\n", e.what());
    return 1;
  }
  return 0;
}
