// mock_mpi_driver.cpp -- runs the fix eph/amd adaptor on one system split over
// the ranks of mpiexec, in the LAMMPS-shaped test harness (TEST INFRASTRUCTURE
// ONLY).  Every rank uses GPU 0.
//
//   mpiexec -n R mock_mpi_driver <in_prefix> <out_prefix> <fix words...>
//
// <in_prefix>.<rank>.bin (native endianness): int32 nlocal, nghost, ntypes,
// nsteps, nnz, nranks; float64 dt, boltz, ftm2v, boxlo[3], boxhi[3],
// mass[ntypes + 1], x[nall][3], v[nall][3]; int32 type[nall], mask[nall],
// tag[nall], recv_first[nranks], recv_n[nranks], send_n[nranks], then the
// send lists of ranks 0..nranks-1; float64 shift[nghost][3]; int32
// off[nlocal + 1], idx[nnz].  The ghosts owned by rank q are the block
// [recv_first[q], recv_first[q] + recv_n[q]); rank q's send list for this rank
// names, in the block's order, its local atoms they are images of.
// Each step follows LAMMPS's Verlet loop for a fix without a pair style:
// initial_integrate, ghost update (x, v of the owners over MPI, x + shift),
// force clear, post_force, final_integrate, end_of_step.
// <out_prefix>.<rank>.bin: int32 tag[nlocal]; float64 x[nlocal][3],
// v[nlocal][3], f[nlocal][3], Ee, mean T_e, array[nlocal][ncols].
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

// the ghosts' x (owner + shift) and v (owner) from their owners' ranks
void update_ghosts(Atom& atom, Comm& comm, const std::vector<double>& shift) {
  const int nl = atom.nlocal;
  std::vector<std::vector<double>> out(comm.nprocs), in(comm.nprocs);
  std::vector<MPI_Request> req;
  for (int q = 0; q < comm.nprocs; ++q) {
    const auto& sl = comm.send_list[q];
    out[q].resize(6 * sl.size() + 1);
    for (size_t k = 0; k < sl.size(); ++k)
      for (int d = 0; d < 3; ++d) {
        out[q][6 * k + d] = atom.x[sl[k]][d];
        out[q][6 * k + 3 + d] = atom.v[sl[k]][d];
      }
    in[q].resize(6 * static_cast<size_t>(comm.recv_n[q]) + 1);
  }
  for (int q = 0; q < comm.nprocs; ++q) {
    if (q == comm.me) continue;
    req.emplace_back();
    MPI_Irecv(in[q].data(), 6 * comm.recv_n[q], MPI_DOUBLE, q, 3, comm.world, &req.back());
    req.emplace_back();
    MPI_Isend(out[q].data(), 6 * static_cast<int>(comm.send_list[q].size()), MPI_DOUBLE, q, 3,
              comm.world, &req.back());
  }
  if (!req.empty()) MPI_Waitall(static_cast<int>(req.size()), req.data(), MPI_STATUSES_IGNORE);
  in[comm.me] = out[comm.me];
  for (int q = 0; q < comm.nprocs; ++q)
    for (int k = 0; k < comm.recv_n[q]; ++k) {
      const int i = comm.recv_first[q] + k, g = i - nl;
      for (int d = 0; d < 3; ++d) {
        atom.x[i][d] = in[q][6 * k + d] + shift[3 * static_cast<size_t>(g) + d];
        atom.v[i][d] = in[q][6 * k + 3 + d];
      }
    }
}
}  // namespace

int main(int argc, char** argv) {
  MPI_Init(&argc, &argv);
  int rc = 0;
  int me = 0, nprocs = 1;
  MPI_Comm_rank(MPI_COMM_WORLD, &me);
  MPI_Comm_size(MPI_COMM_WORLD, &nprocs);
  try {
    if (argc < 4) throw std::runtime_error("usage: mock_mpi_driver <in_prefix> <out_prefix> <fix words...>");
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
    comm.me = me;
    comm.nprocs = nprocs;

    const std::string fin = std::string(argv[1]) + "." + std::to_string(me) + ".bin";
    FILE* in = std::fopen(fin.c_str(), "rb");
    if (!in) throw std::runtime_error("cannot open " + fin);
    int hdr[6];
    get(in, hdr, 6);
    const int nl = hdr[0], ng = hdr[1], nt = hdr[2], nsteps = hdr[3], nnz = hdr[4];
    if (hdr[5] != nprocs) throw std::runtime_error("input made for another rank count");
    const int nall = nl + ng;
    double c[3];
    get(in, c, 3);
    update.dt = c[0]; force.boltz = c[1]; force.ftm2v = c[2];
    get(in, domain.boxlo, 3);
    get(in, domain.boxhi, 3);
    std::vector<double> mass(nt + 1);
    get(in, mass.data(), mass.size());
    memory.grow(atom.x, nall + 1, 3, "x");
    memory.grow(atom.v, nall + 1, 3, "v");
    memory.grow(atom.f, nall + 1, 3, "f");
    get(in, &atom.x[0][0], 3 * static_cast<size_t>(nall));
    get(in, &atom.v[0][0], 3 * static_cast<size_t>(nall));
    std::vector<int> type(nall + 1), mask(nall + 1), tag(nall + 1), off(nl + 1), idx(nnz);
    get(in, type.data(), nall); get(in, mask.data(), nall); get(in, tag.data(), nall);
    comm.recv_first.resize(nprocs);
    comm.recv_n.resize(nprocs);
    std::vector<int> send_n(nprocs);
    get(in, comm.recv_first.data(), nprocs);
    get(in, comm.recv_n.data(), nprocs);
    get(in, send_n.data(), nprocs);
    comm.send_list.resize(nprocs);
    for (int q = 0; q < nprocs; ++q) {
      comm.send_list[q].resize(send_n[q]);
      get(in, comm.send_list[q].data(), send_n[q]);
    }
    std::vector<double> shift(3 * static_cast<size_t>(ng));
    get(in, shift.data(), shift.size());
    get(in, off.data(), off.size());
    get(in, idx.data(), idx.size());
    std::fclose(in);
    atom.type = type.data(); atom.mask = mask.data(); atom.tag = tag.data();
    atom.mass = mass.data();
    atom.nlocal = nl; atom.nghost = ng; atom.ntypes = nt; atom.nmax = nall + 1;
    atom.natoms = nl;
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
    neighbor.lastcall = 0;
    neighbor.ago = 0;

    std::vector<char*> words(argv + 3, argv + argc);
    FixEPHAMD fix(&lmp, static_cast<int>(words.size()), words.data());
    fix.init();
    fix.init_list(0, &list);
    for (int step = 1; step <= nsteps; ++step) {
      update.ntimestep = step;
      fix.initial_integrate(0);
      update_ghosts(atom, comm, shift);
      for (int i = 0; i < nall; ++i) atom.f[i][0] = atom.f[i][1] = atom.f[i][2] = 0.0;
      fix.post_force(0);
      fix.final_integrate();
      fix.end_of_step();
      ++neighbor.ago;
    }
    const double Ee = fix.compute_vector(0), Te = fix.compute_vector(1);
    const std::string fout = std::string(argv[2]) + "." + std::to_string(me) + ".bin";
    FILE* out = std::fopen(fout.c_str(), "wb");
    if (!out) throw std::runtime_error("cannot write " + fout);
    put(out, tag.data(), nl);
    put(out, &atom.x[0][0], 3 * static_cast<size_t>(nl));
    put(out, &atom.v[0][0], 3 * static_cast<size_t>(nl));
    put(out, &atom.f[0][0], 3 * static_cast<size_t>(nl));
    put(out, &Ee, 1);
    put(out, &Te, 1);
    const Fix& base = fix;
    put(out, &base.array_atom[0][0], static_cast<size_t>(nl) * base.size_peratom_cols);
    std::fclose(out);
    memory.destroy(atom.x);
    memory.destroy(atom.v);
    memory.destroy(atom.f);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "mock_mpi_driver (rank %d): %s\n", me, e.what());
    rc = 1;
  }
  if (rc) MPI_Abort(MPI_COMM_WORLD, rc);
  MPI_Finalize();
  return rc;
}
