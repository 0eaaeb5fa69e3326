// Test harness: ranks, the ghost-velocity flag and Comm::forward_comm(Fix *).
// The ghosts of this rank come in blocks, one per owning rank (in rank order,
// this rank's own periodic images included); recv_first[q] / recv_n[q] give the
// block of rank q and send_list[q] the local atoms whose values fill rank q's
// block of ghosts owned here.  forward_comm has the fix pack those values for
// every rank and unpack the blocks it receives, as LAMMPS's Comm does.
#pragma once
#include <vector>
#include "fix.h"
#include "pointers.h"
namespace LAMMPS_NS {
class Comm {
 public:
  int me = 0, nprocs = 1;
  int ghost_velocity = 0;
  std::vector<std::vector<int>> send_list;  // [rank] local indices packed for it
  std::vector<int> recv_first, recv_n;      // [rank] block of ghosts it owns
  MPI_Comm world = MPI_COMM_WORLD;

  void forward_comm(Fix *fix) {
    const int w = fix->comm_forward;
    std::vector<std::vector<double>> out(nprocs), in(nprocs);
    std::vector<MPI_Request> req;
    for (int q = 0; q < nprocs; ++q) {
      out[q].resize(static_cast<size_t>(w) * send_list[q].size() + 1);
      in[q].resize(static_cast<size_t>(w) * recv_n[q] + 1);
      fix->pack_forward_comm(static_cast<int>(send_list[q].size()), send_list[q].data(),
                             out[q].data(), 0, nullptr);
    }
    for (int q = 0; q < nprocs; ++q) {
      if (q == me) continue;
      req.emplace_back();
      MPI_Irecv(in[q].data(), w * recv_n[q], MPI_DOUBLE, q, 7, world, &req.back());
      req.emplace_back();
      MPI_Isend(out[q].data(), w * static_cast<int>(send_list[q].size()), MPI_DOUBLE, q, 7, world,
                &req.back());
    }
    if (!req.empty()) MPI_Waitall(static_cast<int>(req.size()), req.data(), MPI_STATUSES_IGNORE);
    in[me] = out[me];  // this rank's own images
    for (int q = 0; q < nprocs; ++q)
      if (recv_n[q] > 0) fix->unpack_forward_comm(recv_n[q], recv_first[q], in[q].data());
  }
};
}  // namespace LAMMPS_NS
