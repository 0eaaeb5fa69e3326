// Test harness: the LAMMPS object and the pointers every class sees.
#pragma once
#include <mpi.h>
#include "lmptype.h"
namespace LAMMPS_NS {
class Atom;
class Comm;
class Domain;
class Error;
class Force;
class Memory;
class Neighbor;
class Update;
class LAMMPS {
 public:
  MPI_Comm world = MPI_COMM_WORLD;
  Atom *atom = nullptr;
  Comm *comm = nullptr;
  Domain *domain = nullptr;
  Error *error = nullptr;
  Force *force = nullptr;
  Memory *memory = nullptr;
  Neighbor *neighbor = nullptr;
  Update *update = nullptr;
};
class Pointers {
 public:
  explicit Pointers(LAMMPS *ptr)
      : lmp(ptr), world(ptr->world), atom(ptr->atom), comm(ptr->comm), domain(ptr->domain),
        error(ptr->error), force(ptr->force), memory(ptr->memory), neighbor(ptr->neighbor),
        update(ptr->update) {}
  virtual ~Pointers() = default;

 protected:
  LAMMPS *lmp;
  MPI_Comm &world;
  Atom *atom;
  Comm *comm;
  Domain *domain;
  Error *error;
  Force *force;
  Memory *memory;
  Neighbor *neighbor;
  Update *update;
};
}  // namespace LAMMPS_NS
