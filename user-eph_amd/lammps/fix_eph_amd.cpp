/* fix_eph_amd.cpp -- LAMMPS adaptor over the C ABI of eph_amd.h (see
 * fix_eph_amd.h).  Each hook forwards to the context with LAMMPS's own arrays:
 * the context copies x, v, f in and out of HBM and runs the kernels; the
 * atoms, the ghost map and the neighbour list are re-sent after every
 * neighbour-list build, when LAMMPS may have exchanged, sorted or re-imaged
 * atoms.  Per-atom state of the stateful styles comes to LAMMPS's rows before
 * the exchange (pre_exchange), moves with the atoms (copy_arrays) and goes
 * back after the list build.  On one rank the per-step ghost exchange of
 * rho, w and xi (fix_eph.cpp:612, 744, 864, 871) happens inside the context,
 * which holds every ghost's owner.  Over several ranks (each with its own
 * context, ranks may share a GPU) post_force is split where FixEPH exchanges:
 * the local atoms' xi, rho and w go to the ghosts through comm->forward_comm
 * (one exchange serves the XI, RHO and WI states, which only force_prl's
 * stages 2 and 3 read), and end_of_step is split at EPH_FDM's sync_before:
 * the grid's deposits and the tally are summed with MPI_Allreduce
 * (eph_fdm.h:479-491, fix_eph.cpp:402), after which every rank solves the
 * same grid to the same bits (no broadcast needed). */
#include "fix_eph_amd.h"

#include <mpi.h>

#include <algorithm>
#include <cstring>
#include <string>

#include "atom.h"
#include "comm.h"
#include "domain.h"
#include "error.h"
#include "force.h"
#include "memory.h"
#include "neigh_list.h"
#include "neigh_request.h"
#include "neighbor.h"
#include "update.h"

extern "C" {
#include "eph_amd.h"
}

using namespace LAMMPS_NS;
using namespace FixConst;

static_assert(sizeof(tagint) == sizeof(int), "fix eph/amd passes atom tags as int (no -DLAMMPS_BIGBIG)");

namespace {
// style name of the context: the reference name behind "<name>/amd"
std::string reference_style(const char *s) {
  std::string n(s);
  const std::string tail = "/amd";
  if (n.size() > tail.size() && n.compare(n.size() - tail.size(), tail.size(), tail) == 0)
    n.erase(n.size() - tail.size());
  return n;
}
}  // namespace

FixEPHAMD::FixEPHAMD(LAMMPS *lmp, int narg, char **arg) :
    Fix(lmp, narg, arg), ctx(nullptr), list(nullptr), ncols(8), array(nullptr), list_step(-1),
    nstate(0), state(nullptr), state_on_host(false), multi(comm->nprocs > 1)
{
  if (narg < 3) error->all(FLERR, "Illegal fix eph/amd command");
  if (atom->map_style == Atom::MAP_NONE)
    error->all(FLERR, "fix eph/amd needs an atom map (atom_modify map array)");
  // the context parses the same words as fix eph, with the reference style
  const std::string style = reference_style(arg[2]);
  if (style == "eph/coloured/exp" || style == "eph/coloured/exp/v1")
    nstate = 6;
  else if (style == "eph/atomic")
    nstate = 1;
  else if (style != "eph" && style != "eph/coloured")
    error->all(FLERR, "fix eph/amd: unknown style " + style);
  std::vector<const char *> words(arg, arg + narg);
  words[2] = style.c_str();
  const int rc = eph_create(&ctx, narg, words.data(), atom->ntypes, domain->boxlo, domain->boxhi,
                            update->dt, force->boltz, force->ftm2v, /*device=*/0);
  if (rc != EPH_OK) error->all(FLERR, eph_last_error(nullptr));
  check(eph_set_groupbit(ctx, groupbit));
  check(eph_set_masses(ctx, atom->mass, atom->ntypes));
  if (multi) {
    int model = 0;
    check(eph_get_params(ctx, nullptr, &model, nullptr, nullptr, nullptr, nullptr, nullptr));
    if (model != 4 || (style != "eph" && style != "eph/coloured" && style != "eph/coloured/exp"))
      error->all(FLERR, "fix eph/amd over several MPI ranks: styles eph, eph/coloured and "
                        "eph/coloured/exp with model 4");
    comm_forward = EPH_HALO_COLS;  // xi, rho, w per atom
    int dims[3] = {0, 0, 0};
    check(eph_grid_info(eph_context_grid(ctx), dims, nullptr, nullptr));
    dT.resize(static_cast<size_t>(dims[0]) * dims[1] * dims[2]);
  }

  vector_flag = 1;  // Ee and the mean T_e (fix_eph.cpp:406-428)
  size_vector = 2;
  global_freq = 1;
  extvector = 1;
  ncols = eph_peratom_cols(ctx);  // 8, or 12 for eph/atomic
  peratom_flag = 1;
  size_peratom_cols = ncols;
  peratom_freq = 1;
  comm->ghost_velocity = 1;  // fix_eph.cpp:82

  grow_arrays(atom->nmax);
  atom->add_callback(Atom::GROW);
}

FixEPHAMD::~FixEPHAMD()
{
  atom->delete_callback(id, Atom::GROW);
  eph_destroy(ctx);
  memory->destroy(array);
  memory->destroy(state);
}

int FixEPHAMD::setmask()
{
  // eph_setmask(): all four hooks of fix_eph.cpp:293-302, and the state
  // hand-over before atoms move
  int mask = INITIAL_INTEGRATE | POST_FORCE | FINAL_INTEGRATE | END_OF_STEP;
  if (nstate > 0) mask |= PRE_EXCHANGE;
  return mask;
}

void FixEPHAMD::init()
{
  double cutoff = 0.0;
  const int rc = eph_init(ctx, domain->dimension, domain->nonperiodic, domain->triclinic, &cutoff);
  if (rc != EPH_OK) error->all(FLERR, eph_last_error(ctx));
  // a full list with ghosts at the rho(r) cutoff (fix_eph.cpp:273-275)
  auto req = neighbor->add_request(this, NeighConst::REQ_FULL | NeighConst::REQ_GHOST);
  req->set_cutoff(cutoff);
}

void FixEPHAMD::init_list(int, NeighList *ptr)
{
  list = ptr;
}

void FixEPHAMD::reset_dt()
{
  check(eph_reset_dt(ctx, update->dt));
}

void FixEPHAMD::initial_integrate(int)
{
  if (list_step < 0) upload_atoms();  // the first step runs before any post_force
  check(eph_initial_integrate(ctx, &atom->x[0][0], &atom->v[0][0], &atom->f[0][0]));
}

void FixEPHAMD::post_force(int)
{
  if (neighbor->lastcall != list_step) {  // a list build: atoms may have moved in memory
    upload_atoms();
    if (state_on_host) state_to_device();
    upload_list();
    list_step = neighbor->lastcall;
  }
  check(eph_set_timestep(ctx, update->ntimestep));
  if (!multi) {
    check(eph_post_force(ctx, &atom->x[0][0], &atom->v[0][0], &atom->f[0][0]));
    return;
  }
  // xi, rho, w of the locals, then to the ghosts of every rank, then the rest
  const int nlocal = atom->nlocal, nghost = atom->nghost;
  halo.resize(static_cast<size_t>(EPH_HALO_COLS) * std::max(nlocal, 1));
  halo_ghost.resize(static_cast<size_t>(EPH_HALO_COLS) * std::max(nghost, 1));
  check(eph_post_force_local(ctx, &atom->x[0][0], &atom->v[0][0], halo.data()));
  comm->forward_comm(this);
  check(eph_post_force_ghosts(ctx, halo_ghost.data(), &atom->f[0][0]));
}

int FixEPHAMD::pack_forward_comm(int n, int *list_, double *data, int, int *)
{
  int m = 0;
  for (int k = 0; k < n; ++k) {
    const double *row = &halo[static_cast<size_t>(EPH_HALO_COLS) * list_[k]];
    for (int c = 0; c < EPH_HALO_COLS; ++c) data[m++] = row[c];
  }
  return m;
}

void FixEPHAMD::unpack_forward_comm(int n, int first, double *data)
{
  const int nlocal = atom->nlocal;
  int m = 0;
  for (int i = first; i < first + n; ++i) {
    double *row = &halo_ghost[static_cast<size_t>(EPH_HALO_COLS) * (i - nlocal)];
    for (int c = 0; c < EPH_HALO_COLS; ++c) row[c] = data[m++];
  }
}

void FixEPHAMD::final_integrate()
{
  check(eph_final_integrate(ctx, &atom->v[0][0], &atom->f[0][0]));
}

void FixEPHAMD::end_of_step()
{
  if (!multi) {
    check(eph_end_of_step(ctx, &atom->x[0][0], &atom->v[0][0], update->ntimestep));
  } else {
    // this rank's deposits, summed over the ranks, then the same solve on all
    double E = 0.0;
    check(eph_end_of_step_local(ctx, &atom->x[0][0], &atom->v[0][0], dT.data(), &E));
    MPI_Allreduce(MPI_IN_PLACE, dT.data(), static_cast<int>(dT.size()), MPI_DOUBLE, MPI_SUM, world);
    MPI_Allreduce(MPI_IN_PLACE, &E, 1, MPI_DOUBLE, MPI_SUM, world);
    check(eph_end_of_step_reduced(ctx, dT.data(), E, update->ntimestep, comm->me == 0));
  }
  if (atom->nlocal > 0) check(eph_get_peratom(ctx, &array[0][0]));
}

double FixEPHAMD::compute_vector(int i)
{
  double v = 0.0;
  check(eph_compute_vector(ctx, i, &v));
  return v;
}

void FixEPHAMD::post_run()
{
  if (comm->me == 0) check(eph_post_run(ctx));  // fix_eph.cpp:1020, rank 0 writes
}

void FixEPHAMD::grow_arrays(int nmax)
{
  memory->grow(array, nmax, ncols, "eph/amd:array");
  array_atom = array;
  if (nstate > 0) memory->grow(state, nmax, nstate, "eph/amd:state");
}

// before atoms are exchanged or sorted: the context's per-atom state into the
// rows that travel with the atoms
void FixEPHAMD::pre_exchange()
{
  if (nstate > 0 && list_step >= 0) state_to_host();
}

void FixEPHAMD::copy_arrays(int i, int j, int)
{
  for (int k = 0; k < nstate; ++k) state[j][k] = state[i][k];
}

void FixEPHAMD::state_to_host()
{
  const int n = atom->nlocal;
  if (nstate == 6) {
    buf.resize(3 * static_cast<size_t>(n));
    for (int which = 0; which < 2; ++which) {
      if (n) check(eph_get_filter_state(ctx, which, buf.data()));
      for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) state[i][3 * which + d] = buf[3 * i + d];
    }
  } else if (n > 0) {
    buf.resize(n);
    check(eph_get_atomic_energy(ctx, buf.data()));
    for (int i = 0; i < n; ++i) state[i][0] = buf[i];
  }
  state_on_host = true;
}

void FixEPHAMD::state_to_device()
{
  const int n = atom->nlocal;
  if (nstate == 6) {
    buf.resize(3 * static_cast<size_t>(n));
    for (int which = 0; which < 2; ++which) {
      for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) buf[3 * i + d] = state[i][3 * which + d];
      if (n) check(eph_set_filter_state(ctx, which, buf.data()));
    }
  } else if (n > 0) {
    buf.resize(n);
    for (int i = 0; i < n; ++i) buf[i] = state[i][0];
    check(eph_set_atomic_energy(ctx, buf.data()));
  }
  state_on_host = false;
}

void FixEPHAMD::check(int rc)
{
  if (rc != EPH_OK) error->one(FLERR, eph_last_error(ctx));
}

// atoms and the ghost -> owner map; on one rank every ghost is a periodic
// image of a local atom, found through atom->map of its tag.  The hooks send
// the ghosts' own x and v every step, so no image offsets are needed.
void FixEPHAMD::upload_atoms()
{
  const int nlocal = atom->nlocal, nghost = atom->nghost;
  check(eph_set_atoms(ctx, nlocal, nghost, &atom->x[0][0], &atom->v[0][0], atom->type,
                      atom->mask, atom->tag));
  if (multi) return;  // ghost values arrive through forward_comm instead
  owner.resize(nghost);
  for (int g = 0; g < nghost; ++g) {
    const int o = atom->map(atom->tag[nlocal + g]);
    if (o < 0 || o >= nlocal) error->one(FLERR, "fix eph/amd: ghost atom without a local owner");
    owner[g] = o;
  }
  check(eph_set_ghost_map(ctx, nghost, owner.data(), nullptr));
}

// NeighList -> CSR (what transfer_neighbour_list of fix_eph_gpu.cpp:241-280 did)
void FixEPHAMD::upload_list()
{
  const int n = list->inum;
  off.assign(static_cast<size_t>(atom->nlocal) + 1, 0);
  for (int ii = 0; ii < n; ++ii) {
    const int i = list->ilist[ii];
    off[i + 1] = list->numneigh[i];
  }
  for (int i = 0; i < atom->nlocal; ++i) off[i + 1] += off[i];
  idx.resize(off.back());
  for (int ii = 0; ii < n; ++ii) {
    const int i = list->ilist[ii];
    std::copy(list->firstneigh[i], list->firstneigh[i] + list->numneigh[i], idx.begin() + off[i]);
  }
  check(eph_set_neighbors(ctx, atom->nlocal, off.data(), idx.data(), NEIGHMASK));
}
