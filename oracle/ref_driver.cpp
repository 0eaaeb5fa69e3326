// ref_driver.cpp -- golden-vector writer built against the USER-EPH headers.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// the reference's own headers (eph_spline.h, eph_beta.h, eph_linear.h,
// eph_kappa.h, eph_fdm.h, read in place from the reference tree) into
// oracle/_ref/.
// tests/golden/make_golden.py runs it to write the fixtures under
// tests/golden/.  It prints exact values as C99 hex floats (%a).
//
// Commands:
//   spline <dx> <yfile>                       coefficients a b c d per knot
//   beta <file>                               all four tables of every element
//   beta_eval <file> <elem> <which> <xfile>   get_rho_r_sq / get_beta / get_alpha
//   linear <dx> <yfile> <xfile>               operator() and reverse_lookup
//   kappa <file>                              every table of an EPH_kappa
//   fdm <gridfile> <dt> <nsteps> <every> [<efile>]
//        runs nsteps of solve() (energy lines "x y z E" inserted before every
//        solve), printing every cell's T (cell-centre get_T) after each
//        multiple of <every>.
//   fdm_save <gridfile> <dt> <nsteps> <efile|NULL> <state> <heat> <n> <nmore>
//        the same nsteps, then save_state(<state>) and
//        save_temperature(<heat>, <n>) (the files fix eph writes in post_run
//        and every T_freq steps); then a new grid read from <state> runs
//        <nmore> further solves with the same energies and its T is printed
//        as above (the resume from T.restart).

#include <fstream>
#include <mpi.h>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "eph_spline.h"
#include "eph_beta.h"
#include "eph_linear.h"
#include "eph_kappa.h"
#include "eph_fdm.h"

namespace {

struct SplinePeek : Spline {
  explicit SplinePeek(const Spline& s) : Spline(s) {}
  void dump(FILE* o) const {
    std::fprintf(o, "%zu %a\n", c.size(), inv_dx);
    for (const auto& k : c) std::fprintf(o, "%a %a %a %a\n", k.a, k.b, k.c, k.d);
  }
};

struct BetaPeek : Beta {
  explicit BetaPeek(const char* f) : Beta(f) {}
  void dump(FILE* o) const {
    std::fprintf(o, "%zu %a %a %a\n", n_elements, r_cutoff, r_cutoff_sq, rho_cutoff);
    for (size_t e = 0; e < n_elements; ++e) {
      std::fprintf(o, "ELEMENT %zu %s %u\n", e, element_name[e].c_str(), (unsigned)element_number[e]);
      SplinePeek(rho[e]).dump(o);
      SplinePeek(rho_r_sq[e]).dump(o);
      SplinePeek(beta[e]).dump(o);
      SplinePeek(alpha[e]).dump(o);
    }
  }
};

void dump_linear(FILE* o, const EPH_Linear& L) {
  std::fprintf(o, "%zu %a\n", L.y.size(), L.dx);
  for (size_t i = 0; i < L.y.size(); ++i)
    std::fprintf(o, "%a %a\n", L.y[i], i < L.dy.size() ? L.dy[i] : 0.0);
}

std::vector<double> read_values(const char* path) {
  std::vector<double> v;
  std::ifstream f(path);
  double x;
  while (f >> x) v.push_back(x);
  return v;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) { std::fprintf(stderr, "usage: see header\n"); return 2; }
  std::string cmd = argv[1];
  if (cmd == "spline" && argc == 4) {
    double dx = std::strtod(argv[2], nullptr);
    Spline s(dx, read_values(argv[3]));
    SplinePeek(s).dump(stdout);
    return 0;
  }
  if (cmd == "beta" && argc == 3) {
    BetaPeek(argv[2]).dump(stdout);
    return 0;
  }
  if (cmd == "beta_eval" && argc == 6) {
    Beta b(argv[2]);
    size_t e = std::strtoul(argv[3], nullptr, 10);
    int which = std::atoi(argv[4]);
    for (double x : read_values(argv[5])) {
      double y = which == 1 ? b.get_rho_r_sq(e, x) : which == 2 ? b.get_beta(e, x) : b.get_alpha(e, x);
      std::printf("%a\n", y);
    }
    return 0;
  }
  if (cmd == "kappa" && argc == 3) {
    EPH_kappa<> k(argv[2]);
    std::printf("%zu %zu %a %a %a\n", k.n_elements, k.n_pairs, k.r_cutoff, k.r_cutoff_sq, k.T_max);
    for (size_t e = 0; e < k.n_elements; ++e) {
      std::printf("ELEMENT %zu %s %d\n", e, k.element_name[e].c_str(), k.element_number[e]);
      SplinePeek(k.rho_r[e]).dump(stdout);
      SplinePeek(k.rho_r_sq[e]).dump(stdout);
      dump_linear(stdout, k.E_T_atomic[e]);
    }
    for (size_t p = 0; p < k.n_pairs; ++p) {
      std::printf("PAIR %zu\n", p);
      dump_linear(stdout, k.K_T_atomic[p]);
    }
    return 0;
  }
  if (cmd == "linear" && argc == 5) {
    double dx = std::strtod(argv[2], nullptr);
    std::vector<double> y = read_values(argv[3]);
    EPH_Linear L(dx, y.begin(), y.end());
    for (double x : read_values(argv[4])) std::printf("%a %a\n", L(x), L.reverse_lookup(x));
    return 0;
  }
  if (cmd == "fdm" && (argc == 6 || argc == 7)) {
    MPI_Init(&argc, &argv);
    {
      EPH_FDM g(argv[2]);
      g.set_comm(MPI_COMM_WORLD, 0, 1);
      g.set_dt(std::strtod(argv[3], nullptr));
      long nsteps = std::strtol(argv[4], nullptr, 10);
      long every = std::strtol(argv[5], nullptr, 10);
      std::vector<double> ev;
      if (argc == 7) ev = read_values(argv[6]);
      // grid geometry comes from the same file, re-read here for cell centres
      std::ifstream f(argv[2]);
      std::string line;
      for (int k = 0; k < 3; ++k) std::getline(f, line);
      size_t nx, ny, nz, st; double x0, x1, y0, y1, z0, z1;
      f >> nx >> ny >> nz >> st >> x0 >> x1 >> y0 >> y1 >> z0 >> z1;
      double dx = (x1 - x0) / nx, dy = (y1 - y0) / ny, dz = (z1 - z0) / nz;
      for (long s = 1; s <= nsteps; ++s) {
        for (size_t k = 0; k + 3 < ev.size(); k += 4) g.insert_energy(ev[k], ev[k + 1], ev[k + 2], ev[k + 3]);
        g.solve();
        if (s % every == 0) {
          std::printf("STEP %ld\n", s);
          for (size_t k = 0; k < nz; ++k)
            for (size_t j = 0; j < ny; ++j)
              for (size_t i = 0; i < nx; ++i)
                std::printf("%a\n", g.get_T(x0 + (i + 0.5) * dx, y0 + (j + 0.5) * dy, z0 + (k + 0.5) * dz));
        }
      }
    }
    MPI_Finalize();
    return 0;
  }
  if (cmd == "fdm_save" && argc == 10) {
    MPI_Init(&argc, &argv);
    {
      const double dt = std::strtod(argv[3], nullptr);
      const long nsteps = std::strtol(argv[4], nullptr, 10);
      std::vector<double> ev;
      if (std::string(argv[5]) != "NULL") ev = read_values(argv[5]);
      const long nmore = std::strtol(argv[9], nullptr, 10);
      {
        EPH_FDM g(argv[2]);
        g.set_comm(MPI_COMM_WORLD, 0, 1);
        g.set_dt(dt);
        for (long s = 1; s <= nsteps; ++s) {
          for (size_t k = 0; k + 3 < ev.size(); k += 4) g.insert_energy(ev[k], ev[k + 1], ev[k + 2], ev[k + 3]);
          g.solve();
        }
        g.save_state(argv[6]);
        g.save_temperature(argv[7], std::atoi(argv[8]));
      }
      EPH_FDM r(argv[6]);
      r.set_comm(MPI_COMM_WORLD, 0, 1);
      r.set_dt(dt);
      std::ifstream f(argv[6]);
      std::string line;
      for (int k = 0; k < 3; ++k) std::getline(f, line);
      size_t nx, ny, nz, st; double x0, x1, y0, y1, z0, z1;
      f >> nx >> ny >> nz >> st >> x0 >> x1 >> y0 >> y1 >> z0 >> z1;
      const double dx = (x1 - x0) / nx, dy = (y1 - y0) / ny, dz = (z1 - z0) / nz;
      for (long s = 1; s <= nmore; ++s) {
        for (size_t k = 0; k + 3 < ev.size(); k += 4) r.insert_energy(ev[k], ev[k + 1], ev[k + 2], ev[k + 3]);
        r.solve();
      }
      std::printf("STEP %ld\n", nmore);
      for (size_t k = 0; k < nz; ++k)
        for (size_t j = 0; j < ny; ++j)
          for (size_t i = 0; i < nx; ++i)
            std::printf("%a\n", r.get_T(x0 + (i + 0.5) * dx, y0 + (j + 0.5) * dy, z0 + (k + 0.5) * dz));
    }
    MPI_Finalize();
    return 0;
  }
  std::fprintf(stderr, "bad command\n");
  return 2;
}
