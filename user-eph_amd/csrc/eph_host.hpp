// eph_host.hpp -- host-side model data of the fix: spline tables built from a
// .beta file, the piecewise-linear E_e(T) table, and the T_e grid description
// read from a grid file.  Tables are built once on the host (as the
// reference does at construction) and uploaded; every evaluation happens on
// the device.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace eph {

// Cubic in the ABSOLUTE coordinate: value(x) = a + x (b + x (c + x d)).
// The reference stores its Akima interpolant this way (eph_spline.h:111-141);
// the tables must be bit-identical for the kernels to reproduce it.
struct Cubic {
  double a, b, c, d;
};

struct Spline {
  double inv_dx = 0.0;
  std::vector<Cubic> knots;
  bool linear_tail = false;  // top-end tie-break depends on memory past the table
  bool empty() const { return knots.empty(); }
  double operator()(double x) const;  // host evaluation (table checks only)
};

// Akima-type interpolant on uniform knots y[i] at x = i dx, with the
// reference's extrapolated end slopes (eph_spline.h:31-132).
bool build_spline(double dx, const std::vector<double>& y, Spline& out, std::string& err);

// EPH_Linear (eph_linear.h:29-61): y[i] at x = i dx and forward slopes.
struct LinearTable {
  double dx = 1.0;
  std::vector<double> y, dy;
};

// .beta contents (eph_beta.h:39-128): per element rho(r), rho(r^2) resampled,
// beta(rho) and alpha(rho) = interpolant of sqrt(beta) knots.
struct BetaModel {
  std::vector<std::string> names;
  std::vector<int> numbers;
  double r_cutoff = 0, r_cutoff_sq = 0, rho_cutoff = 0;
  int n_rho = 0, n_beta = 0;
  std::vector<Spline> rho, rho_r_sq, beta, alpha;
  double drho = 0;                           // beta knot spacing
  std::vector<std::vector<double>> beta_y;   // beta knot values per element
};
bool load_beta(const std::string& path, BetaModel& m, std::string& err);

// EPH_Linear from uniform samples y at x = i dx: forward slopes (eph_linear.h:29-40)
LinearTable make_linear(double dx, const std::vector<double>& y);
// EPH_Linear::operator() (eph_linear.h:42-49); the top knot, whose slope the
// reference reads past its table, returns the knot value
double linear_forward(const LinearTable& L, double x);

// .kappa contents (eph_kappa.h:55-150) for fix eph/atomic: per element the
// locality rho(r) and its r^2 resampling, the cumulative E(T) table built
// from C(T); then K(T) tables, one per "pair" slot, n_pairs of them with the
// reference's count (n + 1)(n - 1) / 2 (1 for one element).
struct KappaModel {
  std::vector<std::string> names;
  std::vector<int> numbers;
  double r_cutoff = 0, r_cutoff_sq = 0, T_max = 0, dT = 0;
  int n_r = 0, n_T = 0, n_pairs = 0;
  std::vector<Spline> rho_r, rho_r_sq;
  std::vector<LinearTable> E_T, K_T;
};
bool load_kappa(const std::string& path, KappaModel& m, std::string& err);

// Cell flags of the T_e grid (eph_fdm.h:433-437).
enum CellFlag : int8_t { CELL_CONSTANT = 0, CELL_DYNAMIC = 1, CELL_ZERO_DERIVATIVE = 2 };

// Host description of a T_e grid (EPH_FDM state, eph_fdm.h:405-460).
struct GridHost {
  size_t nx = 1, ny = 1, nz = 1;
  double x0 = 0, x1 = 1, y0 = 0, y1 = 1, z0 = 0, z1 = 1;
  double dx = 1, dy = 1, dz = 1, dV = 1;
  size_t steps = 1;
  double dt = 1;
  std::string param_file = "NULL";
  std::vector<double> T, S, rho, C, kappa;
  std::vector<int8_t> flag, tdyn;
  bool has_params = false;
  Spline C_T, kappa_T;   // C_e(T), kappa(T) (eph_fdm.h:95-96)
  LinearTable E_T;       // E_e(T) cumulative table (eph_fdm.h:98-103)
  size_t ntotal() const { return nx * ny * nz; }
  void resize(size_t a, size_t b, size_t c);
  void set_box(double a0, double a1, double b0, double b1, double c0, double c1);
};
bool make_grid_const(size_t nx, size_t ny, size_t nz, const double box[6], double T, double C,
                     double rho, double kappa, GridHost& g, std::string& err);
bool load_grid(const std::string& path, GridHost& g, std::string& err);
bool save_grid_state(const GridHost& g, const std::vector<double>& T, const std::string& path);
bool save_grid_temperature(const GridHost& g, const std::vector<double>& T, const std::string& base,
                           int n);

}  // namespace eph
