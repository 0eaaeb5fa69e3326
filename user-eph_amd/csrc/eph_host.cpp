// eph_host.cpp -- table construction and file formats (host, init time).
#include "eph_host.hpp"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>

namespace eph {

double Spline::operator()(double x) const {
  size_t i = static_cast<size_t>(x * inv_dx);
  if (i >= knots.size()) i = knots.size() - 1;
  const Cubic& k = knots[i];
  return k.a + x * (k.b + x * (k.c + x * k.d));
}

// Secants s_i = (y_{i+1} - y_i)/dx are extended by two linear extrapolations
// below (s_-1 = 2 s_0 - s_1, s_-2 = 2 s_-1 - s_0) and above; the reference's
// upper extension is computed while its top slot still holds 0, so the top
// secant pair is s_{n-1} = 2 s_{n-2}, s_n = 2 s_{n-1} (eph_spline.h:52-58).
// Knot slopes t_i blend s_{i-1}, s_i with Akima weights |s_{i+1} - s_i| and
// |s_{i-1} - s_{i-2}|, with the reference's tie-break order
// (eph_spline.h:81-108).  Each interval's cubic is then solved for in
// absolute x (eph_spline.h:111-126); the same floating-point expression
// order is kept so the coefficients are bit-identical.
bool build_spline(double dx, const std::vector<double>& y, Spline& out, std::string& err) {
  const size_t n = y.size();
  if (!(dx > 0) || n < 4) {
    err = "spline needs dx > 0 and at least 4 knots";
    return false;
  }
  // s[k + 2] holds secant k for k = -2 .. n
  std::vector<double> s(n + 3, 0.0);
  double* sec = s.data() + 2;
  for (size_t i = 0; i + 1 < n; ++i) sec[i] = (y[i + 1] - y[i]) / dx;
  sec[-1] = 2.0 * sec[0] - sec[1];
  sec[-2] = 2.0 * sec[-1] - sec[0];
  sec[n - 1] = 2.0 * sec[n - 2] - 0.0;
  sec[n] = 2.0 * sec[n - 1] - 0.0;

  std::vector<double> t(n);
  out.linear_tail = false;
  for (size_t i = 0; i < n; ++i) {
    const double d_2 = sec[(long)i - 2], d_1 = sec[(long)i - 1], d0 = sec[i];
    const double wr = std::fabs(sec[i + 1] - sec[i]);                 // weight of s_{i-1}
    const double wl = std::fabs(sec[(long)i - 1] - sec[(long)i - 2]);  // weight of s_i
    // At the last knot the reference compares against a value past its table;
    // it is taken to differ from d0 (only matters when d_2 == d_1).
    const bool top = (i == n - 1);
    const bool d0_eq_d1 = top ? false : (d0 == sec[i + 1]);
    if (top && d_2 == d_1) out.linear_tail = true;
    double ti;
    if (d_2 == d_1 && !d0_eq_d1) ti = d_1;
    else if (d0_eq_d1 && d_2 == d_1) ti = d0;
    else if (d_1 == d0) ti = d0;
    else ti = (d_1 * wr + d0 * wl) / (wr + wl);
    t[i] = ti;
  }

  out.inv_dx = 1. / dx;
  out.knots.assign(n, Cubic{0, 0, 0, 0});
  const double h3 = dx * dx * dx;
  for (size_t i = 0; i + 1 < n; ++i) {
    const double xl = i * dx, xl2 = i * dx * xl, xl3 = i * dx * xl2;
    const double xr = (i + 1) * dx, xr2 = (i + 1) * dx * xr, xr3 = (i + 1) * dx * xr2;
    const double tl = t[i], tr = t[i + 1];
    Cubic k;
    k.d = (-tl * xl - tr * xl + tl * xr + tr * xr + 2.0 * y[i] - 2.0 * y[i + 1]) / h3;
    k.c = (-tl + tr + 3.0 * k.d * xl2 - 3.0 * k.d * xr2) / 2.0 / dx;
    k.b = (k.c * xl2 + k.d * xl3 - k.c * xr2 - k.d * xr3 - y[i] + y[i + 1]) / dx;
    k.a = y[i] - k.b * xl - k.c * xl2 - k.d * xl3;
    out.knots[i] = k;
  }
  out.knots[n - 1] = Cubic{y[n - 1], 0.0, 0.0, 0.0};
  return true;
}

bool load_beta(const std::string& path, BetaModel& m, std::string& err) {
  std::ifstream in(path);
  if (!in.is_open()) {
    err = "cannot open .beta file " + path;
    return false;
  }
  std::string line;
  for (int k = 0; k < 3; ++k) std::getline(in, line);  // comment lines
  size_t ne = 0;
  in >> ne;
  if (!in || ne < 1) {
    err = "no elements found in .beta file " + path;
    return false;
  }
  std::getline(in, line);
  std::istringstream names(line);
  m.names.assign(ne, std::string());
  for (size_t e = 0; e < ne; ++e) names >> m.names[e];
  size_t nr = 0, nb = 0;
  double dr = 0, drho = 0;
  in >> nr >> dr >> nb >> drho >> m.r_cutoff;
  if (!in || nr < 4 || nb < 4) {
    err = "malformed spline header in .beta file " + path;
    return false;
  }
  m.n_rho = (int)nr;
  m.n_beta = (int)nb;
  m.r_cutoff_sq = m.r_cutoff * m.r_cutoff;
  m.rho_cutoff = drho * (nb - 1);
  const double dr_sq = m.r_cutoff_sq / (nr - 1);
  m.rho.resize(ne); m.rho_r_sq.resize(ne); m.beta.resize(ne); m.alpha.resize(ne);
  m.drho = drho;
  m.beta_y.assign(ne, std::vector<double>());
  m.numbers.assign(ne, 0);
  std::vector<double> knots_r(nr), knots_b(nb);
  for (size_t e = 0; e < ne; ++e) {
    unsigned short z = 0;
    in >> z;
    m.numbers[e] = z;
    for (size_t j = 0; j < nr; ++j) in >> knots_r[j];
    for (size_t j = 0; j < nb; ++j) in >> knots_b[j];
    if (!in) {
      err = "truncated .beta file " + path;
      return false;
    }
    if (!build_spline(dr, knots_r, m.rho[e], err)) return false;
    // rho(r) resampled on a uniform grid in r^2 (eph_beta.h:108-112)
    std::vector<double> knots_r2(nr);
    for (size_t j = 0; j < nr; ++j) knots_r2[j] = m.rho[e](std::sqrt(j * dr_sq));
    if (!build_spline(dr_sq, knots_r2, m.rho_r_sq[e], err)) return false;
    if (!build_spline(drho, knots_b, m.beta[e], err)) return false;
    m.beta_y[e] = knots_b;
    // alpha knots are sqrt(beta) knots (eph_beta.h:120-124)
    std::vector<double> knots_a(nb);
    for (size_t j = 0; j < nb; ++j) knots_a[j] = std::sqrt(knots_b[j]);
    if (!build_spline(drho, knots_a, m.alpha[e], err)) return false;
  }
  return true;
}

LinearTable make_linear(double dx, const std::vector<double>& y) {
  LinearTable L;
  L.dx = dx;
  L.y = y;
  for (size_t i = 0; i + 1 < y.size(); ++i) L.dy.push_back((y[i + 1] - y[i]) / dx);
  return L;
}

double linear_forward(const LinearTable& L, double x) {
  if (!(x >= 0.0)) return 0.0;
  const double q = x / L.dx;
  if (q >= static_cast<double>(L.y.size())) return 0.0;
  const size_t idx = static_cast<size_t>(q);
  const double delta = x - idx * L.dx;
  if (idx + 1 >= L.y.size()) return L.y[idx];
  return L.y[idx] + L.dy[idx] * delta;
}

bool load_kappa(const std::string& path, KappaModel& m, std::string& err) {
  std::ifstream in(path);
  if (!in.is_open()) {
    err = "cannot open .kappa file " + path;
    return false;
  }
  std::string line;
  for (int k = 0; k < 3; ++k) std::getline(in, line);  // comment lines
  size_t ne = 0;
  in >> ne;
  if (!in || ne < 1) {
    err = "no elements found in .kappa file " + path;
    return false;
  }
  std::getline(in, line);
  std::istringstream names(line);
  m.names.assign(ne, std::string());
  for (size_t e = 0; e < ne; ++e) names >> m.names[e];
  m.n_pairs = ne > 1 ? static_cast<int>((ne + 1) * (ne - 1) / 2) : 1;
  size_t nr = 0, nT = 0;
  double dr = 0;
  in >> nr >> dr >> m.r_cutoff >> nT >> m.dT >> m.T_max;
  if (!in || nr < 4 || nT < 2) {
    err = "malformed header in .kappa file " + path;
    return false;
  }
  m.n_r = (int)nr;
  m.n_T = (int)nT;
  m.r_cutoff_sq = m.r_cutoff * m.r_cutoff;
  const double dr_sq = m.r_cutoff_sq / static_cast<double>(nr - 1);
  m.numbers.assign(ne, 0);
  m.rho_r.resize(ne); m.rho_r_sq.resize(ne); m.E_T.resize(ne); m.K_T.resize(m.n_pairs);
  std::vector<double> r(nr), c(nT);
  for (size_t e = 0; e < ne; ++e) {
    int z = 0;
    in >> z;
    m.numbers[e] = z;
    for (size_t j = 0; j < nr; ++j) in >> r[j];
    if (!in) {
      err = "truncated .kappa file " + path;
      return false;
    }
    if (!build_spline(dr, r, m.rho_r[e], err)) return false;
    for (size_t j = 0; j < nr; ++j) r[j] = m.rho_r[e](std::sqrt(j * dr_sq));
    if (!build_spline(dr_sq, r, m.rho_r_sq[e], err)) return false;
    for (size_t j = 0; j < nT; ++j) in >> c[j];
    if (!in) {
      err = "truncated .kappa file " + path;
      return false;
    }
    // cumulative E(T) from C(T) (eph_kappa.h:132-137)
    c[0] = 0.;
    for (size_t j = 1; j < nT; ++j) c[j] = c[j - 1] + c[j] * m.dT;
    m.E_T[e] = make_linear(m.dT, c);
  }
  for (int k = 0; k < m.n_pairs; ++k) {
    for (size_t j = 0; j < nT; ++j) in >> c[j];
    if (!in) {
      err = "truncated .kappa file " + path;
      return false;
    }
    m.K_T[k] = make_linear(m.dT, c);
  }
  return true;
}

void GridHost::resize(size_t a, size_t b, size_t c) {
  nx = a; ny = b; nz = c;
  const size_t n = a * b * c;
  T.assign(n, 0.0); S.assign(n, 0.0); rho.assign(n, 0.0); C.assign(n, 0.0); kappa.assign(n, 0.0);
  flag.assign(n, CELL_DYNAMIC); tdyn.assign(n, 0);
}

void GridHost::set_box(double a0, double a1, double b0, double b1, double c0, double c1) {
  x0 = a0; x1 = a1; y0 = b0; y1 = b1; z0 = c0; z1 = c1;
  dx = (a1 - a0) / nx; dy = (b1 - b0) / ny; dz = (c1 - c0) / nz;
  dV = dx * dy * dz;
}

bool make_grid_const(size_t nx, size_t ny, size_t nz, const double box[6], double T, double C,
                     double rho, double kappa, GridHost& g, std::string& err) {
  if (nx < 1 || ny < 1 || nz < 1) {
    err = "non-positive grid values";
    return false;
  }
  if (!(box[0] < box[1] && box[2] < box[3] && box[4] < box[5])) {
    err = "grid box has zero or negative extent";
    return false;
  }
  g.resize(nx, ny, nz);
  g.set_box(box[0], box[1], box[2], box[3], box[4], box[5]);
  std::fill(g.T.begin(), g.T.end(), T);
  std::fill(g.rho.begin(), g.rho.end(), rho);
  std::fill(g.C.begin(), g.C.end(), C);
  std::fill(g.kappa.begin(), g.kappa.end(), kappa);
  g.steps = 1;
  g.dt = 1;
  g.param_file = "NULL";
  g.has_params = false;
  return true;
}

static bool read_comment(std::istream& in, std::string& err, const std::string& what) {
  std::string line;
  std::getline(in, line);
  if (!in || line.empty() || line[0] != '#') {
    err = what + ": expected a '#' comment line";
    return false;
  }
  return true;
}

bool load_grid(const std::string& path, GridHost& g, std::string& err) {
  std::ifstream in(path);
  if (!in) {
    err = "cannot open grid file " + path;
    return false;
  }
  for (int k = 0; k < 3; ++k)
    if (!read_comment(in, err, path)) return false;
  size_t a = 0, b = 0, c = 0;
  in >> a >> b >> c;
  if (!in || a < 1 || b < 1 || c < 1) {
    err = "bad grid size in " + path;
    return false;
  }
  g.resize(a, b, c);
  in >> g.steps;
  double bx[6];
  for (double& v : bx) in >> v;
  if (!in || !(bx[0] < bx[1] && bx[2] < bx[3] && bx[4] < bx[5])) {
    err = "bad grid box in " + path;
    return false;
  }
  g.set_box(bx[0], bx[1], bx[2], bx[3], bx[4], bx[5]);
  // The line after the box names the parameter file (or NULL).  Older grid
  // files (Examples/FDM/T.in) have no such line and nine columns per cell,
  // "i j k T S rho C kappa flag", with no T-dependence flag: that layout is
  // recognised by its first cell line and read with the flag at 0.
  std::string rest, next;
  std::getline(in, rest);  // the tokens may also follow the box on its line
  next = rest;
  while (in && next.find_first_not_of(" \t\r") == std::string::npos) std::getline(in, next);
  std::vector<std::string> tok;
  {
    std::istringstream ls(next);
    for (std::string t; ls >> t;) tok.push_back(t);
  }
  bool legacy = false;
  if (tok.size() == 9) {
    legacy = true;
    g.param_file = "NULL";
  } else if (tok.size() == 1) {
    g.param_file = tok[0];
  } else {
    err = "unrecognised line after the box in grid file " + path;
    return false;
  }
  std::istringstream first(legacy ? next : std::string());
  g.has_params = false;
  if (g.param_file != "NULL") {
    // the reference opens the name relative to the working directory; a name
    // relative to the grid file's directory is accepted as a fallback
    std::ifstream pin(g.param_file);
    if (!pin) {
      size_t slash = path.find_last_of('/');
      if (slash != std::string::npos) pin.open(path.substr(0, slash + 1) + g.param_file);
    }
    if (!pin) {
      err = "cannot open parameter file " + g.param_file;
      return false;
    }
    for (int k = 0; k < 3; ++k)
      if (!read_comment(pin, err, g.param_file)) return false;
    size_t n = 0;
    double dT = 0;
    pin >> n >> dT;
    if (!pin || n < 4 || !(dT > 0)) {
      err = "bad header in parameter file " + g.param_file;
      return false;
    }
    std::vector<double> Ce(n), ke(n);
    for (size_t i = 0; i < n; ++i) pin >> Ce[i] >> ke[i];
    if (!pin) {
      err = "truncated parameter file " + g.param_file;
      return false;
    }
    if (!build_spline(dT, Ce, g.C_T, err) || !build_spline(dT, ke, g.kappa_T, err)) return false;
    // E_e(T): running sum of C_e dT with E(0) = 0 (eph_fdm.h:98-103)
    std::vector<double> E(Ce);
    E[0] = 0.;
    for (size_t i = 1; i < E.size(); ++i) E[i] = E[i - 1] + E[i] * dT;
    g.E_T.dx = dT;
    g.E_T.y = E;
    g.E_T.dy.clear();
    for (size_t i = 0; i + 1 < E.size(); ++i) g.E_T.dy.push_back((E[i + 1] - E[i]) / dT);
    g.has_params = true;
  }
  const size_t nt = g.ntotal();
  for (size_t c2 = 0; c2 < nt; ++c2) {
    std::istream& src = (legacy && c2 == 0) ? static_cast<std::istream&>(first) : in;
    long lx, ly, lz;
    src >> lx >> ly >> lz;
    if (!src || lx < 0 || ly < 0 || lz < 0 || (size_t)lx >= g.nx || (size_t)ly >= g.ny ||
        (size_t)lz >= g.nz) {
      err = "bad cell line in grid file " + path;
      return false;
    }
    size_t idx = lx + ly * g.nx + lz * g.nx * g.ny;
    int fl, td = 0;
    src >> g.T[idx] >> g.S[idx] >> g.rho[idx] >> g.C[idx] >> g.kappa[idx] >> fl;
    if (!legacy) src >> td;
    if (!src) {
      err = "truncated grid file " + path;
      return false;
    }
    g.flag[idx] = (int8_t)fl;
    g.tdyn[idx] = (int8_t)td;
  }
  for (size_t c2 = 0; c2 < nt; ++c2)
    if (g.tdyn[c2] && !g.has_params) {
      err = "temperature-dependent cells need a parameter file";
      return false;
    }
  return true;
}

bool save_grid_state(const GridHost& g, const std::vector<double>& T, const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) return false;
  std::fprintf(f, "# A comment\n#\n#\n");
  std::fprintf(f, "%ld %ld %ld %ld\n", (long)g.nx, (long)g.ny, (long)g.nz, (long)g.steps);
  std::fprintf(f, "%.6e %.6e\n", g.x0, g.x1);
  std::fprintf(f, "%.6e %.6e\n", g.y0, g.y1);
  std::fprintf(f, "%.6e %.6e\n", g.z0, g.z1);
  std::fprintf(f, "%s\n", g.param_file.c_str());
  for (size_t k = 0; k < g.nz; ++k)
    for (size_t j = 0; j < g.ny; ++j)
      for (size_t i = 0; i < g.nx; ++i) {
        size_t c = i + j * g.nx + k * g.nx * g.ny;
        std::fprintf(f, "%d %d %d %.6e %.6e %.6e %.6e %.6e %d %d\n", (int)i, (int)j, (int)k, T[c],
                     g.S[c], g.rho[c], g.C[c], g.kappa[c], (int)g.flag[c], (int)g.tdyn[c]);
      }
  std::fclose(f);
  return true;
}

bool save_grid_temperature(const GridHost& g, const std::vector<double>& T, const std::string& base,
                           int n) {
  char name[4096];
  std::snprintf(name, sizeof name, "%s_%06d", base.c_str(), n);
  FILE* f = std::fopen(name, "w");
  if (!f) return false;
  std::fprintf(f, "x y z Te\n");
  for (size_t k = 0; k < g.nz; ++k)
    for (size_t j = 0; j < g.ny; ++j)
      for (size_t i = 0; i < g.nx; ++i) {
        size_t c = i + j * g.nx + k * g.nx * g.ny;
        std::fprintf(f, "%.6e %.6e %.6e %.6e\n", g.x0 + i * g.dx, g.y0 + j * g.dy, g.z0 + k * g.dz,
                     T[c]);
      }
  std::fclose(f);
  return true;
}

}  // namespace eph
