// san_driver.cpp -- host-code run under AddressSanitizer and
// UndefinedBehaviorSanitizer (TEST INFRASTRUCTURE ONLY).
//
// tests/test_oracle.py builds this file with g++ -fsanitize=address,undefined
// together with the oracle (eph_oracle.cpp, one translation unit) and the
// product's host-side readers (user-eph_amd/csrc/eph_host.cpp), then runs it
// on files the test writes:
//
//   san_driver <beta> <kappa> <grid>
//
// It drives every oracle path (fix eph models 1, 2, 4, the coloured styles,
// eph/atomic, the FDM solve and the table readers) on a small periodic
// cluster, and the product's .beta / .kappa / grid readers, so that an
// out-of-range read or undefined operation in either aborts the run.
#include "eph_oracle.cpp"
#include "../user-eph_amd/csrc/eph_host.hpp"

#include <cstdio>

namespace {

struct Cluster {
  int nlocal = 0, nghost = 0;
  std::vector<double> x, v, f;
  std::vector<int> type, mask, tag, owner, off, nbr;
};

// 4 x 4 x 4 cells of a jittered simple cubic lattice, periodic images as
// ghosts, brute-force full lists at 6.5 A
Cluster make_cluster(int ntypes, unsigned seed) {
  Cluster c;
  const int m = 4;
  const double a = 2.6, L = m * a, cut = 6.5;
  std::srand(seed);
  auto rnd = [] { return std::rand() / (double)RAND_MAX - 0.5; };
  for (int k = 0; k < m; ++k)
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) {
        const double p[3] = {(i + 0.5) * a + 0.3 * rnd(), (j + 0.5) * a + 0.3 * rnd(),
                             (k + 0.5) * a + 0.3 * rnd()};
        for (int d = 0; d < 3; ++d) { c.x.push_back(p[d]); c.v.push_back(4.0 * rnd()); }
        c.type.push_back(1 + (int)c.type.size() % ntypes);
      }
  c.nlocal = m * m * m;
  for (int sx = -1; sx <= 1; ++sx)
    for (int sy = -1; sy <= 1; ++sy)
      for (int sz = -1; sz <= 1; ++sz) {
        if (!sx && !sy && !sz) continue;
        for (int i = 0; i < c.nlocal; ++i) {
          const double p[3] = {c.x[3 * i] + sx * L, c.x[3 * i + 1] + sy * L, c.x[3 * i + 2] + sz * L};
          bool in = true;
          for (int d = 0; d < 3; ++d) in = in && p[d] >= -cut && p[d] < L + cut;
          if (!in) continue;
          for (int d = 0; d < 3; ++d) { c.x.push_back(p[d]); c.v.push_back(c.v[3 * i + d]); }
          c.type.push_back(c.type[i]);
          c.owner.push_back(i);
          ++c.nghost;
        }
      }
  const int nall = c.nlocal + c.nghost;
  c.f.assign(3 * (size_t)nall, 0.0);
  c.mask.assign(nall, 1);
  for (int i = 0; i < nall; ++i) c.tag.push_back(i < c.nlocal ? i + 1 : c.tag[c.owner[i - c.nlocal]]);
  c.off.push_back(0);
  for (int i = 0; i < c.nlocal; ++i) {
    for (int j = 0; j < nall; ++j) {
      if (j == i) continue;
      double r2 = 0;
      for (int d = 0; d < 3; ++d) { const double e = c.x[3 * j + d] - c.x[3 * i + d]; r2 += e * e; }
      if (r2 < cut * cut) c.nbr.push_back(j);
    }
    c.off.push_back((int)c.nbr.size());
  }
  return c;
}

int fail(const char* what) {
  std::fprintf(stderr, "san_driver: %s\n", what);
  return 1;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 4) return fail("usage: san_driver <beta> <kappa> <grid>");
  const char* beta = argv[1];
  const char* kappa = argv[2];
  const char* grid = argv[3];
  // the product's host readers
  {
    std::string err;
    eph::BetaModel bm;
    if (!eph::load_beta(beta, bm, err)) return fail(err.c_str());
    eph::KappaModel km;
    if (!eph::load_kappa(kappa, km, err)) return fail(err.c_str());
    eph::GridHost gh;
    if (!eph::load_grid(grid, gh, err)) return fail(err.c_str());
    for (const auto& s : bm.alpha) for (double x = 0; x < 25.0; x += 0.37) (void)s(x);
    for (const auto& L : km.E_T) for (double x = 0; x < 1e4; x += 77.0) (void)eph::linear_forward(L, x);
  }
  // oracle: fix eph styles and models on the cluster, three steps each
  const double lo[3] = {0, 0, 0}, hi[3] = {10.4, 10.4, 10.4};
  struct Case { const char* style; const char* slot; };
  const Case cases[] = {{"eph", "4"}, {"eph", "1"}, {"eph", "2"}, {"eph", "-1"}, {"eph/coloured", "4"},
                        {"eph/coloured/exp", "0.002"}, {"eph/coloured/exp/v1", "0.002"}};
  for (const Case& cs : cases) {
    const char* args[] = {"e", "all", cs.style, "77", "7", cs.slot, "1.0", "3e-06", "0.1", "300.0",
                          "3", "3", "3", "NULL", "0", "Te_out", beta, "A", "B"};
    OFix fix;
    if (!fix.create(19, args, 2, lo, hi, 5e-4, 8.617343e-5, 1.0 / 1.0364269e-4)) return fail(fix.err.c_str());
    fix.mass.assign(3, 63.546);
    Cluster c = make_cluster(2, 11);
    for (int n = 1; n <= 3; ++n) {
      fix.initial_integrate(c.nlocal, c.x.data(), c.v.data(), c.f.data(), c.type.data(), c.mask.data());
      for (int g = 0; g < c.nghost; ++g)
        for (int d = 0; d < 3; ++d) c.v[3 * (c.nlocal + g) + d] = c.v[3 * c.owner[g] + d];
      if (!fix.post_force(c.nlocal, c.nghost, c.x.data(), c.v.data(), c.f.data(), c.type.data(),
                          c.mask.data(), c.off.data(), c.nbr.data(), c.owner.data(), nullptr,
                          c.tag.data(), n))
        return fail(fix.err.c_str());
      fix.final_integrate(c.nlocal, c.v.data(), c.f.data(), c.type.data(), c.mask.data());
      fix.end_of_step(c.nlocal, c.x.data(), c.v.data(), c.type.data(), c.mask.data(), n);
    }
  }
  // oracle: eph/atomic with heat diffusion, three inner loops
  {
    const char* args[] = {"e", "all", "eph/atomic", "77", "7", "400.0", "NULL", "3", "NULL", beta,
                          kappa, "A", "B"};
    OFixAtomic fix;
    if (!fix.create(13, args, 2, 5e-4, 8.617343e-5, 1.0 / 1.0364269e-4)) return fail(fix.err.c_str());
    fix.mass.assign(3, 63.546);
    Cluster c = make_cluster(2, 13);
    for (int n = 1; n <= 3; ++n) {
      if (!fix.post_force(c.nlocal, c.nghost, c.x.data(), c.v.data(), c.f.data(), c.type.data(),
                          c.mask.data(), c.off.data(), c.nbr.data(), c.owner.data(), nullptr,
                          c.tag.data(), n))
        return fail(fix.err.c_str());
      fix.end_of_step(c.nlocal, c.nghost, c.x.data(), c.type.data(), c.mask.data(), c.off.data(),
                      c.nbr.data(), c.owner.data());
    }
  }
  // oracle: the FDM solver on the grid file
  {
    OFDM g;
    std::string err;
    if (!g.load(grid, err)) return fail(err.c_str());
    g.dt = 1e-3;
    for (int n = 0; n < 3; ++n) {
      g.insert_energy(1.0, 1.0, 1.0, 0.5);
      g.solve();
    }
  }
  std::printf("san_driver: ok\n");
  return 0;
}
