// eph_rng.hpp -- counter-based Gaussian noise for the random force.
//
// xi_i(step) = three N(0,1) draws keyed by (seed, atom tag, timestep), so the
// noise of an atom does not depend on which rank or thread owns it and host
// and device produce the same bits.  Generator: Philox4x32-10 (Salmon,
// Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11):
// counter {tag, step_lo, step_hi, block}, key {seed, 0x45504821}.
// Normals: Box-Muller on 53-bit uniforms, with a natural log from the atanh
// series and sin/cos of 2 pi u from quadrant reduction plus Taylor series,
// evaluated only with IEEE +, -, *, / and sqrt, so that every compiler and
// target rounds them identically (build with FMA contraction off).
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define EPH_HD __host__ __device__ inline
#else
#define EPH_HD inline
#endif

namespace eph {

struct Philox4 {
  uint32_t v[4];
};

EPH_HD Philox4 philox4x32_10(Philox4 ctr, uint32_t k0, uint32_t k1) {
  constexpr uint32_t kMulA = 0xD2511F53u, kMulB = 0xCD9E8D57u;
  constexpr uint32_t kWeylA = 0x9E3779B9u, kWeylB = 0xBB67AE85u;
#pragma unroll
  for (int round = 0; round < 10; ++round) {
    const uint64_t pa = static_cast<uint64_t>(kMulA) * ctr.v[0];
    const uint64_t pb = static_cast<uint64_t>(kMulB) * ctr.v[2];
    Philox4 nx;
    nx.v[0] = static_cast<uint32_t>(pb >> 32) ^ ctr.v[1] ^ k0;
    nx.v[1] = static_cast<uint32_t>(pb);
    nx.v[2] = static_cast<uint32_t>(pa >> 32) ^ ctr.v[3] ^ k1;
    nx.v[3] = static_cast<uint32_t>(pa);
    ctr = nx;
    k0 += kWeylA;
    k1 += kWeylB;
  }
  return ctr;
}

// [0,1) with 53 random bits
EPH_HD double uniform53(uint32_t hi, uint32_t lo) {
  return (static_cast<double>(hi >> 5) * 67108864.0 + static_cast<double>(lo >> 6)) *
         (1.0 / 9007199254740992.0);
}

// split x > 0 into m * 2^e, m in [sqrt(1/2), sqrt(2)), through the bit pattern
EPH_HD double split_exponent(double x, int* e) {
  union { double d; uint64_t u; } b;
  b.d = x;
  int ex = static_cast<int>((b.u >> 52) & 0x7ff);
  if (ex == 0) {  // subnormal: scale into the normal range first
    b.d = x * 18014398509481984.0;  // 2^54
    ex = static_cast<int>((b.u >> 52) & 0x7ff) - 54;
  }
  b.u = (b.u & 0x800fffffffffffffull) | (static_cast<uint64_t>(1022) << 52);  // m in [0.5,1)
  int e2 = ex - 1022;
  double m = b.d;
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e2 -= 1;
  }
  *e = e2;
  return m;
}

// ln x for 0 < x <= 1: e ln2 + 2 atanh((m - 1)/(m + 1)), series to s^25
EPH_HD double log_series(double x) {
  int e;
  const double m = split_exponent(x, &e);
  const double s = (m - 1.0) / (m + 1.0);
  const double s2 = s * s;
  double p = 1.0 / 25.0;
  p = 1.0 / 23.0 + s2 * p;
  p = 1.0 / 21.0 + s2 * p;
  p = 1.0 / 19.0 + s2 * p;
  p = 1.0 / 17.0 + s2 * p;
  p = 1.0 / 15.0 + s2 * p;
  p = 1.0 / 13.0 + s2 * p;
  p = 1.0 / 11.0 + s2 * p;
  p = 1.0 / 9.0 + s2 * p;
  p = 1.0 / 7.0 + s2 * p;
  p = 1.0 / 5.0 + s2 * p;
  p = 1.0 / 3.0 + s2 * p;
  const double lm = 2.0 * s + 2.0 * s * (s2 * p);
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  const double de = static_cast<double>(e);
  return de * ln2_hi + (de * ln2_lo + lm);
}

// sin and cos of 2 pi u, u in [0,1)
EPH_HD void sincos_2pi(double u, double* sn, double* cs) {
  const double t = 4.0 * u;
  double q = t + 0.5;
  q = static_cast<double>(static_cast<int64_t>(q));  // floor for q >= 0
  const double f = t - q;                             // [-0.5, 0.5]
  const double x = f * 1.57079632679489661923;
  const double x2 = x * x;
  // Taylor coefficients of sin(x)/x and cos(x) through x^20
  double ps = 1.0 / 51090942171709440000.0, pc = 1.0 / 2432902008176640000.0;
  ps = -1.0 / 121645100408832000.0 + x2 * ps;  pc = -1.0 / 6402373705728000.0 + x2 * pc;
  ps = 1.0 / 355687428096000.0 + x2 * ps;      pc = 1.0 / 20922789888000.0 + x2 * pc;
  ps = -1.0 / 1307674368000.0 + x2 * ps;       pc = -1.0 / 87178291200.0 + x2 * pc;
  ps = 1.0 / 6227020800.0 + x2 * ps;           pc = 1.0 / 479001600.0 + x2 * pc;
  ps = -1.0 / 39916800.0 + x2 * ps;            pc = -1.0 / 3628800.0 + x2 * pc;
  ps = 1.0 / 362880.0 + x2 * ps;               pc = 1.0 / 40320.0 + x2 * pc;
  ps = -1.0 / 5040.0 + x2 * ps;                pc = -1.0 / 720.0 + x2 * pc;
  ps = 1.0 / 120.0 + x2 * ps;                  pc = 1.0 / 24.0 + x2 * pc;
  ps = -1.0 / 6.0 + x2 * ps;                   pc = -1.0 / 2.0 + x2 * pc;
  ps = 1.0 + x2 * ps;                          pc = 1.0 + x2 * pc;
  const double s = x * ps, c = pc;
  switch (static_cast<int>(q) & 3) {
    case 0: *sn = s; *cs = c; break;
    case 1: *sn = c; *cs = -s; break;
    case 2: *sn = -s; *cs = -c; break;
    default: *sn = -c; *cs = s; break;
  }
}

// three N(0,1) numbers for (seed, tag, step)
EPH_HD void gaussian3(uint32_t seed, uint32_t tag, uint64_t step, double out[3]) {
  Philox4 c0{{tag, static_cast<uint32_t>(step), static_cast<uint32_t>(step >> 32), 0u}};
  Philox4 c1{{tag, static_cast<uint32_t>(step), static_cast<uint32_t>(step >> 32), 1u}};
  const Philox4 r0 = philox4x32_10(c0, seed, 0x45504821u);
  const Philox4 r1 = philox4x32_10(c1, seed, 0x45504821u);
  const double u0 = 1.0 - uniform53(r0.v[0], r0.v[1]);  // (0,1]
  const double u1 = uniform53(r0.v[2], r0.v[3]);
  const double u2 = 1.0 - uniform53(r1.v[0], r1.v[1]);
  const double u3 = uniform53(r1.v[2], r1.v[3]);
  const double ra = __builtin_sqrt(-2.0 * log_series(u0));
  const double rb = __builtin_sqrt(-2.0 * log_series(u2));
  double sa, ca, sb, cb;
  sincos_2pi(u1, &sa, &ca);
  sincos_2pi(u3, &sb, &cb);
  out[0] = ra * ca;
  out[1] = ra * sa;
  out[2] = rb * cb;
}

}  // namespace eph
