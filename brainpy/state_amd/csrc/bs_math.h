// bs_math.h -- device arithmetic whose results are fixed bit-for-bit:
// a counter-based RNG (Threefry-2x32, 20 rounds; Salmon et al., SC'11), a
// Poisson sampler by inversion of a host-tabulated CDF (compares only),
// and exp/log/pow from range reduction plus Taylor / atanh series. Every
// operation is an exactly rounded IEEE op (the file is compiled with
// -ffp-contract=off), so a CPU restatement performing the same sequence of
// operations reproduces every bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bs {

// ---------------------------------------------------------------- Threefry
__device__ __forceinline__ uint32_t rotl32(uint32_t x, int r) {
  return (x << r) | (x >> (32 - r));
}

// Threefry-2x32 with 20 rounds: key (k0,k1), counter (c0,c1) -> (x0,x1).
// Rotation schedule and key injection per the Random123 specification.
__device__ __forceinline__ void threefry2x32(uint32_t k0, uint32_t k1, uint32_t c0,
                                             uint32_t c1, uint32_t& o0, uint32_t& o1) {
  const uint32_t ks2 = 0x1BD11BDAu ^ k0 ^ k1;
  const uint32_t ks[3] = {k0, k1, ks2};
  const int R[8] = {13, 15, 26, 6, 17, 29, 16, 24};
  uint32_t x0 = c0 + k0, x1 = c1 + k1;
#pragma unroll
  for (int blk = 0; blk < 5; ++blk) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      x0 += x1;
      x1 = rotl32(x1, R[(blk * 4 + r) & 7]);
      x1 ^= x0;
    }
    x0 += ks[(blk + 1) % 3];
    x1 += ks[(blk + 2) % 3] + (uint32_t)(blk + 1);
  }
  o0 = x0;
  o1 = x1;
}

// Uniform in (0,1): (x + 1/2) * 2^-32, exact in double.
__device__ __forceinline__ double u01(uint32_t x) {
  return ((double)x + 0.5) * 2.3283064365386963e-10;
}

// Poisson(lam) count for generator element `elem` at step `step`, by
// inversion of a tabulated CDF (built on the host by bs_poisson_table):
// lam is split into `parts` equal parts; part p consumes uniform number p of
// the stream threefry(key=(seed, step), ctr=(elem, p/2)), word p%2, as
// u = (x + 1/2) 2^-32 in (0,1), and contributes min{k : u <= cdf[k]} (search
// started from the 64-entry guide table, bounded by the table length).
// Only compares and IEEE ops on exactly representable inputs are used, so
// the count is the same on every IEEE machine.
__device__ __forceinline__ uint32_t poisson_inv(uint32_t seed, uint32_t step, uint32_t elem,
                                                const double* cdf, const unsigned short* guide,
                                                int len, int parts) {
  uint32_t count = 0;
  uint32_t x1 = 0;
  for (int p = 0; p < parts; ++p) {
    uint32_t x;
    if ((p & 1) == 0) {
      threefry2x32(seed, step, elem, (uint32_t)(p >> 1), x, x1);
    } else {
      x = x1;
    }
    const double u = u01(x);
    int k = guide[(int)(u * 64.0)];
    while (k < len - 1 && u > cdf[k]) ++k;
    count += (uint32_t)k;
  }
  return count;
}

// ---------------------------------------------------------------- exp/log
// ln2 split: LN2_HI has 42 significant bits, so k*LN2_HI is exact for |k|<2048.
#define BS_LN2_HI 0x1.62e42fefa3800p-1
#define BS_LN2_LO 0x1.ef35793c76730p-45
#define BS_INV_LN2 0x1.71547652b82fep+0

// exp(x): k = rint(x/ln2), r = (x - k*LN2_HI) - k*LN2_LO, |r| <= 0.347;
// exp(r) by the Taylor series to r^13, then scaled by 2^k. The series is
// evaluated as 1 + (r + r^2 (1/2 + r T(r))) with T(r) = sum_{n=3..13} r^(n-3)/n!
// in Estrin form (pairs, then r^2, r^4, r^8): a dependency chain of about a
// dozen operations instead of Horner's 26, since aeif_cond_alpha's RKF45
// substeps wait on it. Every operation is one IEEE-rounded op (no contraction),
// in exactly this order, so the CPU evaluation gives the same bits.
__device__ __forceinline__ double exp_p(double x) {
  if (x != x) return x;
  if (x > 709.782712893384) return __builtin_inf();
  if (x < -745.1332191019412) return 0.0;
  const double kd = __builtin_rint(x * BS_INV_LN2);
  const int k = (int)kd;
  const double r = (x - kd * BS_LN2_HI) - kd * BS_LN2_LO;
  const double r2 = r * r, r4 = r2 * r2, r8 = r4 * r4;
  const double a0 = 1.0 / 6.0 + 1.0 / 24.0 * r;                // 1/3!, 1/4!
  const double a1 = 1.0 / 120.0 + 1.0 / 720.0 * r;             // 1/5!, 1/6!
  const double a2 = 1.0 / 5040.0 + 1.0 / 40320.0 * r;          // 1/7!, 1/8!
  const double a3 = 1.0 / 362880.0 + 1.0 / 3628800.0 * r;      // 1/9!, 1/10!
  const double a4 = 1.0 / 39916800.0 + 1.0 / 479001600.0 * r;  // 1/11!, 1/12!
  const double b0 = a0 + a1 * r2;
  const double b1 = a2 + a3 * r2;
  const double b2 = a4 + 1.0 / 6227020800.0 * r2;             // 1/13!
  const double t = (b0 + b1 * r4) + b2 * r8;
  const double p = 1.0 + (r + r2 * (0.5 + r * t));
  return __builtin_ldexp(p, k);
}

// log(x) for x > 0: x = m 2^e, m in [sqrt(1/2), sqrt(2)); s = (m-1)/(m+1);
// log(m) = 2 atanh(s) = 2 s (1 + z/3 + z^2/5 + ... + z^11/23), z = s^2,
// the series in Estrin form like exp_p's.
__device__ __forceinline__ double log_p(double x) {
  if (x != x || x < 0.0) return __builtin_nan("");
  if (x == 0.0) return -__builtin_inf();
  if (x == __builtin_inf()) return x;
  int e;
  double m = __builtin_frexp(x, &e);  // m in [0.5, 1)
  if (m < 0.7071067811865476) {
    m = m * 2.0;
    e = e - 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double z = s * s;
  // q = 1 + z/3 + z^2 Q(z), Q(z) = sum_{j=0..9} z^j/(2j+5), Q in Estrin form
  const double z2 = z * z, z4 = z2 * z2, z8 = z4 * z4;
  const double a0 = 1.0 / 5.0 + z * (1.0 / 7.0);
  const double a1 = 1.0 / 9.0 + z * (1.0 / 11.0);
  const double a2 = 1.0 / 13.0 + z * (1.0 / 15.0);
  const double a3 = 1.0 / 17.0 + z * (1.0 / 19.0);
  const double a4 = 1.0 / 21.0 + z * (1.0 / 23.0);
  const double qq = ((a0 + a1 * z2) + (a2 + a3 * z2) * z4) + a4 * z8;
  const double q = 1.0 + z * (1.0 / 3.0 + z * qq);
  const double lm = 2.0 * s * q;
  const double ed = (double)e;
  return ed * BS_LN2_HI + (ed * BS_LN2_LO + lm);
}

// x / c given rc = 1/c (the IEEE quotient, computed once per divisor):
// q = x*rc lies within one ulp of x/c, r = x - q*c is exact with an FMA, and
// q + r*rc rounded once is the correctly rounded quotient (Markstein's
// theorem on FMA division). Three dependent operations, no branch,
// bit-identical to x / c whenever |q| is in [2^-1000, 2^1000]; zero,
// infinite and NaN x give q itself (also the IEEE result). Any other case
// (near-subnormal or huge quotients) sets *bad, and the caller redoes its
// divisions with x / c.
__device__ __forceinline__ double div_rc(double x, double c, double rc, bool& bad) {
  const double q = x * rc;
  const double aq = __builtin_fabs(q);
  const bool ok = aq >= 0x1p-1000 && aq <= 0x1p+1000;
  const bool pass = x == 0.0 || x != x || __builtin_fabs(x) == __builtin_inf();
  bad = bad || (!ok && !pass);
  return ok ? __builtin_fma(__builtin_fma(-q, c, x), rc, q) : q;
}

// x^y for x >= 0 (the only domain used: step-size control, STDP exponents).
__device__ __forceinline__ double pow_p(double x, double y) {
  if (y == 0.0) return 1.0;
  if (y == 1.0) return x;
  if (x == 0.0) return y > 0.0 ? 0.0 : __builtin_inf();
  return exp_p(y * log_p(x));
}

// ---------------------------------------------------------------- sin
// pi/2 in three parts (Cody & Waite): PIO2_1 and PIO2_2 carry 33 significant
// bits, so k * PIO2_1 and k * PIO2_2 are exact for |k| < 2^20; PIO2_3 rounds
// the rest. BS_2_PI is 2/pi rounded.
#define BS_PIO2_1 0x1.921fb54400000p+0
#define BS_PIO2_2 0x1.0b4611a600000p-34
#define BS_PIO2_3 0x1.3198a2e037073p-69
#define BS_2_PI 0x1.45f306dc9c883p-1

// sin(x): k = rint(x 2/pi), r = ((x - k PIO2_1) - k PIO2_2) - k PIO2_3,
// |r| <= pi/4 (plus rounding); by the quadrant k mod 4 the Taylor series of
// sin r (to r^21) or cos r (to r^20), Horner in z = r^2 (truncation below
// 1e-23 relative). Each operation is one IEEE-rounded op in this order.
__device__ __forceinline__ double sin_p(double x) {
  if (!(x - x == 0.0)) return x - x;  // inf and NaN give NaN
  const double kd = __builtin_rint(x * BS_2_PI);
  const double r = ((x - kd * BS_PIO2_1) - kd * BS_PIO2_2) - kd * BS_PIO2_3;
  const double z = r * r;
  const int q = (int)((long long)kd & 3);
  double v;
  if (q & 1) {
    double c = 1.0 / 2432902008176640000.0;   // 1/20!
    c = -1.0 / 6402373705728000.0 + z * c;    // 1/18!
    c = 1.0 / 20922789888000.0 + z * c;       // 1/16!
    c = -1.0 / 87178291200.0 + z * c;         // 1/14!
    c = 1.0 / 479001600.0 + z * c;            // 1/12!
    c = -1.0 / 3628800.0 + z * c;             // 1/10!
    c = 1.0 / 40320.0 + z * c;                // 1/8!
    c = -1.0 / 720.0 + z * c;                 // 1/6!
    c = 1.0 / 24.0 + z * c;                   // 1/4!
    c = -0.5 + z * c;
    v = 1.0 + z * c;
  } else {
    double s = 1.0 / 51090942171709440000.0;  // 1/21!
    s = -1.0 / 121645100408832000.0 + z * s;  // 1/19!
    s = 1.0 / 355687428096000.0 + z * s;      // 1/17!
    s = -1.0 / 1307674368000.0 + z * s;       // 1/15!
    s = 1.0 / 6227020800.0 + z * s;           // 1/13!
    s = -1.0 / 39916800.0 + z * s;            // 1/11!
    s = 1.0 / 362880.0 + z * s;               // 1/9!
    s = -1.0 / 5040.0 + z * s;                // 1/7!
    s = 1.0 / 120.0 + z * s;                  // 1/5!
    s = -1.0 / 6.0 + z * s;                   // 1/3!
    v = r + (r * z) * s;
  }
  return (q & 2) ? -v : v;
}

}  // namespace bs
