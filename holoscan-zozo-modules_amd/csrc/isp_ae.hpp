// isp_ae.hpp - the statistics epilogue with a device-exact log2 / exp2.
//
// finish_stats (reference modules/isp/src/isp_math.hpp:387-425) calls the host
// libm's double log2 and exp2; a GPU libm rounds differently in the last bit.
// This header evaluates log2 and exp2 in double-double arithmetic (Dekker's
// two-sum / two-product with fused multiply-add, "Handbook of Floating-Point
// Arithmetic", Muller et al., ch. 4 and 14) to about 2^-90 relative error and
// rounds once at the end, so the double it returns is the correctly rounded
// log2 / exp2 except where the exact value lies within ~2^-90 of a rounding
// midpoint. The same code runs on the host (tests compare it with glibc over
// millions of arguments) and on the device. Built without contraction: every
// operation below is one IEEE operation, the fma() calls are the only fused ones.
#pragma once

#include <cstdint>
#include <cstring>

#if defined(__HIPCC__)
#define ISP_HD __host__ __device__ inline
#else
#define ISP_HD inline
#endif

namespace isp_amd {

struct dd {
  double hi;
  double lo;
};

ISP_HD double isp_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }

ISP_HD double bits_to_double(std::uint64_t u) {
  double d;
  std::memcpy(&d, &u, sizeof(d));
  return d;
}
ISP_HD std::uint64_t double_to_bits(double d) {
  std::uint64_t u;
  std::memcpy(&u, &d, sizeof(u));
  return u;
}

// |a| >= |b| or a == 0.
ISP_HD dd fast_two_sum(double a, double b) {
  const double s = a + b;
  const double e = b - (s - a);
  return {s, e};
}
ISP_HD dd two_sum(double a, double b) {
  const double s = a + b;
  const double bb = s - a;
  const double e = (a - (s - bb)) + (b - bb);
  return {s, e};
}
ISP_HD dd two_prod(double a, double b) {
  const double p = a * b;
  return {p, isp_fma(a, b, -p)};
}
ISP_HD dd dd_add(dd a, dd b) {
  dd s = two_sum(a.hi, b.hi);
  const dd t = two_sum(a.lo, b.lo);
  s.lo = s.lo + t.hi;
  s = fast_two_sum(s.hi, s.lo);
  s.lo = s.lo + t.lo;
  return fast_two_sum(s.hi, s.lo);
}
ISP_HD dd dd_add_d(dd a, double b) {
  dd s = two_sum(a.hi, b);
  s.lo = s.lo + a.lo;
  return fast_two_sum(s.hi, s.lo);
}
ISP_HD dd dd_neg(dd a) { return {-a.hi, -a.lo}; }
ISP_HD dd dd_mul(dd a, dd b) {
  dd p = two_prod(a.hi, b.hi);
  const double cross = a.hi * b.lo + a.lo * b.hi;
  p.lo = p.lo + cross;
  return fast_two_sum(p.hi, p.lo);
}
ISP_HD dd dd_mul_d(dd a, double b) {
  dd p = two_prod(a.hi, b);
  p.lo = p.lo + a.lo * b;
  return fast_two_sum(p.hi, p.lo);
}
// a / b, three correction steps.
ISP_HD dd dd_div(dd a, dd b) {
  const double q1 = a.hi / b.hi;
  dd r = dd_add(a, dd_neg(dd_mul_d(b, q1)));
  const double q2 = r.hi / b.hi;
  r = dd_add(r, dd_neg(dd_mul_d(b, q2)));
  const double q3 = r.hi / b.hi;
  const dd q = fast_two_sum(q1, q2);
  return dd_add_d(q, q3);
}

// 1/j for j = 1..29 in double-double: hi the double nearest 1/j (what 1.0 / j
// rounds to), lo the double nearest 1/j - hi.
struct dd_bits {
  std::uint64_t hi;
  std::uint64_t lo;
};
inline constexpr dd_bits kRecip[30] = {
    {0, 0},
    {0x3ff0000000000000ULL, 0x0000000000000000ULL},  // 1/1
    {0x3fe0000000000000ULL, 0x0000000000000000ULL},  // 1/2
    {0x3fd5555555555555ULL, 0x3c75555555555555ULL},  // 1/3
    {0x3fd0000000000000ULL, 0x0000000000000000ULL},  // 1/4
    {0x3fc999999999999aULL, 0xbc6999999999999aULL},  // 1/5
    {0x3fc5555555555555ULL, 0x3c65555555555555ULL},  // 1/6
    {0x3fc2492492492492ULL, 0x3c62492492492492ULL},  // 1/7
    {0x3fc0000000000000ULL, 0x0000000000000000ULL},  // 1/8
    {0x3fbc71c71c71c71cULL, 0x3c5c71c71c71c71cULL},  // 1/9
    {0x3fb999999999999aULL, 0xbc5999999999999aULL},  // 1/10
    {0x3fb745d1745d1746ULL, 0xbc4745d1745d1746ULL},  // 1/11
    {0x3fb5555555555555ULL, 0x3c55555555555555ULL},  // 1/12
    {0x3fb3b13b13b13b14ULL, 0xbc53b13b13b13b14ULL},  // 1/13
    {0x3fb2492492492492ULL, 0x3c52492492492492ULL},  // 1/14
    {0x3fb1111111111111ULL, 0x3c31111111111111ULL},  // 1/15
    {0x3fb0000000000000ULL, 0x0000000000000000ULL},  // 1/16
    {0x3fae1e1e1e1e1e1eULL, 0x3c2e1e1e1e1e1e1eULL},  // 1/17
    {0x3fac71c71c71c71cULL, 0x3c4c71c71c71c71cULL},  // 1/18
    {0x3faaf286bca1af28ULL, 0x3c4af286bca1af28ULL},  // 1/19
    {0x3fa999999999999aULL, 0xbc4999999999999aULL},  // 1/20
    {0x3fa8618618618618ULL, 0x3c48618618618618ULL},  // 1/21
    {0x3fa745d1745d1746ULL, 0xbc3745d1745d1746ULL},  // 1/22
    {0x3fa642c8590b2164ULL, 0x3c3642c8590b2164ULL},  // 1/23
    {0x3fa5555555555555ULL, 0x3c45555555555555ULL},  // 1/24
    {0x3fa47ae147ae147bULL, 0xbc2eb851eb851eb8ULL},  // 1/25
    {0x3fa3b13b13b13b14ULL, 0xbc43b13b13b13b14ULL},  // 1/26
    {0x3fa2f684bda12f68ULL, 0x3c42f684bda12f68ULL},  // 1/27
    {0x3fa2492492492492ULL, 0x3c42492492492492ULL},  // 1/28
    {0x3fa1a7b9611a7b96ULL, 0x3c21a7b9611a7b96ULL},  // 1/29
};
ISP_HD dd dd_recip(int j) { return {bits_to_double(kRecip[j].hi), bits_to_double(kRecip[j].lo)}; }

// ln 2 and log2(e) to 106 bits.
ISP_HD dd dd_ln2() { return {bits_to_double(0x3fe62e42fefa39efULL), bits_to_double(0x3c7abc9e3b39803fULL)}; }
ISP_HD dd dd_log2e() { return {bits_to_double(0x3ff71547652b82feULL), bits_to_double(0x3c7777d0ffda0d24ULL)}; }

// log2(x) for finite x > 0 (normal or subnormal).
ISP_HD double isp_log2(double x) {
  std::uint64_t u = double_to_bits(x);
  int e = static_cast<int>((u >> 52) & 0x7ff);
  int bias = 1023;
  if (e == 0) {  // subnormal: scale by 2^64 exactly
    u = double_to_bits(x * 18446744073709551616.0);
    e = static_cast<int>((u >> 52) & 0x7ff);
    bias += 64;
  }
  // m in [1, 2), then folded to [sqrt(1/2), sqrt(2)).
  double m = bits_to_double((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
  int k = e - bias;
  if (m > 1.4142135623730951) {
    m = m * 0.5;
    k = k + 1;
  }
  // log(m) = 2 atanh(s), s = (m - 1) / (m + 1), |s| <= 0.1716.
  const dd num{m - 1.0, 0.0};  // exact (Sterbenz)
  const dd den = two_sum(m, 1.0);
  const dd s = dd_div(num, den);
  const dd t = dd_mul(s, s);
  // 2 atanh(s) / s = 2 (1 + t/3 + t^2/5 + ...): tail t^4/9 .. in double, head in double-double.
  double tail = dd_recip(29).hi;
  for (int j = 27; j >= 9; j -= 2) tail = dd_recip(j).hi + t.hi * tail;
  dd poly{tail, 0.0};
  for (int j = 7; j >= 1; j -= 2) poly = dd_add(dd_recip(j), dd_mul(t, poly));
  const dd ln_m = dd_mul_d(dd_mul(s, poly), 2.0);
  const dd l2 = dd_mul(ln_m, dd_log2e());
  const dd r = dd_add(dd{static_cast<double>(k), 0.0}, l2);
  return r.hi + r.lo;
}

// 2^y for |y| <= 1000 (the statistics need |y| <= 8).
ISP_HD double isp_exp2(double y) {
  const double n = __builtin_floor(y + 0.5);
  const double f = y - n;  // exact, |f| <= 0.5
  // r = f ln 2, then exp(r) = exp(r / 256)^256.
  const dd r = dd_mul_d(dd_ln2(), f);
  const dd q = dd_mul_d(r, 0.00390625);  // exact scaling by 2^-8, |q| <= 0.00136
  // Taylor series by Horner: exp(q) = 1 + q (1 + q/2 (1 + q/3 (1 + ... q/10))).
  dd e{1.0, 0.0};
  for (int j = 10; j >= 1; --j) {
    e = dd_add_d(dd_mul(dd_mul(q, dd_recip(j)), e), 1.0);
  }
  for (int i = 0; i < 8; ++i) e = dd_mul(e, e);
  // Scale by 2^n exactly (n in the normal range here).
  const int ni = static_cast<int>(n);
  const double scale = bits_to_double(static_cast<std::uint64_t>(ni + 1023) << 52);
  return (e.hi + e.lo) * scale;
}

struct AeParams {
  float target;
  float damping;
  std::int64_t min_exposure_ns;
  std::int64_t max_exposure_ns;
  float min_gain;
  float max_gain;
};

// finish_stats (reference isp_math.hpp:387-425) with the log2 / exp2 above.
ISP_HD void finish_stats(const unsigned long long sums[4], long long pixels, const float gains[3],
                         const AeParams& ae, long long exposure_ns, float exposure_gain,
                         float* out) {
  const double n = static_cast<double>(pixels) * 1048576.0;
  const double mean[3] = {static_cast<double>(sums[0]) / n, static_cast<double>(sums[1]) / n,
                          static_cast<double>(sums[2]) / n};
  const double luma = 0.2126 * mean[0] + 0.7152 * mean[1] + 0.0722 * mean[2];
  double error = luma > 0.0 ? isp_log2(static_cast<double>(ae.target) / luma) : 8.0;
  error = error < -8.0 ? -8.0 : error > 8.0 ? 8.0 : error;
  const double scale = isp_exp2(static_cast<double>(ae.damping) * error);
  double next_exposure = 0.0;
  double next_gain = 0.0;
  if (exposure_ns > 0) {
    const double gain = exposure_gain > 0.0F ? static_cast<double>(exposure_gain)
                                             : static_cast<double>(ae.min_gain);
    const double wanted = static_cast<double>(exposure_ns) * gain * scale;
    next_exposure = wanted / static_cast<double>(ae.min_gain);
    const double low = static_cast<double>(ae.min_exposure_ns);
    const double high = static_cast<double>(ae.max_exposure_ns);
    next_exposure = next_exposure < low ? low : next_exposure > high ? high : next_exposure;
    next_gain = wanted / next_exposure;
    const double gmin = static_cast<double>(ae.min_gain);
    const double gmax = static_cast<double>(ae.max_gain);
    next_gain = next_gain < gmin ? gmin : next_gain > gmax ? gmax : next_gain;
  }
  out[0] = static_cast<float>(mean[0]);
  out[1] = static_cast<float>(mean[1]);
  out[2] = static_cast<float>(mean[2]);
  out[3] = static_cast<float>(luma);
  out[4] = static_cast<float>(error);
  out[5] = static_cast<float>(scale);
  out[6] = static_cast<float>(next_exposure);
  out[7] = static_cast<float>(next_gain);
  out[8] = gains[0];
  out[9] = gains[1];
  out[10] = gains[2];
  out[11] = static_cast<float>(static_cast<double>(sums[3]) / static_cast<double>(pixels));
}

}  // namespace isp_amd
