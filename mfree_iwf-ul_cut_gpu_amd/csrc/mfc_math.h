// mfc_math.h — transcendental functions shared bit-for-bit by host C++ and gfx950 device code.
//
// The SPH step calls atan2f/cosf/sinf (artificial stress, actions_gpu.cu:38,76-77),
// pow/log (Johnson-Cook yield, plasticity.cuh:30-46), exp (PSE kernel, kernels.cuh:86)
// and atan2 (tool fillet test, tool_gpu.cu:29-36). Vendor libm results differ in the last
// bit between the host and the GPU, and one differing bit can flip the secant iteration
// count of the radial return. Every function here is therefore written with IEEE
// add/sub/mul/div/sqrt only, evaluated in a fixed order with no FMA contraction (all
// translation units that include this header are compiled with -ffp-contract=off), so the
// host and the GPU produce identical doubles.
//
// Methods (all from the mathematics, not from any vendor library):
//   exp : Cody-Waite reduction x = k ln2 + r (ln2 split in 32+53 bits), |r| <= ln2/2,
//         Taylor series of e^r to r^13, scaling by 2^k through the exponent field.
//   log : x = 2^e m, m in [sqrt(1/2), sqrt(2)), f = m-1, s = f/(2+f),
//         log(1+f) = f - (f^2/2 - s (f^2/2 + R)), R = sum_{k>=1} 2 s^{2k}/(2k+1) (atanh series).
//   pow : exp(y log|x|) with y log|x| formed in double-double (Dekker products), IEEE 754
//         special cases.
//   tanh: em1/(em1+2) with em1 = e^{2|x|}-1 (Taylor series for |x| < 0.55).
//   atan: |x|>1 -> pi/2 - atan(1/x); |x|>tan(pi/12) -> pi/6 + atan((x sqrt3 - 1)/(x + sqrt3));
//         Taylor series of atan to t^29 on |t| <= tan(pi/12).
//   sin/cos: Cody-Waite reduction by pi/2 (three-part split, valid for |x| < 1e6),
//         Taylor series to r^17 / r^18 on |r| <= pi/4.
// float variants (the reference calls atan2f/cosf/sinf with double arguments) are the double
// result rounded once to float.
#pragma once

#include <stdint.h>
#include <math.h>

#if defined(__HIP__)
#define MFC_HD __host__ __device__ inline
#else
#define MFC_HD inline
#endif

namespace mfcm {

MFC_HD uint64_t d2u(double x) { uint64_t u; __builtin_memcpy(&u, &x, 8); return u; }
MFC_HD double u2d(uint64_t u) { double x; __builtin_memcpy(&x, &u, 8); return x; }
MFC_HD bool is_nan(double x) { return x != x; }
MFC_HD bool sign_bit(double x) { return (d2u(x) >> 63) != 0; }
MFC_HD bool is_inf(double x) { return (d2u(x) & 0x7fffffffffffffffULL) == 0x7ff0000000000000ULL; }
MFC_HD double abs_d(double x) { return u2d(d2u(x) & 0x7fffffffffffffffULL); }
// 2^k for k in [-1022, 1023], exact.
MFC_HD double pow2i(int k) { return u2d((uint64_t)(k + 1023) << 52); }

constexpr double LN2_HI  = 0.6931471806019545;        // 0x3fe62e42ff000000 (32 bits)
constexpr double LN2_LO  = -4.2009150726810846e-11;   // ln2 - LN2_HI
constexpr double INV_LN2 = 1.4426950408889634;
constexpr double PI      = 3.141592653589793;
constexpr double PI_LO   = 1.2246467991473532e-16;
constexpr double PIO2    = 1.5707963267948966;
constexpr double PIO2_LO = 6.123233995736766e-17;
constexpr double PIO4    = 0.7853981633974483;
constexpr double PIO6    = 0.5235987755982989;
constexpr double PIO6_LO = -5.360408832255455e-17;
constexpr double TWO_OVER_PI = 0.6366197723675814;
constexpr double PIO2_HI  = 1.5707963267341256;       // 0x3ff921fb54400000 (33 bits)
constexpr double PIO2_MID = 6.077100506303966e-11;    // 0x3dd0b4611a600000 (33 bits)
constexpr double PIO2_LO3 = 2.0222662487959506e-21;
constexpr double SQRT3   = 1.7320508075688772;
constexpr double TAN_PI12 = 0.2679491924311228;
constexpr double SQRT2   = 1.4142135623730951;

// e^x
MFC_HD double exp(double x) {
  if (is_nan(x)) return x;
  if (x > 709.782712893384) return u2d(0x7ff0000000000000ULL);
  if (x < -745.1332191019412) return 0.0;
  double kd = floor(x * INV_LN2 + 0.5);
  int k = (int)kd;
  double r = (x - kd * LN2_HI) - kd * LN2_LO;
  // Taylor coefficients 1/j!, j = 13 .. 2
  double p = 1.6059043836821613e-10;            // 1/13!
  p = 2.08767569878681e-09 + r * p;             // 1/12!
  p = 2.505210838544172e-08 + r * p;            // 1/11!
  p = 2.755731922398589e-07 + r * p;            // 1/10!
  p = 2.7557319223985893e-06 + r * p;           // 1/9!
  p = 2.48015873015873e-05 + r * p;             // 1/8!
  p = 0.0001984126984126984 + r * p;            // 1/7!
  p = 0.001388888888888889 + r * p;             // 1/6!
  p = 0.008333333333333333 + r * p;             // 1/5!
  p = 0.041666666666666664 + r * p;             // 1/4!
  p = 0.16666666666666666 + r * p;              // 1/3!
  p = 0.5 + r * p;                              // 1/2!
  double e = 1.0 + (r + (r * r) * p);
  if (k > 1023) return (e * pow2i(1023)) * pow2i(k - 1023);
  if (k < -1022) return (e * pow2i(k + 54)) * pow2i(-54);
  return e * pow2i(k);
}

// natural logarithm
MFC_HD double log(double x) {
  if (is_nan(x)) return x;
  if (x < 0.0) return u2d(0x7ff8000000000000ULL);
  if (x == 0.0) return -u2d(0x7ff0000000000000ULL);
  if (is_inf(x)) return x;
  int eadj = 0;
  if (x < 2.2250738585072014e-308) { x = x * 18014398509481984.0; eadj = -54; }  // 2^54
  uint64_t u = d2u(x);
  int e = (int)((u >> 52) & 0x7ff) - 1023 + eadj;
  double m = u2d((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
  if (m > SQRT2) { m = m * 0.5; e += 1; }
  double f = m - 1.0;
  double s = f / (2.0 + f);
  double z = s * s;
  // R = sum_{k=1..11} 2/(2k+1) z^k
  double R = 0.08695652173913043;               // 2/23
  R = 0.09523809523809523 + z * R;              // 2/21
  R = 0.10526315789473684 + z * R;              // 2/19
  R = 0.11764705882352941 + z * R;              // 2/17
  R = 0.13333333333333333 + z * R;              // 2/15
  R = 0.15384615384615385 + z * R;              // 2/13
  R = 0.18181818181818182 + z * R;              // 2/11
  R = 0.2222222222222222 + z * R;               // 2/9
  R = 0.2857142857142857 + z * R;               // 2/7
  R = 0.4 + z * R;                              // 2/5
  R = 0.6666666666666666 + z * R;               // 2/3
  R = z * R;
  double hfsq = 0.5 * f * f;
  double dk = (double)e;
  return dk * LN2_HI - ((hfsq - (s * (hfsq + R) + dk * LN2_LO)) - f);
}

MFC_HD bool is_int(double y) { return floor(y) == y; }
MFC_HD bool is_odd_int(double y) {
  if (!is_int(y) || abs_d(y) >= 9007199254740992.0) return false;
  double h = floor(y * 0.5);
  return h * 2.0 != y;
}

// Error-free transformations (Knuth two-sum, Dekker two-product with Veltkamp splitting).
MFC_HD void two_sum(double a, double b, double &s, double &e) {
  s = a + b;
  double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
MFC_HD void split(double a, double &hi, double &lo) {
  double c = 134217729.0 * a;  // 2^27 + 1
  double big = c - a;
  hi = c - big;
  lo = a - hi;
}
MFC_HD void two_prod(double a, double b, double &p, double &e) {
  p = a * b;
  double ah, al, bh, bl;
  split(a, ah, al);
  split(b, bh, bl);
  e = (((ah * bh - p) + ah * bl) + al * bh) + al * bl;
}

// log(x) as an unevaluated sum hi + lo (x finite, > 0), used by pow.
MFC_HD void log_dd(double x, double &hi, double &lo) {
  int eadj = 0;
  if (x < 2.2250738585072014e-308) { x = x * 18014398509481984.0; eadj = -54; }
  uint64_t u = d2u(x);
  int e = (int)((u >> 52) & 0x7ff) - 1023 + eadj;
  double m = u2d((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
  if (m > SQRT2) { m = m * 0.5; e += 1; }
  double f = m - 1.0;
  double s = f / (2.0 + f);
  double z = s * s;
  double R = 0.08695652173913043;
  R = 0.09523809523809523 + z * R;
  R = 0.10526315789473684 + z * R;
  R = 0.11764705882352941 + z * R;
  R = 0.13333333333333333 + z * R;
  R = 0.15384615384615385 + z * R;
  R = 0.18181818181818182 + z * R;
  R = 0.2222222222222222 + z * R;
  R = 0.2857142857142857 + z * R;
  R = 0.4 + z * R;
  R = 0.6666666666666666 + z * R;
  R = z * R;
  double ff, ffe;
  two_prod(f, f, ff, ffe);
  double hfsq = 0.5 * ff, hfe = 0.5 * ffe;
  double corr = s * (hfsq + R) - hfe;   // log(1+f) = f - hfsq + corr
  double dk = (double)e;
  double s1, e1, s2, e2;
  two_sum(dk * LN2_HI, f, s1, e1);
  two_sum(s1, -hfsq, s2, e2);
  double l = ((e1 + e2) + corr) + dk * LN2_LO;
  hi = s2 + l;
  lo = l - (hi - s2);
}

// x^y with IEEE 754 special cases; y log|x| is formed in double-double before exp.
MFC_HD double pow(double x, double y) {
  const double INF = u2d(0x7ff0000000000000ULL);
  if (y == 0.0) return 1.0;
  if (x == 1.0) return 1.0;
  if (is_nan(x) || is_nan(y)) return x + y;
  if (is_inf(y)) {
    double ax = abs_d(x);
    if (ax == 1.0) return 1.0;
    if ((ax > 1.0) == (y > 0.0)) return INF;
    return 0.0;
  }
  if (x == 0.0) {
    bool odd = is_odd_int(y);
    if (y > 0.0) return (odd && sign_bit(x)) ? -0.0 : 0.0;
    return (odd && sign_bit(x)) ? -INF : INF;
  }
  if (is_inf(x)) {
    bool odd = is_odd_int(y);
    if (x > 0.0) return y > 0.0 ? INF : 0.0;
    if (y > 0.0) return odd ? -INF : INF;
    return odd ? -0.0 : 0.0;
  }
  double sgn = 1.0;
  if (x < 0.0) {
    if (!is_int(y)) return u2d(0x7ff8000000000000ULL);
    if (is_odd_int(y)) sgn = -1.0;
    x = -x;
  }
  double lh, ll;
  log_dd(x, lh, ll);
  if (abs_d(y) > 1e290) return sgn * exp(y * lh);
  double p, pe;
  two_prod(y, lh, p, pe);
  double plo = pe + y * ll;
  double e = exp(p);
  if (is_inf(e) || e == 0.0) return sgn * exp(p + plo);
  return sgn * (e + e * plo);
}

// atan(x)
MFC_HD double atan(double x) {
  if (is_nan(x)) return x;
  double a = abs_d(x);
  bool inv = a > 1.0;
  if (inv) a = 1.0 / a;
  bool red = a > TAN_PI12;
  if (red) a = (a * SQRT3 - 1.0) / (a + SQRT3);
  double z = a * a;
  // Q(z) = sum_{k=1..14} (-1)^k z^{k-1}/(2k+1);  atan(a) = a + a z Q(z)
  double q = 0.034482758620689655;              // +1/29
  q = -0.037037037037037035 + z * q;            // -1/27
  q = 0.04 + z * q;                             // +1/25
  q = -0.043478260869565216 + z * q;            // -1/23
  q = 0.047619047619047616 + z * q;             // +1/21
  q = -0.05263157894736842 + z * q;             // -1/19
  q = 0.058823529411764705 + z * q;             // +1/17
  q = -0.06666666666666667 + z * q;             // -1/15
  q = 0.07692307692307693 + z * q;              // +1/13
  q = -0.09090909090909091 + z * q;             // -1/11
  q = 0.1111111111111111 + z * q;               // +1/9
  q = -0.14285714285714285 + z * q;             // -1/7
  q = 0.2 + z * q;                              // +1/5
  q = -0.3333333333333333 + z * q;              // -1/3
  double t = a + a * (z * q);
  if (red) t = PIO6 + (t + PIO6_LO);
  if (inv) t = PIO2 - (t - PIO2_LO);
  return sign_bit(x) ? -t : t;
}

// atan2(y, x) with the IEEE 754 special cases
MFC_HD double atan2(double y, double x) {
  if (is_nan(x) || is_nan(y)) return x + y;
  bool sy = sign_bit(y);
  if (y == 0.0) {
    if (sign_bit(x)) return sy ? -PI : PI;
    return y;
  }
  if (x == 0.0) return sy ? -PIO2 : PIO2;
  if (is_inf(x)) {
    if (is_inf(y)) {
      double t = (x > 0.0) ? PIO4 : 3.0 * PIO4;
      return sy ? -t : t;
    }
    double t = (x > 0.0) ? 0.0 : PI;
    return sy ? -t : t;
  }
  if (is_inf(y)) return sy ? -PIO2 : PIO2;
  double t = atan(abs_d(y / x));
  if (x < 0.0) t = PI - (t - PI_LO);
  return sy ? -t : t;
}

// r reduced argument, q quadrant: x = q pi/2 + r, |r| <= ~pi/4
MFC_HD double reduce_pio2(double x, int &q) {
  if (abs_d(x) <= PIO4) { q = 0; return x; }
  double kd = floor(x * TWO_OVER_PI + 0.5);
  q = ((int)((int64_t)kd & 3));
  return ((x - kd * PIO2_HI) - kd * PIO2_MID) - kd * PIO2_LO3;
}

MFC_HD double sin_poly(double r) {
  double z = r * r;
  // sum_{k=1..8} (-1)^k z^{k-1} /(2k+1)!
  double p = 2.8114572543455206e-15;            // +1/17!
  p = -7.647163731819816e-13 + z * p;           // -1/15!
  p = 1.6059043836821613e-10 + z * p;           // +1/13!
  p = -2.505210838544172e-08 + z * p;           // -1/11!
  p = 2.7557319223985893e-06 + z * p;           // +1/9!
  p = -0.0001984126984126984 + z * p;           // -1/7!
  p = 0.008333333333333333 + z * p;             // +1/5!
  p = -0.16666666666666666 + z * p;             // -1/3!
  return r + r * (z * p);
}

MFC_HD double cos_poly(double r) {
  double z = r * r;
  // cos r = 1 - z/2 + z^2 * sum_{k=2..9} (-1)^k z^{k-2}/(2k)!
  double p = -1.5619206968586225e-16;           // -1/18!
  p = 4.779477332387385e-14 + z * p;            // +1/16!
  p = -1.1470745597729725e-11 + z * p;          // -1/14!
  p = 2.08767569878681e-09 + z * p;             // +1/12!
  p = -2.755731922398589e-07 + z * p;           // -1/10!
  p = 2.48015873015873e-05 + z * p;             // +1/8!
  p = -0.001388888888888889 + z * p;            // -1/6!
  p = 0.041666666666666664 + z * p;             // +1/4!
  double hz = 0.5 * z;
  double w = 1.0 - hz;
  return w + (((1.0 - w) - hz) + (z * z) * p);
}

MFC_HD double sin(double x) {
  if (is_nan(x) || is_inf(x)) return u2d(0x7ff8000000000000ULL);
  int q; double r = reduce_pio2(x, q);
  switch (q) {
    case 0: return sin_poly(r);
    case 1: return cos_poly(r);
    case 2: return -sin_poly(r);
    default: return -cos_poly(r);
  }
}

MFC_HD double cos(double x) {
  if (is_nan(x) || is_inf(x)) return u2d(0x7ff8000000000000ULL);
  int q; double r = reduce_pio2(x, q);
  switch (q) {
    case 0: return cos_poly(r);
    case 1: return -sin_poly(r);
    case 2: return -cos_poly(r);
    default: return sin_poly(r);
  }
}

// e^u - 1 by its Taylor series, for |u| <= 1.2.
MFC_HD double expm1_series(double u) {
  double p = 1.9572941063391263e-20;            // 1/21!
  p = 4.110317623312165e-19 + u * p;            // 1/20!
  p = 8.22063524662433e-18 + u * p;             // 1/19!
  p = 1.5619206968586225e-16 + u * p;           // 1/18!
  p = 2.8114572543455206e-15 + u * p;           // 1/17!
  p = 4.779477332387385e-14 + u * p;            // 1/16!
  p = 7.647163731819816e-13 + u * p;            // 1/15!
  p = 1.1470745597729725e-11 + u * p;           // 1/14!
  p = 1.6059043836821613e-10 + u * p;           // 1/13!
  p = 2.08767569878681e-09 + u * p;             // 1/12!
  p = 2.505210838544172e-08 + u * p;            // 1/11!
  p = 2.755731922398589e-07 + u * p;            // 1/10!
  p = 2.7557319223985893e-06 + u * p;           // 1/9!
  p = 2.48015873015873e-05 + u * p;             // 1/8!
  p = 0.0001984126984126984 + u * p;            // 1/7!
  p = 0.001388888888888889 + u * p;             // 1/6!
  p = 0.008333333333333333 + u * p;             // 1/5!
  p = 0.041666666666666664 + u * p;             // 1/4!
  p = 0.16666666666666666 + u * p;              // 1/3!
  p = 0.5 + u * p;                              // 1/2!
  return u + (u * u) * p;
}

// tanh(x) = sign(x) em1/(em1 + 2), em1 = e^{2|x|} - 1
MFC_HD double tanh(double x) {
  if (is_nan(x)) return x;
  double a = abs_d(x);
  double t;
  if (a > 22.0) t = 1.0;
  else if (a < 0.55) {
    double em1 = expm1_series(2.0 * a);
    t = em1 / (em1 + 2.0);
  } else {
    double e = exp(2.0 * a);
    t = 1.0 - 2.0 / (e + 1.0);
  }
  return sign_bit(x) ? -t : t;
}

// single-precision entry points as called by the reference with double arguments
MFC_HD float atan2f(double y, double x) { return (float)atan2((double)(float)y, (double)(float)x); }
MFC_HD float cosf(double x) { return (float)cos((double)(float)x); }
MFC_HD float sinf(double x) { return (float)sin((double)(float)x); }

}  // namespace mfcm
