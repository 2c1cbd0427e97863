/*
 * bmx_math.h -- sin, cos, acos and pow built from IEEE +, -, *, /, sqrt,
 * fma, floor, frexp and ldexp only (each exactly or correctly rounded), so
 * that a host build and a gfx950 build of the same call give the same bits
 * (with floating-point contraction off).  The segment force
 * law turns orientations (theta, phi) into axes with sin/cos and the update
 * turns axes back with acos (src/chemistry/bmx_cell_interaction_K.H:559-612,
 * src/des/bmx_pc_interaction.cpp:337-429); nearly parallel segments make the
 * closest-point split ill-conditioned, so a one-ulp difference in an axis can
 * move a force by 1e-11 of its size.  Using one implementation on both sides
 * keeps the segment path bit-reproducible.
 *
 * Method: Cody-Waite reduction by pi/2 (three-part constant, the first two
 * parts 33 significant bits so k*C1 and k*C2 are exact for |k| < 2^20),
 * Taylor polynomials of sin (to r^17) and cos (to r^18) on [-pi/4, pi/4] in
 * Horner form; acos through the arcsine series (to y^61) on |y| <= 1/2 with
 * acos(x) = 2 asin(sqrt((1-x)/2)) for x > 1/2 and pi - 2 asin(sqrt((1+x)/2))
 * for x < -1/2.  Coefficients are the exact series terms 1/n! and
 * (2n)! / (4^n (n!)^2 (2n+1)) rounded to double.  Valid for |x| < 1e6;
 * agreement with glibc is checked by the tests to a few ulp.
 */
#ifndef BMX_MATH_H
#define BMX_MATH_H

#include <math.h>

#if defined(__HIPCC__)
#define BMX_MATH_FN __host__ __device__ static inline
#else
#define BMX_MATH_FN static inline
#endif

#define BMX_PIO2_1 0x1.921fb544p+0
#define BMX_PIO2_2 0x1.0b4611a6p-34
#define BMX_PIO2_3 0x1.3198a2e037073p-69
#define BMX_2_OVER_PI 0x1.45f306dc9c883p-1
#define BMX_PIO2 0x1.921fb54442d18p+0
#define BMX_PI 0x1.921fb54442d18p+1

/* sin r for |r| <= pi/4 */
BMX_MATH_FN double bmx_sin_kernel(double r) {
  double z = r * r;
  double p = 2.8114572543455206e-15;   /* 1/17! */
  p = p * z - 7.647163731819816e-13;   /* 1/15! */
  p = p * z + 1.6059043836821613e-10;  /* 1/13! */
  p = p * z - 2.505210838544172e-08;   /* 1/11! */
  p = p * z + 2.7557319223985893e-06;  /* 1/9! */
  p = p * z - 0.0001984126984126984;   /* 1/7! */
  p = p * z + 0.008333333333333333;    /* 1/5! */
  p = p * z - 0.16666666666666666;     /* 1/3! */
  return r + r * z * p;
}

/* cos r for |r| <= pi/4 */
BMX_MATH_FN double bmx_cos_kernel(double r) {
  double z = r * r;
  double p = -1.5619206968586225e-16;  /* 1/18! */
  p = p * z + 4.779477332387385e-14;   /* 1/16! */
  p = p * z - 1.1470745597729725e-11;  /* 1/14! */
  p = p * z + 2.08767569878681e-09;    /* 1/12! */
  p = p * z - 2.755731922398589e-07;   /* 1/10! */
  p = p * z + 2.48015873015873e-05;    /* 1/8! */
  p = p * z - 0.001388888888888889;    /* 1/6! */
  p = p * z + 0.041666666666666664;    /* 1/4! */
  return (1.0 - 0.5 * z) + z * z * p;
}

/* x = k pi/2 + r, |r| <= pi/4 (plus rounding); returns r, *q = k mod 4 */
BMX_MATH_FN double bmx_reduce_pio2(double x, int *q) {
  double k = floor(x * BMX_2_OVER_PI + 0.5);
  double r = ((x - k * BMX_PIO2_1) - k * BMX_PIO2_2) - k * BMX_PIO2_3;
  *q = (int)((long long)k & 3);
  return r;
}

BMX_MATH_FN double bmx_sin(double x) {
  int q;
  double r = bmx_reduce_pio2(x, &q);
  switch (q) {
    case 0: return bmx_sin_kernel(r);
    case 1: return bmx_cos_kernel(r);
    case 2: return -bmx_sin_kernel(r);
    default: return -bmx_cos_kernel(r);
  }
}

BMX_MATH_FN double bmx_cos(double x) {
  int q;
  double r = bmx_reduce_pio2(x, &q);
  switch (q) {
    case 0: return bmx_cos_kernel(r);
    case 1: return -bmx_sin_kernel(r);
    case 2: return -bmx_cos_kernel(r);
    default: return bmx_sin_kernel(r);
  }
}

/* asin y for |y| <= 1/2 */
BMX_MATH_FN double bmx_asin_kernel(double y) {
  double z = y * y;
  double p = 0.0017680811205154183;
  p = p * z + 0.0018622264064031275;
  p = p * z + 0.0019650336162772837;
  p = p * z + 0.0020776610325181676;
  p = p * z + 0.0022014739737101384;
  p = p * z + 0.002338091892111975;
  p = p * z + 0.0024894486782468836;
  p = p * z + 0.00265787063820729;
  p = p * z + 0.002846178401108942;
  p = p * z + 0.0030578216492580306;
  p = p * z + 0.003297059503473485;
  p = p * z + 0.0035692053938259347;
  p = p * z + 0.003880964558837669;
  p = p * z + 0.004240907093679363;
  p = p * z + 0.004660143486915096;
  p = p * z + 0.005153309682319905;
  p = p * z + 0.005740037670841924;
  p = p * z + 0.006447210311889649;
  p = p * z + 0.0073125258735988454;
  p = p * z + 0.008390335809616815;
  p = p * z + 0.009761609529194078;
  p = p * z + 0.011551800896139705;
  p = p * z + 0.01396484375;
  p = p * z + 0.017352764423076924;
  p = p * z + 0.022372159090909092;
  p = p * z + 0.030381944444444444;
  p = p * z + 0.044642857142857144;
  p = p * z + 0.075;
  p = p * z + 0.16666666666666666;
  return y + y * z * p;
}

BMX_MATH_FN double bmx_acos(double x) {
  if (!(x >= -1.0 && x <= 1.0)) return (x - x) / (x - x); /* NaN, as acos */
  if (x <= 0.5 && x >= -0.5) return BMX_PIO2 - bmx_asin_kernel(x);
  if (x > 0.0) return 2.0 * bmx_asin_kernel(sqrt((1.0 - x) * 0.5));
  return BMX_PI - 2.0 * bmx_asin_kernel(sqrt((1.0 + x) * 0.5));
}

/* c^3, the reference's pow(c_length, 3) (K.H:832-834) */
BMX_MATH_FN double bmx_cube(double c) { return c * c * c; }

/* ---- pow -----------------------------------------------------------------
 * x^y for finite x > 0, as exp(y log x) carried in double-double arithmetic
 * (a value hi + lo with |lo| <= ulp(hi)/2; Dekker/Knuth error-free sums and
 * fma products, so every step is IEEE-rounded the same way on both sides).
 * log: x = m 2^e with m in [1/sqrt2, sqrt2), s = (m-1)/(m+1), |s| < 0.172,
 * log m = 2 atanh s = 2s + (2/3)s^3 + (2/5)s^5 + 2 s^7 sum_k s^2k/(2k+7):
 * the first three terms in double-double, the rest (< 1.3e-6) in double.
 * exp: t = k ln2 + r, |r| <= ln2/2, q = r/64; expm1 q = q + q^2/2 (double-
 * double) + q^3 (1/3! + ... + q^5/8!) (double), then six doublings
 * expm1 2a = expm1 a (expm1 a + 2), which keep the relative error.  The
 * error of y log x stays below ~1e-21 and the result is the double nearest
 * the double-double value, so it agrees with a correctly rounded pow except
 * in rare near-halfway cases. */
typedef struct {
  double hi, lo;
} bmx_dd;

BMX_MATH_FN bmx_dd bmx_dd_make(double hi, double lo) {
  bmx_dd r;
  r.hi = hi;
  r.lo = lo;
  return r;
}
/* a + b exactly, any magnitudes */
BMX_MATH_FN bmx_dd bmx_dd_two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  double e = (a - (s - bb)) + (b - bb);
  return bmx_dd_make(s, e);
}
/* a + b exactly, |a| >= |b| */
BMX_MATH_FN bmx_dd bmx_dd_fast_sum(double a, double b) {
  double s = a + b;
  double e = b - (s - a);
  return bmx_dd_make(s, e);
}
BMX_MATH_FN bmx_dd bmx_dd_add(bmx_dd a, bmx_dd b) {
  bmx_dd s = bmx_dd_two_sum(a.hi, b.hi);
  bmx_dd t = bmx_dd_two_sum(a.lo, b.lo);
  double e = s.lo + t.hi;
  bmx_dd u = bmx_dd_fast_sum(s.hi, e);
  e = u.lo + t.lo;
  return bmx_dd_fast_sum(u.hi, e);
}
/* a + b for a double b */
BMX_MATH_FN bmx_dd bmx_dd_add_d(bmx_dd a, double b) {
  bmx_dd s = bmx_dd_two_sum(a.hi, b);
  double e = s.lo + a.lo;
  return bmx_dd_fast_sum(s.hi, e);
}
/* a * b for a double b */
BMX_MATH_FN bmx_dd bmx_dd_mul_d(bmx_dd a, double b) {
  double p = a.hi * b;
  double e = fma(a.hi, b, -p);
  e = e + a.lo * b;
  return bmx_dd_fast_sum(p, e);
}
BMX_MATH_FN bmx_dd bmx_dd_mul(bmx_dd a, bmx_dd b) {
  double p = a.hi * b.hi;
  double e = fma(a.hi, b.hi, -p);
  double c = a.hi * b.lo;
  double d = a.lo * b.hi;
  e = e + (c + d);
  return bmx_dd_fast_sum(p, e);
}
BMX_MATH_FN bmx_dd bmx_dd_div(bmx_dd a, bmx_dd b) {
  double q1 = a.hi / b.hi;
  bmx_dd r = bmx_dd_add(a, bmx_dd_mul_d(b, -q1));
  double q2 = r.hi / b.hi;
  r = bmx_dd_add(r, bmx_dd_mul_d(b, -q2));
  double q3 = r.hi / b.hi;
  bmx_dd q = bmx_dd_fast_sum(q1, q2);
  return bmx_dd_add_d(q, q3);
}

#define BMX_LN2_HI 0x1.62e42fefa39efp-1
#define BMX_LN2_LO 0x1.abc9e3b39803fp-56

/* log x for finite x > 0 (normal range) */
BMX_MATH_FN bmx_dd bmx_log_dd(double x) {
  int ex;
  double m = frexp(x, &ex); /* [0.5, 1) */
  if (m < 0.70710678118654752) {
    m = m * 2.0;
    ex = ex - 1;
  }
  bmx_dd s = bmx_dd_div(bmx_dd_make(m - 1.0, 0.0), bmx_dd_two_sum(m, 1.0));
  bmx_dd z = bmx_dd_mul(s, s);
  bmx_dd s3 = bmx_dd_mul(s, z);
  bmx_dd s5 = bmx_dd_mul(s3, z);
  double q = 0.037037037037037035;
  q = q * z.hi + 0.04;
  q = q * z.hi + 0.043478260869565216;
  q = q * z.hi + 0.047619047619047616;
  q = q * z.hi + 0.05263157894736842;
  q = q * z.hi + 0.058823529411764705;
  q = q * z.hi + 0.06666666666666667;
  q = q * z.hi + 0.07692307692307693;
  q = q * z.hi + 0.09090909090909091;
  q = q * z.hi + 0.1111111111111111;
  q = q * z.hi + 0.14285714285714285;
  double tail = 2.0 * s5.hi * z.hi * q;
  bmx_dd lm = bmx_dd_add(bmx_dd_make(2.0 * s.hi, 2.0 * s.lo),
                         bmx_dd_mul(s3, bmx_dd_make(0.6666666666666666, 3.700743415417188e-17)));
  lm = bmx_dd_add(lm, bmx_dd_mul(s5, bmx_dd_make(0.4, -2.2204460492503132e-17)));
  lm = bmx_dd_add_d(lm, tail);
  double k = (double)ex;
  bmx_dd kl = bmx_dd_fast_sum(k * BMX_LN2_HI, fma(k, BMX_LN2_HI, -(k * BMX_LN2_HI)));
  kl = bmx_dd_add_d(kl, k * BMX_LN2_LO);
  return bmx_dd_add(kl, lm);
}

/* exp t for a double-double t with |t| < 700 */
BMX_MATH_FN double bmx_exp_dd(bmx_dd t) {
  double k = floor(t.hi * 1.4426950408889634 + 0.5);
  bmx_dd kl = bmx_dd_two_sum(-(k * BMX_LN2_HI), -fma(k, BMX_LN2_HI, -(k * BMX_LN2_HI)));
  kl = bmx_dd_add_d(kl, -(k * BMX_LN2_LO));
  bmx_dd r = bmx_dd_add(t, kl);
  bmx_dd q = bmx_dd_make(r.hi * 0x1p-6, r.lo * 0x1p-6);
  double qh = q.hi;
  double p = 2.7557319223985893e-06;
  p = p * qh + 2.48015873015873e-05;
  p = p * qh + 0.0001984126984126984;
  p = p * qh + 0.001388888888888889;
  p = p * qh + 0.008333333333333333;
  p = p * qh + 0.041666666666666664;
  p = p * qh + 0.16666666666666666;
  bmx_dd q2 = bmx_dd_mul(q, q);
  bmx_dd e = bmx_dd_add(q, bmx_dd_make(0.5 * q2.hi, 0.5 * q2.lo));
  e = bmx_dd_add_d(e, qh * qh * qh * p);
  for (int i = 0; i < 6; i++) e = bmx_dd_mul(e, bmx_dd_add_d(e, 2.0));
  e = bmx_dd_add_d(e, 1.0);
  return ldexp(e.hi + e.lo, (int)k);
}

#define BMX_THIRD 0x1.5555555555555p-2 /* the double nearest 1/3 */

/* x^y for y = BMX_THIRD (k = 1) or 2 BMX_THIRD (k = 2, the double nearest
 * 2/3), 2^-1000 < x < 2^1000.  That y is k/3 - k/(3 2^54) exactly, so
 * x^y = cbrt(x)^k exp(eta) with eta = -k ln(x) / (3 2^54), |eta| < 2^-45:
 *   - x = a 2^(3q) with a in [0.5, 4);
 *   - a^(-1/3) to 0.33% from a quadratic in the mantissa, then three Newton
 *     steps z += z (1 - a z^3) / 3 (the relative error e goes to 2 e^2);
 *   - c = a z^2 within a few ulp of cbrt(a), then one correction from the
 *     exact residual a - c^3 (two fma products; a - c^3 is exact by
 *     Sterbenz): cbrt(a) in double-double to about 2^-100;
 *   - ln x to about 1e-10 (atanh series to s^11/11, |s| < 0.172), which is
 *     all eta needs: its error reaches the result at below 2^-88;
 *   - the result cbrt(a)^k (1 + eta) as hi + lo, error below 2^-84 of it.
 * When hi + lo lies within 2^-70 of itself of a rounding boundary the
 * function returns 0 and bmx_pow takes the general double-double path, so
 * either way bmx_pow returns the value that path would (the two agree on
 * every tested argument).  About 1 call in 2^16 falls back. */
BMX_MATH_FN double bmx_pow_third(double x, int k) {
  int ex;
  const double m = frexp(x, &ex); /* [0.5, 1) */
  const int q = (ex + 1200) / 3 - 400; /* floor(ex / 3) */
  const int r = ex - 3 * q;            /* 0, 1, 2 */
  const double a = ldexp(m, r);
  double z = (0.4564207 * m - 1.19161852) * m + 1.73758345; /* m^(-1/3) */
  z = z * (r == 0 ? 1.0 : (r == 1 ? 0.79370052598409973 : 0.62996052494743658));
  for (int i = 0; i < 3; i++) {
    const double t = 1.0 - a * (z * z * z);
    z = z + z * t * BMX_THIRD;
  }
  const double z2 = z * z;
  const double c = a * z2;
  const double p = c * c, pe = fma(c, c, -p);
  const double c3 = p * c, c3e = fma(p, c, -c3);
  const double res = ((a - c3) - c3e) - pe * c;
  const bmx_dd cb = bmx_dd_fast_sum(c, res * z2 * BMX_THIRD);
  /* ln x = e2 ln 2 + ln mm, mm in [0.707, 1.414) */
  double mm = m;
  int e2 = ex;
  if (mm < 0.70710678118654752) {
    mm = mm * 2.0;
    e2 = e2 - 1;
  }
  const double s = (mm - 1.0) / (mm + 1.0), s2 = s * s;
  double sr = 0.090909090909090912; /* 1/11 */
  sr = sr * s2 + 0.1111111111111111;
  sr = sr * s2 + 0.14285714285714285;
  sr = sr * s2 + 0.2;
  sr = sr * s2 + BMX_THIRD;
  sr = sr * s2 + 1.0;
  const double lnx = (double)e2 * 0.69314718055994529 + 2.0 * s * sr;
  double hi, lo;
  if (k == 1) {
    const double eta = lnx * -BMX_THIRD * 0x1p-54;
    hi = cb.hi;
    lo = cb.lo + cb.hi * eta;
  } else {
    const double eta = lnx * (-2.0 * BMX_THIRD) * 0x1p-54;
    hi = cb.hi * cb.hi;
    lo = (fma(cb.hi, cb.hi, -hi) + 2.0 * cb.hi * cb.lo) + hi * eta;
  }
  const bmx_dd v = bmx_dd_fast_sum(hi, lo);
  /* spacing of the doubles above and below v.hi (> 0) */
  int eh;
  const double fh = frexp(v.hi, &eh);
  const double up = ldexp(1.0, eh - 53);
  const double dn = fh == 0.5 ? 0.5 * up : up;
  const double mg = v.hi * 0x1p-70;
  if (v.lo + mg < 0.5 * up && v.lo - mg > -0.5 * dn) return ldexp(v.hi, k * q);
  return 0.0;
}

BMX_MATH_FN double bmx_pow(double x, double y) {
  if (!(x > 0.0) || x - x != 0.0) return (x - x) / (x - x); /* NaN */
  if (x == 1.0 || y == 0.0) return 1.0;
  if ((y == BMX_THIRD || y == 2.0 * BMX_THIRD) && x > 0x1p-1000 && x < 0x1p1000) {
    const double f = bmx_pow_third(x, y == BMX_THIRD ? 1 : 2);
    if (f != 0.0) return f;
  }
  return bmx_exp_dd(bmx_dd_mul_d(bmx_log_dd(x), y));
}

#endif
