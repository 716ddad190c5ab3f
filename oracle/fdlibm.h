// FDLIBM 5.3 sin / cos / sincos for the CPU oracle (test infrastructure; never linked into the product).
//
// The reference evaluates every angle of the Newton step through its FDLIBM port
// (src/GridSim.Core/Numerics/Fdlibm.Trig.cs:133-159 __kernel_sin, :162-197 __kernel_cos, :345-439 rem_pio2,
// :486-637 __kernel_rem_pio2, :639-724 sin/cos/sincos; DetMath.cs:71 SinCos), which is pinned bit for bit to
// Java StrictMath (= FDLIBM 5.3). This file restates Sun's published FDLIBM 5.3 algorithm: plain IEEE double
// arithmetic, no fused multiply-add (build with -ffp-contract=off). The reduction tables are derived from pi by
// oracle/gen_fdlibm_tables.py; the kernel polynomial coefficients are FDLIBM's published minimax constants.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace orc_fdlibm {

static inline int32_t hi_word(double x) { uint64_t u; std::memcpy(&u, &x, 8); return (int32_t)(u >> 32); }
static inline uint32_t lo_word(double x) { uint64_t u; std::memcpy(&u, &x, 8); return (uint32_t)u; }
static inline double from_words(int32_t hi, uint32_t lo) {
  uint64_t u = ((uint64_t)(uint32_t)hi << 32) | lo; double d; std::memcpy(&d, &u, 8); return d;
}
static inline double from_bits(uint64_t u) { double d; std::memcpy(&d, &u, 8); return d; }

// k_sin.c / k_cos.c polynomial coefficients.
static const double S1 = from_bits(0xBFC5555555555549ULL), S2 = from_bits(0x3F8111111110F8A6ULL),
                    S3 = from_bits(0xBF2A01A019C161D5ULL), S4 = from_bits(0x3EC71DE357B1FE7DULL),
                    S5 = from_bits(0xBE5AE5E68A2B9CEBULL), S6 = from_bits(0x3DE5D93A5ACFD57CULL);
static const double C1 = from_bits(0x3FA555555555554CULL), C2 = from_bits(0xBF56C16C16C15177ULL),
                    C3 = from_bits(0x3EFA01A019CB1590ULL), C4 = from_bits(0xBE927E4F809C52ADULL),
                    C5 = from_bits(0x3E21EE9EBDB4B1C4ULL), C6 = from_bits(0xBDA8FAE9BE8838D4ULL);

// e_rem_pio2.c medium-path constants (derived: gen_fdlibm_tables.medium_constants).
static const double invpio2 = from_bits(0x3FE45F306DC9C883ULL);
static const double pio2_1 = from_bits(0x3FF921FB54400000ULL), pio2_1t = from_bits(0x3DD0B4611A626331ULL);
static const double pio2_2 = from_bits(0x3DD0B4611A600000ULL), pio2_2t = from_bits(0x3BA3198A2E037073ULL);
static const double pio2_3 = from_bits(0x3BA3198A2E000000ULL), pio2_3t = from_bits(0x397B839A252049C1ULL);
static const double two24 = 16777216.0, twon24 = 5.9604644775390625e-08;

// High words of the doubles nearest n*pi/2, n = 1..32 (gen_fdlibm_tables.npio2_hw).
static const int32_t npio2_hw[32] = {
    0x3FF921FB, 0x400921FB, 0x4012D97C, 0x401921FB, 0x401F6A7A, 0x4022D97C, 0x4025FDBB, 0x402921FB,
    0x402C463A, 0x402F6A7A, 0x4031475C, 0x4032D97C, 0x40346B9C, 0x4035FDBB, 0x40378FDB, 0x403921FB,
    0x403AB41B, 0x403C463A, 0x403DD85A, 0x403F6A7A, 0x40407E4C, 0x4041475C, 0x4042106C, 0x4042D97C,
    0x4043A28C, 0x40446B9C, 0x404534AC, 0x4045FDBB, 0x4046C6CB, 0x40478FDB, 0x404858EB, 0x404921FB};

// 24-bit chunks of the binary expansion of 2/pi (gen_fdlibm_tables.two_over_pi_chunks).
static const int32_t two_over_pi[66] = {
    0xA2F983, 0x6E4E44, 0x1529FC, 0x2757D1, 0xF534DD, 0xC0DB62, 0x95993C, 0x439041, 0xFE5163, 0xABDEBB, 0xC561B7,
    0x246E3A, 0x424DD2, 0xE00649, 0x2EEA09, 0xD1921C, 0xFE1DEB, 0x1CB129, 0xA73EE8, 0x8235F5, 0x2EBB44, 0x84E99C,
    0x7026B4, 0x5F7E41, 0x3991D6, 0x398353, 0x39F49C, 0x845F8B, 0xBDF928, 0x3B1FF8, 0x97FFDE, 0x05980F, 0xEF2F11,
    0x8B5A0A, 0x6D1F6D, 0x367ECF, 0x27CB09, 0xB74F46, 0x3F669E, 0x5FEA2D, 0x7527BA, 0xC7EBE5, 0xF17B3D, 0x0739F7,
    0x8A5292, 0xEA6BFB, 0x5FB11F, 0x8D5D08, 0x560330, 0x46FC7B, 0x6BABF0, 0xCFBC20, 0x9AF436, 0x1DA9E3, 0x91615E,
    0xE61B08, 0x659985, 0x5F14A0, 0x68408D, 0xFFD880, 0x4D7327, 0x310606, 0x1556CA, 0x73A8C9, 0x60E27B, 0xC08C6B};

// pi/2 in 24-bit pieces (gen_fdlibm_tables.pio2_chunks).
static const double PIo2[8] = {from_bits(0x3FF921FB40000000ULL), from_bits(0x3E74442D00000000ULL),
                               from_bits(0x3CF8469880000000ULL), from_bits(0x3B78CC5160000000ULL),
                               from_bits(0x39F01B8380000000ULL), from_bits(0x387A252040000000ULL),
                               from_bits(0x36E3822280000000ULL), from_bits(0x3569F31D00000000ULL)};

static inline double kernel_sin(double x, double y, int iy) {
  int32_t ix = hi_word(x) & 0x7fffffff;
  if (ix < 0x3e400000) { if ((int)x == 0) return x; }
  double z = x * x;
  double v = z * x;
  double r = S2 + z * (S3 + z * (S4 + z * (S5 + z * S6)));
  if (iy == 0) return x + v * (S1 + z * r);
  return x - ((z * (0.5 * y - v * r) - y) - v * S1);
}

static inline double kernel_cos(double x, double y) {
  int32_t ix = hi_word(x) & 0x7fffffff;
  if (ix < 0x3e400000) { if ((int)x == 0) return 1.0; }
  double z = x * x;
  double r = z * (C1 + z * (C2 + z * (C3 + z * (C4 + z * (C5 + z * C6)))));
  if (ix < 0x3FD33333) return 1.0 - (0.5 * z - (z * r - x * y));
  double qx = ix > 0x3fe90000 ? 0.28125 : from_words(ix - 0x00200000, 0);
  double hz = 0.5 * z - qx;
  double a = 1.0 - qx;
  return a - (hz - (z * r - x * y));
}

// k_rem_pio2.c with prec = 2 (53-bit result in y[0] + y[1]); returns n mod 8.
static inline int kernel_rem_pio2(const double* x, double* y, int e0, int nx) {
  const int jk = 4, jp = 4;  // init_jk[prec = 2]
  int jz, jx, jv, carry, n, iq[20], i, j, k, m, q0, ih;
  double z, fw, f[20], fq[20], q[20];
  jx = nx - 1;
  jv = (e0 - 3) / 24; if (jv < 0) jv = 0;
  q0 = e0 - 24 * (jv + 1);
  j = jv - jx; m = jx + jk;
  for (i = 0; i <= m; i++, j++) f[i] = (j < 0) ? 0.0 : (double)two_over_pi[j];
  for (i = 0; i <= jk; i++) {
    for (j = 0, fw = 0.0; j <= jx; j++) fw += x[j] * f[jx + i - j];
    q[i] = fw;
  }
  jz = jk;
  for (;;) {
    for (i = 0, j = jz, z = q[jz]; j > 0; i++, j--) {
      fw = (double)((int32_t)(twon24 * z));
      iq[i] = (int32_t)(z - two24 * fw);
      z = q[j - 1] + fw;
    }
    z = std::scalbn(z, q0);
    z -= 8.0 * std::floor(z * 0.125);
    n = (int32_t)z;
    z -= (double)n;
    ih = 0;
    if (q0 > 0) {
      i = (iq[jz - 1] >> (24 - q0)); n += i;
      iq[jz - 1] -= i << (24 - q0);
      ih = iq[jz - 1] >> (23 - q0);
    } else if (q0 == 0) {
      ih = iq[jz - 1] >> 23;
    } else if (z >= 0.5) {
      ih = 2;
    }
    if (ih > 0) {
      n += 1; carry = 0;
      for (i = 0; i < jz; i++) {
        j = iq[i];
        if (carry == 0) {
          if (j != 0) { carry = 1; iq[i] = 0x1000000 - j; }
        } else {
          iq[i] = 0xffffff - j;
        }
      }
      if (q0 > 0) {
        if (q0 == 1) iq[jz - 1] &= 0x7fffff;
        else if (q0 == 2) iq[jz - 1] &= 0x3fffff;
      }
      if (ih == 2) {
        z = 1.0 - z;
        if (carry != 0) z -= std::scalbn(1.0, q0);
      }
    }
    if (z == 0.0) {
      j = 0;
      for (i = jz - 1; i >= jk; i--) j |= iq[i];
      if (j == 0) {
        for (k = 1; iq[jk - k] == 0; k++) {}
        for (i = jz + 1; i <= jz + k; i++) {
          f[jx + i] = (double)two_over_pi[jv + i];
          for (j = 0, fw = 0.0; j <= jx; j++) fw += x[j] * f[jx + i - j];
          q[i] = fw;
        }
        jz += k;
        continue;  // recompute
      }
    }
    break;
  }
  if (z == 0.0) {
    jz -= 1; q0 -= 24;
    while (iq[jz] == 0) { jz--; q0 -= 24; }
  } else {
    z = std::scalbn(z, -q0);
    if (z >= two24) {
      fw = (double)((int32_t)(twon24 * z));
      iq[jz] = (int32_t)(z - two24 * fw);
      jz += 1; q0 += 24;
      iq[jz] = (int32_t)fw;
    } else {
      iq[jz] = (int32_t)z;
    }
  }
  fw = std::scalbn(1.0, q0);
  for (i = jz; i >= 0; i--) { q[i] = fw * (double)iq[i]; fw *= twon24; }
  for (i = jz; i >= 0; i--) {
    for (fw = 0.0, k = 0; k <= jp && k <= jz - i; k++) fw += PIo2[k] * q[i + k];
    fq[jz - i] = fw;
  }
  fw = 0.0;
  for (i = jz; i >= 0; i--) fw += fq[i];
  y[0] = (ih == 0) ? fw : -fw;
  fw = fq[0] - fw;
  for (i = 1; i <= jz; i++) fw += fq[i];
  y[1] = (ih == 0) ? fw : -fw;
  return n & 7;
}

// e_rem_pio2.c: x = n*pi/2 + (y[0] + y[1]), |y| <= pi/4 (approximately); returns n.
static inline int rem_pio2(double x, double* y) {
  int32_t hx = hi_word(x);
  int32_t ix = hx & 0x7fffffff;
  if (ix <= 0x3fe921fb) { y[0] = x; y[1] = 0; return 0; }
  if (ix < 0x4002d97c) {
    double z;
    if (hx > 0) {
      z = x - pio2_1;
      if (ix != 0x3ff921fb) { y[0] = z - pio2_1t; y[1] = (z - y[0]) - pio2_1t; }
      else { z -= pio2_2; y[0] = z - pio2_2t; y[1] = (z - y[0]) - pio2_2t; }
      return 1;
    }
    z = x + pio2_1;
    if (ix != 0x3ff921fb) { y[0] = z + pio2_1t; y[1] = (z - y[0]) + pio2_1t; }
    else { z += pio2_2; y[0] = z + pio2_2t; y[1] = (z - y[0]) + pio2_2t; }
    return -1;
  }
  if (ix <= 0x413921fb) {
    double t = std::fabs(x);
    int32_t n = (int32_t)(t * invpio2 + 0.5);
    double fn = (double)n;
    double r = t - fn * pio2_1;
    double w = fn * pio2_1t;
    if (n < 32 && ix != npio2_hw[n - 1]) {
      y[0] = r - w;
    } else {
      int32_t j = ix >> 20;
      y[0] = r - w;
      int32_t i = j - ((hi_word(y[0]) >> 20) & 0x7ff);
      if (i > 16) {
        t = r;
        w = fn * pio2_2;
        r = t - w;
        w = fn * pio2_2t - ((t - r) - w);
        y[0] = r - w;
        i = j - ((hi_word(y[0]) >> 20) & 0x7ff);
        if (i > 49) {
          t = r;
          w = fn * pio2_3;
          r = t - w;
          w = fn * pio2_3t - ((t - r) - w);
          y[0] = r - w;
        }
      }
    }
    y[1] = (r - y[0]) - w;
    if (hx < 0) { y[0] = -y[0]; y[1] = -y[1]; return -n; }
    return n;
  }
  if (ix >= 0x7ff00000) { y[0] = y[1] = x - x; return 0; }
  // large arguments: |x| split into three 24-bit pieces scaled by 2^-e0
  double tx[3];
  int32_t e0 = (ix >> 20) - 1046;
  double z = from_words(ix - e0 * (1 << 20), lo_word(x));   // (e0 may be negative: no left shift of it)
  for (int i = 0; i < 2; i++) { tx[i] = (double)((int32_t)z); z = (z - tx[i]) * two24; }
  tx[2] = z;
  int nx = 3;
  while (tx[nx - 1] == 0.0) nx--;
  int n = kernel_rem_pio2(tx, y, e0, nx);
  if (hx < 0) { y[0] = -y[0]; y[1] = -y[1]; return -n; }
  return n;
}

static inline void sincos(double x, double* s, double* c) {
  int32_t ix = hi_word(x) & 0x7fffffff;
  if (ix <= 0x3fe921fb) { *s = kernel_sin(x, 0.0, 0); *c = kernel_cos(x, 0.0); return; }
  if (ix >= 0x7ff00000) { *s = *c = x - x; return; }
  double y[2];
  int n = rem_pio2(x, y);
  switch (n & 3) {
    case 0: *s = kernel_sin(y[0], y[1], 1); *c = kernel_cos(y[0], y[1]); break;
    case 1: *s = kernel_cos(y[0], y[1]); *c = -kernel_sin(y[0], y[1], 1); break;
    case 2: *s = -kernel_sin(y[0], y[1], 1); *c = -kernel_cos(y[0], y[1]); break;
    default: *s = -kernel_cos(y[0], y[1]); *c = kernel_sin(y[0], y[1], 1); break;
  }
}

static inline double sin(double x) { double s, c; sincos(x, &s, &c); return s; }
static inline double cos(double x) { double s, c; sincos(x, &s, &c); return c; }

}  // namespace orc_fdlibm
