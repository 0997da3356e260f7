/*
 * basics.c -- fixed point, Threefry, units and cell geometry for the oracle.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine, never linked into
 * it.  Restates src/reference/fixed-point.ts, random.ts, engine/units.ts and
 * triclinic.ts of the upstream project.
 */
#include <math.h>
#include <string.h>

#include "oracle.h"

/* ---- fixed point --------------------------------------------------------- */

/* fixedWords: magnitude split, fraction rounded ties-to-even, negated as a
 * two's-complement 64-bit value (src/reference/fixed-point.ts:55-76). */
int64_t orc_to_fixed(float value) {
  double magnitude = fabs((double)value);
  double whole = floor(magnitude);
  double frac = magnitude - whole; /* exact */
  double lo = rint(frac * 4294967296.0); /* default mode: ties to even */
  int64_t m = (int64_t)whole * 4294967296LL + (int64_t)lo;
  return value < 0.0f ? -m : m;
}

void orc_fixed_words(float value, uint32_t *lo, int32_t *hi) {
  int64_t f = orc_to_fixed(value);
  *lo = (uint32_t)((uint64_t)f & 0xffffffffu);
  *hi = (int32_t)(f >> 32);
}

/* fromFixedF32: convert the magnitude word by word in f32, reapply the sign
 * (src/reference/fixed-point.ts:95-102). */
float orc_from_fixed_f32(int64_t fixed) {
  int negative = fixed < 0;
  uint64_t m = negative ? (uint64_t)(-(fixed + 1)) + 1u : (uint64_t)fixed;
  float hi = (float)(uint32_t)(m >> 32);
  float lo = (float)(uint32_t)(m & 0xffffffffu) * 2.3283064365386963e-10f;
  float value = hi + lo;
  return negative ? -value : value;
}

double orc_from_fixed(int64_t fixed) {
  int64_t whole = fixed / 4294967296LL;
  int64_t frac = fixed - whole * 4294967296LL;
  return (double)whole + (double)frac / 4294967296.0;
}

/* ---- Threefry-2x32 (Salmon et al., SC'11; src/reference/random.ts:63-93) -- */

static uint32_t rotl32(uint32_t x, int b) { return (x << b) | (x >> (32 - b)); }

void orc_threefry2x32(const uint32_t counter[2], const uint32_t key[2], int rounds,
                      uint32_t out[2]) {
  static const int rot[8] = {13, 15, 26, 6, 17, 29, 16, 24};
  uint32_t ks[3] = {key[0], key[1], 0x1bd11bdau ^ key[0] ^ key[1]};
  uint32_t x0 = counter[0] + ks[0];
  uint32_t x1 = counter[1] + ks[1];
  for (int r = 0; r < rounds; r++) {
    x0 += x1;
    x1 = rotl32(x1, rot[r % 8]) ^ x0;
    if ((r + 1) % 4 == 0) {
      uint32_t inj = (uint32_t)((r + 1) / 4);
      x0 += ks[inj % 3];
      x1 += ks[(inj + 1) % 3] + inj;
    }
  }
  out[0] = x0;
  out[1] = x1;
}

/* 23 random bits, midpoint of the bucket (src/reference/random.ts:121-122) */
double orc_open_unit_interval(uint32_t bits) {
  return ((double)(bits >> 9) + 0.5) * (1.0 / 8388608.0);
}

void orc_normal_pair(const uint32_t counter[2], const uint32_t key[2], double out[2]) {
  uint32_t bits[2];
  orc_threefry2x32(counter, key, 20, bits);
  double radius = sqrt(-2.0 * log(orc_open_unit_interval(bits[0])));
  double angle = 2.0 * M_PI * orc_open_unit_interval(bits[1]);
  out[0] = radius * cos(angle);
  out[1] = radius * sin(angle);
}

/* counter (2 particle + axis/2, step), key (seed, THERMOSTAT_DOMAIN)
 * (src/reference/random.ts:175-188) */
double orc_thermostat_normal(uint32_t seed, uint32_t step, uint32_t particle, int axis) {
  uint32_t counter[2] = {particle * 2u + (uint32_t)(axis >> 1), step};
  uint32_t key[2] = {seed, ORC_THERMOSTAT_DOMAIN};
  double pair[2];
  orc_normal_pair(counter, key, pair);
  return pair[axis & 1];
}

/* counter (2 particle + axis/2, seed), key (seed, domain)
 * (src/reference/random.ts:155-167) */
double orc_domain_normal(uint32_t domain, uint32_t seed, uint32_t particle, int axis) {
  uint32_t counter[2] = {particle * 2u + (uint32_t)(axis >> 1), seed};
  uint32_t key[2] = {seed, domain};
  double pair[2];
  orc_normal_pair(counter, key, pair);
  return pair[axis & 1];
}

/* ---- units (src/engine/units.ts:53-84) ----------------------------------- */

double orc_ewald_alpha(double cutoff, double tolerance) {
  return (1.0 / cutoff) * sqrt(-log(2.0 * tolerance));
}

int orc_pme_minimum_grid(double alpha, double box_length, double tolerance) {
  int g = (int)ceil((2.0 * alpha * box_length) / (3.0 * pow(tolerance, 0.2)));
  return g < 6 ? 6 : g;
}

int orc_smooth_fft_size(int minimum) {
  for (int n = minimum < 1 ? 1 : minimum;; n++) {
    int r = n;
    static const int primes[4] = {2, 3, 5, 7};
    for (int k = 0; k < 4; k++)
      while (r % primes[k] == 0) r /= primes[k];
    if (r == 1) return n;
  }
}

/* ---- cell geometry (src/reference/triclinic.ts) ------------------------- */

/* Math.round, which resolves ties upward */
static double js_round(double x) { return floor(x + 0.5); }

/* ties to even (src/reference/triclinic.ts:93-125) */
static double reduce_by(double ratio) {
  double nearest = js_round(ratio);
  if (fabs(ratio - trunc(ratio)) == 0.5 && fmod(nearest, 2.0) != 0.0)
    return nearest - (ratio > 0 ? 1.0 : (ratio < 0 ? -1.0 : 0.0));
  return nearest;
}

void orc_reduce_box(const double in[9], double out[9]) {
  double v[9];
  memcpy(v, in, sizeof v);
  double t = reduce_by(v[7] / v[4]);
  for (int k = 0; k < 3; k++) v[6 + k] -= t * v[3 + k];
  t = reduce_by(v[6] / v[0]);
  for (int k = 0; k < 3; k++) v[6 + k] -= t * v[k];
  t = reduce_by(v[3] / v[0]);
  for (int k = 0; k < 3; k++) v[3 + k] -= t * v[k];
  memcpy(out, v, sizeof v);
}

int orc_is_reduced(const double b[9], double tol) {
  if (fabs(b[1]) > tol || fabs(b[2]) > tol || fabs(b[5]) > tol) return 0;
  if (b[0] <= 0 || b[4] <= 0 || b[8] <= 0) return 0;
  return fabs(b[3]) <= 0.5 * b[0] + tol && fabs(b[6]) <= 0.5 * b[0] + tol &&
         fabs(b[7]) <= 0.5 * b[4] + tol;
}

/* f = r . M, M lower triangular (src/reference/triclinic.ts:394-411) */
void orc_inverse_box(const double b[9], double m[9]) {
  double ax = b[0], bx = b[3], by = b[4], cx = b[6], cy = b[7], cz = b[8];
  memset(m, 0, 9 * sizeof(double));
  m[0] = 1.0 / ax;
  m[3] = -bx / (ax * by);
  m[4] = 1.0 / by;
  m[6] = ((bx * cy) / (by * cz) - cx / cz) / ax;
  m[7] = -cy / (by * cz);
  m[8] = 1.0 / cz;
}

void orc_fractional(const double m[9], const double r[3], double f[3]) {
  f[0] = r[0] * m[0] + r[1] * m[3] + r[2] * m[6];
  f[1] = r[0] * m[1] + r[1] * m[4] + r[2] * m[7];
  f[2] = r[0] * m[2] + r[1] * m[5] + r[2] * m[8];
}

static void cross(const double *u, const double *v, double *w) {
  w[0] = u[1] * v[2] - u[2] * v[1];
  w[1] = u[2] * v[0] - u[0] * v[2];
  w[2] = u[0] * v[1] - u[1] * v[0];
}

double orc_box_volume(const double b[9]) {
  double w[3];
  cross(b + 3, b + 6, w);
  return fabs(b[0] * w[0] + b[1] * w[1] + b[2] * w[2]);
}

/* volume over face area (src/reference/triclinic.ts:38-59) */
void orc_perpendicular_widths(const double b[9], double widths[3]) {
  double v = orc_box_volume(b), w[3];
  cross(b + 3, b + 6, w);
  widths[0] = v / sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  cross(b + 6, b, w);
  widths[1] = v / sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  cross(b, b + 3, w);
  widths[2] = v / sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
}

/* reduce against c, then b, then a (src/reference/forces/nonbonded.ts:241-256) */
void orc_nearest_image(const double b[9], double d[3]) {
  double kc = js_round(d[2] / b[8]);
  d[0] -= kc * b[6];
  d[1] -= kc * b[7];
  d[2] -= kc * b[8];
  double kb = js_round(d[1] / b[4]);
  d[0] -= kb * b[3];
  d[1] -= kb * b[4];
  d[0] -= js_round(d[0] / b[0]) * b[0];
}
