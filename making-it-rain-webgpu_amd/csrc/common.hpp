// common.hpp -- shared definitions for the gfx950 molecular-dynamics core.
//
// Number formats (the upstream design, docs/PRECISION.md and
// src/gpu/kernels/fixed64.wgsl:1-93):
//   * positions and forces are Q32.32 two's-complement int64 (scale 2^32), so
//     every accumulation is integer addition: exact and order independent;
//   * velocities, masses and per-pair arithmetic are f32;
//   * f32 -> fixed rounds the fraction to nearest, ties to even, on the
//     magnitude, then negates (`to_fixed`, fixed64.wgsl:63-70);
//   * fixed -> f32 converts the magnitude word by word and reapplies the sign
//     (`from_fixed`, fixed64.wgsl:86-93).
// On gfx950 the reference's two-word carry arithmetic becomes native 64-bit
// integer arithmetic and native 64-bit atomics.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define MD_FIXED_SCALE 4294967296.0f
#define MD_FIXED_SCALE_INV 2.3283064365386963e-10f
#define MD_ONE_4PI_EPS0 138.93545764498457
#define MD_BOLTZMANN 0.008314462618153242
#define MD_THERMOSTAT_DOMAIN 0x9e3779b9u
#define MD_BAROSTAT_DOMAIN 0x6a09e667u
#define MD_VELOCITY_DOMAIN 0xbb67ae85u
#define MD_PME_ORDER 5
#define MD_CLUSTER 8            // atoms per cluster
#define MD_SUPER 64             // atoms per super-cluster (one wavefront: 8 clusters)
#define MD_SEGMENT 8            // list entries per wavefront work unit, at least
#define MD_MAX_CLUSTER_ATOMS 8  // SHAKE cluster limits (constraint-clusters.ts:68-69)
#define MD_MAX_CLUSTER_CONSTRAINTS 8
#define MD_MAX_CONSTRAINT_ITERATIONS 150

// Energy slots on the device, one per OpenMM force group the engine reports.
enum {
  MD_E_BOND = 0,
  MD_E_ANGLE = 1,
  MD_E_TORSION = 2,
  MD_E_PAIR = 3,       // direct-space pairs (erfc Coulomb + LJ)
  MD_E_EXCEPTION = 4,  // exclusion correction + exceptions
  MD_E_RECIPROCAL = 5, // mesh sum
  MD_E_RESTRAINT = 6,
  MD_E_SLOTS = 8
};

// ---- fixed point ----------------------------------------------------------

__host__ __device__ inline int64_t md_to_fixed(float x) {
  float m = fabsf(x);
  float w = floorf(m);
  float f = m - w;  // exact: both non-negative
  uint32_t lo = (uint32_t)rintf(f * MD_FIXED_SCALE);
  int64_t v = (int64_t)(((uint64_t)(uint32_t)w << 32) | lo);
  return x < 0.0f ? -v : v;
}

__host__ __device__ inline float md_from_fixed(int64_t v) {
  bool neg = v < 0;
  uint64_t m = neg ? (uint64_t)0 - (uint64_t)v : (uint64_t)v;
  float value = (float)(uint32_t)(m >> 32) + (float)(uint32_t)(m & 0xffffffffu) * MD_FIXED_SCALE_INV;
  return neg ? -value : value;
}

// f64 reading on the device (exact below 2^21 in magnitude)
__device__ inline double md_fixed_to_double_dev(int64_t v) { return (double)v * 2.3283064365386963e-10; }

// exact f64 reading (host, and device code that must agree with the host)
__host__ __device__ inline double md_fixed_to_double(int64_t v) {
  int64_t whole = v / 4294967296LL;
  int64_t frac = v - whole * 4294967296LL;
  return (double)whole + (double)frac / 4294967296.0;
}

// f64 -> fixed, for uploads of positions (round to nearest, ties to even;
// exact for values representable at 2^-32)
__host__ __device__ inline int64_t md_double_to_fixed(double x) {
  double s = x * 4294967296.0;
  return (int64_t)llrint(s);
}

// ---- Threefry-2x32-20 (Salmon, Moraes, Dror & Shaw, SC'11) ----------------

__host__ __device__ inline uint32_t md_rotl(uint32_t x, int b) { return (x << b) | (x >> (32 - b)); }

__host__ __device__ inline void md_threefry2x32_20(uint32_t c0, uint32_t c1, uint32_t k0, uint32_t k1,
                                                   uint32_t &o0, uint32_t &o1) {
  const uint32_t k2 = 0x1bd11bdau ^ k0 ^ k1;
  uint32_t x0 = c0 + k0, x1 = c1 + k1;
#define MD_TF_ROUND(r) \
  x0 += x1;            \
  x1 = md_rotl(x1, r) ^ x0;
  // rounds 1-4, rotations 13 15 26 6, then inject key 1
  MD_TF_ROUND(13) MD_TF_ROUND(15) MD_TF_ROUND(26) MD_TF_ROUND(6)
  x0 += k1; x1 += k2 + 1u;
  MD_TF_ROUND(17) MD_TF_ROUND(29) MD_TF_ROUND(16) MD_TF_ROUND(24)
  x0 += k2; x1 += k0 + 2u;
  MD_TF_ROUND(13) MD_TF_ROUND(15) MD_TF_ROUND(26) MD_TF_ROUND(6)
  x0 += k0; x1 += k1 + 3u;
  MD_TF_ROUND(17) MD_TF_ROUND(29) MD_TF_ROUND(16) MD_TF_ROUND(24)
  x0 += k1; x1 += k2 + 4u;
  MD_TF_ROUND(13) MD_TF_ROUND(15) MD_TF_ROUND(26) MD_TF_ROUND(6)
  x0 += k2; x1 += k0 + 5u;
#undef MD_TF_ROUND
  o0 = x0;
  o1 = x1;
}

// 23 random bits at the midpoint of their bucket: strictly inside (0,1) and
// exactly representable in f32 (src/reference/random.ts:95-122)
__host__ __device__ inline float md_open_unit(uint32_t bits) {
  return ((float)(bits >> 9) + 0.5f) * 1.1920928955078125e-7f;
}

// ---- cell / box constants passed to kernels by value -----------------------

struct md_box {
  // reduced lower-triangular box rows a=(ax,0,0) b=(bx,by,0) c=(cx,cy,cz)
  int64_t fa[3], fb[3], fc[3];  // fixed-point lattice vectors
  float a[3], b[3], c[3];       // f32 copies
  float inv[9];                 // f = r . M (row-vector convention), lower triangular
  double invd[9];
  int rectangular;
};

// Box arithmetic shared by the host and the device.  Contraction into fused
// multiply-adds is off in these helpers, so both compute the same doubles.
//
// The box constants of a reduced box b (rows a, b, c; 9 doubles).  `bad` is set when a length lies outside the
// fixed-point range (|x| >= 2^31 nm, or not finite).
__host__ __device__ inline void md_box_from_matrix(const double *b, md_box &d, bool &bad) {
#pragma clang fp contract(off)
  for (int k = 0; k < 9; ++k)
    if (!(fabs(b[k]) < 2147483648.0)) bad = true;
  for (int k = 0; k < 3; ++k) {
    d.fa[k] = md_double_to_fixed(b[k]);
    d.fb[k] = md_double_to_fixed(b[3 + k]);
    d.fc[k] = md_double_to_fixed(b[6 + k]);
    d.a[k] = (float)b[k];
    d.b[k] = (float)b[3 + k];
    d.c[k] = (float)b[6 + k];
  }
  const double ax = b[0], bx = b[3], by = b[4], cx = b[6], cy = b[7], cz = b[8];
  const double m[9] = {1.0 / ax, 0, 0, -bx / (ax * by), 1.0 / by, 0, ((bx * cy) / (by * cz) - cx / cz) / ax,
                       -cy / (by * cz), 1.0 / cz};
  for (int k = 0; k < 9; ++k) {
    d.invd[k] = m[k];
    d.inv[k] = (float)m[k];
  }
  d.rectangular = bx == 0 && cx == 0 && cy == 0;
}

__host__ __device__ inline double md_box_volume(const double *b) { return b[0] * b[4] * b[8]; }

// perpendicular widths of the cell along a, b and c
__host__ __device__ inline void md_box_widths(const double *b, double w[3]) {
#pragma clang fp contract(off)
  const double v = md_box_volume(b);
  const double *u[3] = {b + 3, b + 6, b};
  const double *t[3] = {b + 6, b, b + 3};
  for (int k = 0; k < 3; ++k) {
    const double *p = u[k], *q = t[k];
    const double x = p[1] * q[2] - p[2] * q[1], y = p[2] * q[0] - p[0] * q[2], z = p[0] * q[1] - p[1] * q[0];
    w[k] = v / sqrt(x * x + y * y + z * z);
  }
}

// ties to even, as reduceBox (src/reference/triclinic.ts:93-125)
__host__ __device__ inline double md_reduce_by(double ratio) {
  const double nearest = floor(ratio + 0.5);
  if (fabs(ratio - trunc(ratio)) == 0.5 && fmod(nearest, 2.0) != 0.0) return nearest - (ratio > 0 ? 1.0 : -1.0);
  return nearest;
}

// OpenMM's reduced form of a lower-triangular box, in place
__host__ __device__ inline void md_reduce_box(double *v) {
#pragma clang fp contract(off)
  double t = md_reduce_by(v[7] / v[4]);
  for (int k = 0; k < 3; ++k) v[6 + k] -= t * v[3 + k];
  t = md_reduce_by(v[6] / v[0]);
  for (int k = 0; k < 3; ++k) v[6 + k] -= t * v[k];
  t = md_reduce_by(v[3] / v[0]);
  for (int k = 0; k < 3; ++k) v[3 + k] -= t * v[k];
  v[1] = v[2] = v[5] = 0.0;
}

// Nearest image of a fixed-point difference: integer box counts picked in
// f32, subtracted exactly in fixed point, against c, then b, then a -- the
// reduction src/reference/forces/nonbonded.ts:241-256 performs, in the
// integer form of src/gpu/kernels/fixed64.wgsl:233-259.
__device__ inline void md_min_image_fixed(int64_t &dx, int64_t &dy, int64_t &dz, const md_box &bx) {
  float kc = floorf(md_from_fixed(dz) / bx.c[2] + 0.5f);
  if (kc != 0.0f) {
    int64_t k = (int64_t)kc;
    dx -= k * bx.fc[0];
    dy -= k * bx.fc[1];
    dz -= k * bx.fc[2];
  }
  float kb = floorf(md_from_fixed(dy) / bx.b[1] + 0.5f);
  if (kb != 0.0f) {
    int64_t k = (int64_t)kb;
    dx -= k * bx.fb[0];
    dy -= k * bx.fb[1];
  }
  float ka = floorf(md_from_fixed(dx) / bx.a[0] + 0.5f);
  if (ka != 0.0f) dx -= (int64_t)ka * bx.fa[0];
}

// ---- XCD-aware block order ----------------------------------------------------
//
// An MI355X dispatches a launch's workgroups to its 8 XCDs in turn (block b to
// XCD b mod 8), and each XCD has its own L2.  md_xcd_block(b, total) gives XCD
// k one contiguous run of logical blocks instead of every 8th, so blocks that
// read neighbouring data (atoms in spatial order reading the same mesh
// region) share an L2.  A bijection on [0, total) for any total.
__device__ __forceinline__ int md_xcd_block(int b, int total) {
  const int q = total >> 3, r = total & 7, k = b & 7;
  return k * q + min(k, r) + (b >> 3);
}

// ---- wave-level integer reduction (wavefront of 64) -------------------------

__device__ inline int64_t md_wave_sum_i64(int64_t v) {
  for (int off = 32; off > 0; off >>= 1) {
    int64_t o = __shfl_xor(v, off, 64);
    v += o;
  }
  return v;
}

__device__ inline void md_atomic_add_i64(int64_t *p, int64_t v) {
  atomicAdd(reinterpret_cast<unsigned long long *>(p), (unsigned long long)v);
}
