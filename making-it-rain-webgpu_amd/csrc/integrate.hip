// integrate.hip -- the LFMiddle Langevin / Verlet step fused with SETTLE,
// SHAKE and RATTLE and centre-of-mass removal.
//
// Replaces the upstream WGSL phases `lfmiddle_kick`, `lfmiddle_drift`,
// `lfmiddle_finish`, `verlet_kick_drift`, `verlet_finish`
// (src/gpu/kernels/integrate.wgsl:90-203), `settle_cluster`,
// `settle_velocities` (settle.wgsl:74-249), `shake_cluster`, `rattle_cluster`
// (shake.wgsl:124-229) and `momentum`/`subtract_com` (reduce.wgsl:61-105).
// Upstream runs these as seven dispatches per step; here one thread owns one
// constraint cluster (a rigid water, a SHAKE cluster of at most 8 atoms, or
// a single unconstrained atom) and carries it through the whole step in
// registers, so the phases need no global synchronisation between them.
//
// Step semantics follow src/reference/integrator.ts:140-220 (LFMiddle) and
// :68-112 (Verlet), COM removal src/reference/com-motion.ts:38-74 (acts when
// step % frequency == 0, before the forces -- equivalent here because forces
// do not read velocities), and the device number formats of the upstream
// kernels: fixed-point positions with the f32 residual fed back into the next
// displacement (integrate.wgsl:65-71, a deliberate deviation upstream keeps),
// f32 velocities and masses.  Constraint arithmetic runs in f64 on
// differences taken exactly in Q32.32.
#include <cmath>

#include "engine.hpp"
#include "warmstart.hpp"

struct md_int_args {
  int64_t *pos;
  float *vel;
  float *resid;
  int64_t *force;
  const float *inv_mass;
  const float *mass;
  const int64_t *build_pos;
  md_control *ctl;
  uint32_t *group_status;  // the launch group's error word (see k_step_prologue)
  int np;
  float dt, half_dt, inv_dt;
  float vscale, noise_scale, kt;
  uint32_t seed;
  int integrator;  // 0 LFMiddle, 1 Verlet
  int cmm_frequency;
  float inv_total_mass;
  double tol;
  float half_skin2;
  // work items
  const int4 *water_atoms;
  const float2 *water_dist;
  const int *ccl_atoms;
  const int *ccl_natoms;
  const int *ccl_ncons;
  const int4 *ccl_cons;
  const int *free_atoms;
  int n_waters, n_ccl, n_free;
  int n;  // atoms
  // learned SHAKE warm start: the network (null: plain SHAKE), and the
  // sample ring a recording run fills with (features, multipliers)
  const md_ws_net *ws;
  float *rec_x;    // [rec_cap][MD_WS_F]
  float *rec_t;    // [rec_cap][MD_WS_MAXC]
  int64_t rec_first_step;
  int64_t rec_cap;
};

// The three thermostat deviates of one particle: counters (2p, step) and
// (2p+1, step) under key (seed, THERMOSTAT_DOMAIN), Box-Muller, fourth
// deviate dropped (src/reference/random.ts:175-196, random.wgsl:78-83).
__device__ __forceinline__ void md_thermostat_normals(uint32_t seed, uint32_t step, uint32_t p, float *out) {
  uint32_t a0, a1, b0, b1;
  md_threefry2x32_20(2u * p, step, seed, MD_THERMOSTAT_DOMAIN, a0, a1);
  md_threefry2x32_20(2u * p + 1u, step, seed, MD_THERMOSTAT_DOMAIN, b0, b1);
  const float two_pi = 6.283185307179586f;
  float r1 = sqrtf(-2.0f * logf(md_open_unit(a0)));
  float t1 = two_pi * md_open_unit(a1);
  float r2 = sqrtf(-2.0f * logf(md_open_unit(b0)));
  float t2 = two_pi * md_open_unit(b1);
  out[0] = r1 * cosf(t1);
  out[1] = r1 * sinf(t1);
  out[2] = r2 * cosf(t2);
}

// `displace` (integrate.wgsl:67-71): move by `wanted`, carry what to_fixed
// could not represent
__device__ __forceinline__ int64_t md_displace(int64_t from, float wanted, float &left) {
  int64_t taken = md_to_fixed(wanted);
  left = wanted - md_from_fixed(taken);
  return from + taken;
}

template <int MAXA>
struct md_item {
  int n;
  int atom[MAXA];
  float im[MAXA];
  int64_t x[MAXA][3];   // positions at the start of the step
  float v[MAXA][3];
  int64_t tr[MAXA][3];  // trial
  int64_t dc[MAXA][3];  // what the constraints moved the trial by (trial - unconstrained trial)
};

// Most constraints a cluster of MAXA atoms can hold: every pair, at most
// MD_MAX_CLUSTER_CONSTRAINTS.
template <int MAXA>
struct md_max_constraints {
  static constexpr int value = MAXA * (MAXA - 1) / 2 < MD_MAX_CLUSTER_CONSTRAINTS ? MAXA * (MAXA - 1) / 2
                                                                                  : MD_MAX_CLUSTER_CONSTRAINTS;
};

// Private arrays are indexed by compile-time constants only: every loop over
// a cluster's atoms or constraints is unrolled to its bound and guarded, and
// a constraint's runtime atom index selects its atom through a chain of
// selects.  A runtime index into a private array would put the array in
// scratch memory.
__device__ __forceinline__ uint64_t md_bits(int64_t v) { return (uint64_t)v; }
__device__ __forceinline__ uint64_t md_bits(double v) { return (uint64_t)__double_as_longlong(v); }
__device__ __forceinline__ uint32_t md_bits(float v) { return __float_as_uint(v); }
__device__ __forceinline__ void md_unbits(uint64_t b, int64_t &v) { v = (int64_t)b; }
__device__ __forceinline__ void md_unbits(uint64_t b, double &v) { v = __longlong_as_double((long long)b); }
__device__ __forceinline__ void md_unbits(uint32_t b, float &v) { v = __uint_as_float(b); }

// The element `i` of a register array as an OR of masked bit patterns: a
// chain of selects would be folded back into an indexed load.
template <int N, typename T>
__device__ __forceinline__ T md_pick(const T (&a)[N][3], int i, int d) {
  decltype(md_bits(a[0][d])) r = 0;
#pragma unroll
  for (int k = 0; k < N; ++k) r |= md_bits(a[k][d]) & (0 - (decltype(r))(i == k));
  T v;
  md_unbits(r, v);
  return v;
}
template <int N, typename T>
__device__ __forceinline__ T md_pick1(const T (&a)[N], int i) {
  decltype(md_bits(a[0])) r = 0;
#pragma unroll
  for (int k = 0; k < N; ++k) r |= md_bits(a[k]) & (0 - (decltype(r))(i == k));
  T v;
  md_unbits(r, v);
  return v;
}

// kick, and for LFMiddle nothing else before RATTLE
template <int MAXA>
__device__ __forceinline__ void md_load_kick(const md_int_args &A, md_item<MAXA> &it, const float *vcom) {
#pragma unroll
  for (int a = 0; a < MAXA; ++a) {
    if (a >= it.n) break;
    int i = it.atom[a];
    it.im[a] = A.inv_mass[i];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      it.x[a][d] = A.pos[d * A.np + i];
      float v = A.vel[d * A.np + i];
      int64_t f = A.force[d * A.np + i];
      A.force[d * A.np + i] = 0;  // consumed: the next step accumulates afresh
      if (it.im[a] != 0.0f) {
        v -= vcom[d];
        v = v + A.dt * it.im[a] * md_from_fixed(f);
      }
      it.v[a][d] = v;
    }
  }
}

// half drift, thermostat, half drift (LFMiddle) or a whole drift (Verlet)
template <int MAXA>
__device__ __forceinline__ void md_drift(const md_int_args &A, md_item<MAXA> &it, uint32_t step) {
#pragma unroll
  for (int a = 0; a < MAXA; ++a) {
    if (a >= it.n) break;
    int i = it.atom[a];
    if (it.im[a] == 0.0f) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        it.tr[a][d] = it.x[a][d];
        it.dc[a][d] = 0;
      }
      continue;
    }
    if (A.integrator == 1) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        float left = 0.0f;
        it.tr[a][d] = md_displace(it.x[a][d], it.v[a][d] * A.dt + A.resid[d * A.np + i], left);
        A.resid[d * A.np + i] = left;
        it.dc[a][d] = 0;
      }
      continue;
    }
    float dev[3] = {0.0f, 0.0f, 0.0f};
    if (A.noise_scale != 0.0f) md_thermostat_normals(A.seed, step, (uint32_t)i, dev);
    float noise = A.noise_scale * sqrtf(A.kt * it.im[a]);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      float left = A.resid[d * A.np + i];
      int64_t moved = md_displace(it.x[a][d], it.v[a][d] * A.half_dt + left, left);
      float th = A.vscale * it.v[a][d] + noise * dev[d];
      it.v[a][d] = th;
      moved = md_displace(moved, th * A.half_dt + left, left);
      A.resid[d * A.np + i] = left;
      it.tr[a][d] = moved;
      it.dc[a][d] = 0;
    }
  }
}

// velocity from the constraint correction, commit, momentum, displacement
template <int MAXA>
__device__ __forceinline__ void md_finish(const md_int_args &A, md_item<MAXA> &it, double *mom, float &worst, uint32_t &failed) {
#pragma unroll
  for (int a = 0; a < MAXA; ++a) {
    if (a >= it.n) break;
    int i = it.atom[a];
    if (it.im[a] != 0.0f) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        // Verlet: the whole displacement over dt; LFMiddle: the constraint
        // correction over dt added (values selected, not arrays)
        const bool verlet = A.integrator == 1;
        const int64_t moved = verlet ? it.tr[a][d] - it.x[a][d] : it.dc[a][d];
        const float base = verlet ? 0.0f : it.v[a][d];
        it.v[a][d] = base + A.inv_dt * md_from_fixed(moved);
        A.pos[d * A.np + i] = it.tr[a][d];
      }
    }
    float m = A.mass[i];
    float d2 = 0.0f;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      A.vel[d * A.np + i] = it.v[a][d];
      if (!isfinite(it.v[a][d])) failed |= 4u;
      mom[d] += (double)m * it.v[a][d];
      float disp = md_from_fixed(A.pos[d * A.np + i] - A.build_pos[d * A.np + i]);
      d2 += disp * disp;
    }
    worst = fmaxf(worst, d2);
  }
}

// ---- constraint solvers (f64 on exact fixed-point differences) -------------

__device__ __forceinline__ double md_fx(int64_t v) { return (double)v * 2.3283064365386963e-10; }

// RATTLE (src/reference/constraints.ts:195-243): Jacobi, along the current
// bond vectors, divided by the reference squared length
// (false: not converged within MD_MAX_CONSTRAINT_ITERATIONS passes)
template <int MAXA>
__device__ __forceinline__ bool md_rattle(const md_int_args &A, md_item<MAXA> &it, int nc, const int4 *cons) {
  constexpr int MAXC = md_max_constraints<MAXA>::value;
  double rv[MAXC][3], l2[MAXC], mu[MAXC];
  double v[MAXA][3];
#pragma unroll
  for (int a = 0; a < MAXA; ++a)
#pragma unroll
    for (int d = 0; d < 3; ++d) v[a][d] = it.v[a][d];
#pragma unroll
  for (int c = 0; c < MAXC; ++c) {
    l2[c] = 0.0;
    mu[c] = 0.0;
    if (c >= nc) continue;
    int li = cons[c].x, lj = cons[c].y;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      rv[c][d] = md_fx(md_pick(it.x, li, d) - md_pick(it.x, lj, d));
      l2[c] += rv[c][d] * rv[c][d];
    }
    double sum = (double)md_pick1(it.im, li) + (double)md_pick1(it.im, lj);
    mu[c] = sum == 0.0 ? 0.0 : 0.5 / sum;
  }
  double delta[MAXC];
  bool converged = false;
  for (int iter = 0; iter < MD_MAX_CONSTRAINT_ITERATIONS; ++iter) {
    int conv = 0;
#pragma unroll
    for (int c = 0; c < MAXC; ++c) {
      delta[c] = 0.0;
      if (c >= nc) continue;
      int li = cons[c].x, lj = cons[c].y;
      double proj = 0;
#pragma unroll
      for (int d = 0; d < 3; ++d) proj += (md_pick(v, li, d) - md_pick(v, lj, d)) * rv[c][d];
      delta[c] = l2[c] == 0.0 ? 0.0 : -2.0 * mu[c] * proj / l2[c];
      conv += fabs(delta[c]) <= A.tol;
    }
    if (conv == nc) {
      converged = true;
      break;
    }
#pragma unroll
    for (int c = 0; c < MAXC; ++c) {
      if (c >= nc) continue;
      int li = cons[c].x, lj = cons[c].y;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double dr = rv[c][d] * delta[c];
#pragma unroll
        for (int a = 0; a < MAXA; ++a) {
          if (a == li) v[a][d] += dr * it.im[a];
          if (a == lj) v[a][d] -= dr * it.im[a];
        }
      }
    }
  }
#pragma unroll
  for (int a = 0; a < MAXA; ++a)
#pragma unroll
    for (int d = 0; d < 3; ++d) it.v[a][d] = (float)v[a][d];
  return converged || nc == 0;
}

// SHAKE (src/reference/constraints.ts:130-186): Jacobi along the reference
// bond vectors, converged when |r'|^2 lies in [(1-t)^2, (1+t)^2] d^2.
// *passes receives the update passes applied, or -1 when the cluster did not
// converge within MD_MAX_CONSTRAINT_ITERATIONS.
template <int MAXA>
__device__ __forceinline__ void md_shake(const md_int_args &A, md_item<MAXA> &it, int nc, const int4 *cons,
                                         bool warm = false, const float *lam0 = nullptr,
                                         double *lam_total = nullptr, int *passes = nullptr) {
  constexpr int MAXC = md_max_constraints<MAXA>::value;
  double rv[MAXC][3], mu[MAXC], t2[MAXC];
  double t[MAXA][3], moved[MAXA][3];
#pragma unroll
  for (int a = 0; a < MAXA; ++a)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      t[a][d] = md_fx(it.tr[a][d] - it.x[0][d]);
      moved[a][d] = 0.0;
    }
#pragma unroll
  for (int c = 0; c < MAXC; ++c) {
    mu[c] = 0.0;
    t2[c] = 0.0;
    if (c >= nc) continue;
    int li = cons[c].x, lj = cons[c].y;
#pragma unroll
    for (int d = 0; d < 3; ++d) rv[c][d] = md_fx(md_pick(it.x, li, d) - md_pick(it.x, lj, d));
    double sum = (double)md_pick1(it.im, li) + (double)md_pick1(it.im, lj);
    mu[c] = sum == 0.0 ? 0.0 : 0.5 / sum;
    double dist = __hiloint2double(cons[c].w, cons[c].z);
    t2[c] = dist * dist;
  }
  double lower = 1.0 - 2.0 * A.tol + A.tol * A.tol, upper = 1.0 + 2.0 * A.tol + A.tol * A.tol;
  double delta[MAXC];
  int applied = 0;
#pragma unroll
  for (int c = 0; c < MAXC; ++c) delta[c] = 0.0;
  if (warm) {
    // The predicted multipliers (in units of each constraint's reduced-mass
    // factor mu), applied as one update; the passes below correct whatever
    // the prediction missed.  A prediction that leaves the constraints
    // further from their lengths than the unconstrained trial positions is
    // dropped, so a poor network never costs SHAKE more than a cold start.
    auto worst = [&](const double (&p)[MAXA][3]) {
      double w = 0.0;
#pragma unroll
      for (int c = 0; c < MAXC; ++c) {
        if (c >= nc) continue;
        int li = cons[c].x, lj = cons[c].y;
        double x = md_pick(p, li, 0) - md_pick(p, lj, 0), y = md_pick(p, li, 1) - md_pick(p, lj, 1),
               z = md_pick(p, li, 2) - md_pick(p, lj, 2);
        w = fmax(w, fabs((x * x + y * y + z * z) / t2[c] - 1.0));
      }
      return w;
    };
    double tw[MAXA][3];
#pragma unroll
    for (int a = 0; a < MAXA; ++a)
#pragma unroll
      for (int d = 0; d < 3; ++d) tw[a][d] = t[a][d];
    bool finite = true;
#pragma unroll
    for (int c = 0; c < MAXC && c < MD_WS_MAXC; ++c) {
      if (c >= nc) continue;
      delta[c] = mu[c] * (double)lam0[c];
      finite &= isfinite(delta[c]);
    }
    if (finite) {
#pragma unroll
      for (int c = 0; c < MAXC; ++c) {
        if (c >= nc) continue;
        int li = cons[c].x, lj = cons[c].y;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          double dr = rv[c][d] * delta[c];
#pragma unroll
          for (int a = 0; a < MAXA; ++a) {
            if (a == li) tw[a][d] += dr * it.im[a];
            if (a == lj) tw[a][d] -= dr * it.im[a];
          }
        }
      }
    }
    if (finite && worst(tw) < worst(t)) {
#pragma unroll
      for (int a = 0; a < MAXA; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          moved[a][d] = tw[a][d] - t[a][d];
          t[a][d] = tw[a][d];
        }
    } else {
#pragma unroll
      for (int c = 0; c < MAXC; ++c) delta[c] = 0.0;
    }
  }
  double total[MAXC];
#pragma unroll
  for (int c = 0; c < MAXC; ++c) total[c] = delta[c];
  bool converged = nc == 0;
  for (int iter = 0; iter < MD_MAX_CONSTRAINT_ITERATIONS; ++iter) {
    int conv = 0;
#pragma unroll
    for (int c = 0; c < MAXC; ++c) {
      delta[c] = 0.0;
      if (c >= nc) continue;
      int li = cons[c].x, lj = cons[c].y;
      double x = md_pick(t, li, 0) - md_pick(t, lj, 0), y = md_pick(t, li, 1) - md_pick(t, lj, 1),
             z = md_pick(t, li, 2) - md_pick(t, lj, 2);
      double l2 = x * x + y * y + z * z;
      double proj = x * rv[c][0] + y * rv[c][1] + z * rv[c][2];
      delta[c] = proj == 0.0 ? 0.0 : mu[c] * (t2[c] - l2) / proj;
      conv += (l2 >= lower * t2[c] && l2 <= upper * t2[c]);
    }
    if (conv == nc) {
      converged = true;
      break;
    }
    ++applied;
#pragma unroll
    for (int c = 0; c < MAXC; ++c) {
      if (c >= nc) continue;
      int li = cons[c].x, lj = cons[c].y;
      total[c] += delta[c];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double dr = rv[c][d] * delta[c];
#pragma unroll
        for (int a = 0; a < MAXA; ++a) {
          if (a == li) {
            t[a][d] += dr * it.im[a];
            moved[a][d] += dr * it.im[a];
          }
          if (a == lj) {
            t[a][d] -= dr * it.im[a];
            moved[a][d] -= dr * it.im[a];
          }
        }
      }
    }
  }
#pragma unroll
  for (int a = 0; a < MAXA; ++a)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int64_t by = (int64_t)rint(moved[a][d] * 4294967296.0);
      it.tr[a][d] += by;
      it.dc[a][d] += by;
    }
  if (passes) *passes = converged ? applied : -1;
  if (lam_total) {
#pragma unroll
    for (int c = 0; c < MAXC && c < MD_WS_MAXC; ++c) lam_total[c] = mu[c] == 0.0 ? 0.0 : total[c] / mu[c];
  }
}

// The network's inputs for one cluster of up to MD_WS_MAXC constraints, from
// the positions at the start of the step (it.x) and the unconstrained trial
// positions (it.tr): for each constraint c, f[4c..4c+3] = (100 (d^2 - |r'|^2)
// / d^2, 100 (r'.r / d^2 - 1), m_i^-1 / (m_i^-1 + m_j^-1), 1) with r the
// reference and r' the trial bond vector; f[16..21] couple the six pairs of
// constraints that share an atom: +-m_shared^-1 / sqrt(the two inverse-mass
// sums) times the cosine between their reference vectors.  All invariant
// under rotation and translation.
template <int MAXA>
__device__ __forceinline__ void md_ws_features(const md_item<MAXA> &it, int nc, const int4 *cons, float *f) {
#pragma unroll
  for (int k = 0; k < MD_WS_F; ++k) f[k] = 0.0f;
  double rv[MD_WS_MAXC][3], d2[MD_WS_MAXC], isum[MD_WS_MAXC];
  int n = min(nc, MD_WS_MAXC);
#pragma unroll
  for (int c = 0; c < MD_WS_MAXC; ++c) {
    d2[c] = isum[c] = 0.0;
    rv[c][0] = rv[c][1] = rv[c][2] = 0.0;
    if (c >= n) continue;
    int li = cons[c].x, lj = cons[c].y;
    double rp[3], l2 = 0.0, pr = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      rv[c][d] = md_fx(md_pick(it.x, li, d) - md_pick(it.x, lj, d));
      rp[d] = md_fx(md_pick(it.tr, li, d) - md_pick(it.tr, lj, d));
      l2 += rp[d] * rp[d];
      pr += rp[d] * rv[c][d];
    }
    double dist = __hiloint2double(cons[c].w, cons[c].z);
    d2[c] = dist * dist;
    const double imi = md_pick1(it.im, li), imj = md_pick1(it.im, lj);
    isum[c] = imi + imj;
    f[4 * c + 0] = (float)(MD_WS_SCALE * (d2[c] - l2) / d2[c]);
    f[4 * c + 1] = (float)(MD_WS_SCALE * (pr / d2[c] - 1.0));
    f[4 * c + 2] = isum[c] > 0.0 ? (float)(imi / isum[c]) : 0.0f;
    f[4 * c + 3] = 1.0f;
  }
  int pair = 0;
#pragma unroll
  for (int c = 0; c < MD_WS_MAXC; ++c)
#pragma unroll
    for (int e = c + 1; e < MD_WS_MAXC; ++e, ++pair) {
      if (e >= n) continue;
      int ci = cons[c].x, cj = cons[c].y, ei = cons[e].x, ej = cons[e].y;
      double shared = 0.0;
      if (ci == ei) shared += md_pick1(it.im, ci);
      if (cj == ej) shared += md_pick1(it.im, cj);
      if (ci == ej) shared -= md_pick1(it.im, ci);
      if (cj == ei) shared -= md_pick1(it.im, cj);
      if (shared == 0.0 || isum[c] == 0.0 || isum[e] == 0.0) continue;
      double dot = rv[c][0] * rv[e][0] + rv[c][1] * rv[e][1] + rv[c][2] * rv[e][2];
      f[16 + pair] = (float)(shared / sqrt(isum[c] * isum[e]) * dot / sqrt(d2[c] * d2[e]));
    }
}

// SETTLE positions (src/reference/settle.ts:31-196): rigid rotation about the
// centre of mass in the frame of the reference triangle
__device__ __forceinline__ void md_settle_positions(md_item<3> &it, double d1, double d2, const float *mass3) {
  double m0 = mass3[0], m1 = mass3[1], m2 = mass3[2];
  double xp0[3], xp1[3], xp2[3], b0[3], c0[3];
  for (int k = 0; k < 3; ++k) {
    xp0[k] = md_fx(it.tr[0][k] - it.x[0][k]);
    xp1[k] = md_fx(it.tr[1][k] - it.x[1][k]);
    xp2[k] = md_fx(it.tr[2][k] - it.x[2][k]);
    b0[k] = md_fx(it.x[1][k] - it.x[0][k]);
    c0[k] = md_fx(it.x[2][k] - it.x[0][k]);
  }
  double inv_m = 1.0 / (m0 + m1 + m2);
  double com[3], a1[3], b1[3], c1[3];
  for (int k = 0; k < 3; ++k) {
    com[k] = (xp0[k] * m0 + (b0[k] + xp1[k]) * m1 + (c0[k] + xp2[k]) * m2) * inv_m;
    a1[k] = xp0[k] - com[k];
    b1[k] = b0[k] + xp1[k] - com[k];
    c1[k] = c0[k] + xp2[k] - com[k];
  }
  double zd[3] = {b0[1] * c0[2] - b0[2] * c0[1], b0[2] * c0[0] - b0[0] * c0[2], b0[0] * c0[1] - b0[1] * c0[0]};
  double xd[3] = {a1[1] * zd[2] - a1[2] * zd[1], a1[2] * zd[0] - a1[0] * zd[2], a1[0] * zd[1] - a1[1] * zd[0]};
  double yd[3] = {zd[1] * xd[2] - zd[2] * xd[1], zd[2] * xd[0] - zd[0] * xd[2], zd[0] * xd[1] - zd[1] * xd[0]};
  double xl = sqrt(xd[0] * xd[0] + xd[1] * xd[1] + xd[2] * xd[2]);
  double yl = sqrt(yd[0] * yd[0] + yd[1] * yd[1] + yd[2] * yd[2]);
  double zl = sqrt(zd[0] * zd[0] + zd[1] * zd[1] + zd[2] * zd[2]);
  double ex[3], ey[3], ez[3];
  for (int k = 0; k < 3; ++k) {
    ex[k] = xd[k] / xl;
    ey[k] = yd[k] / yl;
    ez[k] = zd[k] / zl;
  }
#define MD_DOT(u, w) ((u)[0] * (w)[0] + (u)[1] * (w)[1] + (u)[2] * (w)[2])
  double xb0d = MD_DOT(ex, b0), yb0d = MD_DOT(ey, b0);
  double xc0d = MD_DOT(ex, c0), yc0d = MD_DOT(ey, c0);
  double za1d = MD_DOT(ez, a1);
  double xb1d = MD_DOT(ex, b1), yb1d = MD_DOT(ey, b1), zb1d = MD_DOT(ez, b1);
  double xc1d = MD_DOT(ex, c1), yc1d = MD_DOT(ey, c1), zc1d = MD_DOT(ez, c1);
  double rc = 0.5 * d2;
  double rb = sqrt(d1 * d1 - rc * rc);
  double ra = rb * (m1 + m2) * inv_m;
  rb -= ra;
  double sinphi = za1d / ra;
  double cosphi = sqrt(1.0 - sinphi * sinphi);
  double sinpsi = (zb1d - zc1d) / (2.0 * rc * cosphi);
  double cospsi = sqrt(1.0 - sinpsi * sinpsi);
  double ya2d = ra * cosphi;
  double xb2d = -rc * cospsi;
  double yb2d = -rb * cosphi - rc * sinpsi * sinphi;
  double yc2d = -rb * cosphi + rc * sinpsi * sinphi;
  double xb2d2 = xb2d * xb2d;
  double hh2 = 4.0 * xb2d2 + (yb2d - yc2d) * (yb2d - yc2d) + (zb1d - zc1d) * (zb1d - zc1d);
  double deltx = 2.0 * xb2d + sqrt(4.0 * xb2d2 - hh2 + d2 * d2);
  xb2d -= deltx * 0.5;
  double alpha = xb2d * (xb0d - xc0d) + yb0d * yb2d + yc0d * yc2d;
  double beta = xb2d * (yc0d - yb0d) + xb0d * yb2d + xc0d * yc2d;
  double gamma = xb0d * yb1d - xb1d * yb0d + xc0d * yc1d - xc1d * yc0d;
  double al2be2 = alpha * alpha + beta * beta;
  double sintheta = (alpha * gamma - beta * sqrt(al2be2 - gamma * gamma)) / al2be2;
  double costheta = sqrt(1.0 - sintheta * sintheta);
  double a3d[3] = {-ya2d * sintheta, ya2d * costheta, za1d};
  double b3d[3] = {xb2d * costheta - yb2d * sintheta, xb2d * sintheta + yb2d * costheta, zb1d};
  double c3d[3] = {-xb2d * costheta - yc2d * sintheta, -xb2d * sintheta + yc2d * costheta, zc1d};
  for (int k = 0; k < 3; ++k) {
    double a3 = ex[k] * a3d[0] + ey[k] * a3d[1] + ez[k] * a3d[2];
    double b3 = ex[k] * b3d[0] + ey[k] * b3d[1] + ez[k] * b3d[2];
    double c3 = ex[k] * c3d[0] + ey[k] * c3d[1] + ez[k] * c3d[2];
    const int64_t t0 = it.x[0][k] + (int64_t)rint((com[k] + a3) * 4294967296.0);
    const int64_t t1 = it.x[1][k] + (int64_t)rint((com[k] + b3 - b0[k]) * 4294967296.0);
    const int64_t t2 = it.x[2][k] + (int64_t)rint((com[k] + c3 - c0[k]) * 4294967296.0);
    it.dc[0][k] += t0 - it.tr[0][k];
    it.dc[1][k] += t1 - it.tr[1][k];
    it.dc[2][k] += t2 - it.tr[2][k];
    it.tr[0][k] = t0;
    it.tr[1][k] = t1;
    it.tr[2][k] = t2;
  }
}

// SETTLE velocities (src/reference/settle.ts:198-284): exact 3x3 impulse solve
__device__ __forceinline__ void md_settle_velocities(md_item<3> &it, const float *mass3) {
  double mA = mass3[0], mB = mass3[1], mC = mass3[2];
  double iA = mA == 0 ? 0 : 1 / mA, iB = mB == 0 ? 0 : 1 / mB, iC = mC == 0 ? 0 : 1 / mC;
  double e[3][3];
  const int from[3] = {0, 1, 2}, to[3] = {1, 2, 0};
  double rel[3];
  for (int q = 0; q < 3; ++q) {
    double x = md_fx(it.x[to[q]][0] - it.x[from[q]][0]);
    double y = md_fx(it.x[to[q]][1] - it.x[from[q]][1]);
    double z = md_fx(it.x[to[q]][2] - it.x[from[q]][2]);
    double l = sqrt(x * x + y * y + z * z);
    e[q][0] = x / l;
    e[q][1] = y / l;
    e[q][2] = z / l;
    rel[q] = 0;
    for (int k = 0; k < 3; ++k) rel[q] += ((double)it.v[to[q]][k] - (double)it.v[from[q]][k]) * e[q][k];
  }
  double *eAB = e[0], *eBC = e[1], *eCA = e[2];
  double vAB = rel[0], vBC = rel[1], vCA = rel[2];
  double cA = -MD_DOT(eAB, eCA), cB = -MD_DOT(eAB, eBC), cC = -MD_DOT(eBC, eCA);
  double s2A = 1 - cA * cA, s2B = 1 - cB * cB, s2C = 1 - cC * cC;
  double mABCinv = 1 / (mA * mB * mC);
  double denom = (((s2A * mB + s2B * mA) * mC + (s2A * mB * mB + 2 * (cA * cB * cC + 1) * mA * mB + s2B * mA * mA)) *
                      mC +
                  s2C * mA * mB * (mA + mB)) *
                 mABCinv;
  double tab = ((cB * cC * mA - cA * mB - cA * mC) * vCA + (cA * cC * mB - cB * mC - cB * mA) * vBC +
                (s2C * mA * mA * mB * mB * mABCinv + (mA + mB + mC)) * vAB) /
               denom;
  double tbc = ((cA * cB * mC - cC * mB - cC * mA) * vCA + (s2A * mB * mB * mC * mC * mABCinv + (mA + mB + mC)) * vBC +
                (cA * cC * mB - cB * mA - cB * mC) * vAB) /
               denom;
  double tca = ((s2B * mA * mA * mC * mC * mABCinv + (mA + mB + mC)) * vCA + (cA * cB * mC - cC * mB - cC * mA) * vBC +
                (cB * cC * mA - cA * mB - cA * mC) * vAB) /
               denom;
  for (int k = 0; k < 3; ++k) {
    it.v[0][k] = (float)((double)it.v[0][k] + (eAB[k] * tab - eCA[k] * tca) * iA);
    it.v[1][k] = (float)((double)it.v[1][k] + (eBC[k] * tbc - eAB[k] * tab) * iB);
    it.v[2][k] = (float)((double)it.v[2][k] + (eCA[k] * tca - eBC[k] * tbc) * iC);
  }
}
#undef MD_DOT

__device__ __forceinline__ void md_com_velocity(const md_int_args &A, uint32_t step, float *vcom) {
  vcom[0] = vcom[1] = vcom[2] = 0.0f;
  if (A.cmm_frequency > 0 && step % (uint32_t)A.cmm_frequency == 0) {
    const int64_t *p = A.ctl->momentum[step & 1];
#pragma unroll
    for (int d = 0; d < 3; ++d) vcom[d] = (float)(md_fixed_to_double_dev(p[d]) * (double)A.inv_total_mass);
  }
}

__device__ __forceinline__ void md_reduce_out(const md_int_args &A, uint32_t step, const double *mom, float worst) {
  // momentum for the next step's COM removal, and the displacement monitor
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    int64_t f = md_wave_sum_i64((int64_t)rint(mom[d] * 4294967296.0));
    if ((threadIdx.x & 63) == 0) md_atomic_add_i64(&A.ctl->momentum[(step + 1) & 1][d], f);
  }
  for (int off = 32; off > 0; off >>= 1) worst = fmaxf(worst, __shfl_xor(worst, off, 64));
  if ((threadIdx.x & 63) == 0) {
    atomicMax(&A.ctl->max_disp2_bits[0], __float_as_uint(worst));
    if (worst > A.half_skin2) A.ctl->rebuild_request[0] = 1;
  }
}

// SHAKE clusters run in workgroups of one wavefront: an ensemble has few
// clusters per replica (~300 at 10k atoms), so 64 per workgroup spreads them
// over four times as many CUs as 256 would, and a slow-converging cluster
// holds up one wave, not four.
#define MD_CLUSTER_BLOCK 64

// kind 0: free atoms, 1: SETTLE waters, 2: SHAKE clusters (MAXA atoms).
// MODE (clusters only): 0 plain SHAKE; 1 plain SHAKE, recording each
// cluster's features and converged multipliers as training samples; 2 SHAKE
// started from the network's predicted multipliers.  Every cluster step adds
// its SHAKE passes to the replica's counters.
template <int KIND, int MAXA, int MODE = 0>
__global__ void __launch_bounds__(256) k_integrate(const md_int_args *__restrict__ T, md_bmap map) {
  int blk;
  const md_int_args &A = T[md_locate(map, blockIdx.x, blk)];
  int t = blk * blockDim.x + threadIdx.x;
  uint32_t step = (uint32_t)A.ctl->step_current;
  float vcom[3];
  md_com_velocity(A, step, vcom);
  double mom[3] = {0.0, 0.0, 0.0};
  float worst = 0.0f;
  int count = KIND == 0 ? A.n_free : (KIND == 1 ? A.n_waters : A.n_ccl);
  const bool live = t < count;
  md_item<MAXA> it;
  it.n = 0;
  int nc = 0;
  uint32_t failed = 0;  // control-block divergence bits this thread raises
  constexpr int MAXC = KIND == 2 ? md_max_constraints<MAXA>::value : 1;
  int4 cons[MAXC];
  if (live) {
    if (KIND == 0) {
      it.n = 1;
      it.atom[0] = A.free_atoms[t];
    } else if (KIND == 1) {
      int4 w = A.water_atoms[t];
      it.n = 3;
      it.atom[0] = w.x;
      it.atom[1] = w.y;
      it.atom[2] = w.z;
    } else {
      it.n = A.ccl_natoms[t];
#pragma unroll
      for (int a = 0; a < MAXA; ++a) it.atom[a] = a < it.n ? A.ccl_atoms[t * MD_MAX_CLUSTER_ATOMS + a] : 0;
      nc = A.ccl_ncons[t];
#pragma unroll
      for (int c = 0; c < MAXC; ++c)
        cons[c] = c < nc ? A.ccl_cons[t * MD_MAX_CLUSTER_CONSTRAINTS + c] : make_int4(0, 0, 0, 0);
    }
    md_load_kick(A, it, vcom);
    if constexpr (KIND == 1) {
      float m3[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) m3[a] = A.mass[it.atom[a]];
      if (A.integrator == 0) md_settle_velocities(it, m3);
      md_drift(A, it, step);
      float2 dd = A.water_dist[t];
      md_settle_positions(it, dd.x, dd.y, m3);
    } else if constexpr (KIND == 2) {
      if (A.integrator == 0 && !md_rattle(A, it, nc, cons)) failed |= 2u;
      md_drift(A, it, step);
    } else {
      md_drift(A, it, step);
    }
  }
  if constexpr (KIND == 2) {
    const bool covered = live && nc > 0 && nc <= MD_WS_MAXC;
    float lam0[MD_WS_MAXC] = {0.0f, 0.0f, 0.0f, 0.0f};
    bool warm = false;
    if constexpr (MODE == 2) {
      // every lane of the workgroup runs the network (its replica's, or none)
      __shared__ float ws_lds[(MD_CLUSTER_BLOCK / 64) * MD_WS_TILE_LDS];
      if (A.ws) {
        float f[MD_WS_F], y[MD_WS_MAXC];
        if (covered)
          md_ws_features(it, nc, cons, f);
        else
          for (int k = 0; k < MD_WS_F; ++k) f[k] = 0.0f;
        md_ws_infer(*A.ws, f, y, ws_lds + (threadIdx.x >> 6) * MD_WS_TILE_LDS);
        if (covered) {
          for (int c = 0; c < MD_WS_MAXC; ++c) lam0[c] = y[c] * (1.0f / MD_WS_SCALE);
          warm = true;
        }
      }
    }
    int passes = 0;
    if (live) {
      double total[MD_WS_MAXC];
      float f[MD_WS_F];
      if (MODE == 1 && covered) md_ws_features(it, nc, cons, f);
      md_shake(A, it, nc, cons, warm, lam0, MODE == 1 ? total : nullptr, &passes);
      if (passes < 0) {
        failed |= 1u;
        passes = MD_MAX_CONSTRAINT_ITERATIONS;
      }
      if (MODE == 1 && covered && !(failed & 1u)) {
        // a cluster-step SHAKE did not converge teaches nothing: its slot
        // keeps the zero features (no constraint present) it was cleared to
        int64_t k = (A.ctl->step_current - A.rec_first_step) * (int64_t)A.n_ccl + t;
        if (k >= 0 && k < A.rec_cap) {
          for (int q = 0; q < MD_WS_F; ++q) A.rec_x[k * MD_WS_F + q] = f[q];
          for (int c = 0; c < MD_WS_MAXC; ++c)
            A.rec_t[k * MD_WS_MAXC + c] = c < nc ? (float)(MD_WS_SCALE * total[c]) : 0.0f;
        }
      }
    }
    // SHAKE passes and clusters of this step, for the warm start's gain
    int64_t ps = md_wave_sum_i64((int64_t)passes), cs = md_wave_sum_i64(live ? 1 : 0);
    if ((threadIdx.x & 63) == 0 && cs) {
      md_atomic_add_i64(&A.ctl->shake_passes, ps);
      md_atomic_add_i64(&A.ctl->shake_clusters, cs);
    }
  }
  if (live) md_finish(A, it, mom, worst, failed);
  if (failed) atomicOr(&A.ctl->diverged, failed);
  md_reduce_out(A, step, mom, worst);
}

// Constraint projection outside a step (src/reference/constraints.ts:254-291):
// WHAT 0 projects the positions onto the constraints with themselves as the
// reference (constrainPositions(x, x)); WHAT 1 removes every constrained
// relative velocity (constrainVelocities).
template <int KIND, int MAXA, int WHAT>
__global__ void __launch_bounds__(256) k_constrain(const md_int_args *__restrict__ T, md_bmap map, double tol) {
  int blk;
  md_int_args A = T[md_locate(map, blockIdx.x, blk)];
  A.tol = tol;
  int t = blk * blockDim.x + threadIdx.x;
  int count = KIND == 1 ? A.n_waters : A.n_ccl;
  if (t >= count) return;
  md_item<MAXA> it;
  int nc = 0;
  constexpr int MAXC = KIND == 2 ? md_max_constraints<MAXA>::value : 1;
  int4 cons[MAXC];
  if (KIND == 1) {
    int4 w = A.water_atoms[t];
    it.n = 3;
    it.atom[0] = w.x;
    it.atom[1] = w.y;
    it.atom[2] = w.z;
  } else {
    it.n = A.ccl_natoms[t];
#pragma unroll
    for (int a = 0; a < MAXA; ++a) it.atom[a] = a < it.n ? A.ccl_atoms[t * MD_MAX_CLUSTER_ATOMS + a] : 0;
    nc = A.ccl_ncons[t];
#pragma unroll
    for (int c = 0; c < MAXC; ++c)
      cons[c] = c < nc ? A.ccl_cons[t * MD_MAX_CLUSTER_CONSTRAINTS + c] : make_int4(0, 0, 0, 0);
  }
#pragma unroll
  for (int a = 0; a < MAXA; ++a) {
    it.im[a] = 0.0f;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      it.x[a][d] = it.tr[a][d] = it.dc[a][d] = 0;
      it.v[a][d] = 0.0f;
    }
    if (a >= it.n) continue;
    int i = it.atom[a];
    it.im[a] = A.inv_mass[i];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      it.x[a][d] = it.tr[a][d] = A.pos[d * A.np + i];
      it.v[a][d] = A.vel[d * A.np + i];
    }
  }
  if constexpr (KIND == 1) {
    float m3[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) m3[a] = A.mass[it.atom[a]];
    if (WHAT == 0) {
      float2 dd = A.water_dist[t];
      md_settle_positions(it, dd.x, dd.y, m3);
    } else {
      md_settle_velocities(it, m3);
    }
  } else {
    if (WHAT == 0)
      md_shake(A, it, nc, cons);
    else
      md_rattle(A, it, nc, cons);
  }
#pragma unroll
  for (int a = 0; a < MAXA; ++a) {
    if (a >= it.n) break;
    int i = it.atom[a];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (WHAT == 0)
        A.pos[d * A.np + i] = it.tr[a][d];
      else
        A.vel[d * A.np + i] = it.v[a][d];
    }
  }
}

// start of a step, one thread per replica: advance the device step counter,
// latch the rebuild request (the displacement test's, or the host's),
// clear the momentum slot this step's integrator fills, and raise the group's
// status word when an earlier step left an error in the control block
__global__ void k_step_prologue(const md_int_args *__restrict__ T, int R, int advance) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  md_control *ctl = T[r].ctl;
  if (ctl->overflow | ctl->diverged | ctl->barostat_error) atomicOr(T[r].group_status, 1u);
  if (advance) {
    ctl->step_current = ctl->step_next;
    ctl->step_next = ctl->step_next + 1;
    int64_t s = ctl->step_current;
    for (int d = 0; d < 3; ++d) ctl->momentum[(s + 1) & 1][d] = 0;
  }
  ctl->rebuild_active = ctl->rebuild_request[0] | ctl->force_pending;
  ctl->rebuild_request[0] = 0;
  ctl->force_pending = 0;
}

// momentum of the current velocities, into the slot the next step reads
__global__ void k_momentum(const md_int_args *__restrict__ T, md_bmap map) {
  int blk;
  const md_int_args &A = T[md_locate(map, blockIdx.x, blk)];
  md_control *ctl = A.ctl;
  int i = blk * blockDim.x + threadIdx.x;
  uint32_t slot = (uint32_t)(ctl->step_next & 1);
  double m[3] = {0, 0, 0};
  if (i < A.n)
    for (int d = 0; d < 3; ++d) m[d] = (double)A.mass[i] * A.vel[d * A.np + i];
  for (int d = 0; d < 3; ++d) {
    int64_t f = md_wave_sum_i64((int64_t)rint(m[d] * 4294967296.0));
    if ((threadIdx.x & 63) == 0) md_atomic_add_i64(&ctl->momentum[slot][d], f);
  }
}

__global__ void k_clear_momentum(const md_int_args *__restrict__ T, int R) {
  int r = blockIdx.x;
  if (r >= R || threadIdx.x >= 3) return;
  md_control *ctl = T[r].ctl;
  ctl->momentum[ctl->step_next & 1][threadIdx.x] = 0;
}

static md_int_args int_args(md_system *s) {
  md_int_args A;
  A.pos = s->pos;
  A.vel = s->vel;
  A.resid = s->resid;
  A.force = s->force;
  A.inv_mass = s->inv_mass;
  A.mass = s->mass_f;
  A.build_pos = s->build_pos;
  A.ctl = s->ctl;
  A.np = s->np;
  double dt = s->opt.time_step;
  A.dt = (float)dt;
  A.half_dt = (float)(0.5 * dt);
  A.inv_dt = (float)(1.0 / dt);
  double vs = std::exp(-dt * s->opt.friction);
  A.vscale = (float)vs;
  A.noise_scale = (float)std::sqrt(1.0 - vs * vs);
  A.kt = (float)(MD_BOLTZMANN * s->opt.temperature);
  A.seed = s->opt.seed;
  A.integrator = s->opt.integrator;
  A.cmm_frequency = s->cmm_frequency;
  A.inv_total_mass = s->total_mass > 0 ? (float)(1.0 / s->total_mass) : 0.0f;
  A.tol = s->opt.constraint_tolerance;
  A.half_skin2 = s->nb.half_skin2;
  A.water_atoms = s->water_atoms;
  A.water_dist = s->water_dist;
  A.ccl_atoms = s->ccl_atoms;
  A.ccl_natoms = s->ccl_natoms;
  A.ccl_ncons = s->ccl_ncons;
  A.ccl_cons = s->ccl_cons;
  A.free_atoms = s->free_atoms;
  A.n_waters = s->n_waters;
  A.n_ccl = s->n_clusters_c;
  A.n_free = s->n_free;
  A.n = s->n;
  A.ws = s->ws_net;
  A.rec_x = s->rec_x;
  A.rec_t = s->rec_t;
  A.rec_first_step = s->rec_first_step;
  A.rec_cap = s->rec_cap;
  return A;
}

struct md_int_group {
  md_int_args *table = nullptr;
  md_launch free_items, waters, clusters, atoms;
  int R = 0;
};

static md_int_group &ig(md_group *g) { return *reinterpret_cast<md_int_group *>(g->integ); }

void md_int_group_setup(md_group *g) {
  auto *G = new md_int_group();
  g->integ = G;
  G->R = (int)g->reps.size();
  std::vector<int> fr, wa, cl, at;
  g->max_cluster_atoms = 0;
  for (md_system *s : g->reps) {
    fr.push_back((s->n_free + 255) / 256);
    wa.push_back((s->n_waters + 255) / 256);
    cl.push_back((s->n_clusters_c + MD_CLUSTER_BLOCK - 1) / MD_CLUSTER_BLOCK);
    at.push_back((s->n + 255) / 256);
    g->max_cluster_atoms = std::max(g->max_cluster_atoms, s->max_cluster_atoms);
  }
  G->free_items = md_make_launch(g, fr);
  G->waters = md_make_launch(g, wa);
  G->clusters = md_make_launch(g, cl);
  G->atoms = md_make_launch(g, at);
  G->table = md_group_alloc<md_int_args>(g, G->R);
  md_int_group_refresh(g);
}

void md_int_group_refresh(md_group *g) {
  std::vector<md_int_args> host;
  int mode = 0;
  for (md_system *s : g->reps) {
    host.push_back(int_args(s));
    host.back().group_status = g->status;
    if (s->rec_cap > 0) mode = 1;
    else if (s->ws_net && mode == 0) mode = 2;
  }
  if (mode != g->ws_mode) {
    g->ws_mode = mode;
    g->graph_stale = true;  // the cluster kernel changes
  }
  MD_HIP(hipMemcpyAsync(ig(g).table, host.data(), sizeof(md_int_args) * host.size(), hipMemcpyHostToDevice,
                        g->stream));
  MD_HIP(hipStreamSynchronize(g->stream));
}

void md_int_group_free(md_group *g) {
  delete &ig(g);
  g->integ = nullptr;
}

void md_step_prologue(md_group *g, bool advance) {
  auto &G = ig(g);
  k_step_prologue<<<(G.R + 63) / 64, 64, 0, g->stream>>>(G.table, G.R, advance ? 1 : 0);
  MD_HIP(hipGetLastError());
}

void md_integrate(md_group *g) {
  auto &G = ig(g);
  hipStream_t st = g->stream;
  const md_int_args *T = G.table;
  if (G.free_items.total) k_integrate<0, 1><<<G.free_items.total, 256, 0, st>>>(T, G.free_items.map);
  if (G.waters.total) k_integrate<1, 3><<<G.waters.total, 256, 0, st>>>(T, G.waters.map);
  if (G.clusters.total) {
    const int mode = g->ws_mode;
    if (g->max_cluster_atoms <= 4) {
      if (mode == 2)
        k_integrate<2, 4, 2><<<G.clusters.total, MD_CLUSTER_BLOCK, 0, st>>>(T, G.clusters.map);
      else if (mode == 1)
        k_integrate<2, 4, 1><<<G.clusters.total, MD_CLUSTER_BLOCK, 0, st>>>(T, G.clusters.map);
      else
        k_integrate<2, 4, 0><<<G.clusters.total, MD_CLUSTER_BLOCK, 0, st>>>(T, G.clusters.map);
    } else {
      if (mode == 2)
        k_integrate<2, 8, 2><<<G.clusters.total, MD_CLUSTER_BLOCK, 0, st>>>(T, G.clusters.map);
      else if (mode == 1)
        k_integrate<2, 8, 1><<<G.clusters.total, MD_CLUSTER_BLOCK, 0, st>>>(T, G.clusters.map);
      else
        k_integrate<2, 8, 0><<<G.clusters.total, MD_CLUSTER_BLOCK, 0, st>>>(T, G.clusters.map);
    }
  }
  MD_HIP(hipGetLastError());
}

void md_constrain(md_group *g, int what, double tolerance) {
  auto &G = ig(g);
  hipStream_t st = g->stream;
  const md_int_args *T = G.table;
  if (G.waters.total) {
    if (what == 0)
      k_constrain<1, 3, 0><<<G.waters.total, 256, 0, st>>>(T, G.waters.map, tolerance);
    else
      k_constrain<1, 3, 1><<<G.waters.total, 256, 0, st>>>(T, G.waters.map, tolerance);
  }
  if (G.clusters.total) {
    int gsz = G.clusters.total;
    if (g->max_cluster_atoms <= 4) {
      if (what == 0)
        k_constrain<2, 4, 0><<<gsz, MD_CLUSTER_BLOCK, 0, st>>>(T, G.clusters.map, tolerance);
      else
        k_constrain<2, 4, 1><<<gsz, MD_CLUSTER_BLOCK, 0, st>>>(T, G.clusters.map, tolerance);
    } else {
      if (what == 0)
        k_constrain<2, 8, 0><<<gsz, MD_CLUSTER_BLOCK, 0, st>>>(T, G.clusters.map, tolerance);
      else
        k_constrain<2, 8, 1><<<gsz, MD_CLUSTER_BLOCK, 0, st>>>(T, G.clusters.map, tolerance);
    }
  }
  MD_HIP(hipGetLastError());
}

void md_momentum(md_group *g) {
  auto &G = ig(g);
  k_clear_momentum<<<G.R, 64, 0, g->stream>>>(G.table, G.R);
  k_momentum<<<G.atoms.total, 256, 0, g->stream>>>(G.table, G.atoms.map);
  MD_HIP(hipGetLastError());
}
