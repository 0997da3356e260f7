/*
 * integrator.c -- LFMiddle and Verlet steps, COM removal, kinetic energy and
 * the Maxwell-Boltzmann draw, in f64.
 *
 * TEST INFRASTRUCTURE ONLY.  Restates src/reference/integrator.ts:68-220,
 * com-motion.ts:48-74, kinetic.ts:31-113 and velocities.ts:78-117.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static double inv_mass(const orc_system *s, int i) { return s->mass[i] == 0 ? 0 : 1 / s->mass[i]; }

/* src/reference/com-motion.ts:48-74 */
void orc_remove_com(const orc_system *s, double *v) {
  double m = 0, px = 0, py = 0, pz = 0;
  for (int i = 0; i < s->n; i++) {
    m += s->mass[i];
    px += s->mass[i] * v[3 * i];
    py += s->mass[i] * v[3 * i + 1];
    pz += s->mass[i] * v[3 * i + 2];
  }
  if (m == 0) return;
  px /= m;
  py /= m;
  pz /= m;
  for (int i = 0; i < s->n; i++) {
    if (s->mass[i] == 0) continue;
    v[3 * i] -= px;
    v[3 * i + 1] -= py;
    v[3 * i + 2] -= pz;
  }
}

double orc_kinetic_energy(const orc_system *s, const double *v) {
  double t = 0;
  for (int i = 0; i < s->n; i++) {
    if (s->mass[i] <= 0) continue;
    t += s->mass[i] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
  }
  return 0.5 * t;
}

/* 3 N_moving - N_constraints - 3 (src/reference/kinetic.ts:104-113) */
int orc_degrees_of_freedom(const orc_system *s) {
  int moving = 0;
  for (int i = 0; i < s->n; i++)
    if (s->mass[i] > 0) moving++;
  return 3 * moving - s->nc - (s->cmm_frequency > 0 ? 3 : 0);
}

static int constrain_positions(const orc_system *s, const orc_split *sp, const double *ref,
                               double *trial, double tol) {
  orc_settle_positions(s, sp, ref, trial);
  int conv;
  return orc_shake(s, sp, ref, trial, tol, ORC_MAX_CONSTRAINT_ITERATIONS, &conv);
}

static int constrain_velocities(const orc_system *s, const orc_split *sp, const double *pos,
                                double *vel, double tol) {
  orc_settle_velocities(s, sp, pos, vel);
  int conv;
  return orc_rattle(s, sp, pos, vel, tol, ORC_MAX_CONSTRAINT_ITERATIONS, &conv);
}

/* v += dt F/m; x' = x + dt v; SHAKE; v = (x' - x)/dt
 * (src/reference/integrator.ts:68-112) */
int orc_verlet_step(const orc_system *s, const orc_split *sp, double *pos, double *vel,
                    const double *f, double dt, double tol) {
  int n = s->n;
  double *prev = malloc(sizeof(double) * 3 * n), *trial = malloc(sizeof(double) * 3 * n);
  memcpy(prev, pos, sizeof(double) * 3 * n);
  for (int i = 0; i < n; i++) {
    double im = inv_mass(s, i);
    for (int a = 0; a < 3; a++) {
      int at = 3 * i + a;
      if (im == 0) {
        trial[at] = prev[at];
        continue;
      }
      vel[at] += im * f[at] * dt;
      trial[at] = prev[at] + vel[at] * dt;
    }
  }
  int it = constrain_positions(s, sp, prev, trial, tol);
  for (int i = 0; i < n; i++) {
    if (inv_mass(s, i) == 0) continue;
    for (int a = 0; a < 3; a++) {
      int at = 3 * i + a;
      vel[at] = (trial[at] - prev[at]) / dt;
      pos[at] = trial[at];
    }
  }
  free(prev);
  free(trial);
  return it;
}

/* kick, RATTLE, half drift, thermostat, half drift, SHAKE, velocity
 * correction (src/reference/integrator.ts:140-220) */
int orc_lfmiddle_step(const orc_system *s, const orc_split *sp, double *pos, double *vel,
                      const double *f, double dt, double temperature, double friction,
                      double tol, uint32_t seed, uint32_t step) {
  int n = s->n;
  double half = 0.5 * dt;
  double kT = ORC_BOLTZMANN * temperature;
  double vscale = exp(-dt * friction);
  double nscale = sqrt(1 - vscale * vscale);
  for (int i = 0; i < n; i++) {
    double im = inv_mass(s, i);
    if (im == 0) continue;
    for (int a = 0; a < 3; a++) vel[3 * i + a] += dt * im * f[3 * i + a];
  }
  constrain_velocities(s, sp, pos, vel, tol);
  double *trial = malloc(sizeof(double) * 3 * n), *unc = malloc(sizeof(double) * 3 * n);
  memcpy(trial, pos, sizeof(double) * 3 * n);
  for (int i = 0; i < n; i++) {
    double im = inv_mass(s, i);
    if (im == 0) continue;
    double noise = nscale * sqrt(kT * im);
    for (int a = 0; a < 3; a++) {
      int at = 3 * i + a;
      trial[at] += vel[at] * half;
      vel[at] = vscale * vel[at] +
                (noise == 0 ? 0 : noise * orc_thermostat_normal(seed, step, (uint32_t)i, a));
      trial[at] += vel[at] * half;
    }
  }
  memcpy(unc, trial, sizeof(double) * 3 * n);
  int it = constrain_positions(s, sp, pos, trial, tol);
  for (int i = 0; i < n; i++) {
    if (inv_mass(s, i) == 0) continue;
    for (int a = 0; a < 3; a++) {
      int at = 3 * i + a;
      vel[at] += (trial[at] - unc[at]) / dt;
      pos[at] = trial[at];
    }
  }
  free(trial);
  free(unc);
  return it;
}

/* src/reference/velocities.ts:78-117 */
void orc_velocities_at_temperature(const orc_system *s, const orc_split *sp,
                                   const double *pos, const double *forces, double temperature,
                                   uint32_t seed, double offset, double *vel) {
  double kT = ORC_BOLTZMANN * temperature;
  memset(vel, 0, sizeof(double) * 3 * s->n);
  for (int i = 0; i < s->n; i++) {
    double m = s->mass[i];
    if (m == 0) continue;
    double scale = sqrt(kT / m);
    for (int a = 0; a < 3; a++) {
      double v = scale * orc_domain_normal(ORC_VELOCITY_DOMAIN, seed, (uint32_t)i, a);
      if (offset != 0 && forces) v -= (offset / m) * forces[3 * i + a];
      vel[3 * i + a] = v;
    }
  }
  constrain_velocities(s, sp, pos, vel, 1e-5);
}

double orc_run_lfmiddle(const orc_system *s, double *pos, double *vel, const double box[9],
                        const int *grid, int steps, double dt, double temperature,
                        double friction, double tol, uint32_t seed, uint32_t first_step) {
  orc_split sp;
  orc_split_constraints(s, &sp);
  double *f = malloc(sizeof(double) * 3 * s->n);
  double e = 0;
  for (int k = 0; k < steps; k++) {
    uint32_t step = first_step + (uint32_t)k;
    if (s->cmm_frequency > 0 && step % (uint32_t)s->cmm_frequency == 0) orc_remove_com(s, vel);
    e = orc_potential(s, pos, box, grid, f, NULL);
    orc_lfmiddle_step(s, &sp, pos, vel, f, dt, temperature, friction, tol, seed, step);
  }
  free(f);
  orc_split_free(&sp);
  return e;
}
