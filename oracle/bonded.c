/*
 * bonded.c -- bonded terms, the position restraint and the whole potential.
 *
 * TEST INFRASTRUCTURE ONLY.  Restates src/reference/forces/bonded.ts:39-227,
 * forces/restraint.ts:38-85 and potential.ts:94-171.  Bonded terms take plain
 * differences with no minimum image, which is the OpenMM convention the
 * reference follows.
 */
#include <math.h>
#include <string.h>

#include "oracle.h"

/* E = 1/2 k (r - r0)^2 (src/reference/forces/bonded.ts:39-67) */
double orc_harmonic_bond(const orc_system *s, const double *p, double *f) {
  double e = 0;
  for (int t = 0; t < s->nb; t++) {
    int a = 3 * s->b1[t], b = 3 * s->b2[t];
    double d[3] = {p[b] - p[a], p[b + 1] - p[a + 1], p[b + 2] - p[a + 2]};
    double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double dr = r - s->blen[t];
    e += 0.5 * s->bk[t] * dr * dr;
    if (f) {
      double sc = s->bk[t] * dr / r;
      for (int k = 0; k < 3; k++) {
        f[a + k] += sc * d[k];
        f[b + k] -= sc * d[k];
      }
    }
  }
  return e;
}

/* E = 1/2 k (theta - theta0)^2 (src/reference/forces/bonded.ts:71-128) */
double orc_harmonic_angle(const orc_system *s, const double *p, double *f) {
  double e = 0;
  for (int t = 0; t < s->na; t++) {
    int a = 3 * s->a1[t], b = 3 * s->a2[t], c = 3 * s->a3[t];
    double v1[3] = {p[a] - p[b], p[a + 1] - p[b + 1], p[a + 2] - p[b + 2]};
    double v2[3] = {p[c] - p[b], p[c + 1] - p[b + 1], p[c + 2] - p[b + 2]};
    double r1 = v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2];
    double r2 = v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2];
    double dot = v1[0] * v2[0] + v1[1] * v2[1] + v1[2] * v2[2];
    double cs = dot / sqrt(r1 * r2);
    cs = cs > 1 ? 1 : (cs < -1 ? -1 : cs);
    double theta = acos(cs);
    double dth = theta - s->atheta[t];
    e += 0.5 * s->ak[t] * dth * dth;
    if (!f) continue;
    double n[3] = {v1[1] * v2[2] - v1[2] * v2[1], v1[2] * v2[0] - v1[0] * v2[2],
                   v1[0] * v2[1] - v1[1] * v2[0]};
    double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (nn < 1e-6) nn = 1e-6;
    double de = s->ak[t] * dth;
    double ta = de / (r1 * nn), tc = -de / (r2 * nn);
    double f1[3] = {-ta * (v1[1] * n[2] - v1[2] * n[1]), -ta * (v1[2] * n[0] - v1[0] * n[2]),
                    -ta * (v1[0] * n[1] - v1[1] * n[0])};
    double f3[3] = {-tc * (v2[1] * n[2] - v2[2] * n[1]), -tc * (v2[2] * n[0] - v2[0] * n[2]),
                    -tc * (v2[0] * n[1] - v2[1] * n[0])};
    for (int k = 0; k < 3; k++) {
      f[a + k] += f1[k];
      f[c + k] += f3[k];
      f[b + k] -= f1[k] + f3[k];
    }
  }
  return e;
}

/* E = k (1 + cos(n phi - phase)) (src/reference/forces/bonded.ts:133-227) */
double orc_periodic_torsion(const orc_system *s, const double *p, double *f) {
  double e = 0;
  for (int t = 0; t < s->nt; t++) {
    int a = 3 * s->t1[t], b = 3 * s->t2[t], c = 3 * s->t3[t], d = 3 * s->t4[t];
    double b1[3], b2[3], b3[3];
    for (int k = 0; k < 3; k++) {
      b1[k] = p[b + k] - p[a + k];
      b2[k] = p[c + k] - p[b + k];
      b3[k] = p[d + k] - p[c + k];
    }
    double n1[3] = {b1[1] * b2[2] - b1[2] * b2[1], b1[2] * b2[0] - b1[0] * b2[2],
                    b1[0] * b2[1] - b1[1] * b2[0]};
    double n2[3] = {b2[1] * b3[2] - b2[2] * b3[1], b2[2] * b3[0] - b2[0] * b3[2],
                    b2[0] * b3[1] - b2[1] * b3[0]};
    double n1sq = n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2];
    double n2sq = n2[0] * n2[0] + n2[1] * n2[1] + n2[2] * n2[2];
    double b2n = sqrt(b2[0] * b2[0] + b2[1] * b2[1] + b2[2] * b2[2]);
    double x = n1[0] * n2[0] + n1[1] * n2[1] + n1[2] * n2[2];
    double y = b2n * (b1[0] * n2[0] + b1[1] * n2[1] + b1[2] * n2[2]);
    double phi = atan2(y, x);
    int per = s->tper[t];
    double k = s->tk[t], ph = s->tphase[t];
    e += k * (1.0 + cos(per * phi - ph));
    if (!f) continue;
    double dedphi = -k * per * sin(per * phi - ph);
    if (n1sq == 0 || n2sq == 0) continue;
    double s1 = dedphi * b2n / n1sq, s4 = -dedphi * b2n / n2sq;
    double f1[3], f4[3];
    for (int q = 0; q < 3; q++) {
      f1[q] = s1 * n1[q];
      f4[q] = s4 * n2[q];
    }
    double b2sq = b2n * b2n;
    double ss = (b1[0] * b2[0] + b1[1] * b2[1] + b1[2] * b2[2]) / b2sq;
    double tt = (b3[0] * b2[0] + b3[1] * b2[1] + b3[2] * b2[2]) / b2sq;
    for (int q = 0; q < 3; q++) {
      double f2 = -(1 + ss) * f1[q] + tt * f4[q];
      double f3 = ss * f1[q] - (1 + tt) * f4[q];
      f[a + q] += f1[q];
      f[b + q] += f2;
      f[c + q] += f3;
      f[d + q] += f4[q];
    }
  }
  return e;
}

/* E = k |minimum_image(x - x0)|^2 (src/reference/forces/restraint.ts:38-85) */
double orc_position_restraint(const orc_system *s, const double *p, const double box_in[9],
                              double *f) {
  double box[9];
  if (orc_is_reduced(box_in, 1e-9))
    memcpy(box, box_in, sizeof box);
  else
    orc_reduce_box(box_in, box);
  double e = 0;
  for (int t = 0; t < s->nr; t++) {
    int i = s->rp[t];
    double d[3] = {p[3 * i] - s->rx0[t], p[3 * i + 1] - s->ry0[t], p[3 * i + 2] - s->rz0[t]};
    orc_nearest_image(box, d);
    double k = s->rk[t];
    e += k * (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (f)
      for (int a = 0; a < 3; a++) f[3 * i + a] -= 2.0 * k * d[a];
  }
  return e;
}

double orc_potential(const orc_system *s, const double *pos, const double box[9],
                     const int *grid_in, double *forces, double groups[ORC_G_COUNT]) {
  double g[ORC_G_COUNT] = {0};
  if (forces) memset(forces, 0, sizeof(double) * 3 * s->n);
  g[ORC_G_BOND] = orc_harmonic_bond(s, pos, forces);
  g[ORC_G_ANGLE] = orc_harmonic_angle(s, pos, forces);
  g[ORC_G_TORSION] = orc_periodic_torsion(s, pos, forces);
  if (s->has_nonbonded) {
    double alpha = orc_alpha_for(s);
    orc_direct dir;
    orc_direct_space(s, pos, box, alpha, 0, forces, &dir);
    g[ORC_G_DIRECT] = dir.total;
    int grid[3];
    if (grid_in && grid_in[0] > 0)
      memcpy(grid, grid_in, sizeof grid);
    else if (s->grid[0] > 0)
      memcpy(grid, s->grid, sizeof grid);
    else
      orc_default_grid(s, box, alpha, grid);
    double volume = orc_box_volume(box);
    g[ORC_G_RECIPROCAL] = orc_pme_reciprocal(s, pos, box, alpha, grid, forces) +
                          orc_self_energy(s, alpha) +
                          orc_net_charge_correction(s, alpha, volume);
  }
  g[ORC_G_RESTRAINT] = orc_position_restraint(s, pos, box, forces);
  double total = 0;
  for (int k = 0; k < ORC_G_COUNT; k++) total += g[k];
  if (groups) memcpy(groups, g, sizeof g);
  return total;
}
