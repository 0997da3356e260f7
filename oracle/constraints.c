/*
 * constraints.c -- constraint clustering, SETTLE, SHAKE and RATTLE in f64.
 *
 * TEST INFRASTRUCTURE ONLY.  Restates src/engine/constraint-clusters.ts:93-249
 * (SETTLE triangles by f32-rounded equal sides, union-find for the rest),
 * src/reference/settle.ts:31-284 (Miyamoto & Kollman, J Comput Chem 13 (1992)
 * 952, in OpenMM's Reference formulation) and src/reference/constraints.ts
 * :130-243 (Jacobi SHAKE/RATTLE along the reference bond vectors).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static int find_root(int *parent, int x) {
  int r = x;
  while (parent[r] != r) r = parent[r];
  while (parent[x] != r) {
    int nx = parent[x];
    parent[x] = r;
    x = nx;
  }
  return r;
}

int orc_split_constraints(const orc_system *s, orc_split *out) {
  int n = s->n, nc = s->nc;
  memset(out, 0, sizeof *out);
  int *count = calloc(n > 0 ? n : 1, sizeof(int));
  char *live = calloc(nc > 0 ? nc : 1, 1);
  for (int c = 0; c < nc; c++) {
    int p1 = s->c1[c], p2 = s->c2[c];
    if (s->mass[p1] == 0.0 && s->mass[p2] == 0.0) continue;
    live[c] = 1;
    count[p1]++;
    count[p2]++;
  }
  /* partners: up to two per atom, only for atoms in exactly two constraints */
  int *pa = malloc(sizeof(int) * 2 * (n > 0 ? n : 1));
  float *pd = malloc(sizeof(float) * 2 * (n > 0 ? n : 1));
  int *np = calloc(n > 0 ? n : 1, sizeof(int));
  for (int c = 0; c < nc; c++) {
    if (!live[c]) continue;
    int p1 = s->c1[c], p2 = s->c2[c];
    if (count[p1] == 2 && count[p2] == 2) {
      float d = (float)s->cd[c];
      if (np[p1] < 2) {
        pa[2 * p1 + np[p1]] = p2;
        pd[2 * p1 + np[p1]] = d;
      }
      np[p1]++;
      if (np[p2] < 2) {
        pa[2 * p2 + np[p2]] = p1;
        pd[2 * p2 + np[p2]] = d;
      }
      np[p2]++;
    }
  }
  out->settle = malloc(sizeof(int) * 3 * (n / 3 + 1));
  out->settle_d = malloc(sizeof(double) * 2 * (n / 3 + 1));
  char *is_settle = calloc(n > 0 ? n : 1, 1);
  for (int p1 = 0; p1 < n; p1++) {
    if (np[p1] != 2) continue;
    int p2 = pa[2 * p1], p3 = pa[2 * p1 + 1];
    if (!(p1 < p2 && p1 < p3)) continue;
    if (np[p2] != 2 || np[p3] != 2) continue;
    /* p2 and p3 must be partners of each other */
    float d23 = -1;
    for (int k = 0; k < 2; k++)
      if (pa[2 * p2 + k] == p3) d23 = pd[2 * p2 + k];
    if (d23 < 0) continue;
    float d12 = pd[2 * p1], d13 = pd[2 * p1 + 1];
    int central, first, second;
    float dist1, dist2;
    if (d12 == d13) {
      central = p1, first = p2, second = p3, dist1 = d12, dist2 = d23;
    } else if (d12 == d23) {
      central = p2, first = p1, second = p3, dist1 = d12, dist2 = d13;
    } else if (d13 == d23) {
      central = p3, first = p1, second = p2, dist1 = d13, dist2 = d12;
    } else {
      continue;
    }
    int k = out->n_settle++;
    out->settle[3 * k] = central;
    out->settle[3 * k + 1] = first;
    out->settle[3 * k + 2] = second;
    out->settle_d[2 * k] = dist1;
    out->settle_d[2 * k + 1] = dist2;
    is_settle[p1] = is_settle[p2] = is_settle[p3] = 1;
  }
  out->iterative = malloc(sizeof(int) * (nc > 0 ? nc : 1));
  int *parent = malloc(sizeof(int) * (n > 0 ? n : 1));
  for (int i = 0; i < n; i++) parent[i] = i;
  for (int c = 0; c < nc; c++) {
    if (!live[c] || is_settle[s->c1[c]]) continue;
    out->iterative[out->n_iterative++] = c;
    int a = find_root(parent, s->c1[c]), b = find_root(parent, s->c2[c]);
    if (a != b) parent[a] = b;
  }
  /* cluster sizes */
  int *atoms = calloc(n > 0 ? n : 1, sizeof(int));
  int *cons = calloc(n > 0 ? n : 1, sizeof(int));
  char *seen = calloc(n > 0 ? n : 1, 1);
  for (int k = 0; k < out->n_iterative; k++) {
    int c = out->iterative[k];
    int r = find_root(parent, s->c1[c]);
    cons[r]++;
    int ends[2] = {s->c1[c], s->c2[c]};
    for (int e = 0; e < 2; e++)
      if (!seen[ends[e]]) {
        seen[ends[e]] = 1;
        atoms[r]++;
      }
  }
  for (int i = 0; i < n; i++) {
    if (atoms[i] > out->max_atoms) out->max_atoms = atoms[i];
    if (cons[i] > out->max_constraints) out->max_constraints = cons[i];
  }
  out->error = out->max_atoms > 8 || out->max_constraints > 8;
  free(count);
  free(live);
  free(pa);
  free(pd);
  free(np);
  free(is_settle);
  free(parent);
  free(atoms);
  free(cons);
  free(seen);
  return out->error;
}

void orc_split_free(orc_split *split) {
  free(split->settle);
  free(split->settle_d);
  free(split->iterative);
  memset(split, 0, sizeof *split);
}

/* rigid-body projection of one water (src/reference/settle.ts:31-196) */
void orc_settle_positions(const orc_system *s, const orc_split *split, const double *ref,
                          double *trial) {
  for (int w = 0; w < split->n_settle; w++) {
    int a0 = 3 * split->settle[3 * w], a1 = 3 * split->settle[3 * w + 1],
        a2 = 3 * split->settle[3 * w + 2];
    double m0 = s->mass[a0 / 3], m1 = s->mass[a1 / 3], m2 = s->mass[a2 / 3];
    double d1 = split->settle_d[2 * w], d2 = split->settle_d[2 * w + 1];
    double xp0[3], xp1[3], xp2[3];
    for (int k = 0; k < 3; k++) {
      xp0[k] = trial[a0 + k] - ref[a0 + k];
      xp1[k] = trial[a1 + k] - ref[a1 + k];
      xp2[k] = trial[a2 + k] - ref[a2 + k];
    }
    double b0[3], c0[3];
    for (int k = 0; k < 3; k++) {
      b0[k] = ref[a1 + k] - ref[a0 + k];
      c0[k] = ref[a2 + k] - ref[a0 + k];
    }
    double inv_m = 1.0 / (m0 + m1 + m2);
    double com[3];
    for (int k = 0; k < 3; k++)
      com[k] = (xp0[k] * m0 + (b0[k] + xp1[k]) * m1 + (c0[k] + xp2[k]) * m2) * inv_m;
    double a1v[3], b1v[3], c1v[3];
    for (int k = 0; k < 3; k++) {
      a1v[k] = xp0[k] - com[k];
      b1v[k] = b0[k] + xp1[k] - com[k];
      c1v[k] = c0[k] + xp2[k] - com[k];
    }
    /* frame: z along the reference normal, x in the plane of the displaced
     * central atom */
    double zd[3] = {b0[1] * c0[2] - b0[2] * c0[1], b0[2] * c0[0] - b0[0] * c0[2],
                    b0[0] * c0[1] - b0[1] * c0[0]};
    double xd[3] = {a1v[1] * zd[2] - a1v[2] * zd[1], a1v[2] * zd[0] - a1v[0] * zd[2],
                    a1v[0] * zd[1] - a1v[1] * zd[0]};
    double yd[3] = {zd[1] * xd[2] - zd[2] * xd[1], zd[2] * xd[0] - zd[0] * xd[2],
                    zd[0] * xd[1] - zd[1] * xd[0]};
    double xl = sqrt(xd[0] * xd[0] + xd[1] * xd[1] + xd[2] * xd[2]);
    double yl = sqrt(yd[0] * yd[0] + yd[1] * yd[1] + yd[2] * yd[2]);
    double zl = sqrt(zd[0] * zd[0] + zd[1] * zd[1] + zd[2] * zd[2]);
    double ex[3], ey[3], ez[3];
    for (int k = 0; k < 3; k++) {
      ex[k] = xd[k] / xl;
      ey[k] = yd[k] / yl;
      ez[k] = zd[k] / zl;
    }
#define DOT(u, v) ((u)[0] * (v)[0] + (u)[1] * (v)[1] + (u)[2] * (v)[2])
    double xb0d = DOT(ex, b0), yb0d = DOT(ey, b0);
    double xc0d = DOT(ex, c0), yc0d = DOT(ey, c0);
    double za1d = DOT(ez, a1v);
    double xb1d = DOT(ex, b1v), yb1d = DOT(ey, b1v), zb1d = DOT(ez, b1v);
    double xc1d = DOT(ex, c1v), yc1d = DOT(ey, c1v), zc1d = DOT(ez, c1v);
    /* canonical triangle */
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
    /* rotation about z that best matches the trial positions */
    double alpha = xb2d * (xb0d - xc0d) + yb0d * yb2d + yc0d * yc2d;
    double beta = xb2d * (yc0d - yb0d) + xb0d * yb2d + xc0d * yc2d;
    double gamma = xb0d * yb1d - xb1d * yb0d + xc0d * yc1d - xc1d * yc0d;
    double al2be2 = alpha * alpha + beta * beta;
    double sintheta = (alpha * gamma - beta * sqrt(al2be2 - gamma * gamma)) / al2be2;
    double costheta = sqrt(1.0 - sintheta * sintheta);
    double a3d[3] = {-ya2d * sintheta, ya2d * costheta, za1d};
    double b3d[3] = {xb2d * costheta - yb2d * sintheta, xb2d * sintheta + yb2d * costheta, zb1d};
    double c3d[3] = {-xb2d * costheta - yc2d * sintheta, -xb2d * sintheta + yc2d * costheta,
                     zc1d};
    for (int k = 0; k < 3; k++) {
      double a3 = ex[k] * a3d[0] + ey[k] * a3d[1] + ez[k] * a3d[2];
      double b3 = ex[k] * b3d[0] + ey[k] * b3d[1] + ez[k] * b3d[2];
      double c3 = ex[k] * c3d[0] + ey[k] * c3d[1] + ez[k] * c3d[2];
      trial[a0 + k] = ref[a0 + k] + com[k] + a3;
      trial[a1 + k] = ref[a1 + k] + com[k] + b3 - b0[k];
      trial[a2 + k] = ref[a2 + k] + com[k] + c3 - c0[k];
    }
  }
}

/* exact 3x3 impulse solve (src/reference/settle.ts:198-284) */
void orc_settle_velocities(const orc_system *s, const orc_split *split, const double *p,
                           double *v) {
  for (int w = 0; w < split->n_settle; w++) {
    int a = 3 * split->settle[3 * w], b = 3 * split->settle[3 * w + 1],
        c = 3 * split->settle[3 * w + 2];
    double mA = s->mass[a / 3], mB = s->mass[b / 3], mC = s->mass[c / 3];
    double iA = mA == 0 ? 0 : 1 / mA, iB = mB == 0 ? 0 : 1 / mB, iC = mC == 0 ? 0 : 1 / mC;
    double eAB[3], eBC[3], eCA[3];
    int from[3] = {a, b, c}, to[3] = {b, c, a};
    double *es[3] = {eAB, eBC, eCA};
    for (int q = 0; q < 3; q++) {
      double x = p[to[q]] - p[from[q]], y = p[to[q] + 1] - p[from[q] + 1],
             z = p[to[q] + 2] - p[from[q] + 2];
      double l = sqrt(x * x + y * y + z * z);
      es[q][0] = x / l;
      es[q][1] = y / l;
      es[q][2] = z / l;
    }
    double rel[3];
    for (int q = 0; q < 3; q++)
      rel[q] = (v[to[q]] - v[from[q]]) * es[q][0] + (v[to[q] + 1] - v[from[q] + 1]) * es[q][1] +
               (v[to[q] + 2] - v[from[q] + 2]) * es[q][2];
    double vAB = rel[0], vBC = rel[1], vCA = rel[2];
    double cA = -DOT(eAB, eCA), cB = -DOT(eAB, eBC), cC = -DOT(eBC, eCA);
    double s2A = 1 - cA * cA, s2B = 1 - cB * cB, s2C = 1 - cC * cC;
    double mABCinv = 1 / (mA * mB * mC);
    double denom = (((s2A * mB + s2B * mA) * mC +
                     (s2A * mB * mB + 2 * (cA * cB * cC + 1) * mA * mB + s2B * mA * mA)) *
                        mC +
                    s2C * mA * mB * (mA + mB)) *
                   mABCinv;
    double tab = ((cB * cC * mA - cA * mB - cA * mC) * vCA +
                  (cA * cC * mB - cB * mC - cB * mA) * vBC +
                  (s2C * mA * mA * mB * mB * mABCinv + (mA + mB + mC)) * vAB) /
                 denom;
    double tbc = ((cA * cB * mC - cC * mB - cC * mA) * vCA +
                  (s2A * mB * mB * mC * mC * mABCinv + (mA + mB + mC)) * vBC +
                  (cA * cC * mB - cB * mA - cB * mC) * vAB) /
                 denom;
    double tca = ((s2B * mA * mA * mC * mC * mABCinv + (mA + mB + mC)) * vCA +
                  (cA * cB * mC - cC * mB - cC * mA) * vBC +
                  (cB * cC * mA - cA * mB - cA * mC) * vAB) /
                 denom;
    for (int k = 0; k < 3; k++) {
      v[a + k] += (eAB[k] * tab - eCA[k] * tca) * iA;
      v[b + k] += (eBC[k] * tbc - eAB[k] * tab) * iB;
      v[c + k] += (eCA[k] * tca - eBC[k] * tbc) * iC;
    }
  }
}

static double inv_mass(const orc_system *s, int i) { return s->mass[i] == 0 ? 0 : 1 / s->mass[i]; }

/* Jacobi SHAKE over the iterative constraints, corrections along the reference
 * vectors, OpenMM's tolerance window (src/reference/constraints.ts:130-186) */
int orc_shake(const orc_system *s, const orc_split *split, const double *ref, double *trial,
              double tol, int max_it, int *converged_out) {
  int count = split->n_iterative;
  if (converged_out) *converged_out = 1;
  if (count == 0) return 0;
  double *rv = malloc(sizeof(double) * 3 * count), *mu = malloc(sizeof(double) * count);
  double *delta = malloc(sizeof(double) * count);
  for (int k = 0; k < count; k++) {
    int c = split->iterative[k];
    int i = 3 * s->c1[c], j = 3 * s->c2[c];
    for (int a = 0; a < 3; a++) rv[3 * k + a] = ref[i + a] - ref[j + a];
    double sum = inv_mass(s, s->c1[c]) + inv_mass(s, s->c2[c]);
    mu[k] = sum == 0 ? 0 : 0.5 / sum;
  }
  double lower = 1 - 2 * tol + tol * tol, upper = 1 + 2 * tol + tol * tol;
  int it = 0, done = 0;
  while (it < max_it) {
    int conv = 0;
    for (int k = 0; k < count; k++) {
      int c = split->iterative[k];
      int i = 3 * s->c1[c], j = 3 * s->c2[c];
      double x = trial[i] - trial[j], y = trial[i + 1] - trial[j + 1], z = trial[i + 2] - trial[j + 2];
      double l2 = x * x + y * y + z * z;
      double t2 = s->cd[c] * s->cd[c];
      double proj = x * rv[3 * k] + y * rv[3 * k + 1] + z * rv[3 * k + 2];
      delta[k] = proj == 0 ? 0 : mu[k] * (t2 - l2) / proj;
      if (l2 >= lower * t2 && l2 <= upper * t2) conv++;
    }
    if (conv == count) {
      done = 1;
      break;
    }
    it++;
    for (int k = 0; k < count; k++) {
      if (delta[k] == 0) continue;
      int c = split->iterative[k];
      int p1 = s->c1[c], p2 = s->c2[c];
      double i1 = inv_mass(s, p1), i2 = inv_mass(s, p2);
      for (int a = 0; a < 3; a++) {
        double dr = rv[3 * k + a] * delta[k];
        trial[3 * p1 + a] += dr * i1;
        trial[3 * p2 + a] -= dr * i2;
      }
    }
  }
  free(rv);
  free(mu);
  free(delta);
  if (converged_out) *converged_out = done;
  return it;
}

/* Jacobi RATTLE, divided by the reference squared length
 * (src/reference/constraints.ts:195-243) */
int orc_rattle(const orc_system *s, const orc_split *split, const double *pos, double *vel,
               double tol, int max_it, int *converged_out) {
  int count = split->n_iterative;
  if (converged_out) *converged_out = 1;
  if (count == 0) return 0;
  double *rv = malloc(sizeof(double) * 3 * count), *mu = malloc(sizeof(double) * count);
  double *l2 = malloc(sizeof(double) * count), *delta = malloc(sizeof(double) * count);
  for (int k = 0; k < count; k++) {
    int c = split->iterative[k];
    int i = 3 * s->c1[c], j = 3 * s->c2[c];
    for (int a = 0; a < 3; a++) rv[3 * k + a] = pos[i + a] - pos[j + a];
    l2[k] = rv[3 * k] * rv[3 * k] + rv[3 * k + 1] * rv[3 * k + 1] + rv[3 * k + 2] * rv[3 * k + 2];
    double sum = inv_mass(s, s->c1[c]) + inv_mass(s, s->c2[c]);
    mu[k] = sum == 0 ? 0 : 0.5 / sum;
  }
  int it = 0, done = 0;
  while (it < max_it) {
    int conv = 0;
    for (int k = 0; k < count; k++) {
      int c = split->iterative[k];
      int i = 3 * s->c1[c], j = 3 * s->c2[c];
      double proj = (vel[i] - vel[j]) * rv[3 * k] + (vel[i + 1] - vel[j + 1]) * rv[3 * k + 1] +
                    (vel[i + 2] - vel[j + 2]) * rv[3 * k + 2];
      delta[k] = l2[k] == 0 ? 0 : -2 * mu[k] * proj / l2[k];
      if (fabs(delta[k]) <= tol) conv++;
    }
    if (conv == count) {
      done = 1;
      break;
    }
    it++;
    for (int k = 0; k < count; k++) {
      if (delta[k] == 0) continue;
      int c = split->iterative[k];
      int p1 = s->c1[c], p2 = s->c2[c];
      double i1 = inv_mass(s, p1), i2 = inv_mass(s, p2);
      for (int a = 0; a < 3; a++) {
        double dr = rv[3 * k + a] * delta[k];
        vel[3 * p1 + a] += dr * i1;
        vel[3 * p2 + a] -= dr * i2;
      }
    }
  }
  free(rv);
  free(mu);
  free(l2);
  free(delta);
  if (converged_out) *converged_out = done;
  return it;
}
