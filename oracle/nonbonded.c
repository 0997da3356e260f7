/*
 * nonbonded.c -- real-space Ewald + Lennard-Jones in f64, for the oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  Restates src/reference/forces/nonbonded.ts:
 * the pair term (:279-321), the exception terms (:330-386), the O(N^2) loop
 * (:394-437), the cell-list loop (:447-575), the dispersion correction
 * (:107-162), the self energy (:610-617) and the net-charge term (:594-607).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

double orc_alpha_for(const orc_system *s) {
  return s->alpha != 0.0 ? s->alpha : orc_ewald_alpha(s->cutoff, s->ewald_tolerance);
}

/* mesh from the box diagonal (src/reference/potential.ts:79-92) */
void orc_default_grid(const orc_system *s, const double box[9], double alpha, int grid[3]) {
  for (int a = 0; a < 3; a++)
    grid[a] = orc_smooth_fft_size(
        orc_pme_minimum_grid(alpha, box[a * 3 + a], s->ewald_tolerance));
}

/* grouped by distinct (sigma, epsilon), averaged over the N(N+1)/2 unordered
 * pairs including i == j (src/reference/forces/nonbonded.ts:107-162) */
double orc_dispersion_correction(const orc_system *s, double volume) {
  int n = s->n;
  if (n == 0) return 0.0;
  double *cs = malloc(sizeof(double) * n), *ce = malloc(sizeof(double) * n);
  double *cnt = malloc(sizeof(double) * n);
  int classes = 0;
  for (int i = 0; i < n; i++) {
    int c;
    for (c = 0; c < classes; c++)
      if (cs[c] == s->sigma[i] && ce[c] == s->epsilon[i]) break;
    if (c == classes) {
      cs[c] = s->sigma[i];
      ce[c] = s->epsilon[i];
      cnt[c] = 0;
      classes++;
    }
    cnt[c] += 1;
  }
  double s6 = 0, s12 = 0;
  for (int c = 0; c < classes; c++)
    for (int d = c; d < classes; d++) {
      double sig = 0.5 * (cs[c] + cs[d]);
      double eps = sqrt(ce[c] * ce[d]);
      double sig6 = pow(sig, 6);
      double mult = c == d ? cnt[c] * (cnt[c] + 1) / 2 : cnt[c] * cnt[d];
      s6 += mult * eps * sig6;
      s12 += mult * eps * sig6 * sig6;
    }
  free(cs);
  free(ce);
  free(cnt);
  double pairs = (double)n * (n + 1) / 2.0;
  double rc = s->cutoff;
  return (8.0 * M_PI * (double)n * n / volume) *
         ((s12 / pairs) / (9.0 * pow(rc, 9)) - (s6 / pairs) / (3.0 * pow(rc, 3)));
}

double orc_dispersion_correction_naive(const orc_system *s, double volume) {
  int n = s->n;
  if (n == 0) return 0.0;
  double s6 = 0, s12 = 0, pairs = 0;
  for (int i = 0; i < n; i++)
    for (int j = i; j < n; j++) {
      double sig = 0.5 * (s->sigma[i] + s->sigma[j]);
      double eps = sqrt(s->epsilon[i] * s->epsilon[j]);
      double sig6 = pow(sig, 6);
      s6 += eps * sig6;
      s12 += eps * sig6 * sig6;
      pairs += 1;
    }
  double rc = s->cutoff;
  return (8.0 * M_PI * (double)n * n / volume) *
         (s12 / pairs / (9.0 * pow(rc, 9)) - s6 / pairs / (3.0 * pow(rc, 3)));
}

double orc_self_energy(const orc_system *s, double alpha) {
  double q2 = 0;
  for (int i = 0; i < s->n; i++) q2 += s->q[i] * s->q[i];
  return -ORC_ONE_4PI_EPS0 * alpha * q2 / sqrt(M_PI);
}

double orc_net_charge_correction(const orc_system *s, double alpha, double volume) {
  double q = 0;
  for (int i = 0; i < s->n; i++) q += s->q[i];
  return -M_PI * q * q * ORC_ONE_4PI_EPS0 / (2.0 * alpha * alpha * volume);
}

/* d/dr erf(x) = 2/sqrt(pi) exp(-x^2) */
static double erf_derivative(double x) { return 2.0 / sqrt(M_PI) * exp(-x * x); }

/* sorted excluded-pair keys, for bsearch */
typedef struct {
  int64_t *keys;
  int count;
} exclusion_set;

static int cmp_i64(const void *a, const void *b) {
  int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
  return x < y ? -1 : x > y;
}

static exclusion_set make_exclusions(const orc_system *s) {
  exclusion_set e;
  e.count = s->ne;
  e.keys = malloc(sizeof(int64_t) * (s->ne > 0 ? s->ne : 1));
  for (int k = 0; k < s->ne; k++) {
    int64_t a = s->e1[k], b = s->e2[k];
    e.keys[k] = (a < b ? a : b) * 2097152LL + (a < b ? b : a);
  }
  qsort(e.keys, e.count, sizeof(int64_t), cmp_i64);
  return e;
}

static int excluded(const exclusion_set *e, int lo, int hi) {
  int64_t key = (int64_t)lo * 2097152LL + hi;
  return bsearch(&key, e->keys, e->count, sizeof(int64_t), cmp_i64) != NULL;
}

typedef struct {
  const orc_system *s;
  const double *pos;
  const double *box;
  double alpha, cutoff2;
  double *forces;
} pair_ctx;

/* one non-excluded pair (src/reference/forces/nonbonded.ts:279-321) */
static int pair_interaction(const pair_ctx *c, int i, int j, double *coul, double *lj) {
  const double *p = c->pos;
  double d[3] = {p[3 * j] - p[3 * i], p[3 * j + 1] - p[3 * i + 1], p[3 * j + 2] - p[3 * i + 2]};
  orc_nearest_image(c->box, d);
  double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  if (r2 >= c->cutoff2) return 0;
  double r = sqrt(r2);
  const orc_system *s = c->s;
  double qq = ORC_ONE_4PI_EPS0 * s->q[i] * s->q[j];
  double ar = c->alpha * r;
  double ec = erfc(ar);
  *coul += qq * ec / r;
  double dedr = qq * (-ec / r2 - c->alpha * erf_derivative(ar) / r);
  double eps = sqrt(s->epsilon[i] * s->epsilon[j]);
  if (eps != 0.0) {
    double sig = 0.5 * (s->sigma[i] + s->sigma[j]);
    double s6 = pow(sig * sig / r2, 3);
    double s12 = s6 * s6;
    *lj += 4.0 * eps * (s12 - s6);
    dedr += (4.0 * eps / r) * (6.0 * s6 - 12.0 * s12);
  }
  if (c->forces) {
    double scale = dedr / r;
    for (int a = 0; a < 3; a++) {
      c->forces[3 * i + a] += scale * d[a];
      c->forces[3 * j + a] -= scale * d[a];
    }
  }
  return 1;
}

/* every exception, at any distance (src/reference/forces/nonbonded.ts:330-386) */
static void exception_terms(const pair_ctx *c, double *corr, double *exc) {
  const orc_system *s = c->s;
  const double *p = c->pos;
  *corr = 0;
  *exc = 0;
  for (int e = 0; e < s->ne; e++) {
    int i = s->e1[e], j = s->e2[e];
    double d[3] = {p[3 * j] - p[3 * i], p[3 * j + 1] - p[3 * i + 1], p[3 * j + 2] - p[3 * i + 2]};
    orc_nearest_image(c->box, d);
    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    double r = sqrt(r2);
    double ar = c->alpha * r;
    double dedr = 0;
    double pp = ORC_ONE_4PI_EPS0 * s->q[i] * s->q[j];
    if (pp != 0.0) {
      double ef = erf(ar);
      *corr -= pp * ef / r;
      dedr += pp * (ef / r2 - c->alpha * erf_derivative(ar) / r);
    }
    double qq = ORC_ONE_4PI_EPS0 * s->eq[e];
    if (qq != 0.0) {
      *exc += qq / r;
      dedr += -qq / r2;
    }
    double eps = s->eeps[e];
    if (eps != 0.0) {
      double sig = s->esig[e];
      double s6 = pow(sig * sig / r2, 3);
      double s12 = s6 * s6;
      *exc += 4.0 * eps * (s12 - s6);
      dedr += (4.0 * eps / r) * (6.0 * s6 - 12.0 * s12);
    }
    if (c->forces && dedr != 0.0) {
      double scale = dedr / r;
      for (int a = 0; a < 3; a++) {
        c->forces[3 * i + a] += scale * d[a];
        c->forces[3 * j + a] -= scale * d[a];
      }
    }
  }
}

void orc_direct_space(const orc_system *s, const double *pos, const double box_in[9],
                      double alpha, int naive, double *forces, orc_direct *out) {
  double box[9];
  if (orc_is_reduced(box_in, 1e-9))
    memcpy(box, box_in, sizeof box);
  else
    orc_reduce_box(box_in, box);
  pair_ctx c = {s, pos, box, alpha, s->cutoff * s->cutoff, forces};
  exclusion_set ex = make_exclusions(s);
  double coul = 0, lj = 0;
  int64_t pairs = 0, geometric = 0;
  int n = s->n;

  double widths[3];
  orc_perpendicular_widths(box, widths);
  int counts[3];
  for (int a = 0; a < 3; a++) {
    counts[a] = (int)floor(widths[a] / s->cutoff);
    if (counts[a] < 1) counts[a] = 1;
  }
  if (counts[0] < 3 || counts[1] < 3 || counts[2] < 3) naive = 1;

  if (naive) {
    for (int i = 0; i < n; i++)
      for (int j = i + 1; j < n; j++) {
        if (excluded(&ex, i, j)) continue;
        double cc = 0, ll = 0;
        if (pair_interaction(&c, i, j, &cc, &ll)) geometric++;
        if (cc != 0.0 || ll != 0.0) pairs++;
        coul += cc;
        lj += ll;
      }
  } else {
    /* fractional bucketing, 27-cell scan, each unordered cell pair once */
    int ncell = counts[0] * counts[1] * counts[2];
    int *cell_of = malloc(sizeof(int) * n);
    int *start = calloc(ncell + 1, sizeof(int));
    int *order = malloc(sizeof(int) * n);
    double inv[9];
    orc_inverse_box(box, inv);
    for (int i = 0; i < n; i++) {
      double f[3];
      orc_fractional(inv, pos + 3 * i, f);
      int cidx[3];
      for (int a = 0; a < 3; a++) {
        double w = f[a] - floor(f[a]);
        int k = (int)floor(w * counts[a]);
        cidx[a] = k == counts[a] ? counts[a] - 1 : k;
      }
      cell_of[i] = (cidx[0] * counts[1] + cidx[1]) * counts[2] + cidx[2];
      start[cell_of[i] + 1]++;
    }
    for (int k = 0; k < ncell; k++) start[k + 1] += start[k];
    int *cursor = malloc(sizeof(int) * ncell);
    memcpy(cursor, start, sizeof(int) * ncell);
    for (int i = 0; i < n; i++) order[cursor[cell_of[i]]++] = i;
    for (int cx = 0; cx < counts[0]; cx++)
      for (int cy = 0; cy < counts[1]; cy++)
        for (int cz = 0; cz < counts[2]; cz++) {
          int cell = (cx * counts[1] + cy) * counts[2] + cz;
          for (int dx = -1; dx <= 1; dx++)
            for (int dy = -1; dy <= 1; dy++)
              for (int dz = -1; dz <= 1; dz++) {
                int nx = (cx + dx + counts[0]) % counts[0];
                int ny = (cy + dy + counts[1]) % counts[1];
                int nz = (cz + dz + counts[2]) % counts[2];
                int nb = (nx * counts[1] + ny) * counts[2] + nz;
                if (nb < cell) continue;
                for (int a = start[cell]; a < start[cell + 1]; a++) {
                  int i = order[a];
                  int from = nb == cell ? a + 1 : start[nb];
                  for (int b = from; b < start[nb + 1]; b++) {
                    int j = order[b];
                    int lo = i < j ? i : j, hi = i < j ? j : i;
                    if (excluded(&ex, lo, hi)) continue;
                    double cc = 0, ll = 0;
                    if (pair_interaction(&c, lo, hi, &cc, &ll)) geometric++;
                    if (cc != 0.0 || ll != 0.0) pairs++;
                    coul += cc;
                    lj += ll;
                  }
                }
              }
        }
    free(cell_of);
    free(start);
    free(order);
    free(cursor);
  }
  free(ex.keys);
  double corr, exc;
  exception_terms(&c, &corr, &exc);
  double disp = s->dispersion ? orc_dispersion_correction(s, orc_box_volume(box)) : 0.0;
  out->coulomb = coul;
  out->lennard_jones = lj;
  out->exclusion_correction = corr;
  out->exceptions = exc;
  out->dispersion = disp;
  out->total = coul + lj + corr + exc + disp;
  out->pairs_within_cutoff = pairs;
  out->pairs_geometric = geometric;
}
