/*
 * pme.c -- smooth PME, B-splines, 3D transforms and the exact Ewald sum, f64.
 *
 * TEST INFRASTRUCTURE ONLY.  Restates src/reference/forces/pme.ts:94-263,
 * forces/bspline.ts:21-132, dft.ts:22-107 and forces/ewald.ts:52-152.  The
 * mixed-radix transform is a recursive decimation-in-time written from the
 * DFT definition; it is checked against the O(n^2) DFT, which is how the
 * reference checks its own Stockham transform (test/fft.test.ts:19-80).
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* order-`order` cardinal B-spline weights at offset w in [0,1) and their
 * derivatives, by the standard recursion from the hat function
 * (src/reference/forces/bspline.ts:55-91) */
void orc_bspline(double w, int order, double *weights, double *derivatives) {
  double lower[16] = {0};
  lower[1] = w;
  lower[0] = 1.0 - w;
  for (int k = 3; k <= order - 1; k++) {
    double den = 1.0 / (k - 1);
    lower[k - 1] = den * w * lower[k - 2];
    for (int j = 1; j < k - 1; j++)
      lower[k - j - 1] = den * ((w + j) * lower[k - j - 2] + (k - j - w) * lower[k - j - 1]);
    lower[0] = den * (1.0 - w) * lower[0];
  }
  if (derivatives) {
    derivatives[0] = -lower[0];
    for (int j = 1; j < order; j++) derivatives[j] = lower[j - 1] - lower[j];
  }
  double den = 1.0 / (order - 1);
  memcpy(weights, lower, sizeof(double) * order);
  weights[order - 1] = den * w * weights[order - 2];
  for (int j = 1; j < order - 1; j++)
    weights[order - j - 1] =
        den * ((w + j) * weights[order - j - 2] + (order - j - w) * weights[order - j - 1]);
  weights[0] = den * (1.0 - w) * weights[0];
}

/* |b(m)|^2 with the neighbour-average guard below 1e-7
 * (src/reference/forces/bspline.ts:102-132) */
void orc_bspline_moduli(int n, int order, double *moduli) {
  double *data = calloc(n, sizeof(double));
  double knots[16];
  orc_bspline(0.0, order, knots, NULL);
  for (int i = 0; i < order && i < n; i++) data[i] = knots[i];
  for (int m = 0; m < n; m++) {
    double re = 0, im = 0;
    for (int k = 0; k < n; k++) {
      double a = 2.0 * M_PI * m * k / n;
      re += data[k] * cos(a);
      im += data[k] * sin(a);
    }
    moduli[m] = re * re + im * im;
  }
  for (int m = 0; m < n; m++)
    if (moduli[m] < 1e-7) moduli[m] = 0.5 * (moduli[(m - 1 + n) % n] + moduli[(m + 1) % n]);
  free(data);
}

/* ---- transforms ------------------------------------------------------ */

static void dft1(const double complex *in, double complex *out, int n, int stride, int sign) {
  for (int k = 0; k < n; k++) {
    double complex acc = 0;
    for (int j = 0; j < n; j++) {
      double a = sign * 2.0 * M_PI * (double)((long)j * k % n) / n;
      acc += in[(long)j * stride] * (cos(a) + I * sin(a));
    }
    out[k] = acc;
  }
}

static int smallest_factor(int n) {
  for (int p = 2; p * p <= n; p++)
    if (n % p == 0) return p;
  return n;
}

/* out[k] = sum_j in[j*stride] e^{sign 2 pi i jk/n}; recursive mixed radix */
static void fft1(const double complex *in, double complex *out, int n, int stride, int sign) {
  int r = smallest_factor(n);
  if (r == n) {
    dft1(in, out, n, stride, sign);
    return;
  }
  int m = n / r;
  double complex *sub = malloc(sizeof(double complex) * n);
  for (int s = 0; s < r; s++) fft1(in + (long)s * stride, sub + (long)s * m, m, stride * r, sign);
  for (int k = 0; k < n; k++) {
    double complex acc = 0;
    for (int s = 0; s < r; s++) {
      double a = sign * 2.0 * M_PI * (double)((long)s * k % n) / n;
      acc += sub[(long)s * m + k % m] * (cos(a) + I * sin(a));
    }
    out[k] = acc;
  }
  free(sub);
}

static void transform3d(double *data, int nx, int ny, int nz, int sign, int direct) {
  double complex *g = (double complex *)data;
  int dims[3] = {nx, ny, nz};
  long strides[3] = {(long)ny * nz, nz, 1};
  int nmax = nx > ny ? nx : ny;
  nmax = nmax > nz ? nmax : nz;
  double complex *line = malloc(sizeof(double complex) * nmax);
  double complex *res = malloc(sizeof(double complex) * nmax);
  long total = (long)nx * ny * nz;
  for (int axis = 2; axis >= 0; axis--) {
    int n = dims[axis];
    long st = strides[axis];
    for (long base = 0; base < total; base++) {
      /* visit every line once: its first element has index 0 along `axis` */
      if ((base / st) % n != 0) continue;
      for (int j = 0; j < n; j++) line[j] = g[base + j * st];
      if (direct)
        dft1(line, res, n, 1, sign);
      else
        fft1(line, res, n, 1, sign);
      for (int j = 0; j < n; j++) g[base + j * st] = res[j];
    }
  }
  free(line);
  free(res);
}

void orc_fft3d(double *data, int nx, int ny, int nz, int sign) {
  transform3d(data, nx, ny, nz, sign, 0);
}

void orc_dft3d(double *data, int nx, int ny, int nz, int sign) {
  transform3d(data, nx, ny, nz, sign, 1);
}

/* ---- reciprocal space --------------------------------------------------- */

typedef struct {
  int base;
  double w[ORC_PME_ORDER], dw[ORC_PME_ORDER];
} spline;

/* grid points floor(u) - order + 1 + k (src/reference/forces/pme.ts:73-84) */
static void spline_for(double fractional, int n, spline *sp) {
  double fraction = fractional - floor(fractional);
  double scaled = fraction * n;
  double cell = floor(scaled);
  orc_bspline(scaled - cell, ORC_PME_ORDER, sp->w, sp->dw);
  sp->base = (int)cell - ORC_PME_ORDER + 1;
}

static int wrap(int i, int n) { return ((i % n) + n) % n; }

double orc_pme_reciprocal(const orc_system *s, const double *pos, const double box_in[9],
                          double alpha, const int grid[3], double *forces) {
  double box[9], inv[9];
  if (orc_is_reduced(box_in, 1e-9))
    memcpy(box, box_in, sizeof box);
  else
    orc_reduce_box(box_in, box);
  orc_inverse_box(box, inv);
  int nx = grid[0], ny = grid[1], nz = grid[2];
  long total = (long)nx * ny * nz;
  double *g = calloc(2 * total, sizeof(double));
  spline *sp = malloc(sizeof(spline) * 3 * (s->n > 0 ? s->n : 1));
  for (int i = 0; i < s->n; i++) {
    double f[3];
    orc_fractional(inv, pos + 3 * i, f);
    spline_for(f[0], nx, &sp[3 * i]);
    spline_for(f[1], ny, &sp[3 * i + 1]);
    spline_for(f[2], nz, &sp[3 * i + 2]);
    double q = s->q[i];
    if (q == 0.0) continue;
    for (int a = 0; a < ORC_PME_ORDER; a++) {
      int gx = wrap(sp[3 * i].base + a, nx);
      double wx = q * sp[3 * i].w[a];
      for (int b = 0; b < ORC_PME_ORDER; b++) {
        int gy = wrap(sp[3 * i + 1].base + b, ny);
        double wxy = wx * sp[3 * i + 1].w[b];
        for (int c = 0; c < ORC_PME_ORDER; c++) {
          int gz = wrap(sp[3 * i + 2].base + c, nz);
          g[2 * (((long)gx * ny + gy) * nz + gz)] += wxy * sp[3 * i + 2].w[c];
        }
      }
    }
  }
  orc_fft3d(g, nx, ny, nz, -1);

  double *mx = malloc(sizeof(double) * nx), *my = malloc(sizeof(double) * ny),
         *mz = malloc(sizeof(double) * nz);
  orc_bspline_moduli(nx, ORC_PME_ORDER, mx);
  orc_bspline_moduli(ny, ORC_PME_ORDER, my);
  orc_bspline_moduli(nz, ORC_PME_ORDER, mz);
  double volume = orc_box_volume(box);
  double prefactor = ORC_ONE_4PI_EPS0 * 2.0 * M_PI / volume;
  double exp_scale = -1.0 / (4.0 * alpha * alpha);
  double energy = 0;
  for (int x = 0; x < nx; x++) {
    int ma = x <= nx / 2.0 ? x : x - nx;
    for (int y = 0; y < ny; y++) {
      int mb = y <= ny / 2.0 ? y : y - ny;
      for (int z = 0; z < nz; z++) {
        long at = 2 * (((long)x * ny + y) * nz + z);
        if (x == 0 && y == 0 && z == 0) {
          g[at] = g[at + 1] = 0;
          continue;
        }
        int mc = z <= nz / 2.0 ? z : z - nz;
        /* k = 2 pi m B^-T (src/reference/triclinic.ts:441-459) */
        double kx = 2.0 * M_PI * ma * inv[0];
        double ky = 2.0 * M_PI * (ma * inv[3] + mb * inv[4]);
        double kz = 2.0 * M_PI * (ma * inv[6] + mb * inv[7] + mc * inv[8]);
        double k2 = kx * kx + ky * ky + kz * kz;
        double mod = mx[x] * my[y] * mz[z];
        double factor = prefactor * exp(k2 * exp_scale) / (k2 * mod);
        double re = g[at], im = g[at + 1];
        energy += factor * (re * re + im * im);
        g[at] = 2.0 * factor * re;
        g[at + 1] = 2.0 * factor * im;
      }
    }
  }
  if (forces) {
    orc_fft3d(g, nx, ny, nz, 1);
    for (int i = 0; i < s->n; i++) {
      double q = s->q[i];
      if (q == 0.0) continue;
      const spline *sx = &sp[3 * i], *sy = &sp[3 * i + 1], *sz = &sp[3 * i + 2];
      double fx = 0, fy = 0, fz = 0;
      for (int a = 0; a < ORC_PME_ORDER; a++) {
        int gx = wrap(sx->base + a, nx);
        for (int b = 0; b < ORC_PME_ORDER; b++) {
          int gy = wrap(sy->base + b, ny);
          for (int c = 0; c < ORC_PME_ORDER; c++) {
            int gz = wrap(sz->base + c, nz);
            double phi = g[2 * (((long)gx * ny + gy) * nz + gz)];
            fx += phi * sx->dw[a] * sy->w[b] * sz->w[c];
            fy += phi * sx->w[a] * sy->dw[b] * sz->w[c];
            fz += phi * sx->w[a] * sy->w[b] * sz->dw[c];
          }
        }
      }
      double df[3] = {-q * fx * nx, -q * fy * ny, -q * fz * nz};
      /* chain rule through B^-1 (src/reference/forces/pme.ts:246-260) */
      for (int v = 0; v < 3; v++)
        forces[3 * i + v] += df[0] * inv[v * 3 + 0] + df[1] * inv[v * 3 + 1] + df[2] * inv[v * 3 + 2];
    }
  }
  free(g);
  free(sp);
  free(mx);
  free(my);
  free(mz);
  return energy;
}

/* half-widths bounded per perpendicular width (src/reference/forces/ewald.ts:52-74) */
double orc_ewald_reciprocal(const double *charge, int n, const double *pos,
                            const double box_in[9], double alpha, double tolerance,
                            double *forces, int64_t *vectors_out) {
  double box[9], inv[9];
  if (orc_is_reduced(box_in, 1e-9))
    memcpy(box, box_in, sizeof box);
  else
    orc_reduce_box(box_in, box);
  orc_inverse_box(box, inv);
  double volume = orc_box_volume(box);
  double target = sqrt(-log(tolerance > 0 ? tolerance : 1e-10));
  int hw[3];
  for (int a = 0; a < 3; a++) {
    int m[3] = {a == 0, a == 1, a == 2};
    double kx = 2.0 * M_PI * m[0] * inv[0];
    double ky = 2.0 * M_PI * (m[0] * inv[3] + m[1] * inv[4]);
    double kz = 2.0 * M_PI * (m[0] * inv[6] + m[1] * inv[7] + m[2] * inv[8]);
    double mag = sqrt(kx * kx + ky * ky + kz * kz);
    hw[a] = (int)ceil(2.0 * alpha * target / mag);
    if (hw[a] < 1) hw[a] = 1;
  }
  double prefactor = 2.0 * M_PI * ORC_ONE_4PI_EPS0 / volume;
  double q4a2 = 1.0 / (4.0 * alpha * alpha);
  double energy = 0;
  int64_t vectors = 0;
  for (int ma = 0; ma <= hw[0]; ma++) {
    for (int mb = (ma == 0 ? 0 : -hw[1]); mb <= hw[1]; mb++) {
      for (int mc = (ma == 0 && mb == 0 ? 1 : -hw[2]); mc <= hw[2]; mc++) {
        double kx = 2.0 * M_PI * ma * inv[0];
        double ky = 2.0 * M_PI * (ma * inv[3] + mb * inv[4]);
        double kz = 2.0 * M_PI * (ma * inv[6] + mb * inv[7] + mc * inv[8]);
        double k2 = kx * kx + ky * ky + kz * kz;
        if (k2 == 0) continue;
        double weight = exp(-k2 * q4a2) / k2;
        if (weight == 0) continue;
        vectors++;
        double re = 0, im = 0;
        for (int i = 0; i < n; i++) {
          double ph = kx * pos[3 * i] + ky * pos[3 * i + 1] + kz * pos[3 * i + 2];
          re += charge[i] * cos(ph);
          im += charge[i] * sin(ph);
        }
        energy += 2.0 * prefactor * weight * (re * re + im * im);
        if (forces) {
          double scale = 4.0 * prefactor * weight;
          for (int i = 0; i < n; i++) {
            double ph = kx * pos[3 * i] + ky * pos[3 * i + 1] + kz * pos[3 * i + 2];
            double term = scale * charge[i] * (sin(ph) * re - cos(ph) * im);
            forces[3 * i] += term * kx;
            forces[3 * i + 1] += term * ky;
            forces[3 * i + 2] += term * kz;
          }
        }
      }
    }
  }
  if (vectors_out) *vectors_out = vectors;
  return energy;
}
