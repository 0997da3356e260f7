// bonded.hip -- bonded terms, nonbonded exceptions and position restraints.
//
// Replaces the upstream WGSL `harmonic_bond`, `harmonic_angle`,
// `periodic_torsion` (src/gpu/kernels/bonded.wgsl:30-127), `direct_exceptions`
// (nonbonded-exceptions.wgsl:73-124) and `position_restraint`
// (restraint.wgsl:80-91) with one term-parallel kernel over a flattened term
// list, ordered by kind so a wavefront rarely diverges.  Physics follows the
// f64 reference: src/reference/forces/bonded.ts:39-227 (plain differences, no
// minimum image, angle normal floored at 1e-6), forces/nonbonded.ts:330-386
// (exclusion correction with the particles' own charges plus the exception's
// own Coulomb and LJ, at any distance, nearest image) and
// forces/restraint.ts:38-85.
//
// The arithmetic is f64: these terms are a small fraction of the work, gfx950
// runs vector FP64 at the FP32 rate, and stiff bonds lose digits in
// `r - r0` in f32.  Differences are taken exactly in Q32.32 first.
#include <cmath>

#include "engine.hpp"

enum { TERM_BOND = 0, TERM_ANGLE = 1, TERM_TORSION = 2, TERM_EXCEPTION = 3, TERM_RESTRAINT = 4 };

struct md_term_args {
  const int4 *atoms;
  const double4 *param;
  const int64_t *restraint_ref;
  const int64_t *pos;
  int64_t *force;
  int64_t *energy;
  int64_t *piece;  // [pieces][3]: each (term, atom) force in Q32.32
  const int *piece_start, *piece_list;  // by atom: [np + 1], then piece numbers
  int np;
  int n_bonds, n_angles, n_torsions, n_exceptions, n_restraints;
  double alpha;
  md_box box;
};

__device__ __forceinline__ int64_t md_d_to_fixed(double v) { return (int64_t)rint(v * 4294967296.0); }

// Piece p of the term list takes one atom's force from one term, stored
// plainly: atoms shared by many terms (a torsion's four atoms sit in a dozen
// other terms) would otherwise serialise their 64-bit atomics at memory.
__device__ __forceinline__ void md_put_piece(const md_term_args &A, int p, double fx, double fy, double fz) {
  int64_t *q = A.piece + 3 * (size_t)p;
  q[0] = md_d_to_fixed(fx);
  q[1] = md_d_to_fixed(fy);
  q[2] = md_d_to_fixed(fz);
}

// The first piece of term t: the terms are ordered bonds, angles, torsions,
// exceptions, restraints, with 2, 3, 4, 2 and 1 atoms, pieces in that order.
__device__ __forceinline__ int md_first_piece(const md_term_args &A, int t) {
  int p = 0, k = t;
  if (k < A.n_bonds) return 2 * k;
  p += 2 * A.n_bonds;
  k -= A.n_bonds;
  if (k < A.n_angles) return p + 3 * k;
  p += 3 * A.n_angles;
  k -= A.n_angles;
  if (k < A.n_torsions) return p + 4 * k;
  p += 4 * A.n_torsions;
  k -= A.n_torsions;
  if (k < A.n_exceptions) return p + 2 * k;
  return p + 2 * A.n_exceptions + (k - A.n_exceptions);
}

__device__ __forceinline__ void md_diff(const md_term_args &A, int from, int to, double *d) {
  const double s = 2.3283064365386963e-10;
  d[0] = (double)(A.pos[to] - A.pos[from]) * s;
  d[1] = (double)(A.pos[A.np + to] - A.pos[A.np + from]) * s;
  d[2] = (double)(A.pos[2 * A.np + to] - A.pos[2 * A.np + from]) * s;
}

template <bool ENERGY>
__global__ void __launch_bounds__(256) k_terms(const md_term_args *__restrict__ T, md_bmap map) {
  int blk;
  const md_term_args &A = T[md_locate(map, blockIdx.x, blk)];
  int t = blk * blockDim.x + threadIdx.x;
  int total = A.n_bonds + A.n_angles + A.n_torsions + A.n_exceptions + A.n_restraints;
  double e_bond = 0, e_angle = 0, e_torsion = 0, e_exc = 0, e_res = 0;
  if (t < total) {
    int4 at = A.atoms[t];
    double4 p = A.param[t];
    const int p0 = md_first_piece(A, t);
    if (t < A.n_bonds) {
      double d[3];
      md_diff(A, at.x, at.y, d);
      double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      double dr = r - p.x;
      e_bond = 0.5 * p.y * dr * dr;
      double sc = p.y * dr / r;
      md_put_piece(A, p0, sc * d[0], sc * d[1], sc * d[2]);
      md_put_piece(A, p0 + 1, -sc * d[0], -sc * d[1], -sc * d[2]);
    } else if (t < A.n_bonds + A.n_angles) {
      double v1[3], v2[3];
      md_diff(A, at.y, at.x, v1);
      md_diff(A, at.y, at.z, v2);
      double r1 = v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2];
      double r2 = v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2];
      double dot = v1[0] * v2[0] + v1[1] * v2[1] + v1[2] * v2[2];
      double cs = fmin(1.0, fmax(-1.0, dot / sqrt(r1 * r2)));
      double dth = acos(cs) - p.x;
      e_angle = 0.5 * p.y * dth * dth;
      double n[3] = {v1[1] * v2[2] - v1[2] * v2[1], v1[2] * v2[0] - v1[0] * v2[2], v1[0] * v2[1] - v1[1] * v2[0]};
      double nn = fmax(sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]), 1e-6);
      double de = p.y * dth;
      double ta = de / (r1 * nn), tc = -de / (r2 * nn);
      double f1[3] = {-ta * (v1[1] * n[2] - v1[2] * n[1]), -ta * (v1[2] * n[0] - v1[0] * n[2]),
                      -ta * (v1[0] * n[1] - v1[1] * n[0])};
      double f3[3] = {-tc * (v2[1] * n[2] - v2[2] * n[1]), -tc * (v2[2] * n[0] - v2[0] * n[2]),
                      -tc * (v2[0] * n[1] - v2[1] * n[0])};
      md_put_piece(A, p0, f1[0], f1[1], f1[2]);
      md_put_piece(A, p0 + 1, -f1[0] - f3[0], -f1[1] - f3[1], -f1[2] - f3[2]);
      md_put_piece(A, p0 + 2, f3[0], f3[1], f3[2]);
    } else if (t < A.n_bonds + A.n_angles + A.n_torsions) {
      double b1[3], b2[3], b3[3];
      md_diff(A, at.x, at.y, b1);
      md_diff(A, at.y, at.z, b2);
      double c3[3];
      md_diff(A, at.z, at.w, c3);
      for (int k = 0; k < 3; ++k) b3[k] = c3[k];
      double n1[3] = {b1[1] * b2[2] - b1[2] * b2[1], b1[2] * b2[0] - b1[0] * b2[2], b1[0] * b2[1] - b1[1] * b2[0]};
      double n2[3] = {b2[1] * b3[2] - b2[2] * b3[1], b2[2] * b3[0] - b2[0] * b3[2], b2[0] * b3[1] - b2[1] * b3[0]};
      double n1sq = n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2];
      double n2sq = n2[0] * n2[0] + n2[1] * n2[1] + n2[2] * n2[2];
      double b2n = sqrt(b2[0] * b2[0] + b2[1] * b2[1] + b2[2] * b2[2]);
      double x = n1[0] * n2[0] + n1[1] * n2[1] + n1[2] * n2[2];
      double y = b2n * (b1[0] * n2[0] + b1[1] * n2[1] + b1[2] * n2[2]);
      double phi = atan2(y, x);
      double per = p.z;
      e_torsion = p.y * (1.0 + cos(per * phi - p.x));
      double dedphi = -p.y * per * sin(per * phi - p.x);
      double f1[3] = {0, 0, 0}, f4[3] = {0, 0, 0}, f2[3] = {0, 0, 0}, f3[3] = {0, 0, 0};
      if (n1sq != 0.0 && n2sq != 0.0) {
        double s1 = dedphi * b2n / n1sq, s4 = -dedphi * b2n / n2sq;
        double b2sq = b2n * b2n;
        double ss = (b1[0] * b2[0] + b1[1] * b2[1] + b1[2] * b2[2]) / b2sq;
        double tt = (b3[0] * b2[0] + b3[1] * b2[1] + b3[2] * b2[2]) / b2sq;
        for (int k = 0; k < 3; ++k) {
          f1[k] = s1 * n1[k];
          f4[k] = s4 * n2[k];
          f2[k] = -(1.0 + ss) * f1[k] + tt * f4[k];
          f3[k] = ss * f1[k] - (1.0 + tt) * f4[k];
        }
      }
      md_put_piece(A, p0, f1[0], f1[1], f1[2]);
      md_put_piece(A, p0 + 1, f2[0], f2[1], f2[2]);
      md_put_piece(A, p0 + 2, f3[0], f3[1], f3[2]);
      md_put_piece(A, p0 + 3, f4[0], f4[1], f4[2]);
    } else if (t < A.n_bonds + A.n_angles + A.n_torsions + A.n_exceptions) {
      // param = (K q_i q_j, K chargeProd, sigma, epsilon)
      int64_t dx = A.pos[at.y] - A.pos[at.x];
      int64_t dy = A.pos[A.np + at.y] - A.pos[A.np + at.x];
      int64_t dz = A.pos[2 * A.np + at.y] - A.pos[2 * A.np + at.x];
      md_min_image_fixed(dx, dy, dz, A.box);
      const double s = 2.3283064365386963e-10;
      double d[3] = {(double)dx * s, (double)dy * s, (double)dz * s};
      double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      double r = sqrt(r2);
      double ar = A.alpha * r;
      double dedr = 0;
      if (p.x != 0.0) {
        double ef = erf(ar);
        e_exc -= p.x * ef / r;
        dedr += p.x * (ef / r2 - A.alpha * (2.0 / sqrt(M_PI)) * exp(-ar * ar) / r);
      }
      if (p.y != 0.0) {
        e_exc += p.y / r;
        dedr -= p.y / r2;
      }
      if (p.w != 0.0) {
        double s6 = p.z * p.z / r2;
        s6 = s6 * s6 * s6;
        double s12 = s6 * s6;
        e_exc += 4.0 * p.w * (s12 - s6);
        dedr += (4.0 * p.w / r) * (6.0 * s6 - 12.0 * s12);
      }
      // every piece is written at every evaluation (zero where the term adds nothing)
      const double sc = dedr != 0.0 ? dedr / r : 0.0;
      md_put_piece(A, p0, sc * d[0], sc * d[1], sc * d[2]);
      md_put_piece(A, p0 + 1, -sc * d[0], -sc * d[1], -sc * d[2]);
    } else {
      int k = t - (A.n_bonds + A.n_angles + A.n_torsions + A.n_exceptions);
      int atom = at.x;
      int64_t dx = A.pos[atom] - A.restraint_ref[k];
      int64_t dy = A.pos[A.np + atom] - A.restraint_ref[A.n_restraints + k];
      int64_t dz = A.pos[2 * A.np + atom] - A.restraint_ref[2 * A.n_restraints + k];
      md_min_image_fixed(dx, dy, dz, A.box);
      const double s = 2.3283064365386963e-10;
      double d[3] = {(double)dx * s, (double)dy * s, (double)dz * s};
      e_res = p.x * (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      md_put_piece(A, p0, -2.0 * p.x * d[0], -2.0 * p.x * d[1], -2.0 * p.x * d[2]);
    }
  }
  if (ENERGY) {
    int64_t v[5] = {md_d_to_fixed(e_bond), md_d_to_fixed(e_angle), md_d_to_fixed(e_torsion), md_d_to_fixed(e_exc),
                    md_d_to_fixed(e_res)};
    const int slot[5] = {MD_E_BOND, MD_E_ANGLE, MD_E_TORSION, MD_E_EXCEPTION, MD_E_RESTRAINT};
    for (int k = 0; k < 5; ++k) {
      int64_t w = md_wave_sum_i64(v[k]);
      if ((threadIdx.x & 63) == 0 && w != 0) md_atomic_add_i64(A.energy + slot[k], w);
    }
  }
}

// Each atom's share of the terms: the sum of its pieces added to its force.
// Kernels on one stream run in order and a thread is the only one holding
// its atom here, so a plain read-add-write is enough; integer sums make the
// order of the pieces irrelevant to the bits.
__global__ void k_terms_gather(const md_term_args *__restrict__ T, md_bmap map) {
  int blk;
  const md_term_args &A = T[md_locate(map, blockIdx.x, blk)];
  const int a = blk * blockDim.x + threadIdx.x;
  if (a >= A.np) return;
  const int b = A.piece_start[a], e = A.piece_start[a + 1];
  if (b == e) return;
  int64_t f[3] = {0, 0, 0};
  for (int k = b; k < e; ++k) {
    const int64_t *q = A.piece + 3 * (size_t)A.piece_list[k];
    f[0] += q[0];
    f[1] += q[1];
    f[2] += q[2];
  }
  for (int d = 0; d < 3; ++d) A.force[d * A.np + a] += f[d];
}

__global__ void k_zero_forces(const md_term_args *__restrict__ T, md_bmap map, int zero_energy) {
  int blk;
  const md_term_args &A = T[md_locate(map, blockIdx.x, blk)];
  int i = blk * blockDim.x + threadIdx.x;
  if (i < 3 * A.np) A.force[i] = 0;
  if (zero_energy && i < MD_E_SLOTS) A.energy[i] = 0;
}

static md_term_args term_args(md_system *s) {
  md_term_args A;
  A.atoms = s->term_atoms;
  A.param = s->term_param;
  A.restraint_ref = s->restraint_ref;
  A.pos = s->pos;
  A.force = s->force;
  A.piece = s->term_piece;
  A.piece_start = s->atom_piece_start;
  A.piece_list = s->atom_piece;
  A.energy = s->energy;
  A.np = s->np;
  A.n_bonds = s->n_bonds;
  A.n_angles = s->n_angles;
  A.n_torsions = s->n_torsions;
  A.n_exceptions = s->n_exceptions;
  A.n_restraints = s->n_restraints;
  A.alpha = s->alpha;
  A.box = s->dbox;
  return A;
}

struct md_bonded_group {
  md_term_args *table = nullptr;
  md_launch zero, terms, atoms;
};

static md_bonded_group &bg(md_group *g) { return *reinterpret_cast<md_bonded_group *>(g->bonded); }

void md_bonded_group_setup(md_group *g) {
  auto *G = new md_bonded_group();
  g->bonded = G;
  std::vector<int> zero, terms, atoms;
  for (md_system *s : g->reps) {
    zero.push_back((3 * s->np + 255) / 256);
    atoms.push_back(s->n_pieces > 0 ? (s->np + 255) / 256 : 0);
    int total = s->n_bonds + s->n_angles + s->n_torsions + s->n_exceptions + s->n_restraints;
    terms.push_back((total + 255) / 256);
  }
  G->zero = md_make_launch(g, zero);
  G->terms = md_make_launch(g, terms);
  G->atoms = md_make_launch(g, atoms);
  G->table = md_group_alloc<md_term_args>(g, g->reps.size());
  md_bonded_group_refresh(g);
}

void md_bonded_group_refresh(md_group *g) {
  std::vector<md_term_args> host;
  for (md_system *s : g->reps) host.push_back(term_args(s));
  MD_HIP(hipMemcpyAsync(bg(g).table, host.data(), sizeof(md_term_args) * host.size(), hipMemcpyHostToDevice,
                        g->stream));
  MD_HIP(hipStreamSynchronize(g->stream));
}

// Replica 0's box from a box in device memory (a barostat move).
__global__ void k_bonded_box(md_term_args *T, const double *box) {
  md_box d{};
  bool bad = false;
  md_box_from_matrix(box, d, bad);
  T[0].box = d;
}

void md_bonded_box_from_device(md_group *g, const double *box) {
  k_bonded_box<<<1, 1, 0, g->stream>>>(bg(g).table, box);
  MD_HIP(hipGetLastError());
}

void md_bonded_group_free(md_group *g) {
  delete &bg(g);
  g->bonded = nullptr;
}

void md_zero_forces(md_group *g, bool energies) {
  auto &G = bg(g);
  k_zero_forces<<<G.zero.total, 256, 0, g->stream>>>(G.table, G.zero.map, energies ? 1 : 0);
  MD_HIP(hipGetLastError());
}

void md_bonded(md_group *g, bool energy) {
  auto &G = bg(g);
  if (G.terms.total == 0) return;
  if (energy)
    k_terms<true><<<G.terms.total, 256, 0, g->stream>>>(G.table, G.terms.map);
  else
    k_terms<false><<<G.terms.total, 256, 0, g->stream>>>(G.table, G.terms.map);
  k_terms_gather<<<G.atoms.total, 256, 0, g->stream>>>(G.table, G.atoms.map);
  MD_HIP(hipGetLastError());
}
