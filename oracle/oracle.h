/*
 * oracle.h -- the CPU f64 oracle for the molecular-dynamics hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, in plain C and double
 * precision, the reference implementation that lives in the upstream project's
 * `src/reference/` (TypeScript).  It is the checker the HIP engine is compared
 * against and the CPU baseline `bench.py` times; nothing in the shipped engine
 * links, loads or calls it.  Each function names the reference file:line it
 * follows.
 *
 * Units are the reference's (OpenMM's): nm, ps, amu, kJ/mol, e, K.
 * Positions and forces are xyz-interleaved double[3N].  Box vectors are nine
 * doubles, row vectors a, b, c, in OpenMM's reduced lower-triangular form.
 */
#ifndef MDCORE_ORACLE_H
#define MDCORE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- constants (src/engine/units.ts:15-25) ---------------------------- */
#define ORC_BOLTZMANN 0.008314462618153242
#define ORC_ONE_4PI_EPS0 138.93545764498457
#define ORC_BAR_TO_KJ_PER_MOL_NM3 0.06022140762081242
#define ORC_PME_ORDER 5
#define ORC_MAX_CONSTRAINT_ITERATIONS 150

/* ---- fixed point (src/reference/fixed-point.ts:55-127) ---------------- */
int64_t orc_to_fixed(float value);
void orc_fixed_words(float value, uint32_t *lo, int32_t *hi);
float orc_from_fixed_f32(int64_t fixed);
double orc_from_fixed(int64_t fixed);

/* ---- counter-based RNG (src/reference/random.ts:63-190) --------------- */
void orc_threefry2x32(const uint32_t counter[2], const uint32_t key[2], int rounds,
                      uint32_t out[2]);
double orc_open_unit_interval(uint32_t bits);
void orc_normal_pair(const uint32_t counter[2], const uint32_t key[2], double out[2]);
double orc_thermostat_normal(uint32_t seed, uint32_t step, uint32_t particle, int axis);
double orc_domain_normal(uint32_t domain, uint32_t seed, uint32_t particle, int axis);
#define ORC_THERMOSTAT_DOMAIN 0x9e3779b9u
#define ORC_BAROSTAT_DOMAIN 0x6a09e667u
#define ORC_VELOCITY_DOMAIN 0xbb67ae85u

/* ---- units and mesh sizes (src/engine/units.ts:53-84) ----------------- */
double orc_ewald_alpha(double cutoff, double tolerance);
int orc_pme_minimum_grid(double alpha, double box_length, double tolerance);
int orc_smooth_fft_size(int minimum);

/* ---- cell geometry (src/reference/triclinic.ts) ----------------------- */
void orc_reduce_box(const double in[9], double out[9]);
int orc_is_reduced(const double box[9], double tolerance);
void orc_inverse_box(const double box[9], double inv[9]);
void orc_perpendicular_widths(const double box[9], double widths[3]);
double orc_box_volume(const double box[9]);
/* nearest image of d (in place), Cartesian sequential reduction c, b, a */
void orc_nearest_image(const double box[9], double d[3]);
void orc_fractional(const double inv[9], const double r[3], double f[3]);

/* ---- the system the potential is evaluated for ------------------------ */
typedef struct {
  int n;
  const double *mass;
  /* constraints */
  int nc;
  const int *c1, *c2;
  const double *cd;
  /* HarmonicBondForce */
  int nb;
  const int *b1, *b2;
  const double *blen, *bk;
  /* HarmonicAngleForce */
  int na;
  const int *a1, *a2, *a3;
  const double *atheta, *ak;
  /* PeriodicTorsionForce */
  int nt;
  const int *t1, *t2, *t3, *t4, *tper;
  const double *tphase, *tk;
  /* NonbondedForce (PME) */
  int has_nonbonded;
  double cutoff, ewald_tolerance, alpha; /* alpha 0 => derived */
  int grid[3];                           /* 0 => derived */
  int dispersion;
  const double *q, *sigma, *epsilon;
  int ne;
  const int *e1, *e2;
  const double *eq, *esig, *eeps;
  /* PositionRestraint: k*periodicdistance(x,y,z,x0,y0,z0)^2 */
  int nr;
  const int *rp;
  const double *rk, *rx0, *ry0, *rz0;
  /* CMMotionRemover frequency, 0 when absent */
  int cmm_frequency;
} orc_system;

/* energy groups, indices into the `groups` output arrays */
enum {
  ORC_G_BOND = 0,
  ORC_G_ANGLE = 1,
  ORC_G_TORSION = 2,
  ORC_G_DIRECT = 3,
  ORC_G_RECIPROCAL = 4,
  ORC_G_RESTRAINT = 5,
  ORC_G_COUNT = 6
};

/* direct-space decomposition (src/reference/forces/nonbonded.ts:73-89) */
typedef struct {
  double coulomb, lennard_jones, exclusion_correction, exceptions, dispersion, total;
  int64_t pairs_within_cutoff; /* reference convention: nonzero contribution */
  int64_t pairs_geometric;     /* every non-excluded pair with r < cutoff */
} orc_direct;

double orc_alpha_for(const orc_system *s);
void orc_default_grid(const orc_system *s, const double box[9], double alpha, int grid[3]);

double orc_dispersion_correction(const orc_system *s, double volume);
double orc_dispersion_correction_naive(const orc_system *s, double volume);
double orc_self_energy(const orc_system *s, double alpha);
double orc_net_charge_correction(const orc_system *s, double alpha, double volume);
/* naive != 0 selects the O(N^2) loop; forces may be NULL */
void orc_direct_space(const orc_system *s, const double *pos, const double box[9],
                      double alpha, int naive, double *forces, orc_direct *out);

/* B-splines and the mesh (src/reference/forces/bspline.ts) */
void orc_bspline(double w, int order, double *weights, double *derivatives);
void orc_bspline_moduli(int n, int order, double *moduli);
/* complex 3D transform, unnormalised; sign -1 forward, +1 inverse */
void orc_fft3d(double *data, int nx, int ny, int nz, int sign);
void orc_dft3d(double *data, int nx, int ny, int nz, int sign);
/* mesh sum only (no self or net-charge term); forces may be NULL */
double orc_pme_reciprocal(const orc_system *s, const double *pos, const double box[9],
                          double alpha, const int grid[3], double *forces);
/* exact Ewald reciprocal sum (src/reference/forces/ewald.ts:82-152) */
double orc_ewald_reciprocal(const double *charge, int n, const double *pos,
                            const double box[9], double alpha, double tolerance,
                            double *forces, int64_t *vectors);

/* bonded (src/reference/forces/bonded.ts:39-227) */
double orc_harmonic_bond(const orc_system *s, const double *pos, double *forces);
double orc_harmonic_angle(const orc_system *s, const double *pos, double *forces);
double orc_periodic_torsion(const orc_system *s, const double *pos, double *forces);
double orc_position_restraint(const orc_system *s, const double *pos, const double box[9],
                              double *forces);

/* everything, by group (src/reference/potential.ts:94-171).  grid NULL or
 * zeros => default mesh.  Returns the total energy. */
double orc_potential(const orc_system *s, const double *pos, const double box[9],
                     const int *grid, double *forces, double groups[ORC_G_COUNT]);

/* ---- constraints (src/engine/constraint-clusters.ts:93-249,
 *                   src/reference/settle.ts, src/reference/constraints.ts) ---- */
typedef struct {
  int n_settle;
  int *settle; /* 3 per cluster: central, first, second */
  double *settle_d; /* 2 per cluster: distance1 (central bonds), distance2 */
  int n_iterative;
  int *iterative; /* constraint indices not handled by SETTLE */
  int max_atoms, max_constraints;
  int error; /* nonzero: a cluster exceeds 8 atoms or 8 constraints */
} orc_split;

int orc_split_constraints(const orc_system *s, orc_split *out);
void orc_split_free(orc_split *split);
void orc_settle_positions(const orc_system *s, const orc_split *split, const double *ref,
                          double *trial);
void orc_settle_velocities(const orc_system *s, const orc_split *split, const double *pos,
                           double *vel);
/* returns iterations; converged flag through *converged */
int orc_shake(const orc_system *s, const orc_split *split, const double *ref, double *trial,
              double tolerance, int max_iterations, int *converged);
int orc_rattle(const orc_system *s, const orc_split *split, const double *pos, double *vel,
               double tolerance, int max_iterations, int *converged);

/* ---- integration (src/reference/integrator.ts, com-motion.ts, kinetic.ts) ---- */
void orc_remove_com(const orc_system *s, double *vel);
double orc_kinetic_energy(const orc_system *s, const double *vel);
int orc_degrees_of_freedom(const orc_system *s);
/* one LFMiddle step given the forces at pos.  Noise from the thermostat
 * stream (seed, step).  Returns SHAKE iterations. */
int orc_lfmiddle_step(const orc_system *s, const orc_split *split, double *pos, double *vel,
                      const double *forces, double dt, double temperature, double friction,
                      double tolerance, uint32_t seed, uint32_t step);
int orc_verlet_step(const orc_system *s, const orc_split *split, double *pos, double *vel,
                    const double *forces, double dt, double tolerance);
/* Maxwell-Boltzmann draw (src/reference/velocities.ts:78-117) */
void orc_velocities_at_temperature(const orc_system *s, const orc_split *split,
                                   const double *pos, const double *forces, double temperature,
                                   uint32_t seed, double time_offset, double *vel);

/* full MD driver used as the CPU baseline: `steps` LFMiddle steps with COM
 * removal and forces recomputed each step.  Returns the last potential. */
double orc_run_lfmiddle(const orc_system *s, double *pos, double *vel, const double box[9],
                        const int *grid, int steps, double dt, double temperature,
                        double friction, double tolerance, uint32_t seed, uint32_t first_step);

#ifdef __cplusplus
}
#endif
#endif
