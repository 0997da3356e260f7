/*
 * mdcore.h -- C ABI of the MI355X (gfx950) molecular-dynamics core.
 *
 * This is the drop-in boundary for the upstream project's WebGPU compute path
 * (pablo-arantes/making-it-rain-webgpu, `src/gpu/`).  The host keeps the
 * `src/engine` API (`ProtocolEngine`, `src/engine/protocol.ts:32-62`); each
 * entry point below names the upstream interface it replaces.  Plain C types
 * only: flat arrays in, flat arrays out, no retained caller pointers (every
 * array is copied on the call), opaque handles own all device memory.
 *
 * Units are OpenMM's: nm, ps, amu, kJ/mol, e, K, bar.  Positions, velocities
 * and forces are xyz-interleaved double[3N].  A box is nine doubles, the row
 * vectors a, b, c; it is reduced to OpenMM's lower-triangular form on entry.
 *
 * Threading: calls on one handle must come from one host thread at a time.
 * Every call is blocking; work is issued on the handle's own HIP stream.
 */
#ifndef MDCORE_H
#define MDCORE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes.  The named JavaScript errors they map to upstream:
 * UnsupportedForce (src/engine/system.ts:178), CellListTooCoarse
 * (src/gpu/nonbonded.ts:54-65), BoxTooSmallForCells (src/gpu/cell-list.ts:41-64),
 * ClusterTooLarge (src/engine/constraint-clusters.ts:71-83), NotSmooth
 * (src/reference/fft.ts:24), CutoffTooLargeForBox (src/reference/triclinic.ts:179),
 * the fixed-point RangeError (src/reference/fixed-point.ts:56-66). */
typedef enum {
  MD_OK = 0,
  MD_ERR_UNSUPPORTED_FORCE = 1,
  MD_ERR_CUTOFF_TOO_LARGE = 2,
  MD_ERR_CLUSTER_TOO_LARGE = 3,
  MD_ERR_FIXED_POINT_OVERFLOW = 4,
  MD_ERR_HIP = 5,
  MD_ERR_OUT_OF_MEMORY = 6,
  MD_ERR_INVALID_ARGUMENT = 7,
  MD_ERR_LIST_OVERFLOW = 8,
  MD_ERR_NOT_SMOOTH = 9,
  MD_ERR_CONSTRAINTS_DIVERGED = 10,
  MD_ERR_FAULT_INJECTED = 11,
  MD_ERR_NO_DEVICE = 12
} md_err;

/* Per-force-group energies returned by md_evaluate, kJ/mol.  Indices follow
 * the group names of src/reference/potential.ts:101-142. */
enum {
  MD_GROUP_BOND = 0,       /* HarmonicBondForce */
  MD_GROUP_ANGLE = 1,      /* HarmonicAngleForce */
  MD_GROUP_TORSION = 2,    /* PeriodicTorsionForce */
  MD_GROUP_DIRECT = 3,     /* NonbondedForce[direct]: pairs + exclusion
                              correction + exceptions + dispersion */
  MD_GROUP_RECIPROCAL = 4, /* NonbondedForce[reciprocal]: mesh + self + net charge */
  MD_GROUP_RESTRAINT = 5,  /* PositionRestraint */
  MD_GROUP_COUNT = 6
};

/* An OpenMM System, flattened (src/engine/system.ts:20-173).  Counts of zero
 * leave the matching arrays unread.  All arrays are copied. */
typedef struct {
  int32_t n_particles;
  const double *masses; /* amu; 0 = frozen */
  double box[9];        /* default periodic box vectors */

  int32_t n_constraints;
  const int32_t *constraint_p1, *constraint_p2;
  const double *constraint_distance;

  int32_t n_bonds; /* HarmonicBondForce: E = k/2 (r - d)^2 */
  const int32_t *bond_p1, *bond_p2;
  const double *bond_length, *bond_k;

  int32_t n_angles; /* HarmonicAngleForce: E = k/2 (theta - a)^2 */
  const int32_t *angle_p1, *angle_p2, *angle_p3;
  const double *angle_theta, *angle_k;

  int32_t n_torsions; /* PeriodicTorsionForce: E = k (1 + cos(n phi - phase)) */
  const int32_t *torsion_p1, *torsion_p2, *torsion_p3, *torsion_p4, *torsion_periodicity;
  const double *torsion_phase, *torsion_k;

  int32_t has_nonbonded; /* NonbondedForce with method PME */
  double cutoff, ewald_tolerance, alpha; /* alpha 0: derived from cutoff and tolerance */
  int32_t grid[3];                       /* 0: derived from the box diagonal */
  int32_t use_dispersion_correction;
  const double *charge, *sigma, *epsilon;
  int32_t n_exceptions;
  const int32_t *exception_p1, *exception_p2;
  const double *exception_chargeprod, *exception_sigma, *exception_epsilon;

  int32_t n_restraints; /* CustomExternalForce k*periodicdistance(x,y,z,x0,y0,z0)^2 */
  const int32_t *restraint_particle;
  const double *restraint_k, *restraint_x0, *restraint_y0, *restraint_z0;

  int32_t cmm_frequency; /* CMMotionRemover frequency; 0 when absent */

  int32_t has_barostat; /* MonteCarloBarostat */
  double barostat_pressure, barostat_temperature;
  int32_t barostat_frequency;
  uint32_t barostat_seed;
} md_system_desc;

/* Run options (RunOptions, src/gpu/loop.ts:117-168). */
typedef struct {
  double time_step;            /* ps */
  double temperature;          /* K */
  double friction;             /* 1/ps */
  double constraint_tolerance; /* relative */
  uint32_t seed;
  int32_t integrator;          /* 0 LangevinMiddle, 1 Verlet (NVE gate) */
  double verlet_skin;          /* nm; < 0 picks the default (0.15, nonbonded.ts:227) */
  int32_t mesh[3];             /* PME mesh override; 0 = from the system or the box */
  int32_t device;              /* HIP device ordinal */
  int32_t max_batch;           /* steps per captured HIP graph (maxBatch,
                                  src/gpu/loop.ts:151-158); 0 = 50, 1 = every
                                  step launched directly, no graph */
} md_run_options;

typedef struct md_system md_system;

/* Streaming System / State ingestion, host only (no device is touched).
 * parseSystem (src/engine/system.ts:521-597) and parseState
 * (src/engine/state.ts:60-152) as one pass over the XML text with no element
 * tree, for systems of hundreds of thousands of atoms.  The same forces are
 * accepted and refused as upstream (MD_ERR_UNSUPPORTED_FORCE with the
 * reason; malformed documents MD_ERR_INVALID_ARGUMENT).  The description's
 * arrays point into the parsed object and live until it is freed. */
typedef struct md_xml_system md_xml_system;
typedef struct md_xml_state md_xml_state;
md_err md_xml_parse_system(const char *text, int64_t length, md_xml_system **out);
md_err md_xml_system_describe(const md_xml_system *system, md_system_desc *desc);
void md_xml_system_free(md_xml_system *system);
md_err md_xml_parse_state(const char *text, int64_t length, md_xml_state **out);
/* positions (and velocities, NULL when the State has none) xyz-interleaved
 * nm and nm/ps; box the row vectors; step and time as written */
md_err md_xml_state_get(const md_xml_state *state, int32_t *particles, const double **positions,
                        const double **velocities, double box[9], int64_t *step, double *time);
void md_xml_state_free(md_xml_state *state);

/* Last error message on this thread ("" when none). */
const char *md_last_error(void);
/* Library version string. */
const char *md_version(void);

/* StepLoop.fromParsed (src/gpu/loop.ts:278-532): validate, pack and upload a
 * System; positions and velocities start at zero until set. */
md_err md_system_create(const md_system_desc *desc, const md_run_options *options,
                        md_system **out);
/* StepLoop.destroy (src/gpu/loop.ts:1113-1115). */
void md_system_destroy(md_system *sys);

/* StepLoop.setPositions / setVelocities (src/gpu/loop.ts:819-855).
 * setPositions clears the integrator's carried residual, as upstream. */
md_err md_set_positions(md_system *sys, const double *positions);
md_err md_set_velocities(md_system *sys, const double *velocities);
/* StepLoop.positions / velocities (src/gpu/loop.ts:900-915). */
md_err md_get_positions(md_system *sys, double *positions);
md_err md_get_velocities(md_system *sys, double *velocities);
/* StepLoop.setBox (src/gpu/loop.ts:1031-1097): every box-dependent constant
 * is recomputed; the mesh size stays fixed, as upstream. */
md_err md_set_box(md_system *sys, const double box[9]);
md_err md_get_box(md_system *sys, double box[9]);

/* StepLoop.mesh (src/gpu/loop.ts:973-981): the PME mesh chosen at creation, all
 * zero without a NonbondedForce. */
md_err md_get_mesh(md_system *sys, int32_t mesh[3]);

/* StepLoop.potential / evaluateForces / report (src/gpu/loop.ts:761-930):
 * forces (kJ/mol/nm, may be NULL) and per-group energies (MD_GROUP_COUNT,
 * may be NULL) at the current positions; returns the total in *total. */
md_err md_evaluate(md_system *sys, double *forces, double *groups, double *total);

/* StepLoop.run (src/gpu/loop.ts:651-685): advance `steps` steps.  Strides are
 * captured once as HIP graphs and replayed; the thermostat's step counter
 * lives on the device. */
md_err md_run(md_system *sys, int64_t steps);

/* The Monte Carlo barostat on the device.  Upstream attempts the move on the
 * host between submissions (StepLoop.attemptBarostat, src/gpu/loop.ts:602-641;
 * barostat.ts:129-274), cutting every batch at an attempt.  Attached here, the
 * whole attempt runs on the device inside md_run's captured strides: the
 * energy, the Threefry BAROSTAT_DOMAIN draws (counter, seed), the isotropic
 * trial box (V' = V + s (2u - 1), lengths times (V'/V)^(1/3), reduced), each
 * molecule's unweighted centroid scaled with its atoms moved rigidly, every
 * box-dependent constant, the second energy, the Metropolis test on
 * w = dE + P dV - N_mol kT ln(V'/V) and the proposal-width adaptation every 10
 * attempts.  The attempt fires after every `frequency` steps counted across
 * runs, as upstream; the PME mesh stays fixed.  A system with a barostat
 * attached cannot join a batch. */
typedef struct {
  double pressure;           /* bar */
  double temperature;        /* K */
  int32_t frequency;         /* steps between attempts; 0 detaches */
  int32_t n_molecules;
  const int32_t *molecule_start; /* [n_molecules + 1], molecule_start[n_molecules] = n_particles */
  const int32_t *molecule_atoms; /* every particle once, ascending within each molecule */
  uint32_t seed;
  /* starting state (a resumed run passes what md_barostat_get returned) */
  double volume_scale;       /* proposal half-width, nm^3; 0: 1% of the volume */
  int64_t draws;             /* Threefry counter */
  int32_t steps_since_attempt;
  int32_t window_attempted, window_accepted;
} md_barostat_desc;
md_err md_barostat_attach(md_system *sys, const md_barostat_desc *desc);

typedef struct {
  int64_t attempts, accepted;      /* since attached */
  int64_t draws;
  int32_t steps_since_attempt;
  int32_t window_attempted, window_accepted;
  double volume_scale;             /* nm^3 */
  int32_t last_accepted;           /* the latest attempt */
  double last_energy_before, last_energy_after, last_work, last_length_scale;  /* kJ/mol */
} md_barostat_state;
md_err md_barostat_get(md_system *sys, md_barostat_state *state);

/* Steps taken so far, and StepLoop.advanceStepCount (src/gpu/loop.ts:554-569). */
md_err md_get_step(md_system *sys, int64_t *step);
md_err md_set_step(md_system *sys, int64_t step);

/* Batched replicas.  No counterpart upstream: StepLoop (src/gpu/loop.ts:208)
 * runs one system per device, and this attaches to its per-system state and
 * step ring (loop.ts:171, 434-445, 715-751).  A batch advances several
 * independent systems of any sizes together: each phase of a step is ONE
 * kernel launch covering every replica (the pair kernel's waves draw list
 * segments of all replicas from one counter), and strides of steps are
 * captured as HIP graphs.  Each replica keeps its own state, options
 * (temperature, seed, ...) and every per-system call above; the batch borrows
 * the systems, which must all use the same device and outlive it.  Calls on
 * a batch and on its systems must not overlap. */
typedef struct md_batch md_batch;
md_err md_batch_create(md_system *const *systems, int32_t count, md_batch **out);
void md_batch_destroy(md_batch *batch);
/* advance every replica by `steps` steps */
md_err md_batch_run(md_batch *batch, int64_t steps);
/* md_profile_steps for the whole batch: per-phase milliseconds of one step
 * (launched directly, no graph), averaged over `steps` steps */
md_err md_batch_profile_steps(md_batch *batch, int32_t steps, double *ms);

/* The pair kernel's autotuner.  No counterpart upstream: the WebGPU loop
 * has fixed workgroup shapes (src/gpu/loop.ts:171).  It times the pair
 * kernel's schedules -- list entries per work unit (8 to 128), resident
 * waves (2-4 per SIMD), and the direct or the compacted evaluation (the
 * in-cutoff pairs of two entries packed densely into full waves) -- on the
 * group's own replicas, `steps` steps each, keeps the fastest, and caches it
 * per (replica count, atom-count bucket, triclinic or not) for launch groups
 * created later in the process.  No schedule changes any bit; the timed
 * steps are ordinary steps. */
typedef struct {
  int32_t segment_entries, pair_waves;                 /* the schedule picked */
  double ms_per_step;                                  /* its mean step, phase by phase */
  int32_t default_segment_entries, default_pair_waves; /* the schedule before tuning */
  double default_ms_per_step;
  int32_t candidates;                                  /* schedules timed */
  int32_t compact_pairs, default_compact_pairs;        /* 1: the compacted evaluation */
} md_schedule;
md_err md_autotune(md_system *sys, int32_t steps, md_schedule *best);
md_err md_batch_autotune(md_batch *batch, int32_t steps, md_schedule *best);
/* Set a schedule directly (segment_entries, pair_waves -- 0 keeps the
 * current -- and compact_pairs of `schedule`; the other fields are ignored). */
md_err md_set_pair_schedule(md_system *sys, const md_schedule *schedule);
md_err md_batch_set_pair_schedule(md_batch *batch, const md_schedule *schedule);

/* Learned SHAKE warm start.  No counterpart upstream: it attaches where
 * `shake_cluster` starts (src/gpu/kernels/shake.wgsl:130-146), before the
 * oracle's multiplier accumulation (src/reference/constraints.ts:147-184).  A
 * small network (32 -> 64 -> 64 -> 4, tanh; BF16 matrix-core operands, FP32
 * accumulation) predicts each H-bond cluster's (up to 4 constraints) converged
 * multipliers from rotation-invariant geometry; SHAKE applies the prediction
 * as its first update and then iterates to the same tolerance as always, so
 * the constraints are satisfied exactly as without it -- only passes are
 * saved.  The network is trained on the device, on samples the engine's own
 * SHAKE produces (md_warmstart_record), with Adam on mean-squared error.
 * Parameters (md_warmstart_param_count floats): W1[32][64] b1[64] W2[64][64]
 * b2[64] W3[64][16] b3[16], outputs 0..3 used. */
typedef struct md_warmstart md_warmstart;
md_err md_warmstart_create(int32_t device, uint32_t seed, int64_t sample_capacity, md_warmstart **out);
void md_warmstart_destroy(md_warmstart *ws);
int32_t md_warmstart_param_count(void);
/* advance `sys` by up to `steps` steps (as many as the sample store holds),
 * recording each cluster's inputs and converged multipliers */
md_err md_warmstart_record(md_warmstart *ws, md_system *sys, int64_t steps);
md_err md_warmstart_sample_count(md_warmstart *ws, int64_t *count);
/* `iterations` Adam steps on minibatches of `batch` recorded samples; *loss
 * (may be NULL) receives the last minibatch's loss */
md_err md_warmstart_train(md_warmstart *ws, int32_t iterations, int32_t batch, double learning_rate, double *loss);
/* start SHAKE from the network's prediction in `sys` (NULL: plain SHAKE) */
md_err md_warmstart_attach(md_system *sys, md_warmstart *ws);
/* read (get) and/or replace (set) the FP32 parameters */
md_err md_warmstart_params(md_warmstart *ws, float *get, const float *set);
/* the network on n given samples (x[n][32], t[n][4]): outputs y[n][16] (may be
 * NULL), the loss gradient (may be NULL) and the loss, without an update */
md_err md_warmstart_evaluate(md_warmstart *ws, const float *x, const float *t, int32_t n, float *y, float *grad,
                             double *loss);
/* SHAKE update passes and cluster-steps so far, all modes */
md_err md_shake_counts(md_system *sys, int64_t *passes, int64_t *clusters);

/* StepLoop's minimiseOnDevice (src/gpu/minimise.ts:33-54) with the whole
 * search on the device: OpenMM's LocalEnergyMinimizer objective (constraint
 * springs stiffened tenfold per round, a give-up that restores the start,
 * then projection to the run's constraint tolerance;
 * src/reference/minimise.ts:164-280) searched by L-BFGS with libLBFGS's
 * backtracking strong-Wolfe line search (src/reference/lbfgs.ts:102-275),
 * memory 6, a 0.5 nm cap on any coordinate's move per trial.  Stops at an
 * RMS gradient per particle <= tolerance (kJ/mol/nm) or after
 * max_iterations (0: no limit).  Leaves the minimised positions in `sys`. */
typedef struct {
  double energy;               /* kJ/mol, at the final positions */
  int32_t iterations, rounds;  /* L-BFGS iterations over all stiffening rounds */
  int32_t gave_up;             /* 1: stiffening stopped helping, the start was restored */
  double max_constraint_error; /* largest |r - d| / d at the end */
} md_minimise_result;
md_err md_minimise(md_system *sys, int32_t max_iterations, double tolerance, md_minimise_result *result);

/* constrainPositions(system, x, x, tol) and constrainVelocities(system, x, v,
 * tol) (src/reference/constraints.ts:254-291) applied on the device to the
 * current positions or velocities: SETTLE for rigid waters, Jacobi SHAKE or
 * RATTLE for the other clusters.  The minimiser and the Maxwell-Boltzmann
 * draw (src/reference/minimise.ts, velocities.ts:78-117) use them. */
md_err md_constrain_positions(md_system *sys, double tolerance);
md_err md_constrain_velocities(md_system *sys, double tolerance);

/* StepLoop.verletFlags (src/gpu/loop.ts:941-960). */
typedef struct {
  int64_t rebuilds;
  int64_t tiles;            /* cluster-pair tiles in the current list */
  int64_t tile_capacity;
  double worst_displacement;/* nm, since the last build */
  int32_t overflowed;
} md_list_flags;
md_err md_list_flags_get(md_system *sys, md_list_flags *flags);

/* StepLoop.profileStep (src/gpu/loop.ts:705-713, src/gpu/profile.ts:68-184):
 * run `steps` steps launched phase by phase with HIP events recorded on the
 * system's stream between phases; `ms[phase]` receives the mean milliseconds
 * per step of each phase (MD_PHASE_COUNT entries). */
enum {
  MD_PHASE_PROLOGUE = 0, /* step counter, rebuild latch */
  MD_PHASE_BUILD = 1,    /* cell sort, clusters, bounding boxes (gated on the device) */
  MD_PHASE_PACK = 2,     /* cluster-frame f32 coordinates */
  MD_PHASE_PAIRS = 3,    /* LJ + erfc-Ewald pair kernel */
  MD_PHASE_BONDED = 4,   /* bonds, angles, torsions, exceptions, restraints */
  MD_PHASE_SPREAD = 5,   /* PME charge spreading */
  MD_PHASE_FFT = 6,      /* PME forward FFT, convolution, inverse FFT */
  MD_PHASE_GATHER = 7,   /* PME force interpolation */
  MD_PHASE_INTEGRATE = 8,/* LFMiddle/Verlet + SETTLE/SHAKE/RATTLE + COM */
  MD_PHASE_SEARCH = 9,   /* cluster-pair search, list starts, segments (gated) */
  MD_PHASE_MASKS = 10,   /* exclusion masks and the list by j-cluster (gated) */
  MD_PHASE_FORCES = 11,  /* pair forces from slots and list entries onto the atoms */
  MD_PHASE_COUNT = 12
};
md_err md_profile_steps(md_system *sys, int32_t steps, double *ms);
const char *md_phase_name(int32_t phase);

/* Work counts the benchmark uses, counted on the device at the current
 * positions: non-excluded pairs inside the cutoff (exact), and the 8x8
 * cluster-pair blocks the pair kernel evaluated (64 pair slots each), so
 * pairs / (64 * tiles) is the fraction of evaluated pair slots that count. */
md_err md_pair_census(md_system *sys, int64_t *pairs_within_cutoff, int64_t *tiles);

/* Diagnostic probes (the upstream `random-probe` kernel, K41, and its
 * fixed-point twin): run the device primitives on host-given inputs. */
md_err md_probe_fixed_point(const float *values, int32_t n, int64_t *fixed, float *back);
md_err md_probe_thermostat(uint32_t seed, uint32_t step, int32_t n, uint32_t *bits, float *normals);
/* No counterpart in the reference: md_probe_fixed_point for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and results; the input
 * and both outputs are staged in a fixed 1 MiB device array that each GPU
 * keeps until the process ends, so n is at most 65536 (more is
 * MD_ERR_INVALID_ARGUMENT).  Calls no hipMalloc/hipFree.  Calls that overlap
 * on one GPU share the array and race. */
md_err md_probe_fixed_point_noalloc(const float *values, int32_t n, int64_t *fixed, float *back);
/* The order in which the neighbour search sorts cells (host only, no
 * device): cells[k] = (x * dims[1] + y) * dims[2] + z of the k-th cell of a
 * dims[0] x dims[1] x dims[2] grid.  Stands in for the counting sort's cell
 * numbering (src/gpu/cell-list.ts:136-155); `cells` holds the grid's size. */
md_err md_probe_cell_order(const int32_t dims[3], int32_t *cells);

/* The pair kernel's real-space Coulomb polynomials for an Ewald splitting
 * parameter alpha (nm^-1) and a cutoff (nm), host only.  With u = alpha^2 r^2
 * the kernel takes F/r per charge product as 1/r^3 + sum_k force[k] v^k and
 * E as 1/r + sum_k energy[k] v^k, v = scale r^2 - 1 (13 terms each, f32);
 * *ok says whether that f32 evaluation is within 1e-6 of 1/r^3 (force) and
 * of 1/r (energy) over [0, cutoff], the errors are in *force_err and
 * *energy_err; where it is not, the kernel evaluates erfc and exp.  Stands in
 * for the screened Coulomb term of the reference's pair kernel
 * (src/gpu/kernels/nonbonded-pairs.wgsl:203-212, erfc_approx in
 * src/gpu/kernels/erfc.wgsl; its bound in test/erf.test.ts:68-89). */
md_err md_probe_ewald_fit(double alpha, double cutoff, float *scale, float *force, float *energy, int32_t *ok,
                          double *force_err, double *energy_err);

/* Fault injection for host-side failure tests: the next `after` kernel
 * launches succeed, then every call returns MD_ERR_FAULT_INJECTED until
 * re-armed with a negative value.  No kernel is altered. */
md_err md_inject_fault(md_system *sys, int32_t after);

#ifdef __cplusplus
}
#endif
#endif
