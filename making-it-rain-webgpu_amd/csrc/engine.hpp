// engine.hpp -- the resident state of one system on one GPU, and the kernel
// launchers each subsystem file provides.
//
// One `md_system` owns every device allocation it uses, made once at creation
// (StepLoop's "no per-step allocation or upload" rule, src/gpu/loop.ts:10-24).
// Per-atom arrays are structure-of-arrays planes of `np` entries (N rounded
// up to 64), so a wavefront touching 64 consecutive atoms reads one
// contiguous span per plane.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/mdcore.h"
#include "common.hpp"

struct md_error : std::exception {
  md_err code;
  std::string message;
  md_error(md_err c, std::string m) : code(c), message(std::move(m)) {}
  const char *what() const noexcept override { return message.c_str(); }
};

// what md_last_error reports on this thread (set by every entry point)
void md_set_last_error(const std::string &message);

#define MD_HIP(call)                                                                \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess)                                                           \
      throw md_error(e_ == hipErrorOutOfMemory ? MD_ERR_OUT_OF_MEMORY : MD_ERR_HIP, \
                     std::string(#call) + ": " + hipGetErrorString(e_));            \
  } while (0)

// Device-resident control words, read and written by the kernels of a step.
struct md_control {
  int64_t step_current;            // the step being integrated (thermostat counter)
  int64_t step_next;               // the step after it; the host sets it before a run
  uint32_t rebuild_request[2];     // [0] set by the integrator's displacement test
  uint32_t rebuild_active;         // latched by the step prologue
  uint32_t overflow;               // the cluster-pair list exceeded its capacity
  int64_t rebuilds;
  uint32_t max_disp2_bits[2];      // [0] worst squared displacement since the build
  uint32_t cluster_radius_bits;    // largest cluster bounding-box half diagonal (f32 bits)
  uint32_t super_radius_bits;      // largest super-cluster radius (f32 bits)
  int64_t momentum[2][3];          // sum m v in Q32.32, by step parity
  int64_t pair_census;             // in-cutoff pairs, when counted
  uint32_t force_pending;          // the host asks for a rebuild at the next step
  uint32_t diverged;               // the integrator's constraint solve failed: bit 0 SHAKE did not
                                   // converge, bit 1 RATTLE did not, bit 2 a non-finite velocity
  int64_t shake_passes;            // SHAKE update passes over all cluster-steps
  int64_t shake_clusters;          // cluster-steps SHAKE has run
  int64_t pair_blocks;             // 8x8 cluster-pair blocks evaluated, when counted
  uint32_t barostat_error;         // a barostat move failed: 1 the trial box is too small for the
                                   // cutoff, 2 a scaled position left the fixed-point range
};

// The per-group energies of the energy slots and the host constants
// (dispersion, self and net-charge terms), and their total -- the same on the
// host (md_evaluate) and on the device (the barostat's Metropolis test).
struct md_energy_consts {
  int has_nonbonded, dispersion;
  double disp_coeff, alpha, q_total, q2_sum;
};
__host__ __device__ inline double md_energy_groups(const int64_t *e, const md_energy_consts &k, double volume,
                                                   double *g) {
#pragma clang fp contract(off)
  for (int i = 0; i < MD_GROUP_COUNT; ++i) g[i] = 0.0;
  g[MD_GROUP_BOND] = md_fixed_to_double(e[MD_E_BOND]);
  g[MD_GROUP_ANGLE] = md_fixed_to_double(e[MD_E_ANGLE]);
  g[MD_GROUP_TORSION] = md_fixed_to_double(e[MD_E_TORSION]);
  g[MD_GROUP_RESTRAINT] = md_fixed_to_double(e[MD_E_RESTRAINT]);
  if (k.has_nonbonded) {
    g[MD_GROUP_DIRECT] = md_fixed_to_double(e[MD_E_PAIR]) + md_fixed_to_double(e[MD_E_EXCEPTION]) +
                         (k.dispersion ? k.disp_coeff / volume : 0.0);
    // self and net-charge terms (nonbonded.ts:594-617)
    const double self = -MD_ONE_4PI_EPS0 * k.alpha * k.q2_sum / sqrt(M_PI);
    const double net = -M_PI * k.q_total * k.q_total * MD_ONE_4PI_EPS0 / (2.0 * k.alpha * k.alpha * volume);
    g[MD_GROUP_RECIPROCAL] = md_fixed_to_double(e[MD_E_RECIPROCAL]) + self + net;
  }
  double total = 0.0;
  for (int i = 0; i < MD_GROUP_COUNT; ++i) total += g[i];
  return total;
}

// The pair kernel's real-space Coulomb polynomials (md_ewald_fit, in nm).
// With u = alpha^2 r^2 the Ewald direct-space force and energy per charge
// product are
//   F/r = 1/r^3 + alpha^3 g(u),  g(u) = ((2/sqrt(pi)) e^-u - erf(sqrt u)/sqrt u)/u
//   E   = 1/r   - alpha h(u),    h(u) = erf(sqrt u)/sqrt u,
// g and h being entire functions of u.  `force` and `energy` hold alpha^3 g
// and -alpha h as polynomials in v = scale r^2 - 1 (power basis, f32).
#define MD_EWALD_TERMS 13
struct md_ewald_poly {
  int ok;  // the f32 evaluation meets the bound on [0, cutoff]
  float scale;
  float force[MD_EWALD_TERMS], energy[MD_EWALD_TERMS];
  double force_err, energy_err;  // max |error| / (1/r^3), / (1/r), evaluated as the kernel does
};
md_ewald_poly md_ewald_fit(double alpha, double cutoff);  // nonbonded.hip

struct md_nb_params {
  int nclusters, nslots;  // 8-atom clusters; slots in whole 64-atom super-clusters
  int ncells, gdim[3];    // sorting cells
  float cutoff2, list_cutoff2, half_skin2;
  md_ewald_poly ewald;
};

// ---- launch groups ------------------------------------------------------------
//
// Every kernel runs over a group of R systems ("replicas") in one launch.  Its
// arguments sit in a device table, one struct per replica, and a block map
// gives each replica a contiguous run of the launch's blocks: replica r owns
// blocks [start[r], start[r+1]).  A block finds its replica by a binary search
// of the map (R = 1: no search).  A system on its own is a group of one; a
// batch of replicas is one group, so each phase of a step is one launch
// however many replicas it carries.  The tables live in device memory, so a
// captured graph stays valid when a replica's box (and its table entry)
// changes.
struct md_bmap {
  const int *start;  // [R + 1]
  int R;
};

__device__ __forceinline__ int md_locate(const md_bmap &m, int blk, int &local) {
  int lo = 0, hi = m.R;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (m.start[mid] <= blk)
      lo = mid;
    else
      hi = mid;
  }
  local = blk - m.start[lo];
  return lo;
}

struct md_system;
struct md_ws_net;
struct md_warmstart;
struct md_min_work;
struct md_baro;

// Host side of a block map: per-replica block counts -> device prefix.
struct md_launch {
  int total = 0;
  md_bmap map{nullptr, 0};
};

struct md_group {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::vector<md_system *> reps;
  md_group **owner = nullptr;  // cleared when the group is released
  int max_cluster_atoms = 0;
  // per-subsystem launch tables (each file's private struct)
  void *nb = nullptr, *bonded = nullptr, *pme = nullptr, *integ = nullptr;
  // captured strides: a whole stride and the last shorter remainder run, each
  // with or without a barostat attempt at its end
  struct captured {
    int steps = 0;
    bool attempt = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
  };
  captured graphs[4];
  int max_batch = 0;  // steps per captured graph (0: the default)
  // set by each step's prologue when a replica's control block holds an
  // error (list or frame overflow, diverged constraints); read after every
  // stride so a failing run stops within one stride
  uint32_t *status = nullptr;
  int ws_mode = 0;          // cluster integrator: 0 plain, 1 recording, 2 learned
  bool graph_stale = false; // the captured stride no longer matches the kernels
  // profiling: events recorded by md_mark while `profiling` is set
  bool profiling = false;
  std::vector<hipEvent_t> prof_events;
  std::vector<int> prof_phase;
  std::vector<void *> allocations;
};

// device memory owned by the group, zero-filled
template <typename T>
T *md_group_alloc(md_group *g, size_t count) {
  void *p = nullptr;
  size_t bytes = count * sizeof(T);
  if (bytes == 0) bytes = 256;
  MD_HIP(hipMalloc(&p, bytes));
  MD_HIP(hipMemsetAsync(p, 0, bytes, g->stream));
  g->allocations.push_back(p);
  return static_cast<T *>(p);
}

// a block map from per-replica block counts (uploaded on the group's stream)
md_launch md_make_launch(md_group *g, const std::vector<int> &blocks);
// a device copy of a host vector, owned by the group; waits for the copy, so
// the caller may free `v` as soon as this returns
template <typename T>
T *md_group_upload(md_group *g, const std::vector<T> &v) {
  T *d = md_group_alloc<T>(g, v.size());
  if (!v.empty()) MD_HIP(hipMemcpyAsync(d, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, g->stream));
  MD_HIP(hipStreamSynchronize(g->stream));
  return d;
}

struct md_system {
  int device = 0;
  hipStream_t stream = nullptr;
  int n = 0, np = 0;
  md_run_options opt{};
  long fault_after = -1;  // host-side fault injection counter (steps)

  // ---- host copies of the System ------------------------------------
  std::vector<double> masses, charge, sigma, epsilon;
  std::vector<int> excl_start, excl_list;  // CSR, both directions
  double box[9] = {0};
  md_box dbox{};
  bool has_nonbonded = false;
  double cutoff = 1.0, ewald_tol = 5e-4, alpha = 0.0;
  bool dispersion = false;
  int grid[3] = {0, 0, 0};
  double disp_coeff = 0;  // dispersion correction times the volume
  double q_total = 0, q2_sum = 0;
  int cmm_frequency = 0;
  double total_mass = 0;
  int dof = 0;
  int n_constraints = 0;

  // ---- per-atom device planes ---------------------------------------
  int64_t *pos = nullptr;    // [3][np] Q32.32 nm
  float *vel = nullptr;      // [3][np] nm/ps
  float *resid = nullptr;    // [3][np] nm carried by the drift
  int64_t *force = nullptr;  // [3][np] Q32.32 kJ/mol/nm
  float *inv_mass = nullptr;
  float *mass_f = nullptr;
  float4 *atom_param = nullptr;  // (q sqrt(K), sigma/2, 2 sqrt(eps), q)
  int64_t *energy = nullptr;     // [MD_E_SLOTS]
  md_control *ctl = nullptr;

  // ---- neighbour list (buffers private to nonbonded.hip in nb_ext) -----
  md_nb_params nb{};
  int *cell_rank = nullptr, *cell_count = nullptr, *cell_cursor = nullptr, *cell_start = nullptr;
  int *atom_cell = nullptr, *cell_member = nullptr;
  int *slot_atom = nullptr, *atom_slot = nullptr;
  int64_t *atom_folded = nullptr, *build_pos = nullptr, *slot_offset = nullptr;
  float2 *slot_lj = nullptr;
  int4 *slot_xyzq = nullptr;      // pair-kernel frame coordinates + charge bits
  float4 *cluster_half = nullptr;
  int *excl_start_d = nullptr, *excl_list_d = nullptr;
  void *nb_ext = nullptr;

  // ---- bonded terms, exceptions, restraints (one flattened list) --------
  int n_bonds = 0, n_angles = 0, n_torsions = 0, n_exceptions = 0, n_restraints = 0;
  int4 *term_atoms = nullptr;
  double4 *term_param = nullptr;
  // the terms' forces as pieces, one Q32.32 triple per (term, atom) in term
  // order (2, 3, 4, 2, 1 atoms for a bond, angle, torsion, exception,
  // restraint), and by atom the pieces to sum: k_terms stores pieces,
  // k_terms_gather adds each atom's sum to its force (no atomics)
  int n_pieces = 0;
  int64_t *term_piece = nullptr;    // [n_pieces][3]
  int *atom_piece_start = nullptr;  // [np + 1]
  int *atom_piece = nullptr;        // [n_pieces]
  int64_t *restraint_ref = nullptr;  // [3][n_restraints] Q32.32

  // ---- PME ------------------------------------------------------------
  int pme_nx = 0, pme_ny = 0, pme_nz = 0, pme_nzc = 0;
  int64_t *pme_grid = nullptr;  // [nx][ny][nz] Q32.32 charge
  float2 *pme_cgrid = nullptr;  // [nx][ny][nz/2+1]
  float *pme_phi = nullptr;     // [nx][ny][nz] dE/dQ
  void *pme_ext = nullptr;

  // ---- constraints / integrator work items -----------------------------
  int n_waters = 0, n_clusters_c = 0, n_free = 0;
  int4 *water_atoms = nullptr;   // (central, first, second, 0)
  float2 *water_dist = nullptr;  // (distance1, distance2)
  int *ccl_atoms = nullptr;      // [n][8]
  int *ccl_natoms = nullptr;
  int *ccl_ncons = nullptr;
  int4 *ccl_cons = nullptr;      // [n][8] (local i, local j, f64 distance bits)
  int *free_atoms = nullptr;
  int max_cluster_atoms = 0;

  // ---- learned SHAKE warm start ------------------------------------------
  const md_ws_net *ws_net = nullptr;  // the network the integrator runs, or none
  float *rec_x = nullptr, *rec_t = nullptr;  // sample ring of a recording run
  int64_t rec_first_step = 0, rec_cap = 0;

  // ---- the device barostat, when attached ------------------------------------
  md_baro *baro = nullptr;

  // ---- the minimiser's work vectors, made on the first md_minimise ----------
  md_min_work *min_work = nullptr;
  std::vector<int> con_p1, con_p2;    // the constraints (host copies)
  std::vector<double> con_d;

  // ---- run state ---------------------------------------------------------
  md_group *self = nullptr;          // this system as a group of one
  std::vector<md_group *> groups;    // every group it is in (its own first)
  int64_t steps_done = 0;
  bool need_rebuild = true;
  bool momentum_dirty = true;

  std::vector<void *> allocations;
};

// zero-filled device memory, freed with the system
template <typename T>
T *md_alloc(md_system *s, size_t count) {
  void *p = nullptr;
  size_t bytes = count * sizeof(T);
  if (bytes == 0) bytes = 256;
  MD_HIP(hipMalloc(&p, bytes));
  MD_HIP(hipMemsetAsync(p, 0, bytes, s->stream));
  s->allocations.push_back(p);
  return static_cast<T *>(p);
}

// Profiling phases (mdcore.h MD_PHASE_*): a launcher calls md_mark when it
// starts a phase; outside md_profile_steps it does nothing.
void md_mark(md_group *g, int phase);

// ---- per-system setup (each in its own .hip file) ---------------------------
void md_nb_setup(md_system *s);                 // nonbonded.hip: buffers of one system
std::vector<int> md_cell_walk(int W, int H, int D);  // nonbonded.hip: cell sort order
void md_nb_free(md_system *s);
int64_t md_nb_entries(md_system *s);
int64_t md_nb_capacity(md_system *s);
void md_pme_setup(md_system *s);                // pme.hip
void md_pme_free(md_system *s);

// ---- group tables: built once per group, refreshed when a box changes -------
void md_nb_group_setup(md_group *g);
void md_nb_group_refresh(md_group *g);
void md_nb_group_free(md_group *g);
void md_bonded_group_setup(md_group *g);
void md_bonded_group_refresh(md_group *g);
void md_bonded_group_free(md_group *g);
void md_pme_group_setup(md_group *g);
void md_pme_group_refresh(md_group *g);
void md_pme_group_free(md_group *g);
void md_int_group_setup(md_group *g);
void md_int_group_refresh(md_group *g);
void md_int_group_free(md_group *g);

// ---- group launchers -----------------------------------------------------------
void md_nb_set_schedule(md_group *g, int seg_len, int waves);  // pair-kernel schedule (autotuner)
void md_nb_get_schedule(md_group *g, int *seg_len, int *waves);
void md_nb_set_compact(md_group *g, bool compact);  // forces through the compacted pair schedule
bool md_nb_get_compact(md_group *g);
void md_nb_build(md_group *g);                  // rebuild kernels, gated per replica
void md_nb_pairs(md_group *g, bool energy);
void md_bonded(md_group *g, bool energy);       // bonded.hip
void md_zero_forces(md_group *g, bool energies);
void md_pme(md_group *g, bool energy);          // pme.hip
// integrate.hip: advance the step counters and latch each replica's rebuild
// request (its own displacement test, or `force_pending` set by the host)
void md_step_prologue(md_group *g, bool advance);
void md_integrate(md_group *g);
void md_momentum(md_group *g);
void md_constrain(md_group *g, int what, double tolerance);  // 0 positions, 1 velocities

// ---- box changes made on the device (the barostat): each subsystem rewrites its
// table entry (replica 0 of the group) from a reduced box in device memory -------
void md_nb_box_from_device(md_group *g, const double *box);
void md_bonded_box_from_device(md_group *g, const double *box);
void md_pme_box_from_device(md_group *g, const double *box);

// ---- the device barostat (barostat.hip) ---------------------------------------
md_energy_consts md_energy_constants(const md_system *s);
void md_baro_attach(md_system *s, const md_barostat_desc *d);
void md_baro_attempt(md_system *s);        // enqueue one attempt on the system's own group
void md_baro_set_box(md_system *s);         // the host's box onto the device state
void md_baro_sync_box(md_system *s);        // the device state's box into s->box (waits)
void md_baro_state(md_system *s, md_barostat_state *out);
void md_baro_release(md_system *s);
int md_baro_frequency(const md_system *s);  // 0 without a barostat
int &md_baro_step(md_system *s);            // steps since the last attempt

// ---- the device-resident minimiser (minimise.hip) -----------------------------
// L-BFGS work vectors: MD_MIN_VECTORS f64 3N-vectors in planes [3][np]
#define MD_MIN_MEMORY 6
#define MD_MIN_VECTORS (6 + 2 * MD_MIN_MEMORY)
struct md_min_work {
  double *vec = nullptr;      // [MD_MIN_VECTORS][3][np]
  double *partial = nullptr;  // reduction partials
  double *scalar = nullptr;   // one reduced scalar
  int *bad = nullptr;         // a coordinate outside the fixed-point range
  int nc = 0;                 // live constraints (springs)
  int *p1 = nullptr, *p2 = nullptr;
  double *dist = nullptr;
  double *spring_f = nullptr;  // [nc][3]: each spring's pull on its first atom
  int *cstart = nullptr;       // [np + 1]: by atom, its springs in clist
  int *clist = nullptr;        // spring c as c (first atom) or -c - 1 (second)
};
md_min_work *md_min_work_create(md_system *s, const std::vector<int> &p1, const std::vector<int> &p2,
                                const std::vector<double> &dist);
double *md_min_vec(md_system *s, md_min_work *w, int k);
double md_min_dot(md_system *s, md_min_work *w, const double *a, const double *b);
double md_min_absmax(md_system *s, md_min_work *w, const double *a);
void md_min_lincomb(md_system *s, double *out, double a, const double *x, double b, const double *y);
void md_min_positions_out(md_system *s, double *x);
bool md_min_positions_in(md_system *s, md_min_work *w, const double *x);
double md_min_gradient(md_system *s, md_min_work *w, const double *x, double kspring, double *g);
double md_min_constraint_error(md_system *s, md_min_work *w, const double *x);

// ---- learned SHAKE warm start (warmstart.hip) ---------------------------------
md_warmstart *md_ws_create(int device, uint32_t seed, int64_t capacity);
void md_ws_destroy(md_warmstart *w);
const md_ws_net *md_ws_device_net(md_warmstart *w);
int md_ws_device(md_warmstart *w);
void md_ws_reserve(md_warmstart *w, int64_t wanted, float **xs, float **ts, int64_t *room);
void md_ws_commit(md_warmstart *w, int64_t added);
int64_t md_ws_count(md_warmstart *w);
void md_ws_get_params(md_warmstart *w, float *out);
void md_ws_set_params(md_warmstart *w, const float *in);
double md_ws_evaluate(md_warmstart *w, const float *x, const float *t, int n, float *y, float *grad);
double md_ws_train(md_warmstart *w, int iterations, int batch, double lr);
