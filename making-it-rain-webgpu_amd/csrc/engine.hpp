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
  uint32_t cluster_reach_bits[2];  // largest cluster radius, super-cluster reach
  int64_t momentum[2][3];          // sum m v in Q32.32, by step parity
  int64_t pair_census;             // in-cutoff pairs, when counted
  int32_t segment_cursor;          // next list segment for the pair kernel's waves
  int32_t pad_;
};

struct md_nb_params {
  int nclusters, nslots;  // 8-atom clusters; slots in whole 64-atom super-clusters
  int ncells, gdim[3];    // sorting cells
  float cutoff2, list_cutoff2, half_skin2;
};

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
  float4 *slot_xyzq = nullptr;
  float4 *cluster_half = nullptr;
  int *excl_start_d = nullptr, *excl_list_d = nullptr;
  void *nb_ext = nullptr;

  // ---- bonded terms, exceptions, restraints (one flattened list) --------
  int n_bonds = 0, n_angles = 0, n_torsions = 0, n_exceptions = 0, n_restraints = 0;
  int4 *term_atoms = nullptr;
  double4 *term_param = nullptr;
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

  // ---- run state ---------------------------------------------------------
  hipGraphExec_t stride_exec = nullptr;
  hipGraph_t stride_graph = nullptr;
  int stride_steps = 0;
  int64_t steps_done = 0;
  bool need_rebuild = true;
  bool momentum_dirty = true;
  // profiling: events recorded by md_mark while `profiling` is set
  bool profiling = false;
  std::vector<hipEvent_t> prof_events;
  std::vector<int> prof_phase;

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
void md_mark(md_system *s, int phase);

// ---- launchers (each in its own .hip file) ----------------------------------
void md_nb_setup(md_system *s);                 // nonbonded.hip
std::vector<int> md_cell_walk(int W, int H, int D);  // nonbonded.hip: cell sort order
void md_nb_free(md_system *s);
void md_nb_build(md_system *s, bool forced);    // rebuild kernels, device-gated
void md_nb_pairs(md_system *s, bool energy, bool census);
int64_t md_nb_entries(md_system *s);
int64_t md_nb_capacity(md_system *s);
void md_bonded(md_system *s, bool energy);      // bonded.hip
void md_zero_forces(md_system *s, bool energies);
void md_pme_setup(md_system *s);                // pme.hip
void md_pme_free(md_system *s);
void md_pme(md_system *s, bool energy);
void md_step_prologue(md_system *s, bool advance, bool force_rebuild);  // integrate.hip
void md_integrate(md_system *s);
void md_momentum(md_system *s);
void md_constrain(md_system *s, int what, double tolerance);  // 0 positions, 1 velocities
