// barostat.hip -- the Monte Carlo barostat on the device.
//
// Upstream attempts the move on the host between GPU submissions
// (StepLoop.attemptBarostat, src/gpu/loop.ts:602-641, with the move of
// src/reference/barostat.ts:129-274 and the molecules of
// src/engine/molecules.ts:71), so every batch of steps is cut at an attempt.
// Here the attempt is a fixed sequence of kernels on the system's stream --
// no host decision, no read-back -- so it sits inside the captured stride
// graphs like any step:
//
//   k_baro_prepare        list rebuild requested (the energies use a fresh list)
//   energy                the step's force kernels with energies on
//   k_baro_propose        E, one Threefry draw, V' = V + s (2u - 1), the
//                         trial box (lengths times (V'/V)^(1/3), reduced)
//   k_baro_scale          per molecule: the unweighted centroid, scaled; atoms
//                         moved rigidly; old positions kept; residual cleared
//   box constants         every subsystem's table from the trial box
//   energy                again, at the trial
//   k_baro_decide         w = dE + P dV - N_mol kT ln(V'/V); accept when
//                         w <= 0 or u' <= e^(-w/kT) (u' drawn only uphill);
//                         the proposal width adapted every 10 attempts
//   k_baro_restore        a rejected trial's positions put back
//   box constants         from the box that won
//   zero forces           the next step starts from clean accumulators
//
// The arithmetic follows the host path (mdcore/barostat.py) operation for
// operation, with contraction into fused multiply-adds off, so a device
// attempt and a host attempt from the same state agree to the last bit of
// every step except the cube root, logarithm and exponential, which the
// device computes with its own library.
#include <cmath>
#include <cstddef>
#include <cstring>
#include <string>
#include <vector>

#include "engine.hpp"

// device-resident state of one system's barostat
struct md_baro_dev {
  double box[9], box_old[9];
  double volume_scale;
  double volume_old, volume_new, delta, length_scale;
  double e_old, e_new, work;
  int64_t draws, attempts, accepted;
  int32_t window_attempted, window_accepted;
  int32_t last_accepted, restore, error;
};

// host side
struct md_baro {
  md_baro_dev *dev = nullptr;
  int64_t *saved = nullptr;  // [3][np] positions before the trial
  int *mol_start = nullptr, *mol_atoms = nullptr;
  int nmol = 0;
  double pressure = 0;  // kJ/mol/nm^3
  double kt = 0;        // kJ/mol
  uint32_t seed = 0;
  int frequency = 0;
  int step = 0;  // steps since the last attempt (the schedule is fixed: host-side)
};

// the device address of the state's current box (no host access to it)
static double *dev_box(md_baro &B) {
  return reinterpret_cast<double *>(reinterpret_cast<char *>(B.dev) + offsetof(md_baro_dev, box));
}
static double *dev_box_old(md_baro &B) {
  return reinterpret_cast<double *>(reinterpret_cast<char *>(B.dev) + offsetof(md_baro_dev, box_old));
}

// src/reference/barostat.ts:48 (the barostat's own expression for a bar)
static const double MD_BAR = 6.02214076e23 * 1e-25;

md_energy_consts md_energy_constants(const md_system *s) {
  md_energy_consts k;
  k.has_nonbonded = s->has_nonbonded ? 1 : 0;
  k.dispersion = s->dispersion ? 1 : 0;
  k.disp_coeff = s->disp_coeff;
  k.alpha = s->alpha;
  k.q_total = s->q_total;
  k.q2_sum = s->q2_sum;
  return k;
}

__device__ inline double md_baro_uniform(md_baro_dev *B, uint32_t seed) {
  uint32_t o0, o1;
  md_threefry2x32_20((uint32_t)B->draws, seed, MD_BAROSTAT_DOMAIN, 0u, o0, o1);
  B->draws += 1;
  return ((double)(o0 >> 9) + 0.5) * 1.1920928955078125e-07;  // 2^-23
}

__global__ void k_baro_prepare(md_control *ctl) { ctl->force_pending = 1; }

__global__ void k_baro_propose(md_baro_dev *B, const int64_t *energy, md_energy_consts ek, uint32_t seed,
                               double cutoff, md_control *ctl) {
#pragma clang fp contract(off)
  double g[MD_GROUP_COUNT];
  const double volume = md_box_volume(B->box);
  B->e_old = md_energy_groups(energy, ek, volume, g);
  const double u = md_baro_uniform(B, seed);
  const double delta = B->volume_scale * 2.0 * (u - 0.5);
  const double new_volume = volume + delta;
  const double s = cbrt(new_volume / volume);
  double trial[9];
  for (int k = 0; k < 9; ++k) trial[k] = B->box[k] * s;
  md_reduce_box(trial);
  md_box d{};
  bool bad = false;
  md_box_from_matrix(trial, d, bad);
  double w[3];
  md_box_widths(trial, w);
  B->volume_old = volume;
  B->volume_new = new_volume;
  B->delta = delta;
  B->length_scale = s;
  B->error = 0;
  if (bad || !(new_volume > 0.0) || (ek.has_nonbonded && cutoff > 0.5 * fmin(w[0], fmin(w[1], w[2])))) {
    B->error = 1;
    ctl->barostat_error = 1;
    return;  // the box stays; the run stops at its next status check
  }
  for (int k = 0; k < 9; ++k) {
    B->box_old[k] = B->box[k];
    B->box[k] = trial[k];
  }
  ctl->force_pending = 1;
}

// one thread per molecule: its centroid (atoms summed in ascending order, as
// numpy's bincount does), moved by (s - 1) times itself, every atom with it
__global__ void k_baro_scale(md_baro_dev *B, const int *mol_start, const int *mol_atoms, int nmol, int64_t *pos,
                             int64_t *saved, float *resid, int np, md_control *ctl) {
#pragma clang fp contract(off)
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nmol || B->error) return;
  const int a0 = mol_start[m], a1 = mol_start[m + 1];
  double c[3] = {0.0, 0.0, 0.0};
  for (int k = a0; k < a1; ++k) {
    const int a = mol_atoms[k];
    for (int d = 0; d < 3; ++d) c[d] += md_fixed_to_double(pos[(size_t)d * np + a]);
  }
  const double size = (double)(a1 - a0), s = B->length_scale;
  double off[3];
  for (int d = 0; d < 3; ++d) {
    const double mean = c[d] / size;
    off[d] = s * mean - mean;
  }
  bool bad = false;
  for (int k = a0; k < a1; ++k) {
    const int a = mol_atoms[k];
    for (int d = 0; d < 3; ++d) {
      const size_t i = (size_t)d * np + a;
      const int64_t old = pos[i];
      const double x = md_fixed_to_double(old) + off[d];
      if (!(fabs(x) < 2147483648.0)) bad = true;
      saved[i] = old;
      pos[i] = bad ? old : md_double_to_fixed(x);
      resid[i] = 0.0f;
    }
  }
  if (bad) ctl->barostat_error = 2;
}

__global__ void k_baro_decide(md_baro_dev *B, const int64_t *energy, md_energy_consts ek, uint32_t seed,
                              double pressure, double kt, int nmol, md_control *ctl) {
#pragma clang fp contract(off)
  ctl->force_pending = 1;
  if (B->error) {
    B->restore = 0;
    B->last_accepted = 0;
    return;
  }
  double g[MD_GROUP_COUNT];
  B->e_new = md_energy_groups(energy, ek, B->volume_new, g);
  const double w = B->e_new - B->e_old + pressure * B->delta - (double)nmol * kt * log(B->volume_new / B->volume_old);
  // an uphill move draws a second deviate; a downhill one does not
  bool accepted = true;
  if (w > 0.0) accepted = !(md_baro_uniform(B, seed) > exp(-w / kt));
  B->work = w;
  B->attempts += 1;
  B->window_attempted += 1;
  if (accepted) {
    B->accepted += 1;
    B->window_accepted += 1;
  }
  if (B->window_attempted >= 10) {
    if (B->window_accepted < 0.25 * B->window_attempted) {
      B->volume_scale /= 1.1;
      B->window_attempted = B->window_accepted = 0;
    } else if (B->window_accepted > 0.75 * B->window_attempted) {
      B->volume_scale = fmin(B->volume_scale * 1.1, B->volume_old * 0.3);
      B->window_attempted = B->window_accepted = 0;
    }
  }
  B->last_accepted = accepted ? 1 : 0;
  B->restore = accepted ? 0 : 1;
  if (!accepted)
    for (int k = 0; k < 9; ++k) B->box[k] = B->box_old[k];
}

__global__ void k_baro_restore(const md_baro_dev *B, int64_t *pos, const int64_t *saved, int n, int np) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !B->restore) return;
  for (int d = 0; d < 3; ++d) pos[(size_t)d * np + i] = saved[(size_t)d * np + i];
}

// the step's force kernels with energies on, into the energy slots
static void baro_energy(md_group *g, bool nb) {
  md_zero_forces(g, true);
  md_step_prologue(g, false);
  if (nb) {
    md_nb_build(g);
    md_nb_pairs(g, true);
  }
  md_bonded(g, true);
  if (nb) md_pme(g, true);
}

static void baro_box_constants(md_group *g, const double *box, bool nb) {
  md_bonded_box_from_device(g, box);
  if (nb) {
    md_nb_box_from_device(g, box);
    md_pme_box_from_device(g, box);
  }
}

void md_baro_attempt(md_system *s) {
  md_baro &B = *s->baro;
  md_group *g = s->self;
  const bool nb = g->nb != nullptr;
  const md_energy_consts ek = md_energy_constants(s);
  k_baro_prepare<<<1, 1, 0, g->stream>>>(s->ctl);
  baro_energy(g, nb);
  k_baro_propose<<<1, 1, 0, g->stream>>>(B.dev, s->energy, ek, B.seed, s->has_nonbonded ? s->cutoff : 0.0, s->ctl);
  k_baro_scale<<<(B.nmol + 63) / 64, 64, 0, g->stream>>>(B.dev, B.mol_start, B.mol_atoms, B.nmol, s->pos, B.saved,
                                                        s->resid, s->np, s->ctl);
  baro_box_constants(g, dev_box(B), nb);
  baro_energy(g, nb);
  k_baro_decide<<<1, 1, 0, g->stream>>>(B.dev, s->energy, ek, B.seed, B.pressure, B.kt, B.nmol, s->ctl);
  k_baro_restore<<<(s->n + 255) / 256, 256, 0, g->stream>>>(B.dev, s->pos, B.saved, s->n, s->np);
  baro_box_constants(g, dev_box(B), nb);
  md_zero_forces(g, false);
  MD_HIP(hipGetLastError());
}

void md_baro_attach(md_system *s, const md_barostat_desc *d) {
  if (d->frequency < 0) throw md_error(MD_ERR_INVALID_ARGUMENT, "a barostat frequency cannot be negative");
  if (d->frequency == 0) {
    md_baro_release(s);
    return;
  }
  if (!std::isfinite(d->pressure) || !(d->temperature > 0) || !std::isfinite(d->temperature))
    throw md_error(MD_ERR_INVALID_ARGUMENT, "the barostat needs a finite pressure and a positive temperature");
  if (!(d->volume_scale >= 0) || d->draws < 0 || d->steps_since_attempt < 0 || d->window_attempted < 0 ||
      d->window_accepted < 0 || d->window_accepted > d->window_attempted)
    throw md_error(MD_ERR_INVALID_ARGUMENT, "bad barostat starting state");
  const int n = s->n, nmol = d->n_molecules;
  if (nmol < 1 || nmol > n || !d->molecule_start || !d->molecule_atoms)
    throw md_error(MD_ERR_INVALID_ARGUMENT, "the barostat needs between 1 and n_particles molecules");
  std::vector<int> start(d->molecule_start, d->molecule_start + nmol + 1);
  std::vector<int> atoms(d->molecule_atoms, d->molecule_atoms + n);
  if (start[0] != 0 || start[nmol] != n)
    throw md_error(MD_ERR_INVALID_ARGUMENT, "molecule_start must run from 0 to n_particles");
  std::vector<char> seen(n, 0);
  for (int m = 0; m < nmol; ++m) {
    if (start[m + 1] <= start[m]) throw md_error(MD_ERR_INVALID_ARGUMENT, "a molecule has no atoms");
    for (int k = start[m]; k < start[m + 1]; ++k) {
      const int a = atoms[k];
      if (a < 0 || a >= n || seen[a])
        throw md_error(MD_ERR_INVALID_ARGUMENT, "molecule_atoms must name every particle exactly once");
      if (k > start[m] && a <= atoms[k - 1])
        throw md_error(MD_ERR_INVALID_ARGUMENT, "molecule_atoms must ascend within each molecule");
      seen[a] = 1;
    }
  }
  for (md_group *g : s->groups)
    if (g != s->self) throw md_error(MD_ERR_INVALID_ARGUMENT, "a system in a batch cannot carry a barostat");
  MD_HIP(hipSetDevice(s->device));
  if (!s->baro) {
    s->baro = new md_baro();
    s->baro->dev = md_alloc<md_baro_dev>(s, 1);
    s->baro->saved = md_alloc<int64_t>(s, 3 * (size_t)s->np);
    s->baro->mol_atoms = md_alloc<int>(s, n);
  }
  md_baro &B = *s->baro;
  if (B.nmol != nmol) B.mol_start = md_alloc<int>(s, nmol + 1);
  B.nmol = nmol;
  B.pressure = d->pressure * MD_BAR;
  B.kt = MD_BOLTZMANN * d->temperature;
  B.seed = d->seed;
  B.frequency = d->frequency;
  B.step = d->steps_since_attempt;
  md_baro_dev h{};
  std::memcpy(h.box, s->box, sizeof h.box);
  std::memcpy(h.box_old, s->box, sizeof h.box_old);
  h.volume_scale = d->volume_scale > 0 ? d->volume_scale : 0.01 * md_box_volume(s->box);
  h.draws = d->draws;
  h.window_attempted = d->window_attempted;
  h.window_accepted = d->window_accepted;
  h.length_scale = 1.0;
  MD_HIP(hipMemcpyAsync(B.mol_start, start.data(), sizeof(int) * start.size(), hipMemcpyHostToDevice, s->stream));
  MD_HIP(hipMemcpyAsync(B.mol_atoms, atoms.data(), sizeof(int) * atoms.size(), hipMemcpyHostToDevice, s->stream));
  MD_HIP(hipMemcpyAsync(B.dev, &h, sizeof h, hipMemcpyHostToDevice, s->stream));
  MD_HIP(hipStreamSynchronize(s->stream));
  s->self->graph_stale = true;  // strides captured before carry no attempts
}

void md_baro_release(md_system *s) {
  if (!s->baro) return;
  // device memory belongs to the system's allocations
  delete s->baro;
  s->baro = nullptr;
  if (s->self) s->self->graph_stale = true;
}

void md_baro_set_box(md_system *s) {
  if (!s->baro) return;
  MD_HIP(hipMemcpyAsync(dev_box(*s->baro), s->box, sizeof s->box, hipMemcpyHostToDevice, s->stream));
  MD_HIP(hipMemcpyAsync(dev_box_old(*s->baro), s->box, sizeof s->box, hipMemcpyHostToDevice, s->stream));
  MD_HIP(hipStreamSynchronize(s->stream));
}

void md_baro_sync_box(md_system *s) {
  if (!s->baro) return;
  double b[9];
  MD_HIP(hipMemcpyAsync(b, dev_box(*s->baro), sizeof b, hipMemcpyDeviceToHost, s->stream));
  MD_HIP(hipStreamSynchronize(s->stream));
  std::memcpy(s->box, b, sizeof b);
  bool bad = false;
  std::memset(&s->dbox, 0, sizeof s->dbox);
  md_box_from_matrix(b, s->dbox, bad);
}

void md_baro_state(md_system *s, md_barostat_state *out) {
  std::memset(out, 0, sizeof *out);
  if (!s->baro) return;
  md_baro_dev h;
  MD_HIP(hipMemcpyAsync(&h, s->baro->dev, sizeof h, hipMemcpyDeviceToHost, s->stream));
  MD_HIP(hipStreamSynchronize(s->stream));
  out->attempts = h.attempts;
  out->accepted = h.accepted;
  out->draws = h.draws;
  out->steps_since_attempt = s->baro->step;
  out->window_attempted = h.window_attempted;
  out->window_accepted = h.window_accepted;
  out->volume_scale = h.volume_scale;
  out->last_accepted = h.last_accepted;
  out->last_energy_before = h.e_old;
  out->last_energy_after = h.e_new;
  out->last_work = h.work;
  out->last_length_scale = h.length_scale;
}

int md_baro_frequency(const md_system *s) { return s->baro ? s->baro->frequency : 0; }
int &md_baro_step(md_system *s) { return s->baro->step; }
