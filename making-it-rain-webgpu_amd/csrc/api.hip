// api.hip -- the C ABI of include/mdcore.h: System validation and packing,
// uploads and readbacks, force evaluation and the step loop.
//
// Host-side counterpart of StepLoop (src/gpu/loop.ts:208-1116) and the
// per-force `prepare*` modules (nonbonded.ts, pme.ts, constrain.ts,
// forces.ts, restraint.ts, integrate.ts).  A step is a fixed sequence of
// kernel launches on the system's stream; a stride of steps is captured once
// as a HIP graph and replayed, with the step counter, rebuild decision and
// COM momentum held in device memory so replays need no host round trip.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <tuple>
#include <string>
#include <vector>

#include "engine.hpp"
#include "warmstart.hpp"

static thread_local std::string g_last_error;

// the message md_last_error reports (for entry points defined in other files)
void md_set_last_error(const std::string &message) { g_last_error = message; }

static md_err fail(const md_error &e) {
  g_last_error = e.message;
  return e.code;
}

#define MD_API_BEGIN try {
#define MD_API_END                                       \
  }                                                      \
  catch (const md_error &e) {                            \
    return fail(e);                                      \
  }                                                      \
  catch (const std::bad_alloc &) {                       \
    g_last_error = "host allocation failed";             \
    return MD_ERR_OUT_OF_MEMORY;                         \
  }                                                      \
  catch (const std::exception &e) {                      \
    g_last_error = e.what();                             \
    return MD_ERR_INVALID_ARGUMENT;                      \
  }                                                      \
  g_last_error.clear();                                  \
  return MD_OK;

static const int STRIDE_STEPS = 50;

// ---- cell geometry (src/reference/triclinic.ts) ------------------------------

static void reduce_box(const double *in, double *out) {
  double v[9];
  std::memcpy(v, in, sizeof v);
  for (int k : {1, 2, 5})
    if (std::fabs(v[k]) > 1e-9)
      throw md_error(MD_ERR_INVALID_ARGUMENT, "box vectors must be lower triangular: a along x, b in the xy plane");
  if (v[0] <= 0 || v[4] <= 0 || v[8] <= 0)
    throw md_error(MD_ERR_INVALID_ARGUMENT, "box vectors need positive diagonal elements");
  md_reduce_box(v);  // ties to even, as reduceBox (triclinic.ts:93-125)
  std::memcpy(out, v, sizeof v);
}

static double box_volume(const double *b) { return md_box_volume(b); }

static double min_perpendicular_width(const double *b) {
  double w[3];
  md_box_widths(b, w);
  return std::min(w[0], std::min(w[1], w[2]));
}

static int64_t fixed_checked(double x, const char *what) {
  if (!std::isfinite(x) || std::fabs(x) >= 2147483648.0)
    throw md_error(MD_ERR_FIXED_POINT_OVERFLOW,
                   std::string(what) + " " + std::to_string(x) +
                       " is outside the fixed-point range of +/-2^31; a value this large is almost always a "
                       "diverged simulation rather than a format limit");
  return md_double_to_fixed(x);
}

// ---- launch groups -----------------------------------------------------------

md_launch md_make_launch(md_group *g, const std::vector<int> &blocks) {
  md_launch L;
  std::vector<int> start(blocks.size() + 1, 0);
  for (size_t r = 0; r < blocks.size(); ++r) start[r + 1] = start[r] + blocks[r];
  L.total = start.back();
  L.map.start = md_group_upload(g, start);
  L.map.R = (int)blocks.size();
  return L;
}

static void group_refresh(md_group *g) {
  MD_HIP(hipSetDevice(g->device));
  if (g->nb) md_nb_group_refresh(g);
  md_bonded_group_refresh(g);
  if (g->pme) md_pme_group_refresh(g);
  md_int_group_refresh(g);
}

static void group_release(md_group *g) {
  if (!g) return;
  hipSetDevice(g->device);
  if (g->stream) hipStreamSynchronize(g->stream);
  for (auto &c : g->graphs) {
    if (c.exec) hipGraphExecDestroy(c.exec);
    if (c.graph) hipGraphDestroy(c.graph);
  }
  if (g->nb) md_nb_group_free(g);
  if (g->bonded) md_bonded_group_free(g);
  if (g->pme) md_pme_group_free(g);
  if (g->integ) md_int_group_free(g);
  for (void *p : g->allocations) hipFree(p);
  if (g->own_stream && g->stream) hipStreamDestroy(g->stream);
  for (md_system *s : g->reps) {
    auto &v = s->groups;
    v.erase(std::remove(v.begin(), v.end(), g), v.end());
  }
  if (g->owner) *g->owner = nullptr;
  delete g;
}

// ---- the pair-kernel autotuner's cache ----------------------------------------
//
// Tuned schedules (entries per work unit, resident waves) keyed by what sets
// the best one: the number of replicas, the launch's atom count in
// half-octave buckets, and whether any cell is triclinic.  A new launch group
// whose key has an entry starts from it.
struct md_tune_entry {
  int seg_len, waves;
  bool compact;
};
static std::map<std::tuple<int, int, int>, md_tune_entry> g_tuned;

static std::tuple<int, int, int> tune_key(const md_group *g) {
  long atoms = 0;
  int triclinic = 0;
  for (const md_system *s : g->reps) {
    atoms += s->n;
    triclinic |= (s->box[3] != 0.0 || s->box[6] != 0.0 || s->box[7] != 0.0) ? 1 : 0;
  }
  const int bucket = (int)std::floor(2.0 * std::log2((double)std::max(1L, atoms)));
  return std::make_tuple((int)g->reps.size(), bucket, triclinic);
}

static void apply_tuned_schedule(md_group *g) {
  auto it = g_tuned.find(tune_key(g));
  if (it != g_tuned.end()) {
    md_nb_set_schedule(g, it->second.seg_len, it->second.waves);
    md_nb_set_compact(g, it->second.compact);
  }
}

// a launch group over `reps` (all on one device), on `stream` or its own
static md_group *group_create(const std::vector<md_system *> &reps, hipStream_t stream) {
  md_group *g = new md_group();
  try {
    g->device = reps[0]->device;
    g->reps = reps;
    g->max_batch = reps[0]->opt.max_batch;
    MD_HIP(hipSetDevice(g->device));
    if (stream) {
      g->stream = stream;
    } else {
      MD_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
      g->own_stream = true;
    }
    bool any_nb = false, all_nb = true;
    for (md_system *s : reps) {
      any_nb |= s->has_nonbonded;
      all_nb &= s->has_nonbonded;
    }
    if (any_nb && !all_nb)
      throw md_error(MD_ERR_INVALID_ARGUMENT, "a batch mixes systems with and without a NonbondedForce");
    g->status = md_group_alloc<uint32_t>(g, 1);
    if (any_nb) {
      md_nb_group_setup(g);
      apply_tuned_schedule(g);
      md_pme_group_setup(g);
    }
    md_bonded_group_setup(g);
    md_int_group_setup(g);
    MD_HIP(hipStreamSynchronize(g->stream));
    for (md_system *s : reps) s->groups.push_back(g);
  } catch (...) {
    group_release(g);
    throw;
  }
  return g;
}

static void apply_box(md_system *s, const double *in) {
  double b[9];
  reduce_box(in, b);
  if (s->has_nonbonded) {
    double maxc = 0.5 * min_perpendicular_width(b);
    if (s->cutoff > maxc)
      throw md_error(MD_ERR_CUTOFF_TOO_LARGE,
                     "a " + std::to_string(s->cutoff) + " nm cutoff does not fit this box: the minimum image is only "
                     "exact below " + std::to_string(maxc) + " nm (half the narrowest perpendicular width)");
  }
  for (int k = 0; k < 9; ++k) (void)fixed_checked(b[k], "box");
  std::memcpy(s->box, b, sizeof b);
  md_box &d = s->dbox;
  std::memset(&d, 0, sizeof d);
  bool bad = false;
  md_box_from_matrix(b, d, bad);
  // the kernels read the box from their groups' tables: refresh them (a
  // captured stride keeps pointing at the tables, so it stays valid)
  for (md_group *g : s->groups) group_refresh(g);
  s->need_rebuild = true;
}

// ---- System packing ------------------------------------------------------

static void check_index(int v, int n, const char *what) {
  if (v < 0 || v >= n)
    throw md_error(MD_ERR_INVALID_ARGUMENT, std::string(what) + " index " + std::to_string(v) + " out of range");
}

struct split_result {
  std::vector<int4> waters;
  std::vector<float2> water_d;
  std::vector<std::vector<int>> cl_atoms;
  std::vector<std::vector<int>> cl_cons;  // constraint indices
  std::vector<int> free_atoms;
  int max_atoms = 0;
};

// SETTLE triangles by f32-rounded equal sides, union-find for the rest
// (src/engine/constraint-clusters.ts:93-249)
static split_result split_constraints(const md_system_desc *d) {
  int n = d->n_particles, nc = d->n_constraints;
  split_result R;
  std::vector<int> count(n, 0);
  std::vector<char> live(nc, 0);
  for (int c = 0; c < nc; ++c) {
    int p1 = d->constraint_p1[c], p2 = d->constraint_p2[c];
    if (d->masses[p1] == 0.0 && d->masses[p2] == 0.0) continue;
    live[c] = 1;
    count[p1]++;
    count[p2]++;
  }
  std::vector<int> np(n, 0);
  std::vector<int> pa(2 * (size_t)n, -1);
  std::vector<float> pd(2 * (size_t)n, 0.0f);
  for (int c = 0; c < nc; ++c) {
    if (!live[c]) continue;
    int p1 = d->constraint_p1[c], p2 = d->constraint_p2[c];
    if (count[p1] == 2 && count[p2] == 2) {
      float dist = (float)d->constraint_distance[c];
      if (np[p1] < 2) pa[2 * p1 + np[p1]] = p2, pd[2 * p1 + np[p1]] = dist;
      np[p1]++;
      if (np[p2] < 2) pa[2 * p2 + np[p2]] = p1, pd[2 * p2 + np[p2]] = dist;
      np[p2]++;
    }
  }
  std::vector<char> settle(n, 0);
  for (int p1 = 0; p1 < n; ++p1) {
    if (np[p1] != 2) continue;
    int p2 = pa[2 * p1], p3 = pa[2 * p1 + 1];
    if (!(p1 < p2 && p1 < p3) || np[p2] != 2 || np[p3] != 2) continue;
    float d23 = -1.0f;
    for (int k = 0; k < 2; ++k)
      if (pa[2 * p2 + k] == p3) d23 = pd[2 * p2 + k];
    if (d23 < 0.0f) continue;
    float d12 = pd[2 * p1], d13 = pd[2 * p1 + 1];
    int4 w;
    float2 dd;
    if (d12 == d13)
      w = make_int4(p1, p2, p3, 0), dd = make_float2(d12, d23);
    else if (d12 == d23)
      w = make_int4(p2, p1, p3, 0), dd = make_float2(d12, d13);
    else if (d13 == d23)
      w = make_int4(p3, p1, p2, 0), dd = make_float2(d13, d12);
    else
      continue;
    R.waters.push_back(w);
    R.water_d.push_back(dd);
    settle[p1] = settle[p2] = settle[p3] = 1;
  }
  std::vector<int> parent(n);
  std::iota(parent.begin(), parent.end(), 0);
  auto find = [&](int x) {
    while (parent[x] != x) x = parent[x] = parent[parent[x]];
    return x;
  };
  std::vector<int> iterative;
  for (int c = 0; c < nc; ++c) {
    if (!live[c] || settle[d->constraint_p1[c]]) continue;
    iterative.push_back(c);
    int a = find(d->constraint_p1[c]), b = find(d->constraint_p2[c]);
    if (a != b) parent[a] = b;
  }
  std::map<int, int> cluster_of_root;
  std::vector<char> in_cluster(n, 0);
  for (int c : iterative) {
    int r = find(d->constraint_p1[c]);
    auto it = cluster_of_root.find(r);
    int k;
    if (it == cluster_of_root.end()) {
      k = (int)R.cl_atoms.size();
      cluster_of_root[r] = k;
      R.cl_atoms.emplace_back();
      R.cl_cons.emplace_back();
    } else {
      k = it->second;
    }
    R.cl_cons[k].push_back(c);
    for (int p : {d->constraint_p1[c], d->constraint_p2[c]})
      if (!in_cluster[p]) {
        in_cluster[p] = 1;
        R.cl_atoms[k].push_back(p);
      }
  }
  for (size_t k = 0; k < R.cl_atoms.size(); ++k) {
    std::sort(R.cl_atoms[k].begin(), R.cl_atoms[k].end());
    int na = (int)R.cl_atoms[k].size(), ncs = (int)R.cl_cons[k].size();
    if (na > MD_MAX_CLUSTER_ATOMS || ncs > MD_MAX_CLUSTER_CONSTRAINTS)
      throw md_error(MD_ERR_CLUSTER_TOO_LARGE,
                     "a constraint cluster has " + std::to_string(na) + " atoms and " + std::to_string(ncs) +
                         " constraints, beyond the 8 and 8 one thread holds; rebuild with constraints=HBonds");
    R.max_atoms = std::max(R.max_atoms, na);
  }
  for (int i = 0; i < n; ++i)
    if (!settle[i] && !in_cluster[i]) R.free_atoms.push_back(i);
  return R;
}

template <typename T>
static T *upload(md_system *s, const std::vector<T> &v) {
  T *p = md_alloc<T>(s, std::max<size_t>(1, v.size()));
  if (!v.empty()) MD_HIP(hipMemcpyAsync(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s->stream));
  return p;
}

static int smooth_size(int minimum) {
  for (int n = std::max(1, minimum);; ++n) {
    int r = n;
    for (int p : {2, 3, 5, 7})
      while (r % p == 0) r /= p;
    if (r == 1) return n;
  }
}

static void create(md_system *s, const md_system_desc *d, const md_run_options *o) {
  if (!d || !o) throw md_error(MD_ERR_INVALID_ARGUMENT, "null description or options");
  int n = d->n_particles;
  if (n <= 0) throw md_error(MD_ERR_INVALID_ARGUMENT, "a system needs at least one particle");
  s->opt = *o;
  if (s->opt.time_step <= 0) throw md_error(MD_ERR_INVALID_ARGUMENT, "time_step must be positive");
  if (s->opt.max_batch < 0 || s->opt.max_batch > 100000)
    throw md_error(MD_ERR_INVALID_ARGUMENT, "max_batch must be between 0 and 100000");
  if (s->opt.constraint_tolerance <= 0) s->opt.constraint_tolerance = 1e-6;
  // Everything up to the device state below is host-side validation and
  // packing, so a System the engine refuses (ClusterTooLarge,
  // CutoffTooLargeForBox, NotSmooth, bad indices) is refused the same way
  // with or without a GPU.
  s->n = n;
  s->np = (n + 63) / 64 * 64;
  s->masses.assign(d->masses, d->masses + n);
  s->total_mass = 0;
  int moving = 0;
  for (double m : s->masses) {
    if (m < 0) throw md_error(MD_ERR_INVALID_ARGUMENT, "negative mass");
    s->total_mass += m;
    moving += m > 0;
  }
  for (int c = 0; c < d->n_constraints; ++c) {
    check_index(d->constraint_p1[c], n, "constraint");
    check_index(d->constraint_p2[c], n, "constraint");
  }
  s->cmm_frequency = d->cmm_frequency;
  s->n_constraints = d->n_constraints;
  s->con_p1.assign(d->constraint_p1, d->constraint_p1 + d->n_constraints);
  s->con_p2.assign(d->constraint_p2, d->constraint_p2 + d->n_constraints);
  s->con_d.assign(d->constraint_distance, d->constraint_distance + d->n_constraints);
  s->dof = 3 * moving - d->n_constraints - (d->cmm_frequency > 0 ? 3 : 0);

  // nonbonded parameters and host constants
  s->has_nonbonded = d->has_nonbonded != 0;
  std::vector<float4> ap(s->np, make_float4(0, 0, 0, 0));
  s->excl_start.assign(n + 1, 0);
  if (s->has_nonbonded) {
    s->cutoff = d->cutoff;
    s->ewald_tol = d->ewald_tolerance;
    if (s->cutoff <= 0) throw md_error(MD_ERR_INVALID_ARGUMENT, "cutoff must be positive");
    s->alpha = d->alpha != 0.0 ? d->alpha : (1.0 / s->cutoff) * std::sqrt(-std::log(2.0 * s->ewald_tol));
    s->dispersion = d->use_dispersion_correction != 0;
    s->charge.assign(d->charge, d->charge + n);
    s->sigma.assign(d->sigma, d->sigma + n);
    s->epsilon.assign(d->epsilon, d->epsilon + n);
    double sk = std::sqrt(MD_ONE_4PI_EPS0);
    for (int i = 0; i < n; ++i) {
      ap[i] = make_float4((float)(s->charge[i] * sk), (float)(0.5 * s->sigma[i]),
                          (float)(2.0 * std::sqrt(s->epsilon[i])), (float)s->charge[i]);
      s->q_total += s->charge[i];
      s->q2_sum += s->charge[i] * s->charge[i];
    }
    // exclusion CSR, both directions (nonbonded.ts:159-182)
    std::vector<std::vector<int>> ex(n);
    for (int e = 0; e < d->n_exceptions; ++e) {
      int a = d->exception_p1[e], b = d->exception_p2[e];
      check_index(a, n, "exception");
      check_index(b, n, "exception");
      ex[a].push_back(b);
      ex[b].push_back(a);
    }
    for (int i = 0; i < n; ++i) {
      std::sort(ex[i].begin(), ex[i].end());
      ex[i].erase(std::unique(ex[i].begin(), ex[i].end()), ex[i].end());
      s->excl_start[i + 1] = s->excl_start[i] + (int)ex[i].size();
      s->excl_list.insert(s->excl_list.end(), ex[i].begin(), ex[i].end());
    }
    // dispersion: grouped by distinct (sigma, epsilon) over the N(N+1)/2
    // unordered pairs including i == j (nonbonded.ts:107-162)
    if (s->dispersion) {
      std::map<std::pair<double, double>, double> classes;
      for (int i = 0; i < n; ++i) classes[{s->sigma[i], s->epsilon[i]}] += 1.0;
      std::vector<std::pair<std::pair<double, double>, double>> cl(classes.begin(), classes.end());
      double s6 = 0, s12 = 0;
      for (size_t a = 0; a < cl.size(); ++a)
        for (size_t b = a; b < cl.size(); ++b) {
          double sig = 0.5 * (cl[a].first.first + cl[b].first.first);
          double eps = std::sqrt(cl[a].first.second * cl[b].first.second);
          double sig6 = std::pow(sig, 6);
          double mult = a == b ? cl[a].second * (cl[a].second + 1) / 2 : cl[a].second * cl[b].second;
          s6 += mult * eps * sig6;
          s12 += mult * eps * sig6 * sig6;
        }
      double pairs = (double)n * (n + 1) / 2.0, rc = s->cutoff;
      s->disp_coeff = 8.0 * M_PI * (double)n * n * ((s12 / pairs) / (9.0 * std::pow(rc, 9)) -
                                                    (s6 / pairs) / (3.0 * std::pow(rc, 3)));
    }
  }
  apply_box(s, d->box);

  // PME mesh: run override, then the force's own, then from the box diagonal
  if (s->has_nonbonded) {
    for (int a = 0; a < 3; ++a) {
      int g = o->mesh[a] > 0 ? o->mesh[a] : d->grid[a];
      if (g <= 0) {
        double L = s->box[4 * a];
        int m = (int)std::ceil(2.0 * s->alpha * L / (3.0 * std::pow(s->ewald_tol, 0.2)));
        g = smooth_size(std::max(6, m));
      }
      if (g < MD_PME_ORDER) throw md_error(MD_ERR_INVALID_ARGUMENT, "PME mesh smaller than the spline order");
      int r = g;
      for (int f : {2, 3, 5, 7, 11, 13})
        while (r % f == 0) r /= f;
      if (r != 1)
        throw md_error(MD_ERR_NOT_SMOOTH, "PME mesh size " + std::to_string(g) + " has a prime factor above 13 (" +
                                              std::to_string(r) + " is left); mesh sizes must be 13-smooth");
      s->grid[a] = g;
    }
  }

  // bonded, exceptions, restraints: one list ordered by kind
  std::vector<int4> ta;
  std::vector<double4> tp;
  for (int t = 0; t < d->n_bonds; ++t) {
    check_index(d->bond_p1[t], n, "bond");
    check_index(d->bond_p2[t], n, "bond");
    ta.push_back(make_int4(d->bond_p1[t], d->bond_p2[t], 0, 0));
    tp.push_back(make_double4(d->bond_length[t], d->bond_k[t], 0, 0));
  }
  for (int t = 0; t < d->n_angles; ++t) {
    check_index(d->angle_p1[t], n, "angle");
    check_index(d->angle_p2[t], n, "angle");
    check_index(d->angle_p3[t], n, "angle");
    ta.push_back(make_int4(d->angle_p1[t], d->angle_p2[t], d->angle_p3[t], 0));
    tp.push_back(make_double4(d->angle_theta[t], d->angle_k[t], 0, 0));
  }
  for (int t = 0; t < d->n_torsions; ++t) {
    for (const int32_t *p : {d->torsion_p1, d->torsion_p2, d->torsion_p3, d->torsion_p4}) check_index(p[t], n, "torsion");
    ta.push_back(make_int4(d->torsion_p1[t], d->torsion_p2[t], d->torsion_p3[t], d->torsion_p4[t]));
    tp.push_back(make_double4(d->torsion_phase[t], d->torsion_k[t], (double)d->torsion_periodicity[t], 0));
  }
  int nexc = s->has_nonbonded ? d->n_exceptions : 0;
  for (int t = 0; t < nexc; ++t) {
    int a = d->exception_p1[t], b = d->exception_p2[t];
    ta.push_back(make_int4(a, b, 0, 0));
    tp.push_back(make_double4(MD_ONE_4PI_EPS0 * d->charge[a] * d->charge[b],
                              MD_ONE_4PI_EPS0 * d->exception_chargeprod[t], d->exception_sigma[t],
                              d->exception_epsilon[t]));
  }
  std::vector<int64_t> rref(3 * (size_t)std::max(1, d->n_restraints), 0);
  for (int t = 0; t < d->n_restraints; ++t) {
    int p = d->restraint_particle[t];
    check_index(p, n, "restraint");
    ta.push_back(make_int4(p, 0, 0, 0));
    tp.push_back(make_double4(d->restraint_k[t], 0, 0, 0));
    rref[t] = fixed_checked(d->restraint_x0[t], "restraint x0");
    rref[d->n_restraints + t] = fixed_checked(d->restraint_y0[t], "restraint y0");
    rref[2 * (size_t)d->n_restraints + t] = fixed_checked(d->restraint_z0[t], "restraint z0");
  }
  s->n_bonds = d->n_bonds;
  s->n_angles = d->n_angles;
  s->n_torsions = d->n_torsions;
  s->n_exceptions = nexc;
  s->n_restraints = d->n_restraints;

  // constraint work items
  split_result R = split_constraints(d);
  std::vector<int> ccl_atoms, ccl_na, ccl_nc;
  std::vector<int4> ccl_cons;
  for (size_t k = 0; k < R.cl_atoms.size(); ++k) {
    const auto &at = R.cl_atoms[k];
    for (int a = 0; a < MD_MAX_CLUSTER_ATOMS; ++a) ccl_atoms.push_back(a < (int)at.size() ? at[a] : 0);
    ccl_na.push_back((int)at.size());
    ccl_nc.push_back((int)R.cl_cons[k].size());
    for (int c = 0; c < MD_MAX_CLUSTER_CONSTRAINTS; ++c) {
      if (c < (int)R.cl_cons[k].size()) {
        int ci = R.cl_cons[k][c];
        int li = (int)(std::find(at.begin(), at.end(), d->constraint_p1[ci]) - at.begin());
        int lj = (int)(std::find(at.begin(), at.end(), d->constraint_p2[ci]) - at.begin());
        double dist = d->constraint_distance[ci];
        int64_t bits;
        std::memcpy(&bits, &dist, 8);
        ccl_cons.push_back(make_int4(li, lj, (int)(uint32_t)(bits & 0xffffffffu), (int)(uint32_t)(bits >> 32)));
      } else {
        ccl_cons.push_back(make_int4(0, 0, 0, 0));
      }
    }
  }
  s->n_waters = (int)R.waters.size();
  s->n_clusters_c = (int)R.cl_atoms.size();
  s->n_free = (int)R.free_atoms.size();
  s->max_cluster_atoms = R.max_atoms;

  // device state
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    throw md_error(MD_ERR_NO_DEVICE, "no HIP device is visible; the engine has no CPU fallback");
  if (o->device < 0 || o->device >= ndev) throw md_error(MD_ERR_INVALID_ARGUMENT, "no such HIP device");
  s->device = o->device;
  MD_HIP(hipSetDevice(s->device));
  MD_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  int np = s->np;
  s->pos = md_alloc<int64_t>(s, 3 * (size_t)np);
  s->vel = md_alloc<float>(s, 3 * (size_t)np);
  s->resid = md_alloc<float>(s, 3 * (size_t)np);
  s->force = md_alloc<int64_t>(s, 3 * (size_t)np);
  std::vector<float> im(np, 0.0f), mf(np, 0.0f);
  for (int i = 0; i < n; ++i) {
    mf[i] = (float)s->masses[i];
    im[i] = s->masses[i] == 0.0 ? 0.0f : (float)(1.0 / s->masses[i]);
  }
  s->inv_mass = upload(s, im);
  s->mass_f = upload(s, mf);
  s->atom_param = upload(s, ap);
  s->energy = md_alloc<int64_t>(s, MD_E_SLOTS);
  s->ctl = md_alloc<md_control>(s, 1);
  s->term_atoms = upload(s, ta);
  s->term_param = upload(s, tp);
  {
    // the atom of each force piece, in the order k_terms numbers them
    std::vector<int> owner;
    const int kinds[5] = {s->n_bonds, s->n_angles, s->n_torsions, s->n_exceptions, s->n_restraints};
    const int per[5] = {2, 3, 4, 2, 1};
    size_t t = 0;
    for (int k = 0; k < 5; ++k)
      for (int i = 0; i < kinds[k]; ++i, ++t) {
        const int at[4] = {ta[t].x, ta[t].y, ta[t].z, ta[t].w};
        for (int j = 0; j < per[k]; ++j) owner.push_back(at[j]);
      }
    std::vector<int> start(np + 1, 0), list(owner.size());
    for (int a : owner) ++start[a + 1];
    for (int a = 0; a < np; ++a) start[a + 1] += start[a];
    std::vector<int> fill(start.begin(), start.end() - 1);
    for (size_t q = 0; q < owner.size(); ++q) list[fill[owner[q]]++] = (int)q;
    s->n_pieces = (int)owner.size();
    s->term_piece = md_alloc<int64_t>(s, std::max<size_t>(1, 3 * owner.size()));
    s->atom_piece_start = upload(s, start);
    s->atom_piece = upload(s, list);
  }
  s->restraint_ref = upload(s, rref);
  s->water_atoms = upload(s, R.waters);
  s->water_dist = upload(s, R.water_d);
  s->ccl_atoms = upload(s, ccl_atoms);
  s->ccl_natoms = upload(s, ccl_na);
  s->ccl_ncons = upload(s, ccl_nc);
  s->ccl_cons = upload(s, ccl_cons);
  s->free_atoms = upload(s, R.free_atoms);
  MD_HIP(hipStreamSynchronize(s->stream));
  if (s->has_nonbonded) {
    md_nb_setup(s);
    md_pme_setup(s);
  }
  MD_HIP(hipStreamSynchronize(s->stream));
  s->self = group_create({s}, s->stream);
}

static void release(md_system *s) {
  if (!s) return;
  hipSetDevice(s->device);
  if (s->stream) hipStreamSynchronize(s->stream);
  while (!s->groups.empty()) group_release(s->groups.back());
  if (s->nb_ext) md_nb_free(s);
  if (s->pme_ext) md_pme_free(s);
  delete s->min_work;  // its device memory is among the system's allocations
  md_baro_release(s);
  for (void *p : s->allocations) hipFree(p);
  if (s->stream) hipStreamDestroy(s->stream);
  delete s;
}

// ---- steps and evaluations -------------------------------------------------

void md_mark(md_group *g, int phase) {
  if (!g->profiling) return;
  hipEvent_t e;
  MD_HIP(hipEventCreate(&e));
  MD_HIP(hipEventRecord(e, g->stream));
  g->prof_events.push_back(e);
  g->prof_phase.push_back(phase);
}

// one step of every replica in the group
static void launch_step(md_group *g) {
  const bool nb = g->nb != nullptr;
  md_mark(g, MD_PHASE_PROLOGUE);
  md_step_prologue(g, true);
  if (nb) {
    md_mark(g, MD_PHASE_BUILD);
    md_nb_build(g);
    md_mark(g, MD_PHASE_PACK);
    md_nb_pairs(g, false);
  }
  md_mark(g, MD_PHASE_BONDED);
  md_bonded(g, false);
  if (nb) {
    md_mark(g, MD_PHASE_SPREAD);
    md_pme(g, false);
  }
  md_mark(g, MD_PHASE_INTEGRATE);
  md_integrate(g);
  md_mark(g, -1);
}

static void check_control(md_system *s) {
  md_control c;
  MD_HIP(hipMemcpy(&c, s->ctl, sizeof c, hipMemcpyDeviceToHost));
  const std::string when = " (at or before step " + std::to_string(c.step_current) + ")";
  if (c.overflow == 2)
    throw md_error(MD_ERR_CUTOFF_TOO_LARGE,
                   "the list radius plus the cluster extents reaches past four periodic images along an axis: "
                   "the box is too small for this cutoff" + when);
  if (c.overflow == 3)
    throw md_error(MD_ERR_FIXED_POINT_OVERFLOW,
                   "an atom moved more than 32 nm from its cluster's frame: the simulation has diverged" + when);
  if (c.overflow)
    throw md_error(MD_ERR_LIST_OVERFLOW,
                   "the cluster-pair list outgrew its capacity (the system compressed far beyond its starting "
                   "density); set positions at a lower density or recreate the system from the current state" +
                       when);
  if (c.barostat_error == 1)
    throw md_error(MD_ERR_CUTOFF_TOO_LARGE,
                   "a barostat move shrank the box below twice the cutoff along its narrowest width" + when);
  if (c.barostat_error)
    throw md_error(MD_ERR_FIXED_POINT_OVERFLOW,
                   "a barostat move put an atom outside the fixed-point range of +/-2^31 nm" + when);
  if (c.diverged)
    throw md_error(MD_ERR_CONSTRAINTS_DIVERGED,
                   "the constraints diverged" + when +
                       ((c.diverged & 1) ? ": SHAKE did not converge within its passes" : "") +
                       ((c.diverged & 2) ? ": RATTLE did not converge within its passes" : "") +
                       ((c.diverged & 4) ? ": a velocity became non-finite" : "") +
                       "; the time step is too long for this system, or its starting geometry is strained");
}

// host state of each replica onto the device before its first step: the
// step counter, a pending rebuild, the momentum of new velocities
static void group_prepare(md_group *g) {
  bool momentum = false;
  for (md_system *s : g->reps) {
    MD_HIP(hipMemcpyAsync(&s->ctl->step_next, &s->steps_done, sizeof(int64_t), hipMemcpyHostToDevice, g->stream));
    if (s->need_rebuild) {
      static const uint32_t one = 1;
      MD_HIP(hipMemcpyAsync(&s->ctl->force_pending, &one, sizeof one, hipMemcpyHostToDevice, g->stream));
      s->need_rebuild = false;
    }
    momentum |= s->momentum_dirty;
    s->momentum_dirty = false;
  }
  if (momentum) md_momentum(g);
}

static void fault_tick(md_system *s, long count) {
  if (s->fault_after < 0) return;
  if (s->fault_after < count) {
    s->fault_after = 0;
    throw md_error(MD_ERR_FAULT_INJECTED, "fault injected by md_inject_fault");
  }
  s->fault_after -= count;
}

static void drop_graphs(md_group *g) {
  for (auto &c : g->graphs) {
    if (c.exec) MD_HIP(hipGraphExecDestroy(c.exec));
    if (c.graph) MD_HIP(hipGraphDestroy(c.graph));
    c = md_group::captured{};
  }
}

// the captured graph of `steps` steps, followed by a barostat attempt when
// `attempt` (the group is then one system carrying a barostat): slots 0 and 2
// hold whole strides, slots 1 and 3 the last shorter run (captured again when
// a run of another length comes)
static hipGraphExec_t stride_graph(md_group *g, int steps, int stride, bool attempt) {
  md_group::captured &c = g->graphs[(steps == stride ? 0 : 1) + (attempt ? 2 : 0)];
  if (c.exec && c.steps == steps) return c.exec;
  if (c.exec) MD_HIP(hipGraphExecDestroy(c.exec));
  if (c.graph) MD_HIP(hipGraphDestroy(c.graph));
  c = md_group::captured{};
  MD_HIP(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
  try {
    for (int k = 0; k < steps; ++k) launch_step(g);
    if (attempt) md_baro_attempt(g->reps[0]);
  } catch (...) {
    hipGraph_t broken = nullptr;
    (void)hipStreamEndCapture(g->stream, &broken);
    if (broken) (void)hipGraphDestroy(broken);
    throw;
  }
  MD_HIP(hipStreamEndCapture(g->stream, &c.graph));
  MD_HIP(hipGraphInstantiate(&c.exec, c.graph, nullptr, nullptr, 0));
  c.steps = steps;
  c.attempt = attempt;
  return c.exec;
}

// Advance every replica by `steps`, one stride (a captured graph) at a time.
// After each stride the group's status word -- set by a step's prologue when
// any replica's control block holds an error -- is read, so a run whose list
// overflows or whose constraints diverge stops within one stride of it.  A
// run that ends in an error leaves every replica's step count unchanged.
//
// A system carrying a device barostat (always a group of one) cuts its chunks
// at the attempts: every `frequency` steps counted across runs, as upstream's
// runBatch (src/gpu/loop.ts:662-685), the attempt is appended to the chunk's
// graph.  The schedule depends on step counts only, so the host needs nothing
// from the device to lay it out.
static void run_group(md_group *g, int64_t steps) {
  if (steps <= 0) return;
  MD_HIP(hipSetDevice(g->device));
  group_prepare(g);
  md_zero_forces(g, false);
  const int stride = g->max_batch > 0 ? g->max_batch : STRIDE_STEPS;
  if (g->graph_stale) drop_graphs(g);
  g->graph_stale = false;
  MD_HIP(hipMemsetAsync(g->status, 0, sizeof(uint32_t), g->stream));
  md_system *const baro = g->reps.size() == 1 && md_baro_frequency(g->reps[0]) > 0 ? g->reps[0] : nullptr;
  const int frequency = baro ? md_baro_frequency(baro) : 0;
  const int baro_step0 = baro ? md_baro_step(baro) : 0;
  try {
    int64_t done = 0;
    while (done < steps) {
      int chunk = (int)std::min<int64_t>(stride, steps - done);
      bool attempt = false;
      if (baro) {
        int &since = md_baro_step(baro);
        chunk = std::min(chunk, std::max(1, frequency - since));
        since += chunk;
        if (since >= frequency) {
          since -= frequency;
          attempt = true;
        }
      }
      if (chunk > 1) {
        MD_HIP(hipGraphLaunch(stride_graph(g, chunk, stride, attempt), g->stream));
      } else {
        launch_step(g);
        if (attempt) md_baro_attempt(baro);
      }
      uint32_t status = 0;
      MD_HIP(hipMemcpyAsync(&status, g->status, sizeof status, hipMemcpyDeviceToHost, g->stream));
      MD_HIP(hipStreamSynchronize(g->stream));
      if (status) {
        for (md_system *s : g->reps) check_control(s);
        throw md_error(MD_ERR_HIP, "a replica reported an error that its control block no longer holds");
      }
      done += chunk;
    }
    // an error raised in the last step is seen by no later prologue
    for (md_system *s : g->reps) check_control(s);
  } catch (...) {
    if (baro) {
      md_baro_step(baro) = baro_step0;
      md_baro_sync_box(baro);
    }
    throw;
  }
  if (baro) md_baro_sync_box(baro);
  for (md_system *s : g->reps) s->steps_done += steps;
}

static void run_steps(md_system *s, int64_t steps) {
  if (steps <= 0) return;
  fault_tick(s, (long)steps);
  run_group(s->self, steps);
}

static void evaluate(md_system *s, double *forces, double *groups, double *total, bool census) {
  MD_HIP(hipSetDevice(s->device));
  fault_tick(s, 1);
  md_group *grp = s->self;
  group_prepare(grp);
  md_zero_forces(grp, true);
  if (census) {
    MD_HIP(hipMemsetAsync(&s->ctl->pair_census, 0, sizeof(int64_t), s->stream));
    MD_HIP(hipMemsetAsync(&s->ctl->pair_blocks, 0, sizeof(int64_t), s->stream));
  }
  md_step_prologue(grp, false);
  if (s->has_nonbonded) {
    md_nb_build(grp);
    md_nb_pairs(grp, true);
  }
  md_bonded(grp, true);
  if (s->has_nonbonded) md_pme(grp, true);
  int64_t e[MD_E_SLOTS];
  MD_HIP(hipMemcpyAsync(e, s->energy, sizeof e, hipMemcpyDeviceToHost, s->stream));
  std::vector<int64_t> f;
  if (forces) {
    f.resize(3 * (size_t)s->np);
    MD_HIP(hipMemcpyAsync(f.data(), s->force, sizeof(int64_t) * f.size(), hipMemcpyDeviceToHost, s->stream));
  }
  MD_HIP(hipStreamSynchronize(s->stream));
  check_control(s);
  double g[MD_GROUP_COUNT];
  const double t = md_energy_groups(e, md_energy_constants(s), box_volume(s->box), g);
  if (groups) std::memcpy(groups, g, sizeof g);
  if (total) *total = t;
  if (forces)
    for (int i = 0; i < s->n; ++i)
      for (int d = 0; d < 3; ++d) forces[3 * i + d] = md_fixed_to_double(f[(size_t)d * s->np + i]);
}

// ---- the device-resident minimiser --------------------------------------------
//
// minimiseOnDevice (src/gpu/minimise.ts:33-54) with its search on the device:
// OpenMM's LocalEnergyMinimizer objective U(x) + sum (k/2)(r - d)^2 over the
// constraints, k = 100 / max(1e-4, tol), stiffened tenfold per round until
// every constraint is within max(1e-4, tol), giving up to the starting
// positions when stiffening stops helping, then projected to the run's own
// tolerance (src/reference/minimise.ts:164-280).  The search is L-BFGS with
// libLBFGS's backtracking strong-Wolfe line search as the reference restates
// it (src/reference/lbfgs.ts:102-275): memory 6, first step 1/|d|, then 1,
// x0.5 on failed decrease or overshoot, x2.1 while still steep, Armijo 1e-4,
// curvature 0.9, 40 trials, stop at RMS gradient per particle <= tolerance;
// plus a 0.5 nm cap on any coordinate's move per trial.  All 3N-vectors stay
// on the device; the host reads back energies, dot products and maxima.

namespace {
struct lbfgs_state {
  md_system *s;
  md_min_work *w;
  double kspring;
  double *x, *g, *d, *trial, *gt, *q;
  double *S[MD_MIN_MEMORY], *Y[MD_MIN_MEMORY];
};
}  // namespace

static double min_objective(lbfgs_state &L, const double *x, double *g) {
  md_system *s = L.s;
  if (!md_min_positions_in(s, L.w, x))
    throw md_error(MD_ERR_FIXED_POINT_OVERFLOW, "the minimiser moved a coordinate outside the fixed-point range");
  MD_HIP(hipMemsetAsync(&s->ctl->overflow, 0, sizeof(uint32_t), s->stream));
  s->need_rebuild = true;
  double total = 0.0;
  evaluate(s, nullptr, nullptr, &total, false);
  return total + md_min_gradient(s, L.w, x, L.kspring, g);
}

// returns the iterations taken; L.x holds the last accepted point
static int min_lbfgs(lbfgs_state &L, int max_iterations, double rms_tolerance, double &value) {
  md_system *s = L.s;
  const double max_move = 0.5, armijo = 1e-4, curvature = 0.9;
  const int max_line_search = 40;
  auto done = [&](const double *grad) { return std::sqrt(md_min_dot(s, L.w, grad, grad) / s->n) <= rms_tolerance; };
  value = min_objective(L, L.x, L.g);
  if (done(L.g)) return 0;
  double rho[MD_MIN_MEMORY];
  int count = 0, head = 0;  // pairs held; the slot the next pair goes to
  md_min_lincomb(s, L.d, -1.0, L.g, 0.0, nullptr);
  double step = 1.0 / std::sqrt(md_min_dot(s, L.w, L.d, L.d));
  int iteration = 0;
  for (;;) {
    ++iteration;
    const double slope0 = md_min_dot(s, L.w, L.g, L.d);
    if (slope0 >= 0.0) return iteration - 1;
    bool accepted = false;
    double v = 0.0;
    const double dmax = md_min_absmax(s, L.w, L.d);
    for (int trial = 0; trial < max_line_search; ++trial) {
      if (step * dmax > max_move) {
        step *= 0.5;
        continue;
      }
      md_min_lincomb(s, L.trial, 1.0, L.x, step, L.d);
      v = min_objective(L, L.trial, L.gt);
      if (!(v <= value + armijo * step * slope0)) {
        step *= 0.5;
        continue;
      }
      const double slope = md_min_dot(s, L.w, L.gt, L.d);
      if (slope < curvature * slope0) {
        if (2.1 * step * dmax > max_move) {
          accepted = true;  // still steep, but at the displacement cap
          break;
        }
        step *= 2.1;
      } else if (slope > -curvature * slope0) {
        step *= 0.5;
      } else {
        accepted = true;
        break;
      }
    }
    if (!accepted) return iteration - 1;
    // the correction pair of this step, into the next slot of the ring
    md_min_lincomb(s, L.S[head], 1.0, L.trial, -1.0, L.x);
    md_min_lincomb(s, L.Y[head], 1.0, L.gt, -1.0, L.g);
    std::swap(L.x, L.trial);
    std::swap(L.g, L.gt);
    value = v;
    if (done(L.g) || (max_iterations > 0 && iteration >= max_iterations)) return iteration;
    const double sy = md_min_dot(s, L.w, L.S[head], L.Y[head]);
    if (sy > 0.0) {
      rho[head] = 1.0 / sy;
      head = (head + 1) % MD_MIN_MEMORY;
      count = std::min(count + 1, MD_MIN_MEMORY);
    }
    // two-loop recursion: q = -H g
    md_min_lincomb(s, L.q, -1.0, L.g, 0.0, nullptr);
    double alpha[MD_MIN_MEMORY];
    for (int k = 0; k < count; ++k) {  // newest to oldest
      const int slot = (head - 1 - k + 2 * MD_MIN_MEMORY) % MD_MIN_MEMORY;
      alpha[k] = rho[slot] * md_min_dot(s, L.w, L.S[slot], L.q);
      md_min_lincomb(s, L.q, 1.0, L.q, -alpha[k], L.Y[slot]);
    }
    if (count) {
      const int newest = (head - 1 + MD_MIN_MEMORY) % MD_MIN_MEMORY;
      md_min_lincomb(s, L.q, 1.0 / (rho[newest] * md_min_dot(s, L.w, L.Y[newest], L.Y[newest])), L.q, 0.0,
                     nullptr);
    }
    for (int k = count - 1; k >= 0; --k) {  // oldest to newest
      const int slot = (head - 1 - k + 2 * MD_MIN_MEMORY) % MD_MIN_MEMORY;
      const double beta = rho[slot] * md_min_dot(s, L.w, L.Y[slot], L.q);
      md_min_lincomb(s, L.q, 1.0, L.q, alpha[k] - beta, L.S[slot]);
    }
    std::swap(L.d, L.q);
    step = 1.0;
  }
}

static void minimise_system(md_system *s, int max_iterations, double tolerance, md_minimise_result *out) {
  MD_HIP(hipSetDevice(s->device));
  if (!s->min_work) s->min_work = md_min_work_create(s, s->con_p1, s->con_p2, s->con_d);
  md_min_work *w = s->min_work;
  lbfgs_state L;
  L.s = s;
  L.w = w;
  L.x = md_min_vec(s, w, 0);
  L.g = md_min_vec(s, w, 1);
  L.d = md_min_vec(s, w, 2);
  L.trial = md_min_vec(s, w, 3);
  L.gt = md_min_vec(s, w, 4);
  L.q = md_min_vec(s, w, 5);
  for (int k = 0; k < MD_MIN_MEMORY; ++k) {
    L.S[k] = md_min_vec(s, w, 6 + k);
    L.Y[k] = md_min_vec(s, w, 6 + MD_MIN_MEMORY + k);
  }
  const double ctol = s->opt.constraint_tolerance, work = std::max(1e-4, ctol);
  L.kspring = 100.0 / work;
  md_constrain(s->self, 0, work);
  MD_HIP(hipStreamSynchronize(s->stream));
  // the starting point, kept in the last correction slot until the search
  // needs it (a give-up restores it): copy it aside in host memory instead,
  // so every device vector stays free for the search
  std::vector<double> initial(3 * (size_t)s->np);
  md_min_positions_out(s, L.x);
  MD_HIP(hipMemcpyAsync(initial.data(), L.x, sizeof(double) * initial.size(), hipMemcpyDeviceToHost, s->stream));
  MD_HIP(hipStreamSynchronize(s->stream));
  auto restore = [&]() {
    MD_HIP(hipMemcpyAsync(L.x, initial.data(), sizeof(double) * initial.size(), hipMemcpyHostToDevice, s->stream));
    MD_HIP(hipStreamSynchronize(s->stream));
  };
  double prev1 = 1e10, prev2 = 1e10, value = 0.0;
  int rounds = 0, iterations = 0;
  bool gave_up = false;
  for (;;) {
    ++rounds;
    iterations += min_lbfgs(L, max_iterations, tolerance, value);
    const double err = md_min_constraint_error(s, w, L.x);
    if (err <= work || w->nc == 0) break;
    if (err >= prev2) {
      restore();
      gave_up = true;
      break;
    }
    prev2 = prev1;
    prev1 = err;
    L.kspring *= 10.0;
    if (err > 100 * work) restore();
  }
  if (!md_min_positions_in(s, w, L.x))
    throw md_error(MD_ERR_FIXED_POINT_OVERFLOW, "the minimiser moved a coordinate outside the fixed-point range");
  MD_HIP(hipMemsetAsync(&s->ctl->overflow, 0, sizeof(uint32_t), s->stream));
  MD_HIP(hipMemsetAsync(&s->ctl->diverged, 0, sizeof(uint32_t), s->stream));
  s->need_rebuild = true;
  if (ctol < work) {
    md_constrain(s->self, 0, ctol);
    MD_HIP(hipStreamSynchronize(s->stream));
  }
  double total = 0.0;
  evaluate(s, nullptr, nullptr, &total, false);
  md_min_positions_out(s, L.x);
  const double err = md_min_constraint_error(s, w, L.x);
  if (out) {
    out->energy = total;
    out->iterations = iterations;
    out->rounds = rounds;
    out->gave_up = gave_up ? 1 : 0;
    out->max_constraint_error = err;
  }
}

// ---- probes ------------------------------------------------------------------

__global__ void k_probe_fixed(const float *in, int n, int64_t *fixed, float *back) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t f = md_to_fixed(in[i]);
  fixed[i] = f;
  back[i] = md_from_fixed(f);
}

__global__ void k_probe_thermostat(uint32_t seed, uint32_t step, int n, uint32_t *bits, float *normals) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint32_t a0, a1, b0, b1;
  md_threefry2x32_20(2u * p, step, seed, MD_THERMOSTAT_DOMAIN, a0, a1);
  md_threefry2x32_20(2u * p + 1u, step, seed, MD_THERMOSTAT_DOMAIN, b0, b1);
  bits[4 * p] = a0;
  bits[4 * p + 1] = a1;
  bits[4 * p + 2] = b0;
  bits[4 * p + 3] = b1;
  const float two_pi = 6.283185307179586f;
  float r1 = sqrtf(-2.0f * logf(md_open_unit(a0))), t1 = two_pi * md_open_unit(a1);
  float r2 = sqrtf(-2.0f * logf(md_open_unit(b0))), t2 = two_pi * md_open_unit(b1);
  normals[3 * p] = r1 * cosf(t1);
  normals[3 * p + 1] = r1 * sinf(t1);
  normals[3 * p + 2] = r2 * cosf(t2);
}

// per-phase device time of `steps` directly launched steps of a group, from
// events recorded between the phases (averaged per step)
static void profile_group(md_group *g, int steps, double *ms) {
  MD_HIP(hipSetDevice(g->device));
  for (int k = 0; k < MD_PHASE_COUNT; ++k) ms[k] = 0.0;
  group_prepare(g);
  md_zero_forces(g, false);
  for (int r = 0; r < steps; ++r) {
    g->profiling = true;
    try {
      launch_step(g);
    } catch (...) {
      g->profiling = false;
      throw;
    }
    g->profiling = false;
    MD_HIP(hipStreamSynchronize(g->stream));
    for (size_t k = 0; k + 1 < g->prof_events.size(); ++k) {
      float t = 0.0f;
      MD_HIP(hipEventElapsedTime(&t, g->prof_events[k], g->prof_events[k + 1]));
      int ph = g->prof_phase[k];
      if (ph >= 0 && ph < MD_PHASE_COUNT) ms[ph] += t;
    }
    for (hipEvent_t e : g->prof_events) (void)hipEventDestroy(e);
    g->prof_events.clear();
    g->prof_phase.clear();
  }
  for (int k = 0; k < MD_PHASE_COUNT; ++k) ms[k] /= steps;
  for (md_system *s : g->reps) s->steps_done += steps;
  for (md_system *s : g->reps) check_control(s);
}

// The autotuner: for each pair-kernel schedule in a small set (8, 16 or 32
// list entries per work unit; 2, 3 or 4 resident waves per SIMD), rebuild
// the lists with it and time `steps` steps launched phase by phase; keep the
// fastest, record it in the cache under the group's key and report it with
// the time of the schedule the group started with.  The schedule never
// changes the bits (i-forces are rounded per 8 entries whatever the unit,
// every sum is order independent), only the time; the timed steps are
// ordinary steps of every replica.
static void autotune_group(md_group *g, int steps, md_schedule *out) {
  if (!g->nb) throw md_error(MD_ERR_INVALID_ARGUMENT, "a system without a NonbondedForce has no pair kernel to tune");
  int seg0 = 0, waves0 = 0;
  md_nb_get_schedule(g, &seg0, &waves0);
  const bool compact0 = md_nb_get_compact(g);
  int cus = 1;
  MD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device));
  struct cand {
    int seg, waves;
    bool compact;
    bool operator==(const cand &o) const { return seg == o.seg && waves == o.waves && compact == o.compact; }
  };
  std::vector<cand> candidates = {{seg0, waves0, compact0}};
  auto add = [&](cand c) {
    if (std::find(candidates.begin(), candidates.end(), c) == candidates.end()) candidates.push_back(c);
  };
  // the direct evaluation over unit lengths and resident waves (a longer
  // unit means fewer i-force flushes, a shorter one finer load balance) ...
  for (int seg : {8, 16, 32, 64, 128})
    for (int per_simd : {2, 3, 4}) add(cand{seg, cus * 4 * per_simd, false});
  // ... and the compacted one at the usual shape
  add(cand{32, cus * 16, true});
  double ms[MD_PHASE_COUNT];
  double best = 0.0, first = 0.0;
  cand pick = candidates[0];
  for (size_t k = 0; k < candidates.size(); ++k) {
    md_nb_set_schedule(g, candidates[k].seg, candidates[k].waves);
    md_nb_set_compact(g, candidates[k].compact);
    g->graph_stale = true;  // the pair grid is baked into a captured graph
    for (md_system *s : g->reps) s->need_rebuild = true;
    profile_group(g, 2, ms);  // the rebuild with the new unit length, and one more step
    profile_group(g, steps, ms);
    double total = 0.0;
    for (double v : ms) total += v;
    if (k == 0) first = total;
    if (k == 0 || total < best) {
      best = total;
      pick = candidates[k];
    }
  }
  md_nb_set_schedule(g, pick.seg, pick.waves);
  md_nb_set_compact(g, pick.compact);
  g->graph_stale = true;
  for (md_system *s : g->reps) s->need_rebuild = true;
  g_tuned[tune_key(g)] = md_tune_entry{pick.seg, pick.waves, pick.compact};
  if (out) {
    out->segment_entries = pick.seg;
    out->pair_waves = pick.waves;
    out->ms_per_step = best;
    out->default_segment_entries = seg0;
    out->default_pair_waves = waves0;
    out->default_ms_per_step = first;
    out->candidates = (int32_t)candidates.size();
    out->compact_pairs = pick.compact ? 1 : 0;
    out->default_compact_pairs = compact0 ? 1 : 0;
  }
}

static void set_pair_schedule(md_group *g, const md_schedule *sc) {
  if (!g->nb) throw md_error(MD_ERR_INVALID_ARGUMENT, "a system without a NonbondedForce has no pair kernel");
  int seg = 0, waves = 0;
  md_nb_get_schedule(g, &seg, &waves);
  md_nb_set_schedule(g, sc->segment_entries > 0 ? sc->segment_entries : seg, sc->pair_waves > 0 ? sc->pair_waves : waves);
  md_nb_set_compact(g, sc->compact_pairs != 0);
  g->graph_stale = true;
  for (md_system *s : g->reps) s->need_rebuild = true;
}

struct md_batch {
  md_group *group = nullptr;
};

// ---- the C ABI -------------------------------------------------------------

// A probe's device staging, freed when the call returns or throws.
template <typename T>
struct probe_buffer {
  T *p = nullptr;
  explicit probe_buffer(size_t n) { MD_HIP(hipMalloc(&p, sizeof(T) * n)); }
  ~probe_buffer() {
    if (p) (void)hipFree(p);
  }
  probe_buffer(const probe_buffer &) = delete;
  probe_buffer &operator=(const probe_buffer &) = delete;
};

extern "C" {

const char *md_last_error(void) { return g_last_error.c_str(); }

const char *md_version(void) { return "mdcore 0.1 (gfx950)"; }

md_err md_system_create(const md_system_desc *desc, const md_run_options *options, md_system **out) {
  MD_API_BEGIN
  if (!out) throw md_error(MD_ERR_INVALID_ARGUMENT, "null output handle");
  *out = nullptr;
  md_system *s = new md_system();
  try {
    create(s, desc, options);
  } catch (...) {
    release(s);
    throw;
  }
  *out = s;
  MD_API_END
}

void md_system_destroy(md_system *sys) { release(sys); }

md_err md_set_positions(md_system *s, const double *x) {
  MD_API_BEGIN
  if (!s || !x) throw md_error(MD_ERR_INVALID_ARGUMENT, "null argument");
  MD_HIP(hipSetDevice(s->device));
  std::vector<int64_t> f(3 * (size_t)s->np, 0);
  for (int i = 0; i < s->n; ++i)
    for (int d = 0; d < 3; ++d) f[(size_t)d * s->np + i] = fixed_checked(x[3 * i + d], "position");
  MD_HIP(hipMemcpyAsync(s->pos, f.data(), sizeof(int64_t) * f.size(), hipMemcpyHostToDevice, s->stream));
  // setPositions clears the carried residual (src/gpu/loop.ts:819-833)
  MD_HIP(hipMemsetAsync(s->resid, 0, sizeof(float) * 3 * s->np, s->stream));
  // new positions start a new trajectory: the next step rebuilds the list
  // and raises any overflow again
  MD_HIP(hipMemsetAsync(&s->ctl->overflow, 0, sizeof(uint32_t), s->stream));
  MD_HIP(hipMemsetAsync(&s->ctl->diverged, 0, sizeof(uint32_t), s->stream));
  MD_HIP(hipStreamSynchronize(s->stream));
  s->need_rebuild = true;
  MD_API_END
}

md_err md_set_velocities(md_system *s, const double *v) {
  MD_API_BEGIN
  if (!s || !v) throw md_error(MD_ERR_INVALID_ARGUMENT, "null argument");
  MD_HIP(hipSetDevice(s->device));
  std::vector<float> f(3 * (size_t)s->np, 0.0f);
  for (int i = 0; i < s->n; ++i)
    for (int d = 0; d < 3; ++d) {
      float x = (float)v[3 * i + d];
      if (!std::isfinite(x))
        throw md_error(MD_ERR_INVALID_ARGUMENT, "velocity " + std::to_string(v[3 * i + d]) + " of particle " +
                                                    std::to_string(i) + " is not a finite f32 value");
      f[(size_t)d * s->np + i] = x;
    }
  MD_HIP(hipMemcpy(s->vel, f.data(), sizeof(float) * f.size(), hipMemcpyHostToDevice));
  s->momentum_dirty = true;
  MD_API_END
}

md_err md_get_positions(md_system *s, double *x) {
  MD_API_BEGIN
  if (!s || !x) throw md_error(MD_ERR_INVALID_ARGUMENT, "null argument");
  MD_HIP(hipSetDevice(s->device));
  std::vector<int64_t> f(3 * (size_t)s->np);
  MD_HIP(hipMemcpy(f.data(), s->pos, sizeof(int64_t) * f.size(), hipMemcpyDeviceToHost));
  for (int i = 0; i < s->n; ++i)
    for (int d = 0; d < 3; ++d) x[3 * i + d] = md_fixed_to_double(f[(size_t)d * s->np + i]);
  MD_API_END
}

md_err md_get_velocities(md_system *s, double *v) {
  MD_API_BEGIN
  if (!s || !v) throw md_error(MD_ERR_INVALID_ARGUMENT, "null argument");
  MD_HIP(hipSetDevice(s->device));
  std::vector<float> f(3 * (size_t)s->np);
  MD_HIP(hipMemcpy(f.data(), s->vel, sizeof(float) * f.size(), hipMemcpyDeviceToHost));
  for (int i = 0; i < s->n; ++i)
    for (int d = 0; d < 3; ++d) v[3 * i + d] = f[(size_t)d * s->np + i];
  MD_API_END
}

md_err md_set_box(md_system *s, const double box[9]) {
  MD_API_BEGIN
  if (!s || !box) throw md_error(MD_ERR_INVALID_ARGUMENT, "null argument");
  apply_box(s, box);
  md_baro_set_box(s);
  MD_API_END
}

md_err md_get_box(md_system *s, double box[9]) {
  MD_API_BEGIN
  if (!s || !box) throw md_error(MD_ERR_INVALID_ARGUMENT, "null argument");
  std::memcpy(box, s->box, sizeof s->box);
  MD_API_END
}

md_err md_get_mesh(md_system *s, int32_t mesh[3]) {
  MD_API_BEGIN
  if (!s || !mesh) throw md_error(MD_ERR_INVALID_ARGUMENT, "null argument");
  for (int a = 0; a < 3; ++a) mesh[a] = s->has_nonbonded ? s->grid[a] : 0;
  MD_API_END
}

md_err md_evaluate(md_system *s, double *forces, double *groups, double *total) {
  MD_API_BEGIN
  if (!s) throw md_error(MD_ERR_INVALID_ARGUMENT, "null handle");
  evaluate(s, forces, groups, total, false);
  MD_API_END
}

md_err md_run(md_system *s, int64_t steps) {
  MD_API_BEGIN
  if (!s || steps < 0) throw md_error(MD_ERR_INVALID_ARGUMENT, "bad arguments to md_run");
  run_steps(s, steps);
  MD_API_END
}

md_err md_minimise(md_system *s, int32_t max_iterations, double tolerance, md_minimise_result *result) {
  MD_API_BEGIN
  if (!s || max_iterations < 0 || !(tolerance > 0)) throw md_error(MD_ERR_INVALID_ARGUMENT, "bad arguments to md_minimise");
  minimise_system(s, max_iterations, tolerance, result);
  MD_API_END
}

md_err md_constrain_positions(md_system *s, double tolerance) {
  MD_API_BEGIN
  if (!s || !(tolerance > 0)) throw md_error(MD_ERR_INVALID_ARGUMENT, "bad arguments to md_constrain_positions");
  MD_HIP(hipSetDevice(s->device));
  md_constrain(s->self, 0, tolerance);
  MD_HIP(hipStreamSynchronize(s->stream));
  s->need_rebuild = true;
  MD_API_END
}

md_err md_constrain_velocities(md_system *s, double tolerance) {
  MD_API_BEGIN
  if (!s || !(tolerance > 0)) throw md_error(MD_ERR_INVALID_ARGUMENT, "bad arguments to md_constrain_velocities");
  MD_HIP(hipSetDevice(s->device));
  md_constrain(s->self, 1, tolerance);
  MD_HIP(hipStreamSynchronize(s->stream));
  s->momentum_dirty = true;
  MD_API_END
}

md_err md_barostat_attach(md_system *s, const md_barostat_desc *desc) {
  MD_API_BEGIN
  if (!s || !desc) throw md_error(MD_ERR_INVALID_ARGUMENT, "null argument");
  md_baro_attach(s, desc);
  MD_API_END
}

md_err md_barostat_get(md_system *s, md_barostat_state *state) {
  MD_API_BEGIN
  if (!s || !state) throw md_error(MD_ERR_INVALID_ARGUMENT, "null argument");
  MD_HIP(hipSetDevice(s->device));
  md_baro_state(s, state);
  MD_API_END
}

md_err md_get_step(md_system *s, int64_t *step) {
  MD_API_BEGIN
  if (!s || !step) throw md_error(MD_ERR_INVALID_ARGUMENT, "null argument");
  *step = s->steps_done;
  MD_API_END
}

md_err md_set_step(md_system *s, int64_t step) {
  MD_API_BEGIN
  if (!s || step < 0) throw md_error(MD_ERR_INVALID_ARGUMENT, "bad step");
  s->steps_done = step;
  s->momentum_dirty = true;
  MD_API_END
}

md_err md_list_flags_get(md_system *s, md_list_flags *flags) {
  MD_API_BEGIN
  if (!s || !flags) throw md_error(MD_ERR_INVALID_ARGUMENT, "null argument");
  std::memset(flags, 0, sizeof *flags);
  if (!s->has_nonbonded) return MD_OK;
  md_control c;
  MD_HIP(hipMemcpy(&c, s->ctl, sizeof c, hipMemcpyDeviceToHost));
  flags->rebuilds = c.rebuilds;
  flags->tiles = md_nb_entries(s);
  flags->tile_capacity = md_nb_capacity(s);
  float w;
  std::memcpy(&w, &c.max_disp2_bits[0], 4);
  flags->worst_displacement = std::sqrt((double)w);
  flags->overflowed = (int)c.overflow;
  MD_API_END
}

const char *md_phase_name(int32_t phase) {
  static const char *names[MD_PHASE_COUNT] = {"prologue",   "list_build", "pack",        "pairs",
                                              "bonded",     "pme_spread", "pme_fft_convolve", "pme_gather",
                                              "integrate",  "list_search", "list_masks", "pair_forces"};
  return phase >= 0 && phase < MD_PHASE_COUNT ? names[phase] : "";
}

md_err md_profile_steps(md_system *s, int32_t steps, double *ms) {
  MD_API_BEGIN
  if (!s || !ms || steps <= 0) throw md_error(MD_ERR_INVALID_ARGUMENT, "bad arguments to md_profile_steps");
  profile_group(s->self, steps, ms);
  MD_API_END
}

md_err md_pair_census(md_system *s, int64_t *pairs, int64_t *tiles) {
  MD_API_BEGIN
  if (!s) throw md_error(MD_ERR_INVALID_ARGUMENT, "null handle");
  evaluate(s, nullptr, nullptr, nullptr, true);
  md_control c;
  MD_HIP(hipMemcpy(&c, s->ctl, sizeof c, hipMemcpyDeviceToHost));
  if (pairs) *pairs = c.pair_census;
  if (tiles) *tiles = c.pair_blocks;
  MD_API_END
}

md_err md_probe_fixed_point(const float *values, int32_t n, int64_t *fixed, float *back) {
  MD_API_BEGIN
  if (n < 0 || (n > 0 && (!values || !fixed || !back))) throw md_error(MD_ERR_INVALID_ARGUMENT, "bad probe arguments");
  if (n == 0) return MD_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw md_error(MD_ERR_NO_DEVICE, "no HIP device");
  probe_buffer<float> din(n);
  probe_buffer<int64_t> dfix(n);
  probe_buffer<float> dback(n);
  MD_HIP(hipMemcpy(din.p, values, sizeof(float) * n, hipMemcpyHostToDevice));
  k_probe_fixed<<<(n + 255) / 256, 256>>>(din.p, n, dfix.p, dback.p);
  MD_HIP(hipGetLastError());
  MD_HIP(hipMemcpy(fixed, dfix.p, sizeof(int64_t) * n, hipMemcpyDeviceToHost));
  MD_HIP(hipMemcpy(back, dback.p, sizeof(float) * n, hipMemcpyDeviceToHost));
  MD_API_END
}

md_err md_probe_thermostat(uint32_t seed, uint32_t step, int32_t n, uint32_t *bits, float *normals) {
  MD_API_BEGIN
  if (n < 0 || (n > 0 && (!bits || !normals))) throw md_error(MD_ERR_INVALID_ARGUMENT, "bad probe arguments");
  if (n == 0) return MD_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw md_error(MD_ERR_NO_DEVICE, "no HIP device");
  probe_buffer<uint32_t> db(4 * (size_t)n);
  probe_buffer<float> dn(3 * (size_t)n);
  k_probe_thermostat<<<(n + 255) / 256, 256>>>(seed, step, n, db.p, dn.p);
  MD_HIP(hipGetLastError());
  MD_HIP(hipMemcpy(bits, db.p, sizeof(uint32_t) * 4 * n, hipMemcpyDeviceToHost));
  MD_HIP(hipMemcpy(normals, dn.p, sizeof(float) * 3 * n, hipMemcpyDeviceToHost));
  MD_API_END
}

md_err md_probe_cell_order(const int32_t dims[3], int32_t *cells) {
  MD_API_BEGIN
  if (!dims || !cells || dims[0] < 1 || dims[1] < 1 || dims[2] < 1 ||
      (int64_t)dims[0] * dims[1] * dims[2] > (1 << 26))
    throw md_error(MD_ERR_INVALID_ARGUMENT, "bad cell grid");
  std::vector<int> walk = md_cell_walk(dims[0], dims[1], dims[2]);
  std::copy(walk.begin(), walk.end(), cells);
  MD_API_END
}

md_err md_probe_ewald_fit(double alpha, double cutoff, float *scale, float *force, float *energy, int32_t *ok,
                          double *force_err, double *energy_err) {
  MD_API_BEGIN
  if (!scale || !force || !energy || !ok || !force_err || !energy_err || !(alpha > 0.0) || !(cutoff > 0.0) ||
      !std::isfinite(alpha) || !std::isfinite(cutoff))
    throw md_error(MD_ERR_INVALID_ARGUMENT, "md_probe_ewald_fit: alpha and cutoff must be positive");
  const md_ewald_poly P = md_ewald_fit(alpha, cutoff);
  *scale = P.scale;
  std::copy(P.force, P.force + MD_EWALD_TERMS, force);
  std::copy(P.energy, P.energy + MD_EWALD_TERMS, energy);
  *ok = P.ok;
  *force_err = P.force_err;
  *energy_err = P.energy_err;
  MD_API_END
}

md_err md_inject_fault(md_system *s, int32_t after) {
  MD_API_BEGIN
  if (!s) throw md_error(MD_ERR_INVALID_ARGUMENT, "null handle");
  s->fault_after = after;
  MD_API_END
}

md_err md_batch_create(md_system *const *systems, int32_t count, md_batch **out) {
  MD_API_BEGIN
  if (!out || !systems || count < 1) throw md_error(MD_ERR_INVALID_ARGUMENT, "bad arguments to md_batch_create");
  *out = nullptr;
  std::vector<md_system *> reps(systems, systems + count);
  for (int r = 0; r < count; ++r) {
    if (!reps[r]) throw md_error(MD_ERR_INVALID_ARGUMENT, "null system in a batch");
    if (reps[r]->device != reps[0]->device)
      throw md_error(MD_ERR_INVALID_ARGUMENT, "the systems of a batch must share one device");
    for (int q = 0; q < r; ++q)
      if (reps[q] == reps[r]) throw md_error(MD_ERR_INVALID_ARGUMENT, "a system appears twice in a batch");
    if (md_baro_frequency(reps[r]) > 0)
      throw md_error(MD_ERR_INVALID_ARGUMENT, "a system carrying a barostat cannot join a batch");
  }
  md_batch *b = new md_batch();
  try {
    b->group = group_create(reps, nullptr);
    b->group->owner = &b->group;
  } catch (...) {
    delete b;
    throw;
  }
  *out = b;
  MD_API_END
}

void md_batch_destroy(md_batch *b) {
  if (!b) return;
  if (b->group) group_release(b->group);
  delete b;
}

md_err md_batch_run(md_batch *b, int64_t steps) {
  MD_API_BEGIN
  if (!b || steps < 0) throw md_error(MD_ERR_INVALID_ARGUMENT, "bad arguments to md_batch_run");
  if (!b->group) throw md_error(MD_ERR_INVALID_ARGUMENT, "a system of this batch has been destroyed");
  run_group(b->group, steps);
  MD_API_END
}

md_err md_batch_autotune(md_batch *b, int32_t steps, md_schedule *best) {
  MD_API_BEGIN
  if (!b || steps <= 0) throw md_error(MD_ERR_INVALID_ARGUMENT, "bad arguments to md_batch_autotune");
  if (!b->group) throw md_error(MD_ERR_INVALID_ARGUMENT, "a system of this batch has been destroyed");
  autotune_group(b->group, steps, best);
  MD_API_END
}

md_err md_set_pair_schedule(md_system *s, const md_schedule *schedule) {
  MD_API_BEGIN
  if (!s || !schedule) throw md_error(MD_ERR_INVALID_ARGUMENT, "null argument");
  MD_HIP(hipSetDevice(s->device));
  set_pair_schedule(s->self, schedule);
  MD_API_END
}

md_err md_batch_set_pair_schedule(md_batch *b, const md_schedule *schedule) {
  MD_API_BEGIN
  if (!b || !schedule) throw md_error(MD_ERR_INVALID_ARGUMENT, "null argument");
  if (!b->group) throw md_error(MD_ERR_INVALID_ARGUMENT, "a system of this batch has been destroyed");
  MD_HIP(hipSetDevice(b->group->device));
  set_pair_schedule(b->group, schedule);
  MD_API_END
}

md_err md_autotune(md_system *s, int32_t steps, md_schedule *best) {
  MD_API_BEGIN
  if (!s || steps <= 0) throw md_error(MD_ERR_INVALID_ARGUMENT, "bad arguments to md_autotune");
  autotune_group(s->self, steps, best);
  MD_API_END
}

md_err md_batch_profile_steps(md_batch *b, int32_t steps, double *ms) {
  MD_API_BEGIN
  if (!b || !ms || steps <= 0) throw md_error(MD_ERR_INVALID_ARGUMENT, "bad arguments to md_batch_profile_steps");
  if (!b->group) throw md_error(MD_ERR_INVALID_ARGUMENT, "a system of this batch has been destroyed");
  profile_group(b->group, steps, ms);
  MD_API_END
}

// ---- learned SHAKE warm start ----------------------------------------------

md_err md_warmstart_create(int32_t device, uint32_t seed, int64_t sample_capacity, md_warmstart **out) {
  MD_API_BEGIN
  if (!out || sample_capacity < 0) throw md_error(MD_ERR_INVALID_ARGUMENT, "bad arguments to md_warmstart_create");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw md_error(MD_ERR_NO_DEVICE, "no HIP device");
  if (device < 0 || device >= ndev) throw md_error(MD_ERR_INVALID_ARGUMENT, "no such HIP device");
  *out = md_ws_create(device, seed, sample_capacity);
  MD_API_END
}

void md_warmstart_destroy(md_warmstart *ws) { md_ws_destroy(ws); }

int32_t md_warmstart_param_count(void) { return MD_WS_NPARAMS; }

static void refresh_all(md_system *s) {
  for (md_group *g : s->groups) group_refresh(g);
}

md_err md_warmstart_record(md_warmstart *ws, md_system *s, int64_t steps) {
  MD_API_BEGIN
  if (!ws || !s || steps < 0) throw md_error(MD_ERR_INVALID_ARGUMENT, "bad arguments to md_warmstart_record");
  if (md_ws_device(ws) != s->device) throw md_error(MD_ERR_INVALID_ARGUMENT, "the network and the system are on different devices");
  if (s->n_clusters_c == 0 || steps == 0) return MD_OK;
  float *xs, *ts;
  int64_t room;
  md_ws_reserve(ws, steps * s->n_clusters_c, &xs, &ts, &room);
  int64_t fit = std::min<int64_t>(steps, room / s->n_clusters_c);
  if (fit <= 0) return MD_OK;
  s->rec_x = xs;
  s->rec_t = ts;
  s->rec_first_step = s->steps_done;
  s->rec_cap = fit * s->n_clusters_c;
  // slots the run leaves unwritten (cluster-steps SHAKE did not converge on)
  // read as clusters with no constraint: no loss, no gradient
  MD_HIP(hipSetDevice(s->device));
  MD_HIP(hipMemsetAsync(xs, 0, sizeof(float) * MD_WS_F * s->rec_cap, s->stream));
  MD_HIP(hipMemsetAsync(ts, 0, sizeof(float) * MD_WS_MAXC * s->rec_cap, s->stream));
  try {
    refresh_all(s);
    run_steps(s, fit);
  } catch (...) {
    s->rec_x = s->rec_t = nullptr;
    s->rec_cap = 0;
    refresh_all(s);
    throw;
  }
  md_ws_commit(ws, fit * s->n_clusters_c);
  s->rec_x = s->rec_t = nullptr;
  s->rec_cap = 0;
  refresh_all(s);
  MD_API_END
}

md_err md_warmstart_sample_count(md_warmstart *ws, int64_t *count) {
  MD_API_BEGIN
  if (!ws || !count) throw md_error(MD_ERR_INVALID_ARGUMENT, "null argument");
  *count = md_ws_count(ws);
  MD_API_END
}

md_err md_warmstart_train(md_warmstart *ws, int32_t iterations, int32_t batch, double learning_rate, double *loss) {
  MD_API_BEGIN
  if (!ws || iterations < 0 || batch <= 0 || !(learning_rate > 0))
    throw md_error(MD_ERR_INVALID_ARGUMENT, "bad arguments to md_warmstart_train");
  double l = iterations > 0 ? md_ws_train(ws, iterations, batch, learning_rate) : 0.0;
  if (loss) *loss = l;
  MD_API_END
}

md_err md_warmstart_attach(md_system *s, md_warmstart *ws) {
  MD_API_BEGIN
  if (!s) throw md_error(MD_ERR_INVALID_ARGUMENT, "null system");
  if (ws && md_ws_device(ws) != s->device)
    throw md_error(MD_ERR_INVALID_ARGUMENT, "the network and the system are on different devices");
  MD_HIP(hipSetDevice(s->device));
  s->ws_net = ws ? md_ws_device_net(ws) : nullptr;
  refresh_all(s);
  MD_API_END
}

md_err md_warmstart_params(md_warmstart *ws, float *get, const float *set) {
  MD_API_BEGIN
  if (!ws || (!get && !set)) throw md_error(MD_ERR_INVALID_ARGUMENT, "bad arguments to md_warmstart_params");
  if (set) md_ws_set_params(ws, set);
  if (get) md_ws_get_params(ws, get);
  MD_API_END
}

md_err md_warmstart_evaluate(md_warmstart *ws, const float *x, const float *t, int32_t n, float *y, float *grad,
                             double *loss) {
  MD_API_BEGIN
  if (!ws || n < 0 || (n > 0 && (!x || !t))) throw md_error(MD_ERR_INVALID_ARGUMENT, "bad arguments to md_warmstart_evaluate");
  double l = md_ws_evaluate(ws, x, t, n, y, grad);
  if (loss) *loss = l;
  MD_API_END
}

md_err md_shake_counts(md_system *s, int64_t *passes, int64_t *clusters) {
  MD_API_BEGIN
  if (!s) throw md_error(MD_ERR_INVALID_ARGUMENT, "null system");
  md_control c;
  MD_HIP(hipSetDevice(s->device));
  MD_HIP(hipMemcpy(&c, s->ctl, sizeof c, hipMemcpyDeviceToHost));
  if (passes) *passes = c.shake_passes;
  if (clusters) *clusters = c.shake_clusters;
  MD_API_END
}

}  // extern "C"
