/*
 * mdcore_node.c -- the N-API binding of libmdcore.so for the upstream
 * TypeScript host (src/engine, Node).  One JavaScript function per C entry
 * point of include/mdcore.h, with typed arrays passed straight through as
 * the const double * / const int32_t * arguments (the library copies them).
 *
 * Errors: a non-zero md_err becomes a thrown (or rejected) JavaScript Error
 * whose `name` is the upstream error class it stands for (UnsupportedForce,
 * CutoffTooLargeForBox, ClusterTooLarge, NotSmooth, ...) and whose `code` is
 * the md_err value, so host code written against the WebGPU engine's errors
 * keeps working.
 *
 * Threading: run() executes md_run on libuv's thread pool (napi_async_work)
 * and returns a Promise, as StepLoop.run does upstream (src/gpu/loop.ts
 * :651-685); every other call blocks.  Calls on one handle must not overlap.
 *
 * Build (no node-gyp needed): cc -shared -fPIC -I/usr/include/node
 *   -I../../include mdcore_node.c -L../mdcore -lmdcore
 *   -Wl,-rpath,'$ORIGIN/../mdcore' -o mdcore.node
 */
#include <node_api.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mdcore.h"

static const char *error_name(md_err code) {
  switch (code) {
    case MD_ERR_UNSUPPORTED_FORCE: return "UnsupportedForce";
    case MD_ERR_CUTOFF_TOO_LARGE: return "CutoffTooLargeForBox";
    case MD_ERR_CLUSTER_TOO_LARGE: return "ClusterTooLarge";
    case MD_ERR_FIXED_POINT_OVERFLOW: return "RangeError";
    case MD_ERR_HIP: return "HipError";
    case MD_ERR_OUT_OF_MEMORY: return "OutOfMemory";
    case MD_ERR_INVALID_ARGUMENT: return "InvalidArgument";
    case MD_ERR_LIST_OVERFLOW: return "ListOverflow";
    case MD_ERR_NOT_SMOOTH: return "NotSmooth";
    case MD_ERR_CONSTRAINTS_DIVERGED: return "ConstraintsDiverged";
    case MD_ERR_FAULT_INJECTED: return "FaultInjected";
    case MD_ERR_NO_DEVICE: return "NoDevice";
    default: return "MdError";
  }
}

static napi_value make_error(napi_env env, md_err code, const char *message) {
  napi_value msg, err, name, num;
  napi_create_string_utf8(env, message, NAPI_AUTO_LENGTH, &msg);
  napi_create_error(env, NULL, msg, &err);
  napi_create_string_utf8(env, error_name(code), NAPI_AUTO_LENGTH, &name);
  napi_set_named_property(env, err, "name", name);
  napi_create_int32(env, (int32_t)code, &num);
  napi_set_named_property(env, err, "code", num);
  return err;
}

/* throw for a failed call; returns 1 when it threw */
static int check(napi_env env, md_err code) {
  if (code == MD_OK) return 0;
  napi_throw(env, make_error(env, code, md_last_error()));
  return 1;
}

static int throw_type(napi_env env, const char *message) {
  napi_throw_type_error(env, NULL, message);
  return 1;
}

/* ---- argument helpers -------------------------------------------------------- */

static int get_args(napi_env env, napi_callback_info info, size_t want, napi_value *argv) {
  size_t argc = want;
  napi_get_cb_info(env, info, &argc, argv, NULL, NULL);
  if (argc < want) return throw_type(env, "missing arguments");
  return 0;
}

static int get_handle(napi_env env, napi_value v, md_system **out) {
  void *p = NULL;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) return throw_type(env, "expected a system handle");
  *out = (md_system *)p;
  return 0;
}

/* a typed array's data and element count, checked for its element type */
static int typed(napi_env env, napi_value v, napi_typedarray_type want, void **data, size_t *length) {
  bool is = false;
  napi_is_typedarray(env, v, &is);
  if (!is) return throw_type(env, want == napi_float64_array ? "expected a Float64Array" : "expected an Int32Array");
  napi_typedarray_type type;
  napi_value buffer;
  size_t offset;
  napi_get_typedarray_info(env, v, &type, length, data, &buffer, &offset);
  if (type != want) return throw_type(env, want == napi_float64_array ? "expected a Float64Array" : "expected an Int32Array");
  return 0;
}

static napi_value new_f64(napi_env env, size_t count, double **data) {
  napi_value buffer, array;
  napi_create_arraybuffer(env, count * sizeof(double), (void **)data, &buffer);
  napi_create_typedarray(env, napi_float64_array, count, buffer, 0, &array);
  return array;
}

static napi_value new_i32(napi_env env, size_t count, int32_t **data) {
  napi_value buffer, array;
  napi_create_arraybuffer(env, count * sizeof(int32_t), (void **)data, &buffer);
  napi_create_typedarray(env, napi_int32_array, count, buffer, 0, &array);
  return array;
}

/* desc.<name> as a number (absent: 0) */
static double num_field(napi_env env, napi_value obj, const char *name) {
  bool has = false;
  napi_has_named_property(env, obj, name, &has);
  if (!has) return 0.0;
  napi_value v;
  double x = 0.0;
  napi_get_named_property(env, obj, name, &v);
  napi_get_value_double(env, v, &x);
  return x;
}

/* desc.<name> as a typed array of at least `count` elements (count 0: may be absent) */
static int array_field(napi_env env, napi_value obj, const char *name, napi_typedarray_type type, size_t count,
                       void **out) {
  *out = NULL;
  if (count == 0) return 0;
  bool has = false;
  napi_has_named_property(env, obj, name, &has);
  if (!has) {
    char msg[160];
    snprintf(msg, sizeof msg, "the system description has no %s", name);
    return throw_type(env, msg);
  }
  napi_value v;
  size_t length = 0;
  napi_get_named_property(env, obj, name, &v);
  if (typed(env, v, type, out, &length)) return 1;
  if (length < count) {
    char msg[160];
    snprintf(msg, sizeof msg, "%s holds %zu values, %zu expected", name, length, count);
    return throw_type(env, msg);
  }
  return 0;
}

#define F64(name, count) \
  if (array_field(env, obj, #name, napi_float64_array, (size_t)(count), (void **)&d.name)) return NULL
#define I32(name, count) \
  if (array_field(env, obj, #name, napi_int32_array, (size_t)(count), (void **)&d.name)) return NULL

/* createSystem(desc, options) -> handle.  `desc` carries md_system_desc's
 * fields under the same names (typed arrays for the arrays); `options`
 * md_run_options's. */
static napi_value js_create(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (get_args(env, info, 2, argv)) return NULL;
  napi_value obj = argv[0], opt = argv[1];
  md_system_desc d;
  memset(&d, 0, sizeof d);
  d.n_particles = (int32_t)num_field(env, obj, "n_particles");
  F64(masses, d.n_particles);
  {
    double *box = NULL;
    if (array_field(env, obj, "box", napi_float64_array, 9, (void **)&box)) return NULL;
    memcpy(d.box, box, sizeof d.box);
  }
  d.n_constraints = (int32_t)num_field(env, obj, "n_constraints");
  I32(constraint_p1, d.n_constraints);
  I32(constraint_p2, d.n_constraints);
  F64(constraint_distance, d.n_constraints);
  d.n_bonds = (int32_t)num_field(env, obj, "n_bonds");
  I32(bond_p1, d.n_bonds);
  I32(bond_p2, d.n_bonds);
  F64(bond_length, d.n_bonds);
  F64(bond_k, d.n_bonds);
  d.n_angles = (int32_t)num_field(env, obj, "n_angles");
  I32(angle_p1, d.n_angles);
  I32(angle_p2, d.n_angles);
  I32(angle_p3, d.n_angles);
  F64(angle_theta, d.n_angles);
  F64(angle_k, d.n_angles);
  d.n_torsions = (int32_t)num_field(env, obj, "n_torsions");
  I32(torsion_p1, d.n_torsions);
  I32(torsion_p2, d.n_torsions);
  I32(torsion_p3, d.n_torsions);
  I32(torsion_p4, d.n_torsions);
  I32(torsion_periodicity, d.n_torsions);
  F64(torsion_phase, d.n_torsions);
  F64(torsion_k, d.n_torsions);
  d.has_nonbonded = (int32_t)num_field(env, obj, "has_nonbonded");
  d.cutoff = num_field(env, obj, "cutoff");
  d.ewald_tolerance = num_field(env, obj, "ewald_tolerance");
  d.alpha = num_field(env, obj, "alpha");
  {
    int32_t *grid = NULL;
    bool has = false;
    napi_has_named_property(env, obj, "grid", &has);
    if (has) {
      if (array_field(env, obj, "grid", napi_int32_array, 3, (void **)&grid)) return NULL;
      memcpy(d.grid, grid, sizeof d.grid);
    }
  }
  d.use_dispersion_correction = (int32_t)num_field(env, obj, "use_dispersion_correction");
  size_t nb = d.has_nonbonded ? (size_t)d.n_particles : 0;
  F64(charge, nb);
  F64(sigma, nb);
  F64(epsilon, nb);
  d.n_exceptions = (int32_t)num_field(env, obj, "n_exceptions");
  I32(exception_p1, d.n_exceptions);
  I32(exception_p2, d.n_exceptions);
  F64(exception_chargeprod, d.n_exceptions);
  F64(exception_sigma, d.n_exceptions);
  F64(exception_epsilon, d.n_exceptions);
  d.n_restraints = (int32_t)num_field(env, obj, "n_restraints");
  I32(restraint_particle, d.n_restraints);
  F64(restraint_k, d.n_restraints);
  F64(restraint_x0, d.n_restraints);
  F64(restraint_y0, d.n_restraints);
  F64(restraint_z0, d.n_restraints);
  d.cmm_frequency = (int32_t)num_field(env, obj, "cmm_frequency");
  d.has_barostat = (int32_t)num_field(env, obj, "has_barostat");
  d.barostat_pressure = num_field(env, obj, "barostat_pressure");
  d.barostat_temperature = num_field(env, obj, "barostat_temperature");
  d.barostat_frequency = (int32_t)num_field(env, obj, "barostat_frequency");
  d.barostat_seed = (uint32_t)num_field(env, obj, "barostat_seed");

  md_run_options o;
  memset(&o, 0, sizeof o);
  o.time_step = num_field(env, opt, "time_step");
  o.temperature = num_field(env, opt, "temperature");
  o.friction = num_field(env, opt, "friction");
  o.constraint_tolerance = num_field(env, opt, "constraint_tolerance");
  o.seed = (uint32_t)num_field(env, opt, "seed");
  o.integrator = (int32_t)num_field(env, opt, "integrator");
  {
    bool has = false;
    napi_has_named_property(env, opt, "verlet_skin", &has);
    o.verlet_skin = has ? num_field(env, opt, "verlet_skin") : -1.0;
  }
  o.device = (int32_t)num_field(env, opt, "device");
  o.max_batch = (int32_t)num_field(env, opt, "max_batch");

  md_system *sys = NULL;
  if (check(env, md_system_create(&d, &o, &sys))) return NULL;
  napi_value handle;
  napi_create_external(env, sys, NULL, NULL, &handle);
  return handle;
}

static napi_value js_destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  md_system *s;
  if (get_args(env, info, 1, argv) || get_handle(env, argv[0], &s)) return NULL;
  md_system_destroy(s);
  return NULL;
}

/* setPositions / setVelocities(handle, Float64Array[3N]) */
static napi_value set_xyz(napi_env env, napi_callback_info info, md_err (*fn)(md_system *, const double *)) {
  napi_value argv[2];
  md_system *s;
  double *data;
  size_t length;
  if (get_args(env, info, 2, argv) || get_handle(env, argv[0], &s) ||
      typed(env, argv[1], napi_float64_array, (void **)&data, &length))
    return NULL;
  check(env, fn(s, data));
  return NULL;
}
static napi_value js_set_positions(napi_env env, napi_callback_info info) {
  return set_xyz(env, info, md_set_positions);
}
static napi_value js_set_velocities(napi_env env, napi_callback_info info) {
  return set_xyz(env, info, md_set_velocities);
}

/* positions / velocities(handle, n) -> Float64Array[3N] */
static napi_value get_xyz(napi_env env, napi_callback_info info, md_err (*fn)(md_system *, double *)) {
  napi_value argv[2];
  md_system *s;
  int32_t n = 0;
  if (get_args(env, info, 2, argv) || get_handle(env, argv[0], &s)) return NULL;
  napi_get_value_int32(env, argv[1], &n);
  double *data;
  napi_value out = new_f64(env, 3 * (size_t)n, &data);
  if (check(env, fn(s, data))) return NULL;
  return out;
}
static napi_value js_positions(napi_env env, napi_callback_info info) { return get_xyz(env, info, md_get_positions); }
static napi_value js_velocities(napi_env env, napi_callback_info info) {
  return get_xyz(env, info, md_get_velocities);
}

static napi_value js_set_box(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  md_system *s;
  double *data;
  size_t length;
  if (get_args(env, info, 2, argv) || get_handle(env, argv[0], &s) ||
      typed(env, argv[1], napi_float64_array, (void **)&data, &length))
    return NULL;
  if (length != 9) {
    throw_type(env, "a box is nine numbers, the row vectors a, b, c");
    return NULL;
  }
  check(env, md_set_box(s, data));
  return NULL;
}

static napi_value js_box(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  md_system *s;
  if (get_args(env, info, 1, argv) || get_handle(env, argv[0], &s)) return NULL;
  double *data;
  napi_value out = new_f64(env, 9, &data);
  if (check(env, md_get_box(s, data))) return NULL;
  return out;
}

static napi_value js_mesh(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  md_system *s;
  if (get_args(env, info, 1, argv) || get_handle(env, argv[0], &s)) return NULL;
  int32_t *data;
  napi_value out = new_i32(env, 3, &data);
  if (check(env, md_get_mesh(s, data))) return NULL;
  return out;
}

/* evaluate(handle, n) -> {total, groups: Float64Array[6], forces: Float64Array[3N]} */
static napi_value js_evaluate(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  md_system *s;
  int32_t n = 0;
  if (get_args(env, info, 2, argv) || get_handle(env, argv[0], &s)) return NULL;
  napi_get_value_int32(env, argv[1], &n);
  double *forces, *groups, total = 0.0;
  napi_value f = new_f64(env, 3 * (size_t)n, &forces);
  napi_value g = new_f64(env, MD_GROUP_COUNT, &groups);
  if (check(env, md_evaluate(s, forces, groups, &total))) return NULL;
  napi_value out, t;
  napi_create_object(env, &out);
  napi_create_double(env, total, &t);
  napi_set_named_property(env, out, "total", t);
  napi_set_named_property(env, out, "groups", g);
  napi_set_named_property(env, out, "forces", f);
  return out;
}

static napi_value js_step(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  md_system *s;
  if (get_args(env, info, 1, argv) || get_handle(env, argv[0], &s)) return NULL;
  int64_t step = 0;
  if (check(env, md_get_step(s, &step))) return NULL;
  napi_value out;
  napi_create_int64(env, step, &out);
  return out;
}

static napi_value js_set_step(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  md_system *s;
  if (get_args(env, info, 2, argv) || get_handle(env, argv[0], &s)) return NULL;
  int64_t step = 0;
  napi_get_value_int64(env, argv[1], &step);
  check(env, md_set_step(s, step));
  return NULL;
}

/* attachBarostat(handle, {pressure, temperature, frequency, moleculeStart:
 * Int32Array, moleculeAtoms: Int32Array, seed, volumeScale, draws,
 * stepsSinceAttempt, windowAttempted, windowAccepted}): md_barostat_attach,
 * so run() performs the Monte Carlo moves on the device (frequency 0
 * detaches).  The molecules are upstream's findMolecules
 * (src/engine/molecules.ts:71) as a CSR table. */
static napi_value js_attach_barostat(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  md_system *s;
  if (get_args(env, info, 2, argv) || get_handle(env, argv[0], &s)) return NULL;
  napi_value obj = argv[1];
  md_barostat_desc d;
  memset(&d, 0, sizeof d);
  d.pressure = num_field(env, obj, "pressure");
  d.temperature = num_field(env, obj, "temperature");
  d.frequency = (int32_t)num_field(env, obj, "frequency");
  d.seed = (uint32_t)num_field(env, obj, "seed");
  d.volume_scale = num_field(env, obj, "volumeScale");
  d.draws = (int64_t)num_field(env, obj, "draws");
  d.steps_since_attempt = (int32_t)num_field(env, obj, "stepsSinceAttempt");
  d.window_attempted = (int32_t)num_field(env, obj, "windowAttempted");
  d.window_accepted = (int32_t)num_field(env, obj, "windowAccepted");
  if (d.frequency > 0) {
    void *start = NULL, *atoms = NULL;
    size_t nstart = 0, natoms = 0;
    napi_value v;
    napi_get_named_property(env, obj, "moleculeStart", &v);
    if (typed(env, v, napi_int32_array, &start, &nstart)) return NULL;
    napi_get_named_property(env, obj, "moleculeAtoms", &v);
    if (typed(env, v, napi_int32_array, &atoms, &natoms)) return NULL;
    if (nstart < 2) {
      throw_type(env, "moleculeStart needs at least two entries");
      return NULL;
    }
    d.n_molecules = (int32_t)(nstart - 1);
    d.molecule_start = (const int32_t *)start;
    d.molecule_atoms = (const int32_t *)atoms;
  }
  check(env, md_barostat_attach(s, &d));
  return NULL;
}

/* barostatState(handle) -> md_barostat_state's fields */
static napi_value js_barostat_state(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  md_system *s;
  if (get_args(env, info, 1, argv) || get_handle(env, argv[0], &s)) return NULL;
  md_barostat_state st;
  if (check(env, md_barostat_get(s, &st))) return NULL;
  napi_value out, v;
  napi_create_object(env, &out);
#define PUT_NUM(key, value)                   \
  napi_create_double(env, (double)(value), &v); \
  napi_set_named_property(env, out, key, v)
  PUT_NUM("attempts", st.attempts);
  PUT_NUM("accepted", st.accepted);
  PUT_NUM("draws", st.draws);
  PUT_NUM("stepsSinceAttempt", st.steps_since_attempt);
  PUT_NUM("windowAttempted", st.window_attempted);
  PUT_NUM("windowAccepted", st.window_accepted);
  PUT_NUM("volumeScale", st.volume_scale);
  PUT_NUM("lastAccepted", st.last_accepted);
  PUT_NUM("lastEnergyBefore", st.last_energy_before);
  PUT_NUM("lastEnergyAfter", st.last_energy_after);
  PUT_NUM("lastWork", st.last_work);
  PUT_NUM("lastLengthScale", st.last_length_scale);
#undef PUT_NUM
  return out;
}

/* run(handle, steps) -> Promise<void>, md_run on the thread pool */
typedef struct {
  napi_async_work work;
  napi_deferred deferred;
  md_system *sys;
  int64_t steps;
  md_err rc;
  char message[512];
} run_job;

static void run_execute(napi_env env, void *data) {
  run_job *j = (run_job *)data;
  j->rc = md_run(j->sys, j->steps);
  if (j->rc != MD_OK) {
    /* md_last_error is per thread: take it on the thread that failed */
    strncpy(j->message, md_last_error(), sizeof j->message - 1);
    j->message[sizeof j->message - 1] = 0;
  }
}

static void run_complete(napi_env env, napi_status status, void *data) {
  run_job *j = (run_job *)data;
  if (j->rc == MD_OK) {
    napi_value undefined;
    napi_get_undefined(env, &undefined);
    napi_resolve_deferred(env, j->deferred, undefined);
  } else {
    napi_reject_deferred(env, j->deferred, make_error(env, j->rc, j->message));
  }
  napi_delete_async_work(env, j->work);
  free(j);
}

static napi_value js_run(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  md_system *s;
  if (get_args(env, info, 2, argv) || get_handle(env, argv[0], &s)) return NULL;
  run_job *j = (run_job *)calloc(1, sizeof *j);
  j->sys = s;
  napi_get_value_int64(env, argv[1], &j->steps);
  napi_value promise, name;
  napi_create_promise(env, &j->deferred, &promise);
  napi_create_string_utf8(env, "md_run", NAPI_AUTO_LENGTH, &name);
  napi_create_async_work(env, NULL, name, run_execute, run_complete, j, &j->work);
  napi_queue_async_work(env, j->work);
  return promise;
}

static napi_value js_version(napi_env env, napi_callback_info info) {
  napi_value out;
  napi_create_string_utf8(env, md_version(), NAPI_AUTO_LENGTH, &out);
  return out;
}

/* probeCellOrder(Int32Array[3]) -> Int32Array of the cell sort order */
static napi_value js_probe_cell_order(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  int32_t *dims;
  size_t length;
  if (get_args(env, info, 1, argv) || typed(env, argv[0], napi_int32_array, (void **)&dims, &length)) return NULL;
  if (length != 3) {
    throw_type(env, "expected three grid dimensions");
    return NULL;
  }
  long cells = (long)dims[0] * dims[1] * dims[2];
  int32_t *data;
  napi_value out = new_i32(env, cells > 0 ? (size_t)cells : 0, &data);
  if (check(env, md_probe_cell_order(dims, data))) return NULL;
  return out;
}

#define EXPORT(name, fn)                                    \
  do {                                                      \
    napi_value f_;                                          \
    napi_create_function(env, name, NAPI_AUTO_LENGTH, fn, NULL, &f_); \
    napi_set_named_property(env, exports, name, f_);        \
  } while (0)

static napi_value init(napi_env env, napi_value exports) {
  EXPORT("version", js_version);
  EXPORT("probeCellOrder", js_probe_cell_order);
  EXPORT("createSystem", js_create);
  EXPORT("destroySystem", js_destroy);
  EXPORT("setPositions", js_set_positions);
  EXPORT("setVelocities", js_set_velocities);
  EXPORT("positions", js_positions);
  EXPORT("velocities", js_velocities);
  EXPORT("setBox", js_set_box);
  EXPORT("box", js_box);
  EXPORT("mesh", js_mesh);
  EXPORT("evaluate", js_evaluate);
  EXPORT("step", js_step);
  EXPORT("setStep", js_set_step);
  EXPORT("run", js_run);
  EXPORT("attachBarostat", js_attach_barostat);
  EXPORT("barostatState", js_barostat_state);
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
