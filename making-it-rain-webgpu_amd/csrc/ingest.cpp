// ingest.cpp -- streaming System/State XML ingestion (host only).
//
// Replaces the upstream JavaScript scanner `parseSystem` / `parseState`
// (src/engine/system.ts:521-597, src/engine/state.ts:60-152) for systems of
// hundreds of thousands of atoms: one pass over the document, no element
// tree, every value parsed straight into the flat arrays md_system_desc
// points at.  The accepted set and the refusals are the reference's: the
// seven force kinds (HarmonicBond/Angle, PeriodicTorsion, Nonbonded with PME,
// CMMotionRemover, MonteCarloBarostat, and a CustomExternalForce holding
// exactly k*periodicdistance(x,y,z,x0,y0,z0)^2 with per-particle k, x0, y0,
// z0); anything else, parameter offsets, LJ-PME, a switching function or a
// restraint with global parameters is MD_ERR_UNSUPPORTED_FORCE with the
// reason.  Several forces of one harmonic kind are concatenated.
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mdcore.h"

void md_set_last_error(const std::string &message);

namespace {

struct parse_error {
  md_err code;
  std::string message;
};

[[noreturn]] void fail(md_err code, const std::string &message) { throw parse_error{code, message}; }

struct attr {
  const char *name;
  size_t name_len;
  const char *value;
  size_t value_len;
};

// A start (or empty) element: its name and attributes, pointing into the text.
struct element {
  const char *name = nullptr;
  size_t name_len = 0;
  std::vector<attr> attrs;
  bool empty = false;  // "<X ... />"

  bool is(const char *n) const { return name_len == std::strlen(n) && std::memcmp(name, n, name_len) == 0; }
  const attr *find(const char *n) const {
    const size_t l = std::strlen(n);
    for (const attr &a : attrs)
      if (a.name_len == l && std::memcmp(a.name, n, l) == 0) return &a;
    return nullptr;
  }
  std::string text(const char *n) const {
    const attr *a = find(n);
    if (!a) return std::string();
    std::string out;
    out.reserve(a->value_len);
    for (size_t i = 0; i < a->value_len; ++i) {  // the five predefined entities
      const char *p = a->value + i;
      const size_t left = a->value_len - i;
      if (*p == '&') {
        if (left >= 4 && !std::memcmp(p, "&lt;", 4)) { out += '<'; i += 3; continue; }
        if (left >= 4 && !std::memcmp(p, "&gt;", 4)) { out += '>'; i += 3; continue; }
        if (left >= 5 && !std::memcmp(p, "&amp;", 5)) { out += '&'; i += 4; continue; }
        if (left >= 6 && !std::memcmp(p, "&quot;", 6)) { out += '"'; i += 5; continue; }
        if (left >= 6 && !std::memcmp(p, "&apos;", 6)) { out += '\''; i += 5; continue; }
      }
      out += *p;
    }
    return out;
  }
  double num(const char *n, bool required = true, double fallback = 0.0) const {
    const attr *a = find(n);
    if (!a) {
      if (required) fail(MD_ERR_INVALID_ARGUMENT, "<" + std::string(name, name_len) + "> has no attribute " + n);
      return fallback;
    }
    char buf[64];
    const size_t l = a->value_len < sizeof buf - 1 ? a->value_len : sizeof buf - 1;
    std::memcpy(buf, a->value, l);
    buf[l] = 0;
    char *end = nullptr;
    const double v = std::strtod(buf, &end);
    if (end == buf) fail(MD_ERR_INVALID_ARGUMENT, "attribute " + std::string(n) + " is not a number");
    return v;
  }
  int integer(const char *n, bool required = true, int fallback = 0) const {
    return (int)num(n, required, (double)fallback);
  }
  bool flag(const char *n) const {
    const attr *a = find(n);
    if (!a) return false;
    return !(a->value_len == 1 && a->value[0] == '0') && !(a->value_len == 5 && !std::memcmp(a->value, "false", 5)) &&
           !(a->value_len == 5 && !std::memcmp(a->value, "False", 5));
  }
};

// the first `pat` in [p, stop), or nullptr (the text need not be NUL-terminated)
const char *find_seq(const char *p, const char *stop, const char *pat) {
  const size_t l = std::strlen(pat);
  for (; p + l <= stop; ++p)
    if (std::memcmp(p, pat, l) == 0) return p;
  return nullptr;
}

// One pass over the document: calls start(e) for every start or empty
// element and end(name, len) for every end tag (and after an empty element).
template <typename Start, typename End>
void scan(const char *text, size_t length, Start start, End end) {
  const char *p = text, *stop = text + length;
  element e;
  while (p < stop) {
    const char *lt = static_cast<const char *>(std::memchr(p, '<', (size_t)(stop - p)));
    if (!lt) break;
    p = lt + 1;
    if (p >= stop) break;
    if (*p == '?') {  // declaration
      const char *q = find_seq(p, stop, "?>");
      if (!q) fail(MD_ERR_INVALID_ARGUMENT, "an unterminated XML declaration");
      p = q + 2;
      continue;
    }
    if (*p == '!') {  // comment
      const char *q = find_seq(p, stop, "-->");
      if (!q) fail(MD_ERR_INVALID_ARGUMENT, "an unterminated comment");
      p = q + 3;
      continue;
    }
    if (*p == '/') {  // end tag
      const char *n = ++p;
      while (p < stop && *p != '>' && *p != ' ' && *p != '\t' && *p != '\n' && *p != '\r') ++p;
      end(n, (size_t)(p - n));
      while (p < stop && *p != '>') ++p;
      ++p;
      continue;
    }
    e.name = p;
    while (p < stop && *p != '>' && *p != '/' && *p != ' ' && *p != '\t' && *p != '\n' && *p != '\r') ++p;
    e.name_len = (size_t)(p - e.name);
    e.attrs.clear();
    e.empty = false;
    for (;;) {
      while (p < stop && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
      if (p >= stop) fail(MD_ERR_INVALID_ARGUMENT, "an unterminated element");
      if (*p == '/') {
        e.empty = true;
        p += 2;
        if (p > stop) fail(MD_ERR_INVALID_ARGUMENT, "an unterminated element");
        break;
      }
      if (*p == '>') {
        ++p;
        break;
      }
      attr a;
      a.name = p;
      while (p < stop && *p != '=' && *p != ' ') ++p;
      a.name_len = (size_t)(p - a.name);
      while (p < stop && *p != '"' && *p != '\'') ++p;
      if (p >= stop) fail(MD_ERR_INVALID_ARGUMENT, "an attribute without a value");
      const char quote = *p++;
      a.value = p;
      const char *q = static_cast<const char *>(std::memchr(p, quote, (size_t)(stop - p)));
      if (!q) fail(MD_ERR_INVALID_ARGUMENT, "an unterminated attribute value");
      a.value_len = (size_t)(q - p);
      p = q + 1;
      e.attrs.push_back(a);
    }
    start(e);
    if (e.empty) end(e.name, e.name_len);
  }
}

bool name_is(const char *n, size_t l, const char *want) { return l == std::strlen(want) && !std::memcmp(n, want, l); }

}  // namespace

// The parsed System: owns every array md_system_desc points into.
struct md_xml_system {
  int n = 0;
  double box[9] = {0};
  int rows = 0;
  std::vector<double> masses;
  std::vector<int32_t> c1, c2;
  std::vector<double> cd;
  std::vector<int32_t> b1, b2;
  std::vector<double> bl, bk;
  std::vector<int32_t> a1, a2, a3;
  std::vector<double> at, ak;
  std::vector<int32_t> t1, t2, t3, t4, tn;
  std::vector<double> tp, tk;
  bool nonbonded = false;
  double cutoff = 0, tolerance = 0, alpha = 0;
  int grid[3] = {0, 0, 0};
  bool dispersion = false;
  std::vector<double> q, sig, eps;
  std::vector<int32_t> e1, e2;
  std::vector<double> eq, esig, eeps;
  std::vector<int32_t> rp;
  std::vector<double> rk, rx, ry, rz;
  int cmm = 0;
  bool barostat = false;
  double pressure = 0, temperature = 0;
  int frequency = 0;
  uint32_t seed = 0;
};

struct md_xml_state {
  std::vector<double> x, v;
  bool has_v = false;
  double box[9] = {0};
  int rows = 0;
  int64_t step = 0;
  double time = 0;
};

static std::string squeeze(const std::string &s) {
  std::string out;
  for (char c : s)
    if (c != ' ' && c != '\t' && c != '\n' && c != '\r' && c != ';') out += c;
  return out;
}

static void read_box(const element &e, double *box, int &rows) {
  if (!e.is("A") && !e.is("B") && !e.is("C"))
    fail(MD_ERR_INVALID_ARGUMENT, "a box vector element must be A, B or C, not " + std::string(e.name, e.name_len));
  const int r = e.is("A") ? 0 : e.is("B") ? 1 : 2;
  box[3 * r] = e.num("x");
  box[3 * r + 1] = e.num("y");
  box[3 * r + 2] = e.num("z");
  rows |= 1 << r;
}

static void parse_system(const char *text, size_t length, md_xml_system &S) {
  std::vector<std::string> stack;
  std::string force;          // the current <Force type>
  std::vector<std::string> params;  // CustomExternalForce per-particle parameter names
  bool seen_system = false;
  scan(
      text, length,
      [&](const element &e) {
        const std::string name(e.name, e.name_len);
        const std::string parent = stack.empty() ? "" : stack.back();
        if (stack.empty()) {
          if (!e.is("System")) fail(MD_ERR_INVALID_ARGUMENT, "no <System> element");
          seen_system = true;
        } else if (parent == "PeriodicBoxVectors" && stack.size() == 2) {
          read_box(e, S.box, S.rows);
        } else if (parent == "Particles" && stack.size() == 2 && e.is("Particle")) {
          S.masses.push_back(e.num("mass"));
        } else if (parent == "Constraints" && e.is("Constraint")) {
          S.c1.push_back(e.integer("p1"));
          S.c2.push_back(e.integer("p2"));
          S.cd.push_back(e.num("d"));
        } else if (parent == "Forces" && e.is("Force")) {
          force = e.text("type");
          params.clear();
          if (force == "NonbondedForce") {
            if (S.nonbonded) fail(MD_ERR_UNSUPPORTED_FORCE, "NonbondedForce is not implemented twice in one System");
            const int method = e.integer("method");
            if (method == 5) fail(MD_ERR_UNSUPPORTED_FORCE, "NonbondedForce: LJ-PME is not implemented");
            if (method != 4)
              fail(MD_ERR_UNSUPPORTED_FORCE, "NonbondedForce method " + std::to_string(method) +
                                                 ": this engine implements PME (4)");
            if (e.flag("useSwitchingFunction"))
              fail(MD_ERR_UNSUPPORTED_FORCE, "NonbondedForce with a switching function is not implemented");
            S.nonbonded = true;
            S.cutoff = e.num("cutoff");
            S.tolerance = e.num("ewaldTolerance");
            S.alpha = e.num("alpha", false, 0.0);
            S.grid[0] = e.integer("nx", false);
            S.grid[1] = e.integer("ny", false);
            S.grid[2] = e.integer("nz", false);
            S.dispersion = e.flag("dispersionCorrection");
          } else if (force == "CMMotionRemover") {
            S.cmm = e.integer("frequency", false, 1);
          } else if (force == "MonteCarloBarostat") {
            S.barostat = true;
            S.pressure = e.num("pressure");
            S.temperature = e.num("temperature");
            S.frequency = e.integer("frequency", false, 25);
            S.seed = (uint32_t)(int64_t)e.num("randomSeed", false, 0.0);
          } else if (force == "CustomExternalForce") {
            const std::string expr = e.text("energy");
            if (squeeze(expr) != "k*periodicdistance(x,y,z,x0,y0,z0)^2")
              fail(MD_ERR_UNSUPPORTED_FORCE, "CustomExternalForce: its energy expression is \"" + expr +
                                                 "\", and this engine implements only "
                                                 "\"k*periodicdistance(x,y,z,x0,y0,z0)^2\"");
          } else if (force != "HarmonicBondForce" && force != "HarmonicAngleForce" &&
                     force != "PeriodicTorsionForce") {
            fail(MD_ERR_UNSUPPORTED_FORCE, force + " is not implemented by this engine");
          }
        } else if (!force.empty()) {
          if (force == "HarmonicBondForce" && e.is("Bond")) {
            S.b1.push_back(e.integer("p1"));
            S.b2.push_back(e.integer("p2"));
            S.bl.push_back(e.num("d"));
            S.bk.push_back(e.num("k"));
          } else if (force == "HarmonicAngleForce" && e.is("Angle")) {
            S.a1.push_back(e.integer("p1"));
            S.a2.push_back(e.integer("p2"));
            S.a3.push_back(e.integer("p3"));
            S.at.push_back(e.num("a"));
            S.ak.push_back(e.num("k"));
          } else if (force == "PeriodicTorsionForce" && e.is("Torsion")) {
            S.t1.push_back(e.integer("p1"));
            S.t2.push_back(e.integer("p2"));
            S.t3.push_back(e.integer("p3"));
            S.t4.push_back(e.integer("p4"));
            S.tn.push_back(e.integer("periodicity"));
            S.tp.push_back(e.num("phase"));
            S.tk.push_back(e.num("k"));
          } else if (force == "NonbondedForce") {
            if (parent == "ParticleOffsets" || parent == "ExceptionOffsets")
              fail(MD_ERR_UNSUPPORTED_FORCE, "NonbondedForce: it carries parameter offsets, which make the energy "
                                             "depend on a global parameter");
            if (parent == "Particles" && e.is("Particle")) {
              S.q.push_back(e.num("q"));
              S.sig.push_back(e.num("sig"));
              S.eps.push_back(e.num("eps"));
            } else if (parent == "Exceptions" && e.is("Exception")) {
              S.e1.push_back(e.integer("p1"));
              S.e2.push_back(e.integer("p2"));
              S.eq.push_back(e.num("q"));
              S.esig.push_back(e.num("sig"));
              S.eeps.push_back(e.num("eps"));
            }
          } else if (force == "CustomExternalForce") {
            if (parent == "GlobalParameters" && e.is("Parameter"))
              fail(MD_ERR_UNSUPPORTED_FORCE, "CustomExternalForce: it has global parameters (" + e.text("name") + ")");
            if (parent == "PerParticleParameters" && e.is("Parameter")) params.push_back(e.text("name"));
            if (parent == "Particles" && e.is("Particle")) {
              auto column = [&](const char *want) {
                for (size_t k = 0; k < params.size(); ++k)
                  if (params[k] == want) return e.num(("param" + std::to_string(k + 1)).c_str());
                fail(MD_ERR_UNSUPPORTED_FORCE,
                     std::string("CustomExternalForce: it has no per-particle parameter \"") + want + "\"");
              };
              S.rp.push_back(e.integer("index"));
              S.rk.push_back(column("k"));
              S.rx.push_back(column("x0"));
              S.ry.push_back(column("y0"));
              S.rz.push_back(column("z0"));
            }
          }
        }
        if (!e.empty) stack.push_back(name);
      },
      [&](const char *n, size_t l) {
        if (name_is(n, l, "Force")) force.clear();
        if (!stack.empty() && name_is(n, l, stack.back().c_str())) stack.pop_back();
      });
  if (!seen_system) fail(MD_ERR_INVALID_ARGUMENT, "no <System> element");
  if (S.rows != 7) fail(MD_ERR_INVALID_ARGUMENT, "no <PeriodicBoxVectors> with A, B and C");
  S.n = (int)S.masses.size();
  if (S.nonbonded && (int)S.q.size() != S.n)
    fail(MD_ERR_INVALID_ARGUMENT, "NonbondedForce has " + std::to_string(S.q.size()) + " particles but the System has " +
                                      std::to_string(S.n));
}

static void parse_state(const char *text, size_t length, md_xml_state &T) {
  std::vector<std::string> stack;
  bool seen = false;
  scan(
      text, length,
      [&](const element &e) {
        const std::string parent = stack.empty() ? "" : stack.back();
        if (stack.empty()) {
          if (!e.is("State")) fail(MD_ERR_INVALID_ARGUMENT, "no <State> element");
          seen = true;
          T.step = (int64_t)e.num("stepCount", false, 0.0);
          T.time = e.num("time", false, 0.0);
        } else if (parent == "PeriodicBoxVectors") {
          read_box(e, T.box, T.rows);
        } else if (parent == "Positions" && e.is("Position")) {
          T.x.push_back(e.num("x"));
          T.x.push_back(e.num("y"));
          T.x.push_back(e.num("z"));
        } else if (parent == "Velocities" && e.is("Velocity")) {
          T.has_v = true;
          T.v.push_back(e.num("x"));
          T.v.push_back(e.num("y"));
          T.v.push_back(e.num("z"));
        } else if (e.is("Velocities")) {
          T.has_v = true;
        }
        if (!e.empty) stack.push_back(std::string(e.name, e.name_len));
      },
      [&](const char *n, size_t l) {
        if (!stack.empty() && name_is(n, l, stack.back().c_str())) stack.pop_back();
      });
  if (!seen) fail(MD_ERR_INVALID_ARGUMENT, "no <State> element");
  if (T.rows != 7) fail(MD_ERR_INVALID_ARGUMENT, "no <PeriodicBoxVectors> with A, B and C");
  if (T.has_v && T.v.size() != T.x.size())
    fail(MD_ERR_INVALID_ARGUMENT, "velocities has " + std::to_string(T.v.size() / 3) + " entries but there are " +
                                      std::to_string(T.x.size() / 3) + " positions");
}

extern "C" {

md_err md_xml_parse_system(const char *text, int64_t length, md_xml_system **out) {
  if (!text || length < 0 || !out) {
    md_set_last_error("bad arguments to md_xml_parse_system");
    return MD_ERR_INVALID_ARGUMENT;
  }
  *out = nullptr;
  md_xml_system *S = new md_xml_system();
  try {
    parse_system(text, (size_t)length, *S);
  } catch (const parse_error &e) {
    delete S;
    md_set_last_error(e.message);
    return e.code;
  } catch (const std::bad_alloc &) {
    delete S;
    md_set_last_error("host allocation failed");
    return MD_ERR_OUT_OF_MEMORY;
  }
  *out = S;
  md_set_last_error(std::string());
  return MD_OK;
}

md_err md_xml_system_describe(const md_xml_system *S, md_system_desc *d) {
  if (!S || !d) {
    md_set_last_error("null argument");
    return MD_ERR_INVALID_ARGUMENT;
  }
  std::memset(d, 0, sizeof *d);
  d->n_particles = S->n;
  d->masses = S->masses.data();
  std::memcpy(d->box, S->box, sizeof d->box);
  d->n_constraints = (int32_t)S->c1.size();
  d->constraint_p1 = S->c1.data();
  d->constraint_p2 = S->c2.data();
  d->constraint_distance = S->cd.data();
  d->n_bonds = (int32_t)S->b1.size();
  d->bond_p1 = S->b1.data();
  d->bond_p2 = S->b2.data();
  d->bond_length = S->bl.data();
  d->bond_k = S->bk.data();
  d->n_angles = (int32_t)S->a1.size();
  d->angle_p1 = S->a1.data();
  d->angle_p2 = S->a2.data();
  d->angle_p3 = S->a3.data();
  d->angle_theta = S->at.data();
  d->angle_k = S->ak.data();
  d->n_torsions = (int32_t)S->t1.size();
  d->torsion_p1 = S->t1.data();
  d->torsion_p2 = S->t2.data();
  d->torsion_p3 = S->t3.data();
  d->torsion_p4 = S->t4.data();
  d->torsion_periodicity = S->tn.data();
  d->torsion_phase = S->tp.data();
  d->torsion_k = S->tk.data();
  d->has_nonbonded = S->nonbonded ? 1 : 0;
  d->cutoff = S->cutoff;
  d->ewald_tolerance = S->tolerance;
  d->alpha = S->alpha;
  std::memcpy(d->grid, S->grid, sizeof d->grid);
  d->use_dispersion_correction = S->dispersion ? 1 : 0;
  d->charge = S->q.data();
  d->sigma = S->sig.data();
  d->epsilon = S->eps.data();
  d->n_exceptions = (int32_t)S->e1.size();
  d->exception_p1 = S->e1.data();
  d->exception_p2 = S->e2.data();
  d->exception_chargeprod = S->eq.data();
  d->exception_sigma = S->esig.data();
  d->exception_epsilon = S->eeps.data();
  d->n_restraints = (int32_t)S->rp.size();
  d->restraint_particle = S->rp.data();
  d->restraint_k = S->rk.data();
  d->restraint_x0 = S->rx.data();
  d->restraint_y0 = S->ry.data();
  d->restraint_z0 = S->rz.data();
  d->cmm_frequency = S->cmm;
  d->has_barostat = S->barostat ? 1 : 0;
  d->barostat_pressure = S->pressure;
  d->barostat_temperature = S->temperature;
  d->barostat_frequency = S->frequency;
  d->barostat_seed = S->seed;
  md_set_last_error(std::string());
  return MD_OK;
}

void md_xml_system_free(md_xml_system *S) { delete S; }

md_err md_xml_parse_state(const char *text, int64_t length, md_xml_state **out) {
  if (!text || length < 0 || !out) {
    md_set_last_error("bad arguments to md_xml_parse_state");
    return MD_ERR_INVALID_ARGUMENT;
  }
  *out = nullptr;
  md_xml_state *T = new md_xml_state();
  try {
    parse_state(text, (size_t)length, *T);
  } catch (const parse_error &e) {
    delete T;
    md_set_last_error(e.message);
    return e.code;
  } catch (const std::bad_alloc &) {
    delete T;
    md_set_last_error("host allocation failed");
    return MD_ERR_OUT_OF_MEMORY;
  }
  *out = T;
  md_set_last_error(std::string());
  return MD_OK;
}

md_err md_xml_state_get(const md_xml_state *T, int32_t *n, const double **positions, const double **velocities,
                        double box[9], int64_t *step, double *time) {
  if (!T) {
    md_set_last_error("null state");
    return MD_ERR_INVALID_ARGUMENT;
  }
  if (n) *n = (int32_t)(T->x.size() / 3);
  if (positions) *positions = T->x.data();
  if (velocities) *velocities = T->has_v ? T->v.data() : nullptr;
  if (box) std::memcpy(box, T->box, sizeof T->box);
  if (step) *step = T->step;
  if (time) *time = T->time;
  md_set_last_error(std::string());
  return MD_OK;
}

void md_xml_state_free(md_xml_state *T) { delete T; }

}  // extern "C"
