// Node.js N-API addon over the C ABI (include/rainmd.h): the thin binding the
// JavaScript GpuSimulation facade (simulation.js beside this file) calls.
//
// Every export is a direct call of one rmd_* entry point. Handles travel as
// napi_external values whose finalizers free them. The long-running calls
// (initialise, run, minimise, scaleVolume, attemptVolumeChange) run as
// napi_async_work on the libuv pool and return Promises, so the facade keeps
// the reference's async method shapes (webgpu/simulation.js:412-1070) and the
// event loop never blocks. A failing call rejects or throws an Error whose
// `code` is the RMD_E_* status and whose message is rmd_last_error(); the
// facade maps codes to the reference's error classes.
//
// Written against the N-API C interface of Node 12 (node_api.h), built with
// g++ by rainmd_amd/build.py.
#include <node_api.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rainmd.h"

namespace {

#define NAPI_CALL(env, call)                                   \
  do {                                                         \
    if ((call) != napi_ok) {                                   \
      napi_throw_error((env), nullptr, "N-API call failed: " #call); \
      return nullptr;                                          \
    }                                                          \
  } while (0)

napi_value makeError(napi_env env, int code, const std::string &message) {
  napi_value msg, err, codeValue;
  napi_create_string_utf8(env, message.c_str(), message.size(), &msg);
  napi_create_error(env, nullptr, msg, &err);
  napi_create_int32(env, code, &codeValue);
  napi_set_named_property(env, err, "code", codeValue);
  return err;
}

// Throws for a failing status; returns true when the call succeeded.
bool check(napi_env env, int status, const rmd_ctx *ctx) {
  if (status == RMD_OK) return true;
  const char *m = ctx ? rmd_last_error(ctx) : rmd_last_global_error();
  napi_throw(env, makeError(env, status, m ? m : "rainmd error"));
  return false;
}

// ---- argument helpers ----

napi_value undefinedValue(napi_env env) {
  napi_value u;
  napi_get_undefined(env, &u);
  return u;
}

bool isNullish(napi_env env, napi_value v) {
  napi_valuetype t;
  napi_typeof(env, v, &t);
  return t == napi_undefined || t == napi_null;
}

double getNumber(napi_env env, napi_value v, double fallback) {
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t != napi_number) return fallback;
  double d;
  napi_get_value_double(env, v, &d);
  return d;
}

double prop(napi_env env, napi_value obj, const char *name, double fallback) {
  if (isNullish(env, obj)) return fallback;
  bool has = false;
  napi_has_named_property(env, obj, name, &has);
  if (!has) return fallback;
  napi_value v;
  napi_get_named_property(env, obj, name, &v);
  return getNumber(env, v, fallback);
}

bool propValue(napi_env env, napi_value obj, const char *name, napi_value *out) {
  if (isNullish(env, obj)) return false;
  bool has = false;
  napi_has_named_property(env, obj, name, &has);
  if (!has) return false;
  napi_get_named_property(env, obj, name, out);
  return !isNullish(env, *out);
}

// Bytes of a Buffer, a typed array or a string (copied into `storage`).
bool getBytes(napi_env env, napi_value v, std::string &storage) {
  bool isBuffer = false;
  napi_is_buffer(env, v, &isBuffer);
  if (isBuffer) {
    void *data;
    size_t length;
    napi_get_buffer_info(env, v, &data, &length);
    storage.assign(static_cast<const char *>(data), length);
    return true;
  }
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t == napi_string) {
    size_t length;
    napi_get_value_string_utf8(env, v, nullptr, 0, &length);
    storage.resize(length + 1);
    napi_get_value_string_utf8(env, v, &storage[0], length + 1, &length);
    storage.resize(length);
    return true;
  }
  return false;
}

// A typed array's data as `T` with its element count; false if the value is
// not a typed array of that element type.
template <typename T>
bool getTyped(napi_env env, napi_value v, napi_typedarray_type want, T **data, size_t *count) {
  bool isTyped = false;
  napi_is_typedarray(env, v, &isTyped);
  if (!isTyped) return false;
  napi_typedarray_type type;
  void *raw;
  napi_value buffer;
  size_t offset;
  napi_get_typedarray_info(env, v, &type, count, &raw, &buffer, &offset);
  if (type != want) return false;
  *data = static_cast<T *>(raw);
  return true;
}

// Doubles from a Float64Array, a Float32Array or a plain Array.
bool getDoubles(napi_env env, napi_value v, std::vector<double> &out) {
  double *d;
  float *f;
  size_t n;
  if (getTyped(env, v, napi_float64_array, &d, &n)) {
    out.assign(d, d + n);
    return true;
  }
  if (getTyped(env, v, napi_float32_array, &f, &n)) {
    out.assign(f, f + n);
    return true;
  }
  bool isArray = false;
  napi_is_array(env, v, &isArray);
  if (!isArray) return false;
  uint32_t len;
  napi_get_array_length(env, v, &len);
  out.resize(len);
  for (uint32_t k = 0; k < len; ++k) {
    napi_value e;
    napi_get_element(env, v, k, &e);
    out[k] = getNumber(env, e, 0.0);
  }
  return true;
}

bool getInts(napi_env env, napi_value v, std::vector<int32_t> &out) {
  std::vector<double> d;
  if (!getDoubles(env, v, d)) {
    int32_t *i;
    uint32_t *u;
    size_t n;
    if (getTyped(env, v, napi_int32_array, &i, &n)) {
      out.assign(i, i + n);
      return true;
    }
    if (getTyped(env, v, napi_uint32_array, &u, &n)) {
      out.assign(u, u + n);
      return true;
    }
    return false;
  }
  out.resize(d.size());
  for (size_t k = 0; k < d.size(); ++k) out[k] = (int32_t)d[k];
  return true;
}

template <typename T>
napi_value typedCopy(napi_env env, const T *src, size_t count, napi_typedarray_type type) {
  napi_value buffer, array;
  void *data;
  napi_create_arraybuffer(env, count * sizeof(T), &data, &buffer);
  if (count && src) std::memcpy(data, src, count * sizeof(T));
  napi_create_typedarray(env, type, count, buffer, 0, &array);
  return array;
}

napi_value f64(napi_env env, const double *src, size_t n) { return typedCopy(env, src, n, napi_float64_array); }
napi_value f32(napi_env env, const float *src, size_t n) { return typedCopy(env, src, n, napi_float32_array); }
napi_value i32(napi_env env, const int32_t *src, size_t n) { return typedCopy(env, src, n, napi_int32_array); }

void setNumber(napi_env env, napi_value obj, const char *name, double value) {
  napi_value v;
  napi_create_double(env, value, &v);
  napi_set_named_property(env, obj, name, v);
}
void setBool(napi_env env, napi_value obj, const char *name, bool value) {
  napi_value v;
  napi_get_boolean(env, value, &v);
  napi_set_named_property(env, obj, name, v);
}
void setValue(napi_env env, napi_value obj, const char *name, napi_value v) { napi_set_named_property(env, obj, name, v); }

napi_value newObject(napi_env env) {
  napi_value o;
  napi_create_object(env, &o);
  return o;
}

// ---- handles ----

template <typename T>
void finalizeHandle(napi_env, void *data, void *hint);

template <>
void finalizeHandle<rmd_topology>(napi_env, void *data, void *) {
  rmd_topology_free(static_cast<rmd_topology *>(data));
}
template <>
void finalizeHandle<rmd_coordinates>(napi_env, void *data, void *) {
  rmd_coordinates_free(static_cast<rmd_coordinates *>(data));
}
template <>
void finalizeHandle<rmd_system>(napi_env, void *data, void *) {
  rmd_system_free(static_cast<rmd_system *>(data));
}

// A context handle: the external holds this box so destroy() can run early
// and the finalizer does nothing afterwards.
struct CtxBox {
  rmd_ctx *ctx = nullptr;
  int busy = 0;
  bool owned = true;  // false: a replica of a batch (the batch owns it and this box)
};
void finalizeCtx(napi_env, void *data, void *) {
  CtxBox *box = static_cast<CtxBox *>(data);
  if (!box->owned) return;
  if (box->ctx) rmd_destroy(box->ctx);
  delete box;
}

template <typename T>
napi_value wrap(napi_env env, T *handle) {
  napi_value v;
  napi_create_external(env, handle, finalizeHandle<T>, nullptr, &v);
  return v;
}

template <typename T>
T *unwrap(napi_env env, napi_value v) {
  void *data = nullptr;
  if (isNullish(env, v) || napi_get_value_external(env, v, &data) != napi_ok) return nullptr;
  return static_cast<T *>(data);
}

CtxBox *unwrapCtx(napi_env env, napi_value v) {
  CtxBox *box = unwrap<CtxBox>(env, v);
  if (!box || !box->ctx) {
    napi_throw(env, makeError(env, RMD_E_INVALID, "the simulation was released"));
    return nullptr;
  }
  return box;
}

// A context that no asynchronous call is using: the synchronous entry points
// enqueue on the context's stream and read its host state, which a run,
// minimise or move on a worker thread (or its batch's run) owns until its
// promise settles.
CtxBox *unwrapIdleCtx(napi_env env, napi_value v) {
  CtxBox *box = unwrapCtx(env, v);
  if (box && box->busy) {
    napi_throw(env, makeError(env, RMD_E_INVALID, "another call on this simulation is still running"));
    return nullptr;
  }
  return box;
}

napi_value args(napi_env env, napi_callback_info info, size_t want, napi_value *argv) {
  size_t argc = want;
  for (size_t k = 0; k < want; ++k) napi_get_undefined(env, &argv[k]);
  napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr);
  return nullptr;
}

// ---- host side ----

napi_value ParsePrmtop(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  args(env, info, 1, argv);
  std::string text;
  if (!getBytes(env, argv[0], text)) {
    napi_throw_type_error(env, nullptr, "parsePrmtop expects a string or a Buffer");
    return nullptr;
  }
  rmd_topology *t = nullptr;
  if (!check(env, rmd_topology_parse_prmtop(text.data(), text.size(), &t), nullptr)) return nullptr;
  return wrap(env, t);
}

napi_value ParseInpcrd(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  args(env, info, 2, argv);
  std::string text;
  if (!getBytes(env, argv[0], text)) {
    napi_throw_type_error(env, nullptr, "parseInpcrd expects a string or a Buffer");
    return nullptr;
  }
  const int expected = (int)getNumber(env, argv[1], 0);
  rmd_coordinates *c = nullptr;
  if (!check(env, rmd_coordinates_parse_inpcrd(text.data(), text.size(), expected, &c), nullptr)) return nullptr;
  return wrap(env, c);
}

napi_value CoordinatesView(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  args(env, info, 1, argv);
  rmd_coordinates *c = unwrap<rmd_coordinates>(env, argv[0]);
  rmd_coordinates_view v{};
  if (!check(env, rmd_coordinates_view_get(c, &v), nullptr)) return nullptr;
  napi_value o = newObject(env);
  setNumber(env, o, "atomCount", v.atom_count);
  setValue(env, o, "positions", f64(env, v.positions, 3 * (size_t)v.atom_count));
  if (v.velocities) setValue(env, o, "velocities", f64(env, v.velocities, 3 * (size_t)v.atom_count));
  if (v.has_box) {
    setValue(env, o, "boxLengths", f64(env, v.box_lengths, 3));
    setValue(env, o, "boxAngles", f64(env, v.box_angles, 3));
    double box[9];
    if (!check(env, rmd_box_vectors_from_lengths_and_angles(v.box_lengths, v.box_angles, box), nullptr)) return nullptr;
    setValue(env, o, "boxVectors", f64(env, box, 9));
  }
  setNumber(env, o, "time", v.time);
  return o;
}

napi_value BoxVectors(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  args(env, info, 2, argv);
  std::vector<double> l, a;
  if (!getDoubles(env, argv[0], l) || !getDoubles(env, argv[1], a) || l.size() != 3 || a.size() != 3) {
    napi_throw_type_error(env, nullptr, "boxVectors expects lengths[3] and angles[3]");
    return nullptr;
  }
  double box[9];
  if (!check(env, rmd_box_vectors_from_lengths_and_angles(l.data(), a.data(), box), nullptr)) return nullptr;
  return f64(env, box, 9);
}

int constraintMode(napi_env env, napi_value options) {
  napi_value v;
  if (!propValue(env, options, "constraints", &v)) return RMD_CONSTRAINTS_HBONDS;
  std::string s;
  if (!getBytes(env, v, s)) return RMD_CONSTRAINTS_HBONDS;
  if (s == "none" || s == "None") return RMD_CONSTRAINTS_NONE;
  if (s == "AllBonds" || s == "allbonds") return RMD_CONSTRAINTS_ALLBONDS;
  if (s == "HBonds" || s == "hbonds") return RMD_CONSTRAINTS_HBONDS;
  return -1;
}

napi_value BuildSystem(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  args(env, info, 3, argv);
  rmd_topology *t = unwrap<rmd_topology>(env, argv[0]);
  rmd_coordinates *c = unwrap<rmd_coordinates>(env, argv[1]);
  rmd_build_options o;
  rmd_build_options_default(&o);
  const int mode = constraintMode(env, argv[2]);
  if (mode < 0) {
    napi_throw(env, makeError(env, RMD_E_SYSTEM, "constraints must be none, HBonds or AllBonds (HAngles is refused)"));
    return nullptr;
  }
  o.constraints = mode;
  napi_value v;
  if (propValue(env, argv[2], "rigidWater", &v)) {
    bool b = true;
    napi_get_value_bool(env, v, &b);
    o.rigid_water = b ? 1 : 0;
  }
  o.cutoff = prop(env, argv[2], "cutoff", o.cutoff);
  std::vector<double> box;
  if (propValue(env, argv[2], "boxVectors", &v) && getDoubles(env, v, box) && box.size() == 9) {
    o.has_box_vectors = 1;
    std::memcpy(o.box_vectors, box.data(), sizeof o.box_vectors);
  }
  rmd_system *s = nullptr;
  if (!check(env, rmd_system_build(t, c, &o, &s), nullptr)) return nullptr;
  return wrap(env, s);
}

napi_value BuildSolvated(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  args(env, info, 3, argv);
  std::string prmtop, inpcrd;
  if (!getBytes(env, argv[0], prmtop) || !getBytes(env, argv[1], inpcrd)) {
    napi_throw_type_error(env, nullptr, "buildSolvated expects the template prmtop and solute inpcrd");
    return nullptr;
  }
  rmd_solvated_options o;
  rmd_solvated_options_default(&o);
  o.water_count = (int32_t)prop(env, argv[2], "waterCount", o.water_count);
  o.box_edge = prop(env, argv[2], "boxEdge", o.box_edge);
  o.seed = (uint32_t)prop(env, argv[2], "seed", o.seed);
  napi_value v;
  if (propValue(env, argv[2], "includeSolute", &v)) {
    bool b = true;
    napi_get_value_bool(env, v, &b);
    o.include_solute = b ? 1 : 0;
  }
  o.solute_clearance = prop(env, argv[2], "clearance", o.solute_clearance);
  o.jitter = prop(env, argv[2], "jitter", o.jitter);
  o.cutoff = prop(env, argv[2], "cutoff", o.cutoff);
  const int mode = constraintMode(env, argv[2]);
  if (mode < 0) {
    napi_throw(env, makeError(env, RMD_E_SYSTEM, "constraints must be none, HBonds or AllBonds"));
    return nullptr;
  }
  o.constraints = mode;
  rmd_system *s = nullptr;
  if (!check(env, rmd_system_build_solvated(prmtop.data(), prmtop.size(), inpcrd.data(), inpcrd.size(), &o, &s),
             nullptr))
    return nullptr;
  return wrap(env, s);
}

// The same system in the live box of a finished stage (web/app.js:354-364).
napi_value ReboxSystem(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  args(env, info, 2, argv);
  rmd_system *sys = unwrap<rmd_system>(env, argv[0]);
  std::vector<double> box;
  if (!getDoubles(env, argv[1], box) || box.size() != 9) {
    napi_throw_type_error(env, nullptr, "reboxSystem expects boxVectors[9]");
    return nullptr;
  }
  rmd_system *s = nullptr;
  if (!check(env, rmd_system_rebox(sys, box.data(), &s), nullptr)) return nullptr;
  return wrap(env, s);
}

// soluteHeavyAtoms (shared/topology/restraints.js:46-49).
napi_value SoluteHeavyAtoms(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  args(env, info, 1, argv);
  rmd_system *sys = unwrap<rmd_system>(env, argv[0]);
  int32_t count = 0;
  if (!check(env, rmd_system_solute_heavy_atoms(sys, nullptr, 0, &count), nullptr)) return nullptr;
  std::vector<int32_t> atoms(count > 0 ? count : 1);
  if (!check(env, rmd_system_solute_heavy_atoms(sys, atoms.data(), (int32_t)atoms.size(), &count), nullptr))
    return nullptr;
  return i32(env, atoms.data(), (size_t)count);
}

napi_value SystemView(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  args(env, info, 1, argv);
  rmd_system *s = unwrap<rmd_system>(env, argv[0]);
  rmd_system_view v{};
  if (!check(env, rmd_system_view_get(s, &v), nullptr)) return nullptr;
  const size_t n = v.atom_count, t2 = (size_t)v.type_count * v.type_count;
  napi_value o = newObject(env);
  setNumber(env, o, "atomCount", v.atom_count);
  setNumber(env, o, "typeCount", v.type_count);
  setNumber(env, o, "moleculeCount", v.molecule_count);
  setNumber(env, o, "residueCount", v.residue_count);
  setNumber(env, o, "waterCount", v.water_count);
  setNumber(env, o, "degreesOfFreedom", v.degrees_of_freedom);
  setNumber(env, o, "cutoff", v.cutoff);
  setBool(env, o, "isPeriodic", v.is_periodic != 0);
  setValue(env, o, "boxVectors", f64(env, v.box, 9));
  setValue(env, o, "charges", f64(env, v.charges, n));
  setValue(env, o, "masses", f64(env, v.masses, n));
  setValue(env, o, "atomTypes", i32(env, v.atom_types, n));
  setValue(env, o, "atomicNumbers", i32(env, v.atomic_numbers, n));
  setValue(env, o, "residueOfAtom", i32(env, v.residue_of_atom, n));
  setValue(env, o, "moleculeOfAtom", i32(env, v.molecule_of_atom, n));
  setValue(env, o, "moleculeStart", i32(env, v.molecule_start, (size_t)v.molecule_count + 1));
  setValue(env, o, "moleculeAtoms", i32(env, v.molecule_atoms, n));
  setValue(env, o, "ljA", f64(env, v.lj_a, t2));
  setValue(env, o, "ljB", f64(env, v.lj_b, t2));
  setNumber(env, o, "bondCount", v.bond_count);
  setNumber(env, o, "angleCount", v.angle_count);
  setNumber(env, o, "torsionCount", v.torsion_count);
  setNumber(env, o, "constraintCount", v.constraint_count);
  setNumber(env, o, "exceptionCount", v.exception_count);
  setValue(env, o, "constraintA", i32(env, v.constraint_a, v.constraint_count));
  setValue(env, o, "constraintB", i32(env, v.constraint_b, v.constraint_count));
  setValue(env, o, "constraintDistance", f64(env, v.constraint_distance, v.constraint_count));
  if (v.positions) setValue(env, o, "positions", f64(env, v.positions, 3 * n));
  return o;
}

napi_value PmeParameters(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  args(env, info, 2, argv);
  std::vector<double> box;
  if (!getDoubles(env, argv[0], box) || box.size() != 9) {
    napi_throw_type_error(env, nullptr, "pmeParameters expects boxVectors[9]");
    return nullptr;
  }
  const double cutoff = prop(env, argv[1], "cutoff", 1.0);
  const double tolerance = prop(env, argv[1], "tolerance", 5e-4);
  const double alpha = prop(env, argv[1], "alpha", 0.0);
  int32_t grid[3] = {0, 0, 0};
  napi_value g;
  std::vector<int32_t> gv;
  if (propValue(env, argv[1], "grid", &g) && getInts(env, g, gv) && gv.size() == 3)
    for (int k = 0; k < 3; ++k) grid[k] = gv[k];
  rmd_pme_parameters p{};
  if (!check(env, rmd_pme_parameters_derive(box.data(), cutoff, tolerance, alpha, grid, &p), nullptr)) return nullptr;
  napi_value o = newObject(env);
  setNumber(env, o, "alpha", p.alpha);
  setValue(env, o, "grid", i32(env, p.grid, 3));
  setNumber(env, o, "order", p.order);
  setNumber(env, o, "tolerance", p.tolerance);
  return o;
}

// ---- the simulation context ----

void readOptions(napi_env env, napi_value opt, rmd_options &o, std::vector<int32_t> &restrained,
                 std::vector<double> &reference);

napi_value Create(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  args(env, info, 2, argv);
  rmd_system *s = unwrap<rmd_system>(env, argv[0]);
  if (!s) {
    napi_throw_type_error(env, nullptr, "create expects a built system");
    return nullptr;
  }
  rmd_options o;
  std::vector<int32_t> restrained;
  std::vector<double> reference;
  readOptions(env, argv[1], o, restrained, reference);
  rmd_ctx *ctx = nullptr;
  if (!check(env, rmd_create(s, &o, &ctx), nullptr)) return nullptr;
  CtxBox *box = new CtxBox();
  box->ctx = ctx;
  napi_value out;
  napi_create_external(env, box, finalizeCtx, nullptr, &out);
  return out;
}

// The reference's constructor options (webgpu/simulation.js:82-180) into
// rmd_options; the restraint arrays live in the caller's vectors.
void readOptions(napi_env env, napi_value opt, rmd_options &o, std::vector<int32_t> &restrained,
                 std::vector<double> &reference) {
  rmd_options_default(&o);
  o.timestep = prop(env, opt, "timestep", o.timestep);
  o.temperature = prop(env, opt, "temperature", o.temperature);
  o.friction = prop(env, opt, "friction", o.friction);
  o.seed = (uint32_t)prop(env, opt, "seed", o.seed);
  o.tolerance = prop(env, opt, "tolerance", o.tolerance);
  o.max_constraint_iterations = (int32_t)prop(env, opt, "maxConstraintIterations", o.max_constraint_iterations);
  o.rebuild_interval = (int32_t)prop(env, opt, "rebuildInterval", o.rebuild_interval);
  napi_value v;
  if (propValue(env, opt, "adaptiveRebuild", &v)) {
    bool b = true;
    napi_get_value_bool(env, v, &b);
    o.adaptive_rebuild = b ? 1 : 0;
  }
  if (propValue(env, opt, "pressure", &v)) {
    o.has_pressure = 1;
    o.pressure = getNumber(env, v, 1.0);
  }
  o.barostat_interval = (int32_t)prop(env, opt, "barostatInterval", o.barostat_interval);
  if (propValue(env, opt, "dispersionCorrection", &v)) {
    bool b = true;
    napi_get_value_bool(env, v, &b);
    o.dispersion_correction = b ? 1 : 0;
  }
  o.restraint_force = prop(env, opt, "restraintForce", 0.0);
  if (propValue(env, opt, "restraintAtoms", &v)) getInts(env, v, restrained);
  if (propValue(env, opt, "restraintReference", &v)) getDoubles(env, v, reference);
  o.restraint_count = (int32_t)restrained.size();
  o.restraint_atoms = restrained.empty() ? nullptr : restrained.data();
  o.restraint_reference = reference.empty() ? nullptr : reference.data();
  o.skin = prop(env, opt, "skin", o.skin);
  o.neighbour_slack = prop(env, opt, "neighbourSlack", o.neighbour_slack);
  o.max_neighbours = (int32_t)prop(env, opt, "maxNeighbours", o.max_neighbours);
  o.ewald_tolerance = prop(env, opt, "ewaldTolerance", o.ewald_tolerance);
  o.alpha = prop(env, opt, "alpha", o.alpha);
  std::vector<int32_t> grid;
  if (propValue(env, opt, "grid", &v) && getInts(env, v, grid) && grid.size() == 3)
    for (int k = 0; k < 3; ++k) o.grid[k] = grid[k];
  o.device = (int32_t)prop(env, opt, "device", o.device);
  o.lanes_per_atom = (int32_t)prop(env, opt, "lanesPerAtom", o.lanes_per_atom);
  o.cell_reach = (int32_t)prop(env, opt, "cellReach", o.cell_reach);
  o.search_mode = (int32_t)prop(env, opt, "searchMode", o.search_mode);
  o.pme_bins = (int32_t)prop(env, opt, "pmeBins", o.pme_bins);
}

napi_value Destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  args(env, info, 1, argv);
  CtxBox *box = unwrap<CtxBox>(env, argv[0]);
  if (box && box->ctx) {
    if (!box->owned) {
      napi_throw(env, makeError(env, RMD_E_INVALID, "a batch replica is released with its batch"));
      return nullptr;
    }
    if (box->busy) {
      napi_throw(env, makeError(env, RMD_E_INVALID, "release() while a call is still running"));
      return nullptr;
    }
    rmd_destroy(box->ctx);
    box->ctx = nullptr;
  }
  return undefinedValue(env);
}

napi_value SetState(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  args(env, info, 4, argv);
  CtxBox *box = unwrapIdleCtx(env, argv[0]);
  if (!box) return nullptr;
  std::vector<double> pos, vel;
  if (!getDoubles(env, argv[1], pos)) {
    napi_throw_type_error(env, nullptr, "setState expects positions");
    return nullptr;
  }
  const bool hasVel = !isNullish(env, argv[2]) && getDoubles(env, argv[2], vel);
  const int64_t step = (int64_t)prop(env, argv[3], "step", 0);
  napi_value sn;
  bool supplied = false;
  if (propValue(env, argv[3], "suppliedNoise", &sn)) napi_get_value_bool(env, sn, &supplied);
  if (!check(env, rmd_set_state(box->ctx, pos.data(), hasVel ? vel.data() : nullptr, step, supplied ? 1 : 0), box->ctx))
    return nullptr;
  return undefinedValue(env);
}

napi_value SetNoise(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  args(env, info, 2, argv);
  CtxBox *box = unwrapIdleCtx(env, argv[0]);
  if (!box) return nullptr;
  std::vector<double> d;
  if (!getDoubles(env, argv[1], d)) {
    napi_throw_type_error(env, nullptr, "setNoise expects 3N numbers");
    return nullptr;
  }
  std::vector<float> f(d.begin(), d.end());
  if (!check(env, rmd_set_noise(box->ctx, f.data()), box->ctx)) return nullptr;
  return undefinedValue(env);
}

napi_value RequestStop(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  args(env, info, 1, argv);
  CtxBox *box = unwrapCtx(env, argv[0]);
  if (!box) return nullptr;
  rmd_request_stop(box->ctx);
  return undefinedValue(env);
}

napi_value EnergyObject(napi_env env, const rmd_energy_terms &e) {
  napi_value o = newObject(env);
  setNumber(env, o, "step", (double)e.step);
  setNumber(env, o, "time", e.time);
  setNumber(env, o, "bonded", e.bonded);
  setNumber(env, o, "harmonicBond", e.harmonic_bond);
  setNumber(env, o, "harmonicAngle", e.harmonic_angle);
  setNumber(env, o, "periodicTorsion", e.periodic_torsion);
  setNumber(env, o, "lennardJones", e.lennard_jones);
  setNumber(env, o, "dispersion", e.dispersion);
  setNumber(env, o, "coulombDirect", e.coulomb_direct);
  setNumber(env, o, "coulombReciprocal", e.coulomb_reciprocal);
  setNumber(env, o, "coulombSelf", e.coulomb_self);
  setNumber(env, o, "coulomb", e.coulomb);
  setNumber(env, o, "potential", e.potential);
  setNumber(env, o, "kinetic", e.kinetic);
  setNumber(env, o, "total", e.total);
  setNumber(env, o, "positionRestraint", e.position_restraint);
  setNumber(env, o, "restrainedAtoms", e.restrained_atoms);
  setNumber(env, o, "temperature", e.temperature);
  setNumber(env, o, "kineticForTemperature", e.kinetic_for_temperature);
  napi_value sampled;
  const char *at = e.kinetic_sampled_in_step ? "before the closing half kick" : "at the current velocities";
  napi_create_string_utf8(env, at, NAPI_AUTO_LENGTH, &sampled);
  setValue(env, o, "kineticSampledAt", sampled);
  setNumber(env, o, "volume", e.volume);
  setValue(env, o, "boxVectors", f64(env, e.box, 9));
  return o;
}

napi_value Energy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  args(env, info, 1, argv);
  CtxBox *box = unwrapIdleCtx(env, argv[0]);
  if (!box) return nullptr;
  rmd_energy_terms e{};
  if (!check(env, rmd_energy(box->ctx, &e), box->ctx)) return nullptr;
  return EnergyObject(env, e);
}

// The context's atom count, which sizes every per-atom read; a count the
// caller passes must agree with it.
bool atomCount(napi_env env, CtxBox *box, napi_value given, size_t &n) {
  const int32_t count = rmd_atom_count(box->ctx);
  if (count < 0) {
    napi_throw(env, makeError(env, count, "the simulation has no atom count"));
    return false;
  }
  if (!isNullish(env, given) && (size_t)getNumber(env, given, count) != (size_t)count) {
    napi_throw(env, makeError(env, RMD_E_INVALID, "the atom count given does not match the simulation's"));
    return false;
  }
  n = (size_t)count;
  return true;
}

napi_value ReadPositions(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  args(env, info, 2, argv);
  CtxBox *box = unwrapIdleCtx(env, argv[0]);
  if (!box) return nullptr;
  size_t n = 0;
  if (!atomCount(env, box, argv[1], n)) return nullptr;
  std::vector<float> p(3 * n);
  if (!check(env, rmd_read_positions(box->ctx, p.data()), box->ctx)) return nullptr;
  return f32(env, p.data(), p.size());
}

napi_value ReadState(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  args(env, info, 2, argv);
  CtxBox *box = unwrapIdleCtx(env, argv[0]);
  if (!box) return nullptr;
  size_t n = 0;
  if (!atomCount(env, box, argv[1], n)) return nullptr;
  std::vector<float> p(3 * n), v(3 * n), f(3 * n);
  double b[9], volume = 0;
  if (!check(env, rmd_read_state(box->ctx, p.data(), v.data(), f.data(), b, &volume), box->ctx)) return nullptr;
  napi_value o = newObject(env);
  setValue(env, o, "positions", f32(env, p.data(), p.size()));
  setValue(env, o, "velocities", f32(env, v.data(), v.size()));
  setValue(env, o, "forces", f32(env, f.data(), f.size()));
  setValue(env, o, "boxVectors", f64(env, b, 9));
  setNumber(env, o, "volume", volume);
  return o;
}

napi_value Diagnostics(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  args(env, info, 1, argv);
  CtxBox *box = unwrapIdleCtx(env, argv[0]);
  if (!box) return nullptr;
  rmd_diagnostics d{};
  if (!check(env, rmd_diagnostics_get(box->ctx, &d), box->ctx)) return nullptr;
  napi_value o = newObject(env);
  setNumber(env, o, "maxShakeIterations", d.max_shake_iterations);
  setBool(env, o, "shakeHitIterationCap", d.shake_hit_iteration_cap != 0);
  setBool(env, o, "rattleHitIterationCap", d.rattle_hit_iteration_cap != 0);
  setNumber(env, o, "neighbourRequired", d.neighbour_required);
  setNumber(env, o, "neighbourCapacity", d.neighbour_capacity);
  setNumber(env, o, "worstNeighbourRequired", d.worst_neighbour_required);
  setNumber(env, o, "neighbourGrowths", d.neighbour_growths);
  setNumber(env, o, "skin", d.skin);
  setNumber(env, o, "worstDisplacement", d.worst_displacement);
  setNumber(env, o, "rebuilds", (double)d.rebuilds);
  setNumber(env, o, "steps", (double)d.steps);
  setValue(env, o, "cellCounts", i32(env, d.cell_counts, 3));
  setNumber(env, o, "volume", d.volume);
  napi_value baro = newObject(env);
  setNumber(env, baro, "attempts", d.barostat_attempts);
  setNumber(env, baro, "accepted", d.barostat_accepted);
  setNumber(env, baro, "refusedForCutoff", d.barostat_refused);
  setNumber(env, baro, "volumeScale", d.volume_scale);
  setValue(env, o, "barostat", baro);
  setNumber(env, o, "constraintTolerance", d.constraint_tolerance);
  setValue(env, o, "pmeGrid", i32(env, d.pme_grid, 3));
  setNumber(env, o, "pmeAlpha", d.pme_alpha);
  setNumber(env, o, "lanesPerAtom", d.lanes_per_atom);
  setNumber(env, o, "waterUnits", d.water_units);
  setNumber(env, o, "clusterUnits", d.cluster_units);
  setNumber(env, o, "freeUnits", d.free_units);
  setBool(env, o, "pmePlanes", d.pme_planes != 0);
  setNumber(env, o, "cellReach", d.cell_reach);
  setBool(env, o, "searchByCell", d.search_by_cell != 0);
  setNumber(env, o, "stepsSinceBarostat", (double)d.steps_since_barostat);
  setBool(env, o, "pmeXFolded", d.pme_x_folded != 0);
  return o;
}

// ---- async calls ----

enum AsyncKind { A_INITIALISE, A_RUN, A_MINIMISE, A_SCALE, A_MOVE };

struct AsyncCall {
  AsyncKind kind;
  CtxBox *box;
  napi_ref boxRef = nullptr;
  napi_deferred deferred = nullptr;
  napi_async_work work = nullptr;
  int status = RMD_OK;
  std::string error;
  // arguments and results
  int64_t steps = 0;
  int32_t block = 0;
  int64_t stepOut = 0;
  rmd_min_options minOptions{};
  rmd_min_result minResult{};
  double scale = 1.0, potential = 0.0;
  rmd_move move{};
};

void executeAsync(napi_env, void *data) {
  AsyncCall *call = static_cast<AsyncCall *>(data);
  rmd_ctx *ctx = call->box->ctx;
  switch (call->kind) {
    case A_INITIALISE: call->status = rmd_initialise(ctx); break;
    case A_RUN: call->status = rmd_run(ctx, call->steps, call->block, &call->stepOut); break;
    case A_MINIMISE: call->status = rmd_minimise(ctx, &call->minOptions, &call->minResult); break;
    case A_SCALE: call->status = rmd_scale_volume(ctx, call->scale, &call->potential); break;
    case A_MOVE: call->status = rmd_attempt_volume_change(ctx, &call->move); break;
  }
  if (call->status != RMD_OK) {
    const char *m = rmd_last_error(ctx);
    call->error = m ? m : "rainmd error";
  }
}

void completeAsync(napi_env env, napi_status, void *data) {
  AsyncCall *call = static_cast<AsyncCall *>(data);
  call->box->busy = 0;
  if (call->status != RMD_OK) {
    napi_reject_deferred(env, call->deferred, makeError(env, call->status, call->error));
  } else {
    napi_value result = undefinedValue(env);
    switch (call->kind) {
      case A_INITIALISE: break;
      case A_RUN: napi_create_double(env, (double)call->stepOut, &result); break;
      case A_MINIMISE: {
        result = newObject(env);
        const rmd_min_result &r = call->minResult;
        setNumber(env, result, "iterations", r.iterations);
        setNumber(env, result, "rejected", r.rejected);
        setNumber(env, result, "potential", r.potential);
        setNumber(env, result, "maxForce", r.max_force);
        setNumber(env, result, "maxStep", r.max_step);
        setBool(env, result, "stalled", r.stalled != 0);
        setBool(env, result, "converged", r.converged != 0);
        setNumber(env, result, "energyResolution", r.energy_resolution);
        break;
      }
      case A_SCALE: napi_create_double(env, call->potential, &result); break;
      case A_MOVE: {
        const rmd_move &m = call->move;
        if (!m.attempted) {
          napi_get_null(env, &result);
          break;
        }
        result = newObject(env);
        setBool(env, result, "accepted", m.accepted != 0);
        if (m.refused) setBool(env, result, "refused", true);
        setNumber(env, result, "weight", m.weight);
        setNumber(env, result, "deltaVolume", m.delta_volume);
        setNumber(env, result, "lengthScale", m.length_scale);
        setNumber(env, result, "oldPotential", m.old_potential);
        setNumber(env, result, "newPotential", m.new_potential);
        setNumber(env, result, "volume", m.volume);
        setNumber(env, result, "volumeScale", m.volume_scale);
        setNumber(env, result, "acceptanceRate", m.acceptance_rate);
        break;
      }
    }
    napi_resolve_deferred(env, call->deferred, result);
  }
  napi_delete_async_work(env, call->work);
  if (call->boxRef) napi_delete_reference(env, call->boxRef);
  delete call;
}

napi_value startAsync(napi_env env, napi_value handle, AsyncCall *call, const char *name) {
  CtxBox *box = unwrapCtx(env, handle);
  if (!box) {
    delete call;
    return nullptr;
  }
  if (box->busy) {
    delete call;
    napi_throw(env, makeError(env, RMD_E_INVALID, "another call on this simulation is still running"));
    return nullptr;
  }
  call->box = box;
  box->busy = 1;
  napi_value promise, resource;
  napi_create_promise(env, &call->deferred, &promise);
  napi_create_reference(env, handle, 1, &call->boxRef);
  napi_create_string_utf8(env, name, NAPI_AUTO_LENGTH, &resource);
  napi_create_async_work(env, nullptr, resource, executeAsync, completeAsync, call, &call->work);
  napi_queue_async_work(env, call->work);
  return promise;
}

napi_value Initialise(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  args(env, info, 1, argv);
  AsyncCall *call = new AsyncCall();
  call->kind = A_INITIALISE;
  return startAsync(env, argv[0], call, "rainmd.initialise");
}

napi_value Run(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  args(env, info, 3, argv);
  AsyncCall *call = new AsyncCall();
  call->kind = A_RUN;
  call->steps = (int64_t)getNumber(env, argv[1], 0);
  call->block = (int32_t)getNumber(env, argv[2], 0);
  return startAsync(env, argv[0], call, "rainmd.run");
}

napi_value Minimise(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  args(env, info, 2, argv);
  AsyncCall *call = new AsyncCall();
  call->kind = A_MINIMISE;
  rmd_min_options_default(&call->minOptions);
  call->minOptions.max_iterations = (int32_t)prop(env, argv[1], "maxIterations", call->minOptions.max_iterations);
  call->minOptions.force_tolerance = prop(env, argv[1], "forceTolerance", call->minOptions.force_tolerance);
  call->minOptions.max_step = prop(env, argv[1], "maxStep", call->minOptions.max_step);
  call->minOptions.max_displacement = prop(env, argv[1], "maxDisplacement", call->minOptions.max_displacement);
  return startAsync(env, argv[0], call, "rainmd.minimise");
}

napi_value ScaleVolume(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  args(env, info, 2, argv);
  AsyncCall *call = new AsyncCall();
  call->kind = A_SCALE;
  call->scale = getNumber(env, argv[1], 1.0);
  return startAsync(env, argv[0], call, "rainmd.scaleVolume");
}

napi_value AttemptVolumeChange(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  args(env, info, 1, argv);
  AsyncCall *call = new AsyncCall();
  call->kind = A_MOVE;
  return startAsync(env, argv[0], call, "rainmd.attemptVolumeChange");
}

// ---- replica batches (rmd_batch_*) ----
// A batch handle owns the boxes of its replicas' contexts: a replica view
// (batchReplica) is a non-owning external of such a box, usable through every
// context entry point above while the batch lives and is not running. The
// JavaScript facade keeps the batch alive from each view.

struct BatchBox {
  rmd_batch *batch = nullptr;
  int busy = 0;
  std::vector<CtxBox *> members;
};
void finalizeBatch(napi_env, void *data, void *) {
  BatchBox *b = static_cast<BatchBox *>(data);
  if (b->batch) rmd_batch_destroy(b->batch);
  for (CtxBox *m : b->members) delete m;
  delete b;
}
void finalizeView(napi_env, void *, void *) {}

BatchBox *unwrapBatch(napi_env env, napi_value v) {
  BatchBox *b = unwrap<BatchBox>(env, v);
  if (!b || !b->batch) {
    napi_throw(env, makeError(env, RMD_E_INVALID, "the batch was released"));
    return nullptr;
  }
  return b;
}

bool checkBatch(napi_env env, int status, rmd_batch *b) {
  if (status == RMD_OK) return true;
  const char *m = rmd_batch_last_error(b);
  napi_throw(env, makeError(env, status, m ? m : "rainmd error"));
  return false;
}

bool batchIdle(napi_env env, BatchBox *b) {
  if (!b->busy) return true;
  napi_throw(env, makeError(env, RMD_E_INVALID, "another call on this batch is still running"));
  return false;
}

// batchCreate(systems[], options, seeds[] | null)
napi_value BatchCreate(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  args(env, info, 3, argv);
  bool isArray = false;
  napi_is_array(env, argv[0], &isArray);
  if (!isArray) {
    napi_throw_type_error(env, nullptr, "batchCreate expects an array of built systems");
    return nullptr;
  }
  uint32_t count = 0;
  napi_get_array_length(env, argv[0], &count);
  std::vector<const rmd_system *> systems(count);
  for (uint32_t k = 0; k < count; ++k) {
    napi_value e;
    napi_get_element(env, argv[0], k, &e);
    systems[k] = unwrap<rmd_system>(env, e);
    if (!systems[k]) {
      napi_throw_type_error(env, nullptr, "batchCreate expects an array of built systems");
      return nullptr;
    }
  }
  rmd_options o;
  std::vector<int32_t> restrained;
  std::vector<double> reference;
  readOptions(env, argv[1], o, restrained, reference);
  std::vector<int32_t> given;
  std::vector<uint32_t> seeds;
  if (!isNullish(env, argv[2]) && getInts(env, argv[2], given)) {
    if (given.size() != count) {
      napi_throw(env, makeError(env, RMD_E_INVALID, "batchCreate needs one seed per system"));
      return nullptr;
    }
    for (int32_t v : given) seeds.push_back((uint32_t)v);
  }
  rmd_batch *batch = nullptr;
  if (!check(env, rmd_batch_create(systems.data(), (int32_t)count, &o, seeds.empty() ? nullptr : seeds.data(), &batch),
             nullptr))
    return nullptr;
  BatchBox *b = new BatchBox();
  b->batch = batch;
  for (uint32_t r = 0; r < count; ++r) {
    CtxBox *m = new CtxBox();
    m->ctx = rmd_batch_replica(batch, (int32_t)r);
    m->owned = false;
    b->members.push_back(m);
  }
  napi_value out;
  napi_create_external(env, b, finalizeBatch, nullptr, &out);
  return out;
}

napi_value BatchDestroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  args(env, info, 1, argv);
  BatchBox *b = unwrap<BatchBox>(env, argv[0]);
  if (b && b->batch) {
    if (!batchIdle(env, b)) return nullptr;
    for (CtxBox *m : b->members)
      if (m->busy) {
        napi_throw(env, makeError(env, RMD_E_INVALID, "release() while a call on a replica is still running"));
        return nullptr;
      }
    rmd_batch_destroy(b->batch);
    b->batch = nullptr;
    for (CtxBox *m : b->members) m->ctx = nullptr;
  }
  return undefinedValue(env);
}

napi_value BatchReplica(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  args(env, info, 2, argv);
  BatchBox *b = unwrapBatch(env, argv[0]);
  if (!b) return nullptr;
  const double r = getNumber(env, argv[1], -1);
  if (!(r >= 0 && r < (double)b->members.size())) {
    napi_throw(env, makeError(env, RMD_E_INVALID, "no such replica"));
    return nullptr;
  }
  napi_value out;
  napi_create_external(env, b->members[(size_t)r], finalizeView, nullptr, &out);
  return out;
}

napi_value BatchInfo(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  args(env, info, 1, argv);
  BatchBox *b = unwrapBatch(env, argv[0]);
  if (!b) return nullptr;
  rmd_batch_info i;
  if (!checkBatch(env, rmd_batch_info_get(b->batch, &i), b->batch)) return nullptr;
  napi_value o = newObject(env);
  setNumber(env, o, "replicas", i.replicas);
  setNumber(env, o, "atoms", (double)i.atoms);
  setNumber(env, o, "blocksRun", (double)i.blocks_run);
  setNumber(env, o, "stepWorkgroups", (double)i.step_workgroups);
  setNumber(env, o, "nonbondedWorkgroups", (double)i.nonbonded_workgroups);
  return o;
}

napi_value BatchSwitch(napi_env env, napi_callback_info info, int (*fn)(rmd_batch *, int32_t)) {
  napi_value argv[2];
  args(env, info, 2, argv);
  BatchBox *b = unwrapBatch(env, argv[0]);
  if (!b || !batchIdle(env, b)) return nullptr;
  bool on = true;
  napi_get_value_bool(env, argv[1], &on);
  if (!checkBatch(env, fn(b->batch, on ? 1 : 0), b->batch)) return nullptr;
  return undefinedValue(env);
}
napi_value BatchSetGraphs(napi_env env, napi_callback_info info) {
  return BatchSwitch(env, info, rmd_batch_set_graphs);
}
napi_value BatchSetConcurrentForces(napi_env env, napi_callback_info info) {
  return BatchSwitch(env, info, rmd_batch_set_concurrent_forces);
}

napi_value BatchRequestStop(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  args(env, info, 1, argv);
  BatchBox *b = unwrap<BatchBox>(env, argv[0]);
  if (b && b->batch) rmd_batch_request_stop(b->batch);
  return undefinedValue(env);
}

// batchScreenConfigure(batch, {enabled, trainingMoves, trainingSteps, learningRate, seed})
napi_value BatchScreenConfigure(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  args(env, info, 2, argv);
  BatchBox *b = unwrapBatch(env, argv[0]);
  if (!b || !batchIdle(env, b)) return nullptr;
  rmd_screen_options o;
  rmd_screen_options_default(&o);
  napi_value v;
  if (propValue(env, argv[1], "enabled", &v)) {
    bool e = true;
    napi_get_value_bool(env, v, &e);
    o.enabled = e ? 1 : 0;
  }
  o.training_moves = (int32_t)prop(env, argv[1], "trainingMoves", o.training_moves);
  o.training_steps = (int32_t)prop(env, argv[1], "trainingSteps", o.training_steps);
  o.learning_rate = (float)prop(env, argv[1], "learningRate", o.learning_rate);
  o.seed = (uint32_t)prop(env, argv[1], "seed", o.seed);
  if (!checkBatch(env, rmd_batch_screen_configure(b->batch, &o), b->batch)) return nullptr;
  return undefinedValue(env);
}

napi_value BatchScreenStats(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  args(env, info, 1, argv);
  BatchBox *b = unwrapBatch(env, argv[0]);
  if (!b) return nullptr;
  rmd_screen_stats st;
  if (!checkBatch(env, rmd_batch_screen_stats_get(b->batch, &st), b->batch)) return nullptr;
  napi_value o = newObject(env);
  setNumber(env, o, "phase", st.phase);
  setNumber(env, o, "samples", (double)st.samples);
  setNumber(env, o, "moves", (double)st.moves);
  setNumber(env, o, "exactEvaluations", (double)st.exact_evaluations);
  setNumber(env, o, "screenedOut", (double)st.screened_out);
  setNumber(env, o, "secondStageRejections", (double)st.second_stage_rejections);
  setNumber(env, o, "accepted", (double)st.accepted);
  setNumber(env, o, "trainingLoss", st.training_loss);
  setNumber(env, o, "residualRms", st.residual_rms);
  setNumber(env, o, "estimateRms", st.estimate_rms);
  return o;
}

// initialise and run, off the JavaScript thread; the batch and its replicas
// are busy until the promise settles
enum BatchKind { B_INITIALISE, B_RUN };
struct BatchCall {
  BatchKind kind;
  BatchBox *box = nullptr;
  napi_ref boxRef = nullptr;
  napi_deferred deferred = nullptr;
  napi_async_work work = nullptr;
  int status = RMD_OK;
  std::string error;
  int64_t steps = 0;
  int32_t block = 0;
  int64_t stepOut = 0;
};

void setBatchBusy(BatchBox *b, int busy) {
  b->busy = busy;
  for (CtxBox *m : b->members) m->busy = busy;
}

void executeBatch(napi_env, void *data) {
  BatchCall *call = static_cast<BatchCall *>(data);
  rmd_batch *batch = call->box->batch;
  if (call->kind == B_INITIALISE)
    call->status = rmd_batch_initialise(batch);
  else
    call->status = rmd_batch_run(batch, call->steps, call->block, &call->stepOut);
  if (call->status != RMD_OK) {
    const char *m = rmd_batch_last_error(batch);
    call->error = m ? m : "rainmd error";
  }
}

void completeBatch(napi_env env, napi_status, void *data) {
  BatchCall *call = static_cast<BatchCall *>(data);
  setBatchBusy(call->box, 0);
  if (call->status != RMD_OK) {
    napi_reject_deferred(env, call->deferred, makeError(env, call->status, call->error));
  } else {
    napi_value result = undefinedValue(env);
    if (call->kind == B_RUN) napi_create_double(env, (double)call->stepOut, &result);
    napi_resolve_deferred(env, call->deferred, result);
  }
  napi_delete_async_work(env, call->work);
  if (call->boxRef) napi_delete_reference(env, call->boxRef);
  delete call;
}

napi_value startBatch(napi_env env, napi_value handle, BatchCall *call, const char *name) {
  BatchBox *b = unwrapBatch(env, handle);
  bool idle = b && batchIdle(env, b);
  if (idle)
    for (CtxBox *m : b->members)
      if (m->busy) {
        napi_throw(env, makeError(env, RMD_E_INVALID, "a call on a replica of this batch is still running"));
        idle = false;
        break;
      }
  if (!idle) {
    delete call;
    return nullptr;
  }
  call->box = b;
  setBatchBusy(b, 1);
  napi_value promise, resource;
  napi_create_promise(env, &call->deferred, &promise);
  napi_create_reference(env, handle, 1, &call->boxRef);
  napi_create_string_utf8(env, name, NAPI_AUTO_LENGTH, &resource);
  napi_create_async_work(env, nullptr, resource, executeBatch, completeBatch, call, &call->work);
  napi_queue_async_work(env, call->work);
  return promise;
}

napi_value BatchInitialise(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  args(env, info, 1, argv);
  BatchCall *call = new BatchCall();
  call->kind = B_INITIALISE;
  return startBatch(env, argv[0], call, "rainmd.batchInitialise");
}

// batchRun(batch, steps, block) -> the replicas' step afterwards
napi_value BatchRun(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  args(env, info, 3, argv);
  BatchCall *call = new BatchCall();
  call->kind = B_RUN;
  call->steps = (int64_t)getNumber(env, argv[1], 0);
  call->block = (int32_t)getNumber(env, argv[2], 0);
  return startBatch(env, argv[0], call, "rainmd.batchRun");
}

napi_value DeviceInfo(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  args(env, info, 1, argv);
  rmd_device_info d{};
  const int st = rmd_device_info_get((int32_t)getNumber(env, argv[0], 0), &d);
  napi_value o = newObject(env);
  setBool(env, o, "ok", st == RMD_OK);
  setNumber(env, o, "deviceCount", d.device_count);
  napi_value s;
  napi_create_string_utf8(env, d.name, NAPI_AUTO_LENGTH, &s);
  setValue(env, o, "name", s);
  napi_create_string_utf8(env, d.arch, NAPI_AUTO_LENGTH, &s);
  setValue(env, o, "arch", s);
  setNumber(env, o, "computeUnits", d.compute_units);
  return o;
}

napi_value AbiVersion(napi_env env, napi_callback_info) {
  napi_value v;
  napi_create_int32(env, rmd_abi_version(), &v);
  return v;
}

napi_value Init(napi_env env, napi_value exports) {
  struct {
    const char *name;
    napi_callback fn;
  } table[] = {
      {"abiVersion", AbiVersion},
      {"deviceInfo", DeviceInfo},
      {"parsePrmtop", ParsePrmtop},
      {"parseInpcrd", ParseInpcrd},
      {"coordinatesView", CoordinatesView},
      {"boxVectors", BoxVectors},
      {"buildSystem", BuildSystem},
      {"buildSolvated", BuildSolvated},
      {"systemView", SystemView},
      {"reboxSystem", ReboxSystem},
      {"soluteHeavyAtoms", SoluteHeavyAtoms},
      {"pmeParameters", PmeParameters},
      {"create", Create},
      {"destroy", Destroy},
      {"setState", SetState},
      {"setNoise", SetNoise},
      {"requestStop", RequestStop},
      {"initialise", Initialise},
      {"run", Run},
      {"minimise", Minimise},
      {"scaleVolume", ScaleVolume},
      {"attemptVolumeChange", AttemptVolumeChange},
      {"batchCreate", BatchCreate},
      {"batchDestroy", BatchDestroy},
      {"batchReplica", BatchReplica},
      {"batchInfo", BatchInfo},
      {"batchSetGraphs", BatchSetGraphs},
      {"batchSetConcurrentForces", BatchSetConcurrentForces},
      {"batchRequestStop", BatchRequestStop},
      {"batchScreenConfigure", BatchScreenConfigure},
      {"batchScreenStats", BatchScreenStats},
      {"batchInitialise", BatchInitialise},
      {"batchRun", BatchRun},
      {"energy", Energy},
      {"readPositions", ReadPositions},
      {"readState", ReadState},
      {"diagnostics", Diagnostics},
  };
  for (const auto &entry : table) {
    napi_value fn;
    NAPI_CALL(env, napi_create_function(env, entry.name, NAPI_AUTO_LENGTH, entry.fn, nullptr, &fn));
    NAPI_CALL(env, napi_set_named_property(env, exports, entry.name, fn));
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
