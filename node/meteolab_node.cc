// N-API binding of the C ABI in include/meteolab.h for a Node host (the reference's host runtime is Node:
// src/simulation.worker.ts drives AtmosphereModel.step and the GPU plugin GpuDynamics, src/gpu/dynamics.ts:250-504).
// Plain functions over the ABI; node/meteolab.js builds the host classes on them. An ABI error becomes a thrown Error
// carrying ml_last_error()'s message.
#include <node_api.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../include/meteolab.h"

namespace {

#define NAPI_OK(env, call)                                                     \
  do {                                                                         \
    if ((call) != napi_ok) { napi_throw_error(env, nullptr, #call " failed"); return nullptr; } \
  } while (0)

napi_value undefinedOf(napi_env env) { napi_value u; napi_get_undefined(env, &u); return u; }

bool args(napi_env env, napi_callback_info info, size_t want, napi_value *out) {
  size_t n = want;
  if (napi_get_cb_info(env, info, &n, out, nullptr, nullptr) != napi_ok || n < want) {
    napi_throw_type_error(env, nullptr, "missing arguments");
    return false;
  }
  return true;
}

ml_ctx *ctxOf(napi_env env, napi_value v) {
  void *p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) { napi_throw_type_error(env, nullptr, "not a meteolab context"); return nullptr; }
  return static_cast<ml_ctx *>(p);
}

bool intOf(napi_env env, napi_value v, int32_t &out) {
  if (napi_get_value_int32(env, v, &out) != napi_ok) { napi_throw_type_error(env, nullptr, "expected a number"); return false; }
  return true;
}

bool numberOf(napi_env env, napi_value v, double &out) {
  if (napi_get_value_double(env, v, &out) != napi_ok) { napi_throw_type_error(env, nullptr, "expected a number"); return false; }
  return true;
}

double prop(napi_env env, napi_value obj, const char *name, double fallback) {
  bool has = false;
  if (napi_has_named_property(env, obj, name, &has) != napi_ok || !has) return fallback;
  napi_value v; double d = fallback;
  if (napi_get_named_property(env, obj, name, &v) == napi_ok) napi_get_value_double(env, v, &d);
  return d;
}
// Elements 0..n-1 of the array property `name` into out (zeros where it or an element is absent).
void propArray(napi_env env, napi_value obj, const char *name, double *out, uint32_t n) {
  for (uint32_t k = 0; k < n; k++) out[k] = 0;
  bool has = false;
  if (napi_has_named_property(env, obj, name, &has) != napi_ok || !has) return;
  napi_value arr; if (napi_get_named_property(env, obj, name, &arr) != napi_ok) return;
  for (uint32_t k = 0; k < n; k++) {
    napi_value v; bool el = false;
    if (napi_has_element(env, arr, k, &el) == napi_ok && el && napi_get_element(env, arr, k, &v) == napi_ok) napi_get_value_double(env, v, &out[k]);
  }
}

// A typed array's data of the expected element type and at least `count` elements.
template <class T> bool typed(napi_env env, napi_value v, napi_typedarray_type type, size_t count, T *&data) {
  napi_typedarray_type t; size_t len; void *raw; napi_value buf; size_t off;
  if (napi_get_typedarray_info(env, v, &t, &len, &raw, &buf, &off) != napi_ok || t != type || len < count) {
    napi_throw_type_error(env, nullptr, "typed array of the wrong type or length");
    return false;
  }
  data = static_cast<T *>(raw);
  return true;
}

napi_value status(napi_env env, ml_ctx *ctx, int rc) {
  if (rc < 0) { napi_throw_error(env, nullptr, (std::string("meteolab: ") + ml_last_error(ctx)).c_str()); return nullptr; }
  napi_value r; napi_create_int32(env, rc, &r); return r;
}

ml_grid gridOf(napi_env env, napi_value g) {
  ml_grid o{};
  o.nx = int(prop(env, g, "nx", 0)); o.ny = int(prop(env, g, "ny", 0)); o.nz = int(prop(env, g, "nz", 0));
  o.width = prop(env, g, "width", 0); o.depth = prop(env, g, "depth", 0); o.height = prop(env, g, "height", 0); o.bottom = prop(env, g, "bottom", 0);
  return o;
}

// ctxCreate(device, precision) -> context
napi_value CtxCreate(napi_env env, napi_callback_info info) {
  napi_value a[2]; if (!args(env, info, 2, a)) return nullptr;
  int32_t device, precision; if (!intOf(env, a[0], device) || !intOf(env, a[1], precision)) return nullptr;
  ml_ctx *ctx = nullptr;
  if (ml_ctx_create(device, precision, &ctx) != 0) { napi_throw_error(env, nullptr, "meteolab: no usable HIP device"); return nullptr; }
  napi_value r; NAPI_OK(env, napi_create_external(env, ctx, nullptr, nullptr, &r));
  return r;
}

napi_value CtxDestroy(napi_env env, napi_callback_info info) {
  napi_value a[1]; if (!args(env, info, 1, a)) return nullptr;
  ml_ctx *ctx = ctxOf(env, a[0]); if (!ctx) return nullptr;
  ml_ctx_destroy(ctx);
  return undefinedOf(env);
}

napi_value DeviceName(napi_env env, napi_callback_info info) {
  napi_value a[1]; if (!args(env, info, 1, a)) return nullptr;
  ml_ctx *ctx = ctxOf(env, a[0]); if (!ctx) return nullptr;
  char buf[256] = {0};
  if (ml_device_name(ctx, buf, sizeof buf) < 0) return status(env, ctx, -1);
  napi_value r; NAPI_OK(env, napi_create_string_utf8(env, buf, NAPI_AUTO_LENGTH, &r));
  return r;
}

// poolCreate(ctx, grid, capacity, electrified) -> pool id
napi_value PoolCreate(napi_env env, napi_callback_info info) {
  napi_value a[4]; if (!args(env, info, 4, a)) return nullptr;
  ml_ctx *ctx = ctxOf(env, a[0]); if (!ctx) return nullptr;
  ml_grid g = gridOf(env, a[1]);
  int32_t cap, el; if (!intOf(env, a[2], cap) || !intOf(env, a[3], el)) return nullptr;
  int pool = -1;
  if (napi_value r = status(env, ctx, ml_pool_create(ctx, &g, cap, el, &pool)); !r) return nullptr;
  napi_value r; napi_create_int32(env, pool, &r); return r;
}

napi_value PoolDestroy(napi_env env, napi_callback_info info) {
  napi_value a[2]; if (!args(env, info, 2, a)) return nullptr;
  ml_ctx *ctx = ctxOf(env, a[0]); if (!ctx) return nullptr;
  int32_t pool; if (!intOf(env, a[1], pool)) return nullptr;
  return status(env, ctx, ml_pool_destroy(ctx, pool));
}

napi_value SlotBytes(napi_env env, napi_callback_info info) {
  napi_value a[2]; if (!args(env, info, 2, a)) return nullptr;
  ml_ctx *ctx = ctxOf(env, a[0]); if (!ctx) return nullptr;
  ml_grid g = gridOf(env, a[1]);
  int64_t bytes = ml_slot_bytes(ctx, &g);
  if (bytes < 0) return status(env, ctx, int(bytes));
  napi_value r; napi_create_double(env, double(bytes), &r); return r;
}

// poolGrow(ctx, pool, capacity): ml_pool_grow, the pool's runs kept.
napi_value PoolGrow(napi_env env, napi_callback_info info) {
  napi_value a[3]; if (!args(env, info, 3, a)) return nullptr;
  ml_ctx *ctx = ctxOf(env, a[0]); if (!ctx) return nullptr;
  int32_t pool, cap; if (!intOf(env, a[1], pool) || !intOf(env, a[2], cap)) return nullptr;
  return status(env, ctx, ml_pool_grow(ctx, pool, cap));
}

// memInfo(ctx) -> [free, total] device bytes (ml_mem_info).
napi_value MemInfo(napi_env env, napi_callback_info info) {
  napi_value a[1]; if (!args(env, info, 1, a)) return nullptr;
  ml_ctx *ctx = ctxOf(env, a[0]); if (!ctx) return nullptr;
  int64_t fr = 0, tot = 0;
  if (!status(env, ctx, ml_mem_info(ctx, &fr, &tot))) return nullptr;
  napi_value r, v0, v1;
  napi_create_array_with_length(env, 2, &r);
  napi_create_double(env, double(fr), &v0); napi_create_double(env, double(tot), &v1);
  napi_set_element(env, r, 0, v0); napi_set_element(env, r, 1, v1);
  return r;
}

// ctxQuery(ctx) -> true when the context's work has finished (ml_ctx_query), without waiting.
napi_value CtxQuery(napi_env env, napi_callback_info info) {
  napi_value a[1]; if (!args(env, info, 1, a)) return nullptr;
  ml_ctx *ctx = ctxOf(env, a[0]); if (!ctx) return nullptr;
  const int r = ml_ctx_query(ctx);
  if (r < 0) return status(env, ctx, r);
  napi_value v; napi_get_boolean(env, r == 1, &v); return v;
}

// runCreate(ctx, pool, profile, config) -> run handle. `profile` and `config` are plain objects with the fields of
// ml_profile / ml_run_config.
napi_value RunCreate(napi_env env, napi_callback_info info) {
  napi_value a[4]; if (!args(env, info, 4, a)) return nullptr;
  ml_ctx *ctx = ctxOf(env, a[0]); if (!ctx) return nullptr;
  int32_t pool; if (!intOf(env, a[1], pool)) return nullptr;
  ml_profile p{};
  p.theta0 = prop(env, a[2], "theta0", 300); p.dtheta = prop(env, a[2], "dtheta", 43); p.ztr = prop(env, a[2], "ztr", 12000);
  p.ttr = prop(env, a[2], "ttr", 213); p.rh_drop = prop(env, a[2], "rh_drop", .75); p.rh_top = prop(env, a[2], "rh_top", .25);
  p.qv_max = prop(env, a[2], "qv_max", .014); p.shear = prop(env, a[2], "shear", 25); p.shear_depth = prop(env, a[2], "shear_depth", 3000);
  propArray(env, a[2], "wind_u", p.wind_u, 4); propArray(env, a[2], "wind_v", p.wind_v, 4);
  ml_run_config c{};
  c.seed = int(prop(env, a[3], "seed", 1)); c.latitude = prop(env, a[3], "latitude", 45); c.turbulence = prop(env, a[3], "turbulence", .55);
  c.tropopause_km = prop(env, a[3], "tropopause_km", 12); c.roughness = prop(env, a[3], "roughness", .03);
  c.frame_u = prop(env, a[3], "frame_u", 0); c.frame_v = prop(env, a[3], "frame_v", 0);
  c.bubble_amp = prop(env, a[3], "bubble_amp", 2); c.bubble_x = prop(env, a[3], "bubble_x", 0); c.bubble_y = prop(env, a[3], "bubble_y", 0);
  c.bubble_z = prop(env, a[3], "bubble_z", 1400); c.bubble_rh = prop(env, a[3], "bubble_rh", 10000); c.bubble_rz = prop(env, a[3], "bubble_rz", 1400);
  c.noise_seed = int(prop(env, a[3], "noise_seed", c.seed)); c.noise_amp = prop(env, a[3], "noise_amp", .1); c.noise_top = prop(env, a[3], "noise_top", 1000);
  int run = -1;
  if (!status(env, ctx, ml_run_create(ctx, pool, &p, &c, &run))) return nullptr;
  napi_value r; napi_create_int32(env, run, &r); return r;
}

// runRelease(ctx, run) / refreshRotation(ctx, run): entry points that take a run and nothing else.
template <int (*F)(ml_ctx *, int)> napi_value RunOp(napi_env env, napi_callback_info info) {
  napi_value a[2]; if (!args(env, info, 2, a)) return nullptr;
  ml_ctx *ctx = ctxOf(env, a[0]); if (!ctx) return nullptr;
  int32_t run; if (!intOf(env, a[1], run)) return nullptr;
  return status(env, ctx, F(ctx, run));
}

// upload(ctx, run, field, Float32Array) / download(ctx, run, field, Float32Array); the same for ground columns.
template <bool UP, bool COLUMN> napi_value Transfer(napi_env env, napi_callback_info info) {
  napi_value a[4]; if (!args(env, info, 4, a)) return nullptr;
  ml_ctx *ctx = ctxOf(env, a[0]); if (!ctx) return nullptr;
  int32_t run, which; if (!intOf(env, a[1], run) || !intOf(env, a[2], which)) return nullptr;
  // The array must hold a whole field (or column) of the run's grid; node/meteolab.js allocates them so.
  float *data; if (!typed(env, a[3], napi_float32_array, 1, data)) return nullptr;
  int rc;
  if (COLUMN) rc = UP ? ml_run_upload_column(ctx, run, which, data) : ml_run_download_column(ctx, run, which, data);
  else rc = UP ? ml_run_upload(ctx, run, which, data) : ml_run_download(ctx, run, which, data);
  return status(env, ctx, rc);
}

// step(ctx, Int32Array runs, steps, dt): ml_step of a batch of runs of one pool (asynchronous; sync() waits).
napi_value Step(napi_env env, napi_callback_info info) {
  napi_value a[4]; if (!args(env, info, 4, a)) return nullptr;
  ml_ctx *ctx = ctxOf(env, a[0]); if (!ctx) return nullptr;
  napi_typedarray_type t; size_t len; void *raw; napi_value buf; size_t off;
  if (napi_get_typedarray_info(env, a[1], &t, &len, &raw, &buf, &off) != napi_ok || t != napi_int32_array) { napi_throw_type_error(env, nullptr, "runs: Int32Array"); return nullptr; }
  int32_t steps; double dt; if (!intOf(env, a[2], steps) || !numberOf(env, a[3], dt)) return nullptr;
  return status(env, ctx, ml_step(ctx, static_cast<const int *>(raw), int(len), steps, dt));
}

// setSwaths(ctx, run, on): ml_run_set_swaths.
napi_value SetSwaths(napi_env env, napi_callback_info info) {
  napi_value a[3]; if (!args(env, info, 3, a)) return nullptr;
  ml_ctx *ctx = ctxOf(env, a[0]); if (!ctx) return nullptr;
  int32_t run, on; if (!intOf(env, a[1], run) || !intOf(env, a[2], on)) return nullptr;
  return status(env, ctx, ml_run_set_swaths(ctx, run, on));
}

// probabilityMap(ctx, Int32Array runs, column, threshold, Float32Array out): ml_probability_map.
napi_value ProbabilityMap(napi_env env, napi_callback_info info) {
  napi_value a[5]; if (!args(env, info, 5, a)) return nullptr;
  ml_ctx *ctx = ctxOf(env, a[0]); if (!ctx) return nullptr;
  napi_typedarray_type t; size_t len; void *raw; napi_value buf; size_t off;
  if (napi_get_typedarray_info(env, a[1], &t, &len, &raw, &buf, &off) != napi_ok || t != napi_int32_array) { napi_throw_type_error(env, nullptr, "runs: Int32Array"); return nullptr; }
  int32_t column; double threshold; if (!intOf(env, a[2], column) || !numberOf(env, a[3], threshold)) return nullptr;
  napi_typedarray_type t2; size_t len2; void *raw2;
  if (napi_get_typedarray_info(env, a[4], &t2, &len2, &raw2, &buf, &off) != napi_ok || t2 != napi_float32_array) { napi_throw_type_error(env, nullptr, "out: Float32Array of the layer"); return nullptr; }
  return status(env, ctx, ml_probability_map(ctx, static_cast<const int *>(raw), int(len), column, threshold, static_cast<float *>(raw2)));
}

// modelDiagnostics(ctx, Int32Array runs, Float64Array out): ml_model_diagnostics, ML_MD_COUNT values a run.
napi_value ModelDiagnostics(napi_env env, napi_callback_info info) {
  napi_value a[3]; if (!args(env, info, 3, a)) return nullptr;
  ml_ctx *ctx = ctxOf(env, a[0]); if (!ctx) return nullptr;
  napi_typedarray_type t; size_t len; void *raw; napi_value buf; size_t off;
  if (napi_get_typedarray_info(env, a[1], &t, &len, &raw, &buf, &off) != napi_ok || t != napi_int32_array) { napi_throw_type_error(env, nullptr, "runs: Int32Array"); return nullptr; }
  napi_typedarray_type t2; size_t len2; void *raw2;
  if (napi_get_typedarray_info(env, a[2], &t2, &len2, &raw2, &buf, &off) != napi_ok || t2 != napi_float64_array || len2 < len * ML_MD_COUNT) { napi_throw_type_error(env, nullptr, "out: Float64Array of runs x ML_MD_COUNT"); return nullptr; }
  return status(env, ctx, ml_model_diagnostics(ctx, static_cast<const int *>(raw), int(len), static_cast<double *>(raw2)));
}

// dynamics / physics / mix / afterStep (ctx, Int32Array runs, dt): one part of a step of a batch (ml_dynamics,
// ml_physics, ml_mix, ml_after_step; asynchronous).
template <int (*F)(ml_ctx *, const int *, int, double)> napi_value Part(napi_env env, napi_callback_info info) {
  napi_value a[3]; if (!args(env, info, 3, a)) return nullptr;
  ml_ctx *ctx = ctxOf(env, a[0]); if (!ctx) return nullptr;
  napi_typedarray_type t; size_t len; void *raw; napi_value buf; size_t off;
  if (napi_get_typedarray_info(env, a[1], &t, &len, &raw, &buf, &off) != napi_ok || t != napi_int32_array) { napi_throw_type_error(env, nullptr, "runs: Int32Array"); return nullptr; }
  double dt; if (!numberOf(env, a[2], dt)) return nullptr;
  return status(env, ctx, F(ctx, static_cast<const int *>(raw), int(len), dt));
}

// fallParts(ctx, run, Int32Array(8)) -> number of falls written (ml_fall_inputs).
napi_value FallParts(napi_env env, napi_callback_info info) {
  napi_value a[3]; if (!args(env, info, 3, a)) return nullptr;
  ml_ctx *ctx = ctxOf(env, a[0]); if (!ctx) return nullptr;
  int32_t run; if (!intOf(env, a[1], run)) return nullptr;
  napi_typedarray_type t; size_t len; void *raw; napi_value buf; size_t off;
  if (napi_get_typedarray_info(env, a[2], &t, &len, &raw, &buf, &off) != napi_ok || t != napi_int32_array) { napi_throw_type_error(env, nullptr, "Int32Array"); return nullptr; }
  return status(env, ctx, ml_fall_inputs(ctx, run, static_cast<int *>(raw), int(len)));
}

napi_value Sync(napi_env env, napi_callback_info info) {
  napi_value a[1]; if (!args(env, info, 1, a)) return nullptr;
  ml_ctx *ctx = ctxOf(env, a[0]); if (!ctx) return nullptr;
  return status(env, ctx, ml_sync(ctx));
}

// diag(ctx, run, Float64Array(ML_D_COUNT)), levels(ctx, run, Float64Array(7 nz)), pressure(ctx, run,
// Float64Array((nz - 1) nx ny)).
template <int (*F)(ml_ctx *, int, double *)> napi_value Doubles(napi_env env, napi_callback_info info) {
  napi_value a[3]; if (!args(env, info, 3, a)) return nullptr;
  ml_ctx *ctx = ctxOf(env, a[0]); if (!ctx) return nullptr;
  int32_t run; if (!intOf(env, a[1], run)) return nullptr;
  double *out; if (!typed(env, a[2], napi_float64_array, 1, out)) return nullptr;
  return status(env, ctx, F(ctx, run, out));
}

// flashes(ctx, run, Float64Array(8 max)) -> count written.
napi_value Flashes(napi_env env, napi_callback_info info) {
  napi_value a[3]; if (!args(env, info, 3, a)) return nullptr;
  ml_ctx *ctx = ctxOf(env, a[0]); if (!ctx) return nullptr;
  int32_t run; if (!intOf(env, a[1], run)) return nullptr;
  napi_typedarray_type t; size_t len; void *raw; napi_value buf; size_t off;
  if (napi_get_typedarray_info(env, a[2], &t, &len, &raw, &buf, &off) != napi_ok || t != napi_float64_array) { napi_throw_type_error(env, nullptr, "Float64Array"); return nullptr; }
  return status(env, ctx, ml_run_flashes(ctx, run, static_cast<double *>(raw), int(len / 8)));
}

napi_value SetTime(napi_env env, napi_callback_info info) {
  napi_value a[3]; if (!args(env, info, 3, a)) return nullptr;
  ml_ctx *ctx = ctxOf(env, a[0]); if (!ctx) return nullptr;
  int32_t run; double t; if (!intOf(env, a[1], run) || !numberOf(env, a[2], t)) return nullptr;
  return status(env, ctx, ml_run_set_time(ctx, run, t));
}

// setTotals(ctx, run, chargeGround, flashesIC, flashesCG, persisted): a saved state's header totals (ml_run_set_totals).
napi_value SetTotals(napi_env env, napi_callback_info info) {
  napi_value a[6]; if (!args(env, info, 6, a)) return nullptr;
  ml_ctx *ctx = ctxOf(env, a[0]); if (!ctx) return nullptr;
  int32_t run; double q, ic, cg, held;
  if (!intOf(env, a[1], run) || !numberOf(env, a[2], q) || !numberOf(env, a[3], ic) || !numberOf(env, a[4], cg) || !numberOf(env, a[5], held)) return nullptr;
  return status(env, ctx, ml_run_set_totals(ctx, run, q, int64_t(ic), int64_t(cg), held));
}

// intervene(ctx, run, kind, x, y, strength, thermalRadius, Float64Array out[2]): ml_run_intervene.
napi_value Intervene(napi_env env, napi_callback_info info) {
  napi_value a[8]; if (!args(env, info, 8, a)) return nullptr;
  ml_ctx *ctx = ctxOf(env, a[0]); if (!ctx) return nullptr;
  int32_t run, kind; double x, y, strength, radius;
  if (!intOf(env, a[1], run) || !intOf(env, a[2], kind) || !numberOf(env, a[3], x) || !numberOf(env, a[4], y) ||
      !numberOf(env, a[5], strength) || !numberOf(env, a[6], radius)) return nullptr;
  double *out; if (!typed(env, a[7], napi_float64_array, 2, out)) return nullptr;
  return status(env, ctx, ml_run_intervene(ctx, run, kind, x, y, strength, radius, out));
}

// setAdded(ctx, run, heat, vapour, count): a saved state's header.added (ml_run_set_added).
napi_value SetAdded(napi_env env, napi_callback_info info) {
  napi_value a[5]; if (!args(env, info, 5, a)) return nullptr;
  ml_ctx *ctx = ctxOf(env, a[0]); if (!ctx) return nullptr;
  int32_t run; double heat, vapour, count;
  if (!intOf(env, a[1], run) || !numberOf(env, a[2], heat) || !numberOf(env, a[3], vapour) || !numberOf(env, a[4], count)) return nullptr;
  return status(env, ctx, ml_run_set_added(ctx, run, heat, vapour, count));
}

napi_value Init(napi_env env, napi_value exports) {
  struct { const char *name; napi_callback fn; } fns[] = {
    {"ctxCreate", CtxCreate}, {"ctxDestroy", CtxDestroy}, {"deviceName", DeviceName}, {"poolCreate", PoolCreate},
    {"poolDestroy", PoolDestroy}, {"slotBytes", SlotBytes}, {"runCreate", RunCreate}, {"runRelease", RunOp<ml_run_release>},
    {"upload", Transfer<true, false>}, {"download", Transfer<false, false>}, {"uploadColumn", Transfer<true, true>},
    {"downloadColumn", Transfer<false, true>}, {"step", Step}, {"sync", Sync}, {"diag", Doubles<ml_run_diag>},
    {"levels", Doubles<ml_run_levels>}, {"pressure", Doubles<ml_run_pressure>}, {"flashes", Flashes}, {"setTime", SetTime}, {"dynamics", Part<ml_dynamics>},
    {"physics", Part<ml_physics>}, {"mix", Part<ml_mix>}, {"afterStep", Part<ml_after_step>}, {"fallParts", FallParts},
    {"setTotals", SetTotals}, {"rotation", Doubles<ml_run_rotation>}, {"refreshRotation", RunOp<ml_run_refresh_rotation>},
    {"modelDiagnostics", ModelDiagnostics}, {"setSwaths", SetSwaths}, {"probabilityMap", ProbabilityMap},
    {"intervene", Intervene}, {"added", Doubles<ml_run_added>}, {"setAdded", SetAdded}, {"poolGrow", PoolGrow},
    {"memInfo", MemInfo}, {"ctxQuery", CtxQuery},
  };
  for (auto &f : fns) {
    napi_value v;
    if (napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.fn, nullptr, &v) != napi_ok) return nullptr;
    if (napi_set_named_property(env, exports, f.name, v) != napi_ok) return nullptr;
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
