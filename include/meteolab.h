/*
 * meteolab: the StormLab model step on MI355X (gfx950), C ABI.
 *
 * The drop-in boundary of the GPU step. The reference's step is AtmosphereModel.step(dt) (src/core/model.ts:873-877)
 * on the model's Float32Array fields, and its GPU plugin is GpuDynamics (src/gpu/dynamics.ts:250-504: create, step,
 * mix, matches, destroy, errors, adapter). Here a context owns one device and stream; runs of one grid size share a
 * pool of slots (their state stays resident in HBM); a step advances a batch of runs of one pool by whole model steps.
 *
 * Conventions: every function returns 0 on success or a negative ML_E_* code; ml_last_error() gives the message.
 * Plain pointers and sizes only; host buffers are the caller's. Field ids follow the carried order of
 * AtmosphereModel.carried() (model.ts:1319-1335), then the per-node arrays a step recomputes. Node index
 * i = x + nx (y + ny z), z = 0 at the ground. One context per host thread; calls on a context are stream-ordered and
 * ml_sync() waits for them.
 */
#ifndef METEOLAB_H
#define METEOLAB_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum {
  ML_OK = 0, ML_E_ARG = -1, ML_E_HIP = -2, ML_E_FULL = -3, ML_E_STATE = -4, ML_E_UNSUPPORTED = -5
};

/* Field ids (carried order, model.ts:1319-1335, then fieldZ, the fall speeds and the condensate). */
enum {
  ML_RAIN = 0, ML_SNOW, ML_GRAUPEL, ML_HAIL, ML_U, ML_V, ML_W, ML_THETA, ML_Q, ML_CLOUD, ML_ICE, ML_ICENUMBER, ML_COLD,
  ML_CHARGE_ICE, ML_CHARGE_GRAUPEL, ML_CHARGE_RAIN, ML_CHARGE_AIR,
  ML_CARRIED = 17,
  ML_FIELDZ = 17, ML_FALL_RAIN, ML_FALL_SNOW, ML_FALL_GRAUPEL, ML_FALL_HAIL, ML_FALL_CHARGE_ICE, ML_FALL_CHARGE_GRAUPEL,
  ML_CONDENSATE, ML_FIELDS
};
/* Ground columns (STATE_COLUMNS, src/core/state.ts:10). */
/* Then an ensemble member's swaths (src/ensemble/member.ts:19-21): the strongest surface wind (m/s) and 2-5 km
 * updraught helicity (m2/s2) over each ground column, kept while ml_run_set_swaths is on. */
enum { ML_PRECIPITATION = 0, ML_HAIL_GROUND, ML_HAIL_SIZE, ML_GUST_MAX, ML_UH_MAX, ML_COLUMNS };
/* Per-level base state (environment.ts:124-130), float64, nz each, in this order. */
enum { ML_LEVEL_P = 0, ML_LEVEL_EXNER, ML_LEVEL_THETA, ML_LEVEL_Q, ML_LEVEL_U, ML_LEVEL_V, ML_LEVEL_RHO, ML_LEVELS = 7 };
/* Diagnostics of a run (model.ts:1672-1696 and the run metrics), float64. */
enum {
  ML_D_TIME = 0, ML_D_MAXW, ML_D_MINW, ML_D_MAXTHETAP, ML_D_RAINTOTAL, ML_D_WATER, ML_D_CHARGE, ML_D_CLIPPED,
  ML_D_NEGFILLED, ML_D_CHARGEGROUND, ML_D_FLASHES_IC, ML_D_FLASHES_CG, ML_D_FIRSTFLASH, ML_D_ERROR, ML_D_COUNT
};
/* ML_D_ERROR: bits of conditions a step met and does not handle, 0 when none: 1, a lightning front with more prunes
 * than the step keeps; 2, a water fall needing more than 16 sedimentation sub-steps in one step (its rain beyond the
 * 16th does not reach the gauge). */

/* AtmosphereModel.diagnostics() of a run (ml_model_diagnostics; ModelDiagnostics, model.ts:161-180), float64. */
enum {
  ML_MD_UPDRAFT = 0, ML_MD_DOWNDRAFT, ML_MD_RAINRATE, ML_MD_RAINTOTAL, ML_MD_CLOUDTOP, ML_MD_THERMALTOP, ML_MD_MAXCLOUD,
  ML_MD_COLDMAX, ML_MD_CORES, ML_MD_SHEAR06, ML_MD_OUTFLOW, ML_MD_OUTFLOWLOW, ML_MD_GUST, ML_MD_HAIL, ML_MD_CLIPPED,
  ML_MD_COUNT
};
/* Rotation diagnostics of a run (ml_run_rotation). */
enum {
  ML_ROT_UH = 0, ML_ROT_X, ML_ROT_Y, ML_ROT_ANTICYCLONIC, ML_ROT_PERSISTED, ML_ROT_UH01, ML_ROT_UH03, ML_ROT_LOW,
  ML_ROT_CORE_ZETA, ML_ROT_CORE_UH, ML_ROT_CORE_ANTI, ML_ROT_COUNT
};

/* Grid (src/core/grid.ts GridOptions, createGrid :99-121): periodic nx x ny, nz levels stretched geometrically from
 * `bottom` m at the ground to `height` (bottom = 0: uniform). */
typedef struct { int nx, ny, nz; double width, depth, height, bottom; } ml_grid;

/* Analytic sounding (src/core/profiles.ts:61-73 weismanKlemp): theta = theta0 + dtheta (z/ztr)^1.25 below ztr,
 * isothermal ttr above; RH = 1 - rh_drop (z/ztr)^1.25 (rh_top above), vapour capped at qv_max; wind
 * u = shear tanh(z / shear_depth), v = 0, plus the wind errors wind_u, wind_v at 0, 3, 6 and 10 km, linear between those
 * heights and held below and above them (the sounding wind of perturbConfig, src/ensemble/member.ts:55-57; zero for
 * the plain sounding). */
typedef struct { double theta0, dtheta, ztr, ttr, rh_drop, rh_top, qv_max, shear, shear_depth, wind_u[4], wind_v[4]; } ml_profile;

/* A run's settings (SimConfig values the step reads, config.ts:15-188) and its initial thermal and noise: a cos^2
 * bubble of bubble_amp K (model.ts:94 ELLIPSOID form) and uniform +-noise_amp K theta noise below noise_top from
 * mulberry32(noise_seed) in node order. */
typedef struct {
  int seed;
  double latitude, turbulence, tropopause_km, roughness, frame_u, frame_v;
  double bubble_amp, bubble_x, bubble_y, bubble_z, bubble_rh, bubble_rz;
  int noise_seed;
  double noise_amp, noise_top;
} ml_run_config;

typedef struct ml_ctx ml_ctx;

/* A context on device `device`. precision 32: float32 arithmetic (the production kernels); 64: the same kernels in
 * float64 (the precision study). Replaces GpuDynamics.create's device setup (gpu/dynamics.ts:273-278). */
int ml_ctx_create(int device, int precision, ml_ctx **out);
void ml_ctx_destroy(ml_ctx *ctx);
/* Message of the last error on this context (GpuDynamics.errors, gpu/dynamics.ts:252). */
const char *ml_last_error(ml_ctx *ctx);
/* The device's name (GpuDynamics.adapter, gpu/dynamics.ts:251). */
int ml_device_name(ml_ctx *ctx, char *out, int size);

/* A pool of `capacity` run slots of one grid (a size class); electrified 1 carries the charges and runs the lightning
 * (SimConfig.lightning, model.ts:1316). *pool receives its id. */
int ml_pool_create(ml_ctx *ctx, const ml_grid *grid, int capacity, int electrified, int *pool);
/* Device bytes one slot of a pool of this grid takes (the serving loop's admission by HBM footprint). */
int64_t ml_slot_bytes(ml_ctx *ctx, const ml_grid *grid);
/* Free slots left in a pool. */
int ml_pool_free(ml_ctx *ctx, int pool);
/* Frees a pool and its device memory (waits for the context's stream); its run handles become invalid and its id may
 * be reused by a later pool. No counterpart in the reference (its runs are garbage-collected objects). */
int ml_pool_destroy(ml_ctx *ctx, int pool);

/* A run in a free slot of `pool`: its base state from the sounding (environment.ts:45-73, in float64 on the device) and
 * its initial state (model.ts:399-451 single-storm start with the config's thermal and noise). *run receives a handle.
 * Replaces new AtmosphereModel(config, grid, profile) for the step's state. */
int ml_run_create(ml_ctx *ctx, int pool, const ml_profile *profile, const ml_run_config *config, int *run);
/* `count` runs at once, as ml_run_create of profiles[k] and configs[k] (runs[k] receives run k's handle): one launch of
 * each start-up kernel for all of them and one wait. Replaces the new AtmosphereModel of each run a batch front-end
 * starts (the ensemble's runPool, src/ensemble/pool.ts:22-47, one per worker); ML_E_FULL, and no run made, when the
 * pool has fewer free slots. */
int ml_runs_create(ml_ctx *ctx, int pool, int count, const ml_profile *profiles, const ml_run_config *configs, int *runs);
/* Grows a pool to `capacity` slots (at least its own): its runs keep their slots, handles and state, the new slots are
 * free. Waits for the context's work; the next step of each batch is captured again. On an error the pool is left as it
 * was. No counterpart in the reference (its runs are garbage-collected objects); the serving loop's pools grow with
 * the requests admitted. */
int ml_pool_grow(ml_ctx *ctx, int pool, int capacity);
/* Free and total device memory in bytes (hipMemGetInfo): what the serving loop admits runs against, at
 * ml_slot_bytes a run. No counterpart in the reference. */
int ml_mem_info(ml_ctx *ctx, int64_t *free_bytes, int64_t *total_bytes);
/* 1 when all the work given to the context so far has finished, 0 while some is still running, without waiting: the
 * promise of GpuDynamics.step settling (gpu/dynamics.ts:419, 494) as a poll, so that one host thread keeps several
 * contexts' steps in flight. */
int ml_ctx_query(ml_ctx *ctx);
/* Frees the run's slot. */
int ml_run_release(ml_ctx *ctx, int run);

/* Copies one field (n floats) / ground column (nx ny floats) between the host and a run (the model-hook round trip:
 * dynamicsInputs / setCarried, model.ts:994-1026; state layout of src/core/state.ts). */
int ml_run_upload(ml_ctx *ctx, int run, int field, const float *host);
int ml_run_download(ml_ctx *ctx, int run, int field, float *host);
int ml_run_upload_column(ml_ctx *ctx, int run, int column, const float *host);
int ml_run_download_column(ml_ctx *ctx, int run, int column, float *host);
/* The run's base state, ML_LEVELS x nz float64. */
int ml_run_levels(ml_ctx *ctx, int run, double *out);
/* The pressure of the run's last projection (model.ts:358 pressure; the solver's cells), (nz - 1) x nx x ny float64 in
 * cell order c = x + nx (y + ny k), cell k between levels k and k + 1: what GpuDynamics.mix returns as `pressure`
 * (gpu/dynamics.ts:356-395) and endGpuStep copies into the model (model.ts:1145-1150). Stream-synchronising. */
int ml_run_pressure(ml_ctx *ctx, int run, double *out);
/* The run's diagnostics, ML_D_COUNT float64 (stream-synchronising). */
int ml_run_diag(ml_ctx *ctx, int run, double *out);
/* The run's most recent flashes (at most 64 are kept), oldest first: up to `max` records of 8 float64 each: time (s),
 * CG (1) or IC (0), start node, channel nodes, charge (C: neutralized by an IC, lowered to the ground by a CG),
 * potential (V), charge moment change (C m) (lightning.ts Flash, :57-72). Returns the number written. */
int ml_run_flashes(ml_ctx *ctx, int run, double *out, int max);
/* Model time of the run, s (AtmosphereModel.time). */
int ml_run_set_time(ml_ctx *ctx, int run, double time);
/* The totals a saved state's header carries (applyState, src/core/state.ts:84-95): the charge carried to the ground
 * (C, header.chargeGround), the IC and CG flash counts (header.flashCount) and how long the mesocyclone criterion has
 * held (s, header.persisted, rotation.persisted). The flash log and the first-flash time are left as they are (a new
 * run has none: applyState forgets the flashes after the state's time). */
int ml_run_set_totals(ml_ctx *ctx, int run, double charge_ground, int64_t flashes_ic, int64_t flashes_cg, double persisted);
/* Interventions (InterventionKind, src/core/intervene.ts:13). */
enum { ML_INT_THERMAL = 0, ML_INT_WARM, ML_INT_COOL, ML_INT_MOISTEN, ML_INT_DRY, ML_INT_COLDPOOL, ML_INT_VORTEX, ML_INTERVENTIONS };
/* intervene(m, {kind, x, y, strength}) (src/core/intervene.ts:38-68; the worker's 'intervene' message,
 * src/simulation.worker.ts:219, and injectBubble, model.ts:482): the kind's patch (INTERVENTIONS, intervene.ts:22-30)
 * centred at (x, y) m from the domain's corner, `strength` times its standard size, added to the run's fields as they
 * are. thermal_radius: the thermal's e-folding radius in km (SimConfig.bubbleRadius; <= 0: its default 4.2). out
 * (may be NULL): the heat (J) and vapour (kg) put in, also added to the run's counts (ml_run_added). Between steps
 * only; stream-synchronising. */
int ml_run_intervene(ml_ctx *ctx, int run, int kind, double x, double y, double strength, double thermal_radius, double *out);
/* AtmosphereModel.added (model.ts:227-228): heat (J), vapour (kg) put in by interventions, and how many, 3 float64. */
int ml_run_added(ml_ctx *ctx, int run, double *out);
/* Sets those counts (applyState's header.added, src/core/state.ts:78). */
int ml_run_set_added(ml_ctx *ctx, int run, double heat, double vapour, double count);
/* The rotation diagnostics of the run's last step (updateRotation, model.ts:1609-1648; RotationState :142-158),
 * ML_ROT_COUNT float64: the largest updraught helicity 2-5 km (m2/s2) and its column (x, y), the largest
 * anticyclonic one, how long the mesocyclone criterion has held (s), the largest 0-1 km and 0-3 km helicity and the
 * column of the 0-1 km one, the most cyclonic updraught core's mean vorticity (s-1) and its strongest helicity, and the
 * most anticyclonic core's mean vorticity. */
int ml_run_rotation(ml_ctx *ctx, int run, double *out);
/* Recomputes the rotation diagnostics from the run's fields as they are, keeping how long the criterion has held
 * (refreshRotation, model.ts:1607; applyState calls it, src/core/state.ts:73-74). */
int ml_run_refresh_rotation(ml_ctx *ctx, int run);
/* Whether each step of the run folds its surface wind and 2-5 km updraught helicity into the ML_GUST_MAX and ML_UH_MAX
 * columns over the ground under the moving domain (addSwaths, src/ensemble/member.ts:85-100, which runMember calls after
 * every step, :111-114). Off for a new run; the columns start at 0. */
int ml_run_set_swaths(ml_ctx *ctx, int run, int on);
/* probabilityMap (src/ensemble/combine.ts:70-75) over `count` runs of one pool: per ground column (layer float32 into
 * out) the share of the runs whose ground column `column` (ML_PRECIPITATION as the rain total, ML_HAIL_SIZE,
 * ML_GUST_MAX, ML_UH_MAX) reached `threshold` there (for helicity the caller scales the threshold to the grid,
 * uhScale). */
int ml_probability_map(ml_ctx *ctx, const int *runs, int count, int column, double threshold, float *out);
/* A numerics check with no counterpart in the reference: the transport kernels' WENO5 face value (weno,
 * src/core/advection.ts:388-399, minus c) of each of n stencils (a, b, c, d, e; 5n float32, flow from a towards e),
 * computed on the device in the context's arithmetic (float32 or float64), into out (n float64). Allocates its own
 * device buffers; stream-synchronising. */
int ml_check_weno(ml_ctx *ctx, const float *stencils, int n, double *out);
/* AtmosphereModel.diagnostics() (model.ts:1672-1696) of `count` runs of one pool from their fields as they are,
 * ML_MD_COUNT float64 a run into out (count * ML_MD_COUNT): the largest up- and downdraught (m/s), the heaviest
 * precipitation reaching the ground (mm/h, surfaceflux.ts:22-37 at the ground nodes) and the largest rain total (mm),
 * the cloud top and the top of w > 0.6 m/s (km), the most cloud water + ice (kg/kg), the largest cold-pool indicator
 * below 1.5 km, the cloudy updraught cores (countCores, model.ts:1650-1670), the environment's 0-6 km shear (m/s), the
 * outflow's largest virtual-temperature deficit and the largest within 2 km of the low-level rotation (K, rkw.ts:72-90),
 * the strongest ground wind (m/s), the largest hailstones at the ground (mm), and the wind components clipped since the
 * start. One launch for the runs, the ensemble's sampling of every member at once. */
int ml_model_diagnostics(ml_ctx *ctx, const int *runs, int count, double *out);
/* ml_model_diagnostics without allocating: the same kernel and the same results, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). No counterpart in
 * the reference. Its scratch (the diagnostics rows and the run list) is a static array in the library's device memory,
 * one per device, held until the process ends; it calls no hipMalloc, hipFree or other allocating routine. At most
 * ML_NOALLOC_RUNS runs a call (more is ML_E_ARG). Every call on one device shares that device's array: calls that
 * overlap on one device (from any context or thread) race. Stream-synchronising. */
enum { ML_NOALLOC_RUNS = 1024 };
int ml_model_diagnostics_noalloc(ml_ctx *ctx, const int *runs, int count, double *out);

/* Advances `count` runs of one pool by `steps` model steps of dt s each: AtmosphereModel.step(dt) then time += dt
 * (model.ts:873-877 and its callers, e.g. tests/bench/threads.ts:38-42). One launch per kernel covers all the runs;
 * the launch sequence of a batch is captured once into a hipGraph and replayed. Asynchronous: ml_sync waits. */
int ml_step(ml_ctx *ctx, const int *runs, int count, int steps, double dt);
/* Scheduling of ml_step (no counterpart in the reference, whose step is serial): with on = 1 (the default) a step's
 * discharge and conduction (the serial lightning, a block per run) run on a stream of their own beside the next step's
 * sedimentation of water and transport of all but the charges, which read and write none of what the lightning does;
 * the charges' falls and transport wait for it. A call still ends with every step complete, and the results are those
 * of on = 0, which runs the step's parts in line. */
int ml_ctx_set_overlap(ml_ctx *ctx, int on);

/* The parts of one step, split where the reference's GPU step splits it (src/gpu/step.ts:12-31), for a host that runs
 * some parts itself. ml_dynamics, ml_physics, ml_mix and ml_after_step in that order are one ml_step (the same
 * kernels, eagerly launched). Asynchronous; the runs' state stays on the device between the calls.
 *   ml_dynamics:   beginStep's fall speeds, hail swath and sedimentation, the face projection and the three RK3 stages
 *                  with buoyancy and the pressure in the stages: GpuDynamics.step(fields, env, dt, timed, falls)
 *                  (gpu/dynamics.ts:419); the ground deposits go into the run's precipitation column.
 *   ml_physics:    physicsStep (model.ts:1116-1127): the positivity fix with the sums of the last ml_dynamics, then
 *                  the column physics. A host that changed the fields after ml_dynamics runs its own physicsStep.
 *   ml_mix:        mixStep, the projection and clipWind: GpuDynamics.mix(fields, mixingInputs, dt) (:356) and
 *                  endGpuStep (model.ts:1145-1150); the condensate is the run's `condensate` field.
 *   ml_after_step: afterStep's lightning (model.ts:1130-1133), then time += dt. */
int ml_dynamics(ml_ctx *ctx, const int *runs, int count, double dt);
int ml_physics(ml_ctx *ctx, const int *runs, int count, double dt);
int ml_mix(ml_ctx *ctx, const int *runs, int count, double dt);
int ml_after_step(ml_ctx *ctx, const int *runs, int count, double dt);
/* The sub-steps of each fall of the run's last ml_dynamics (or step): fallInputs(dt)'s `parts` (model.ts:992-998,
 * fallParts :575-580), in the order rain, snow, graupel, hail, then the charges of ice, graupel and rain when the
 * run is electrified; the fall speeds are the fields fallRain .. fallChargeGraupel. Returns the number of falls
 * (stream-synchronising); `max` below it is ML_E_ARG. */
int ml_fall_inputs(ml_ctx *ctx, int run, int *parts, int max);
int ml_sync(ml_ctx *ctx);
/* The launch shape of a batch's transport kernel, shape[4]: shape[0] the layout (0 strips of a level with shape[1]
 * field groups per block, 1 shape[3] whole levels of one field per block, 2 the same with each level in a padded LDS
 * plane), shape[2] 1 if each run's blocks are kept on one XCD; *ms the stage-0 transport time it measured. The layout
 * is fixed by the grid (2 where its planes fit a block's LDS, else 1), since the layouts round differently in float32;
 * levels per block and the block order, which leave the arithmetic unchanged, are autotuned by timing on the device
 * when the batch is first stepped. No counterpart in the reference. */
int ml_batch_tuning(ml_ctx *ctx, const int *runs, int count, int *shape, float *ms);
/* Sets a batch's transport launch shape (as ml_batch_tuning reports it) in place of the autotuned one; the next
 * ml_step recaptures the step. No counterpart in the reference. */
int ml_batch_set_tuning(ml_ctx *ctx, const int *runs, int count, const int *shape);
/* ml_step of `steps` steps with a HIP event on the context's stream after each: ms[k] receives the device time of
 * step k (the graph replay), for throughput and latency measurement. Synchronous. */
int ml_step_timed(ml_ctx *ctx, const int *runs, int count, int steps, double dt, double *ms);
/* A plain device copy of `bytes` (read + write), `iters` times: the achievable HBM bandwidth in GB/s into *gbps, and a
 * kernel of known traffic for calibrating the HBM byte counters. No counterpart in the reference. */
int ml_copy_probe(ml_ctx *ctx, int64_t bytes, int iters, double *gbps);

/* Kernel timing: the batch's step launched kernel by kernel (no graph) `steps` times with a HIP event after every
 * kernel; ms[i] receives launch i's device time averaged over the steps and `names` (size bytes) the kernels' names,
 * '\n'-separated, in launch order. Returns the number of launches per step (at most `max` are written). */
int ml_profile_step(ml_ctx *ctx, const int *runs, int count, int steps, double dt, double *ms, int max, char *names, int size);

#ifdef __cplusplus
}
#endif
#endif
