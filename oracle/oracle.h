/*
 * StormLab step oracle: C ABI of the CPU restatement.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the HIP step is compared with; it is loaded by tests/,
 * by __graft_entry__.smoke() and by bench.py's CPU-baseline leg, never by the product path.
 *
 * Field storage is float32 (the reference's Float32Array fields), arithmetic is float64 (the reference's JS numbers),
 * in the order of operations of the reference (src/core/model.ts AtmosphereModel.step and what it calls).
 */
#ifndef STORMLAB_ORACLE_H
#define STORMLAB_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Grid (src/core/grid.ts GridOptions): periodic nx x ny, nz levels stretched from `bottom` m at the ground (0 = uniform). */
typedef struct { int nx, ny, nz; double width, depth, height, bottom; } orc_grid;

/*
 * Analytic sounding (src/core/profiles.ts weismanKlemp): theta = theta0 + dtheta (z/ztr)^1.25 below ztr, isothermal ttr
 * above; RH = 1 - rh_drop (z/ztr)^1.25 (rh_top above); vapour capped at qv_max; wind u = shear tanh(z / shear_depth), v = 0,
 * plus the wind errors wind_u, wind_v at 0, 3, 6 and 10 km, linear between and held beyond (src/ensemble/member.ts:55-57).
 */
typedef struct { double theta0, dtheta, ztr, ttr, rh_drop, rh_top, qv_max, shear, shear_depth, wind_u[4], wind_v[4]; } orc_profile;

/*
 * One run: SimConfig values that matter at the baseline (src/core/fixtures.ts SUMMER_DAY with single: true) and the
 * initial perturbation written after construction: a cos^2 thermal of amplitude bubble_amp centred at (bubble_x,
 * bubble_y, bubble_z) with radii (bubble_rh, bubble_rz), plus uniform +-noise_amp K theta noise below noise_top from
 * mulberry32(noise_seed) in node order.
 */
typedef struct {
  int seed;                 /* SimConfig.seed (the lightning's random numbers use seed + 1) */
  double latitude, turbulence, tropopause_km, roughness;
  double frame_u, frame_v;  /* SimConfig.frameMotion */
  double bubble_amp, bubble_x, bubble_y, bubble_z, bubble_rh, bubble_rz;
  int noise_seed; double noise_amp, noise_top;
} orc_run;

/* Field ids: the carried order of AtmosphereModel.carried() (model.ts:1319-1335), then the extra per-node arrays. */
enum {
  ORC_RAIN = 0, ORC_SNOW, ORC_GRAUPEL, ORC_HAIL, ORC_U, ORC_V, ORC_W, ORC_THETA, ORC_Q, ORC_CLOUD, ORC_ICE, ORC_ICENUMBER,
  ORC_COLD, ORC_CHARGE_ICE, ORC_CHARGE_GRAUPEL, ORC_CHARGE_RAIN, ORC_CHARGE_AIR,
  ORC_CARRIED = 17,
  ORC_FIELDZ = 17, ORC_FALL_RAIN, ORC_FALL_SNOW, ORC_FALL_GRAUPEL, ORC_FALL_HAIL, ORC_FALL_CHARGE_ICE, ORC_FALL_CHARGE_GRAUPEL,
  ORC_FIELDS = 24
};
/* Column ids (STATE_COLUMNS, src/core/state.ts:10). */
/* Then an ensemble member's swaths (src/ensemble/member.ts:19-21): the strongest surface wind and 2-5 km helicity. */
enum { ORC_PRECIPITATION = 0, ORC_HAIL_GROUND, ORC_HAIL_SIZE, ORC_GUST_MAX, ORC_UH_MAX, ORC_COLUMNS };

/* Diagnostics vector (orc_diag): */
enum {
  ORC_D_TIME = 0, ORC_D_MAXW, ORC_D_MAXTHETAP, ORC_D_RAINTOTAL, ORC_D_FIRSTFLASH, ORC_D_FLASHES_IC, ORC_D_FLASHES_CG,
  ORC_D_CLIPPED, ORC_D_NEGFILLED, ORC_D_CHARGEGROUND, ORC_D_MINW, ORC_D_WATER, ORC_D_COUNT
};

void *orc_create(const orc_grid *grid, const orc_profile *profile, const orc_run *run);
void orc_destroy(void *model);
/* n steps of dt seconds (AtmosphereModel.step, then time += dt as the callers do). */
void orc_step(void *model, double dt, int n);
/* Switches lightning (discharge) on (1) or off (0); on by default (SimConfig.lightning). */
void orc_set_lightning(void *model, int on);
int orc_get_field(void *model, int field, float *out);
int orc_set_field(void *model, int field, const float *in);
int orc_get_column(void *model, int column, float *out);
int orc_set_column(void *model, int column, const float *in);
/* Base state per level: p, exner, theta, q, u, v, rho (7 x nz float64). */
void orc_base_state(void *model, double *out);
void orc_grid_levels(void *model, double *zs, double *hz);
void orc_diag(void *model, double *out);
/* The model's projection pressure (cells) after the last step, (nz-1) * layer float64. */
void orc_pressure(void *model, double *out);
void orc_set_time(void *model, double t);
/* intervene (intervene.ts:38-68): kind 0-6 = thermal, warm, cool, moisten, dry, coldPool, vortex at (x, y) m from the
 * domain's corner; thermal_radius (km, <= 0: 4.2) is the thermal's e-folding radius (SimConfig.bubbleRadius). out[2]:
 * the heat (J) and vapour (kg) put in. 0, or -1 for an unknown kind. */
int orc_intervene(void *model, int kind, double x, double y, double strength, double thermal_radius, double *out);
/* AtmosphereModel.added: heat (J), vapour (kg), count. */
void orc_added(void *model, double *out);
/* Rotation diagnostics of the last step (updateRotation, model.ts:1609-1648), ORC_ROT float64 in this order. */
enum {
  ORC_ROT_UH = 0, ORC_ROT_X, ORC_ROT_Y, ORC_ROT_ANTI, ORC_ROT_PERSISTED, ORC_ROT_UH01, ORC_ROT_UH03, ORC_ROT_LOW,
  ORC_ROT_CORE_ZETA, ORC_ROT_CORE_UH, ORC_ROT_CORE_ANTI, ORC_ROT
};
void orc_rotation(void *model, double *out);
/* updateRotation(dt) on the fields as they are; keep != 0 then restores how long the rotation had held
 * (refreshRotation, model.ts:1607). */
void orc_update_rotation(void *model, double dt, int keep);
/* AtmosphereModel.diagnostics() (model.ts:1672-1696; ModelDiagnostics :161-180), ORC_MD_COUNT float64 in this order. */
enum {
  ORC_MD_UPDRAFT = 0, ORC_MD_DOWNDRAFT, ORC_MD_RAINRATE, ORC_MD_RAINTOTAL, ORC_MD_CLOUDTOP, ORC_MD_THERMALTOP,
  ORC_MD_MAXCLOUD, ORC_MD_COLDMAX, ORC_MD_CORES, ORC_MD_SHEAR06, ORC_MD_OUTFLOW, ORC_MD_OUTFLOWLOW, ORC_MD_GUST,
  ORC_MD_HAIL, ORC_MD_CLIPPED, ORC_MD_COUNT
};
void orc_model_diagnostics(void *model, double *out);
/* runMember's addSwaths after every step of orc_step (src/ensemble/member.ts:85-114) when on. */
void orc_set_swaths(void *model, int on);
/* addSwaths once, with the model's time and the last updateRotation's helicity columns; those columns (layer float32,
 * -1 before any updateRotation). */
void orc_add_swaths(void *model);
int orc_uh_column(void *model, float *out);
/* outflowDeficit (rkw.ts:72-80): the ground columns' deficits (layer float32, may be null) and the warm air's thv;
 * returns the largest deficit. deficitNear (rkw.ts:82-90): the largest of `deficit` within 2 km of `column`. */
double orc_outflow_deficit(void *model, float *deficit, double *warm);
double orc_deficit_near(void *model, const float *deficit, int column);

/* Known-answer surface of the pieces (for tests). */
double orc_weno(double a, double b, double c, double d, double e);
double orc_upwind5(double a, double b, double c, double d, double e);
void orc_fft(int n, double *re, double *im, int sign);
/* Forward 2-D real DFT of `layers` layers into the Hermitian half spectra re/im (nx*ny*layers each), and back. */
void orc_hdft_forward(int nx, int ny, int layers, const double *src, double *re, double *im);
void orc_hdft_inverse(int nx, int ny, int layers, double *re, double *im, double *dst);
double orc_qsat(double tc, double p);
double orc_qsat_ice(double tc, double p);
double orc_fall_speed(double rain, double rho, double rho_ground);
double orc_snow_fall_speed(double snow, double rho, double rho_ground);
double orc_graupel_fall_speed(double g, double rho);
double orc_hail_fall_speed(double h, double rho);
/* rainProcesses on one node (q, cloud, rain as float32 in/out); returns the evaporated rain. */
double orc_rain_processes(float *q, float *cloud, float *rain, double tk, double p, double rho, double dt);
double orc_saturation_adjust(float *theta, float *q, float *cloud, double exner, double p);
double orc_takahashi(double tc, double lwc);
double orc_takahashi_collision(double tc, double lwc, double d, double dv);
int orc_flash_limit(double dt);
double orc_stretch_ratio(double bottom, double height, int intervals);
/* Pressure projection on its own: random-test surface (u, v, w float32 in/out, rho per level); returns max |div| after. */
double orc_project(const orc_grid *grid, const double *rho, float *u, float *v, float *w, double dt, double *div_before);
/* Face projection + RK3 advection of one field (FluxTransport.setVelocity + advect), for the transport tests. */
void orc_advect(const orc_grid *grid, const double *rho, const float *u, const float *v, const float *w,
                float *q, double dt, int steps, int monotone);
uint32_t orc_mulberry32(uint32_t *state);
/* Electric potential of a node charge density (ElectricSolver.solve): rho, phi are n float64. */
void orc_electric_solve(const orc_grid *grid, const double *rho, double *phi);

/* The column physics' pieces on one node, float32 fields in/out (ice.ts, graupel.ts, hail.ts, charge.ts). */
double orc_sat_adjust_mixed(float *theta, float *q, float *cloud, float *ice, double exner, double p, float *number /* may be null */);
void orc_freeze_rain(float *theta, float *rain, float *graupel, double exner);
double orc_ice_processes(float *theta, float *q, float *ice, float *snow, float *number, double exner, double p, double rho, double dt);
void orc_settle_crystals(float *ice, float *number);
double orc_snow_processes(float *theta, float *q, float *cloud, float *ice, float *snow, float *rain, double exner, double p, double rho, double rho_ground, double dt);
double orc_graupel_processes(float *theta, float *q, float *cloud, float *ice, float *snow, float *rain, float *graupel, double exner, double p, double rho, double dt);
double orc_hail_processes(float *theta, float *q, float *cloud, float *ice, float *snow, float *rain, float *graupel, float *hail, double exner, double p, double rho, double dt);
double orc_hail_max_diameter(double h, double rho);
/* wetGrowing (hail.ts:55-60) of graupel (rimer 0) or hail (1). */
int orc_wet_growing(int rimer, double q, double tc, double qv, double p, double rho, double cloud, double rain, double ice, double snow);
double orc_cooper_crystals(double tc);
double orc_splintering(double tc);
double orc_fragments_per_kg(double lambda);
double orc_koenig_growth(double tc, double m);
double orc_crystal_size(double qi, double n, double rho);
/* Constants the tests use: 0 LF, 1 LS, 2 CP, 3 KAPPA, 4 SPLINTERS, 5 CRYSTAL_MASS, 6 LV. */
double orc_constant(int which);
/* chargingRate (charge.ts:138-158, Takahashi scheme); have_crystals 0: the reference's `crystals` undefined. */
double orc_charging_rate(double tc, double rho, double cloud, double ice, double snow, double graupel, double hail, double snow_fall, int wet_g, int wet_h, double crystals, int have_crystals);
double orc_inductive_rate(double tc, double rho, double cloud, double graupel, double hail, double ez, double rimer_charge, int wet_g, int wet_h);
/* transferCharge (charge.ts:215-227): charge[4] in place from the masses before[4] and after[4]. */
void orc_transfer_charge(double *charge, const double *before, const double *after);
double orc_crystal_number(double tc, double qi, double rho);
double orc_crystal_diameter(double qi, double rho, double tc);

/* Turbulence.viscosity (explicit K; anisotropic 1 the baseline's, 0 the class default kh = km; condensate null: dry)
 * and Turbulence.mix on a grid. */
void orc_viscosity(const orc_grid *grid, const float *u, const float *v, const float *w, const float *theta, const double *theta_env, double cs, double dt,
                   const float *q, const float *condensate, const double *exner, double *km, double *kh, int anisotropic);
void orc_mix(const orc_grid *grid, const double *rho, const double *km, const double *kh, float *a, const double *base, double factor, double dt);
/* fallParts and sediment (model.ts:565-598) of a model's field with the speeds in field `fall`; deposit 1 lands it in
 * the precipitation column (under the moving ground at the model's time), 0 counts it as charge to the ground. */
int orc_fall_parts(void *model, int field, int fall, double dt);
int orc_sediment(void *model, int field, int fall, double dt, int deposit);
double orc_charge_ground(void *model);

/* Lightning (lightning.ts) on its own: a grid, the air density per level and the seed. */
void *orc_lightning_create(const orc_grid *grid, const double *air, int seed);
void orc_lightning_destroy(void *lightning);
void orc_lightning_solve(void *lightning, float *ice, float *graupel, float *rain, float *air);
/* strongest(): the largest |E| / E_init and its node. */
double orc_lightning_strongest(void *lightning, int *node);
/* discharge(carriers, time, limit): the flashes, 8 float64 each: time, CG, start node, channel nodes, potential, charge,
 * charge moment, ground node (-1 for an IC); returns their number. orc_lightning_channel: flash k's channel nodes, their
 * parents and polarities. */
int orc_lightning_discharge(void *lightning, float *ice, float *graupel, float *rain, float *air, double time, int limit, double *records, int max);
int orc_lightning_channel(void *lightning, int flash, int *nodes, int *parent, int *polarity, int max);
double orc_lightning_conduct(void *lightning, float *air_charge, double dt);
void orc_lightning_set_cloudy(void *lightning, const float *cloudy);
void orc_lightning_set_corona(void *lightning, int on);
void orc_lightning_phi(void *lightning, double *out);
void orc_lightning_totals(void *lightning, double *from_ground, double *conducted);

#ifdef __cplusplus
}
#endif
#endif
