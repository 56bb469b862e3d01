// Data layout of the StormLab step in HBM: a pool of fixed-size run slots per size class.
//
// Every kernel sees the pool base, the slot stride and the list of active slots; a run's arrays live at fixed
// offsets inside its slot, so a batch of runs is one launch whatever runs it holds, and a run keeps its slot (its
// state stays resident) from admission to release.
#pragma once
#include <cstddef>
#include <cstdint>

namespace ml {

// Field ids: the carried order of AtmosphereModel.carried() (src/core/model.ts:1319-1335), then the per-node arrays
// a step recomputes (fall speeds: model.ts:1220-1235; fieldZ: model.ts:1550-1551).
enum : int {
  F_RAIN = 0, F_SNOW, F_GRAUPEL, F_HAIL, F_U, F_V, F_W, F_THETA, F_Q, F_CLOUD, F_ICE, F_ICENUMBER, F_COLD,
  F_CHARGE_ICE, F_CHARGE_GRAUPEL, F_CHARGE_RAIN, F_CHARGE_AIR, N_CARRIED,
  F_FIELDZ = N_CARRIED, F_FALL_RAIN, F_FALL_SNOW, F_FALL_GRAUPEL, F_FALL_HAIL, F_FALL_CI, F_FALL_CG, N_FIELDS
};
// Ground columns (STATE_COLUMNS, src/core/state.ts:10).
enum : int { C_PRECIPITATION = 0, C_HAIL_GROUND, C_HAIL_SIZE, N_COLUMNS };
// Per-level tables of a run (double, nz each): the base state (src/core/environment.ts:124-130) and what the step
// derives from it.
enum : int {
  L_P = 0, L_EXNER, L_THETA, L_Q, L_U, L_V, L_RHO, L_UENV, L_VENV, L_LW, L_EINIT, L_CAPACITANCE, N_LEVELS
};
// Per-run parameters (double).
enum : int {
  P_FRAME_U = 0, P_FRAME_V, P_CORIOLIS, P_CS, P_DRAG, P_SPONGE_START, P_SEED, P_BUBBLE_AMP, P_BUBBLE_X, P_BUBBLE_Y,
  P_BUBBLE_Z, P_BUBBLE_RH, P_BUBBLE_RZ, P_NOISE_SEED, P_NOISE_AMP, P_NOISE_TOP, P_QVMAX, P_SHEAR, P_SHEAR_DEPTH,
  P_THETA0, P_DTHETA, P_ZTR, P_TTR, P_RH_DROP, P_RH_TOP, N_PARAMS = 32
};
// Per-run scalars (8-byte words) written by the kernels of a step.
enum : int {
  S_COURANT = 0,           // 7 words: the largest Courant number of each fall (uint64 bits of a double)
  S_FILL = 8,              // 7 x 3 doubles: removed, kept, negative of the positivity fix (fillNegative)
  S_STRONGEST = 32,        // uint64 key: |E|/E_init squared (float bits) << 32 | ~node (model.ts discharge)
  S_CLIPPED = 33,          // int64 velocity components clipped this step
  S_NEGATIVE = 34,         // double: the negative mass the positivity fix removed this step
  N_SCALARS = 48
};
// The falls of a step in order (model.ts:928-932): water into the gauge, then charge with its carrier's speed.
constexpr int N_FALLS = 7;

struct Dims {
  int nx, ny, nz, layer, n, cells, mh, nm;  // nm: modes kept, (nx/2 + 1) * ny
  double dx, dy, dz, height, width, depth;
};

// Byte offsets of a run's arrays inside its slot; `real` is the byte size of the arithmetic type.
struct Offsets {
  size_t fields;    // float [N_FIELDS][n]
  size_t tmpA;      // float [N_CARRIED][n]  RK3 stage buffer / mixed fields
  size_t tmpB;      // float [N_CARRIED][n]  RK3 stage buffer
  size_t columns;   // float [N_COLUMNS][layer]
  size_t mx, my, mz, div, phi;   // R [n]   face mass fluxes, their divergence, the face potential
  size_t cells;     // R [(nz-1) layer]     projection right-hand side / electric right-hand side
  size_t pressure;  // R [(nz-1) layer]     projection pressure (model.ts:247)
  size_t spec;      // R [2][nz][nm]        Hermitian half spectra, interleaved re/im
  size_t km, kh;    // R [n]   eddy viscosities
  size_t ephi;      // R [n]   electric potential
  size_t levels;    // double [N_LEVELS][nz]
  size_t params;    // double [N_PARAMS]
  size_t scalars;   // 8-byte words [N_SCALARS]
  size_t chargeOut; // double [layer] charge carried to the ground per column (diagnostic)
  size_t stride;    // bytes per slot
};

}  // namespace ml
