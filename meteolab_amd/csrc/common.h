// Layout of the StormLab step in HBM and the parameters every kernel receives.
//
// Runs of one grid size live in a pool of slots. Each array kind is one allocation [slot][...], so a batch of runs is
// one launch over the list of active slots (blockIdx.y or .z indexes the list) and a run keeps its slot, and its
// state stays resident, from admission to release. Fields are float32 (the reference's Float32Array fields); the
// intermediate arrays (face mass fluxes, spectra, eddy viscosities, potentials) have the arithmetic type R of the
// kernels: float in production, double in the float64 variant of the same kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace ml {

// Field ids: the carried order of AtmosphereModel.carried() (src/core/model.ts:1319-1335), then the per-node arrays a
// step recomputes: the vertical electric field (model.ts:1550-1551) and the fall speeds (model.ts:1220-1235).
enum : int {
  F_RAIN = 0, F_SNOW, F_GRAUPEL, F_HAIL, F_U, F_V, F_W, F_THETA, F_Q, F_CLOUD, F_ICE, F_ICENUMBER, F_COLD,
  F_CHARGE_ICE, F_CHARGE_GRAUPEL, F_CHARGE_RAIN, F_CHARGE_AIR, N_CARRIED,
  F_FIELDZ = N_CARRIED, F_FALL_RAIN, F_FALL_SNOW, F_FALL_GRAUPEL, F_FALL_HAIL, F_FALL_CI, F_FALL_CG, F_CONDENSATE,
  // Internal: cloud + ice + snow (float32) as the lightning sees it, written with the electric solve's right-hand side.
  F_CLOUDY,
  N_FIELDS
};
// Ground columns (STATE_COLUMNS, src/core/state.ts:10).
// Ground columns: STATE_COLUMNS (state.ts:10), then an ensemble member's swaths (ensemble/member.ts:19-21 gustMax, uhMax).
enum : int { C_PRECIPITATION = 0, C_HAIL_GROUND, C_HAIL_SIZE, C_GUST_MAX, C_UH_MAX, N_COLUMNS };
// Per-level tables of a run (double, nz each): the base state (src/core/environment.ts:124-130) and what the step
// derives from it.
enum : int {
  L_P = 0, L_EXNER, L_THETA, L_Q, L_U, L_V, L_RHO, L_UENV, L_VENV, L_LW, L_EINIT, L_CAPACITANCE,
  // The level coefficients of the vertical operators of the face projection (F0) and the pressure (P1): a mode's row
  // is these with its horizontal symbols (modeSolve).
  L_F0LO, L_F0UP, L_F0RH, L_P1A, L_P1B, L_P1C, L_P1D, L_P1E, N_LEVELS
};
// Per-run parameters (double).
enum : int {
  P_FRAME_U = 0, P_FRAME_V, P_CORIOLIS, P_CS, P_DRAG, P_SPONGE_START, P_SEED, P_BUBBLE_AMP, P_BUBBLE_X, P_BUBBLE_Y,
  P_BUBBLE_Z, P_BUBBLE_RH, P_BUBBLE_RZ, P_NOISE_SEED, P_NOISE_AMP, P_NOISE_TOP, P_QVMAX, P_SHEAR, P_SHEAR_DEPTH,
  P_THETA0, P_DTHETA, P_ZTR, P_TTR, P_RH_DROP, P_RH_TOP, P_ELECTRIFIED, P_TIME,
  P_SWATHS,  // nonzero: each step folds its surface wind and updraught helicity into the swath columns
  P_WIND_U,  // 4 words each: the sounding's wind errors at 0, 3, 6 and 10 km (m/s)
  P_WIND_V = P_WIND_U + 4,
  N_PARAMS = 40
};
static_assert(P_WIND_V + 4 <= N_PARAMS, "per-run parameters overflow");
// The sounding's wind error at height z from its errors d at 0, 3, 6 and 10 km: linear between those heights, held
// below and above them (perturbConfig's `at`, src/ensemble/member.ts:56).
__host__ __device__ inline double windError(const double *d, double z) {
  const double at[4] = {0, 3000, 6000, 10000};
  int k = 0;
  while (k < 2 && z >= at[k + 1]) k++;
  const double f = (z - at[k]) / (at[k + 1] - at[k]);
  return d[k] + (d[k + 1] - d[k]) * (f < 0 ? 0.0 : f > 1 ? 1.0 : f);
}
// Per-run scalars (8-byte words). Per step: cleared by k_begin; the rest accumulate over the run.
enum : int {
  S_COURANT = 0,         // 7 words: the largest Courant number of each fall (float bits, atomicMax on uint32)
  S_FILL = 8,            // 7 x 3 doubles: removed, kept, negative of the positivity fix of each conserved field
  S_CHARGED = 29,        // uint64: nonzero when a node of the run holds charge at the step's electric solve
  S_STEP_END = 30,       // [0, S_STEP_END) is cleared at the start of each step
  S_CLIPPED = 32,        // int64: velocity components clipped since the start (model.ts:1586-1595)
  S_NEGATIVE = 33,       // double: negative mass removed by the positivity fix since the start (model.ts:1119)
  S_CHARGE_GROUND = 34,  // double: charge carried to the ground by falling carriers (model.ts:584-587)
  S_FLASH_IC = 35, S_FLASH_CG = 36,  // int64 flash counts (model.ts:1541 flashCount)
  S_FIRST_FLASH = 37,    // double: time of the first flash, s (-1: none yet)
  S_ERROR = 38,          // int64: nonzero when a kernel of the run met a condition it does not handle
  S_FLASH_ID = 39,       // int64: the lightning's stamp counter (lightning.ts flashId)
  S_FROM_GROUND = 40,    // double: charge the CG flashes took from the ground (lightning.ts fromGround)
  S_FLASH_LOG = 41,      // int64: flashes recorded in the run's flash log (a ring of FLASH_LOG entries)
  S_STRONGEST = 42,      // uint64 key of the largest (|E|/E_init)^2 (float bits << 32 | ~node) of the step's electric
                         // solve; cleared with its right-hand side, read by the discharge (which may run during the
                         // next step's dynamics)
  S_ROT = 48,            // the rotation diagnostics (model.ts:142-158 RotationState), ROT_* from here
  S_ADDED = 59,          // 3 doubles: heat (J), vapour (kg) put in by interventions and their count (model.ts:227-228)
  S_LAST_ADDED = 62,     // 2 doubles: the heat and vapour of the run's last intervention
  N_SCALARS = 64
};
enum : int { ROT_UH = 0, ROT_X, ROT_Y, ROT_ANTI, ROT_PERSISTED, ROT_UH01, ROT_UH03, ROT_LOW, ROT_CORE_ZETA, ROT_CORE_UH, ROT_CORE_ANTI, N_ROT };
static_assert(S_ROT + N_ROT <= S_ADDED, "the rotation diagnostics overlap the intervention counts");
// Lightning channel limits (lightning.ts:32-37) and the per-run flash log.
constexpr int MAX_CHANNEL = 3000, STROKES = 4, FLASH_LOG = 64, FLASH_FIELDS = 8;
// The lightning's per-node int arrays and per-front-entry arrays of a slot.
enum : int { LM_MARK = 0, LM_FAILED, LM_SEEN, LM_POS, LM_FNODE, LM_FLINK, LM_FFLAG, N_LMARK };
enum : int { LN_FSCALE = 0, LN_FPHI, LN_FWEIGHT, LN_NSCALE, N_LNODE };
// A flash log entry: time, CG (1) or IC (0), start node, channel nodes, charge (C), potential (V), charge moment (C m).
enum : int { FL_TIME = 0, FL_CG, FL_START, FL_NODES, FL_CHARGE, FL_POTENTIAL, FL_MOMENT };
// The falls of a step in order (model.ts:928-932, sedimentStep): water into the gauge, then charge with its carrier's
// speed.
constexpr int N_FALLS = 7;
// Sedimentation sub-steps of a water fall whose ground-node values a step keeps for the rain gauge (fallParts: large
// hail through the 100 m ground levels takes 2 at dt = 3 s). A fall that needs more flags S_ERROR (ERR_LAND_PARTS).
constexpr int LAND_PARTS = 16;
// S_ERROR bits.
enum : long long { ERR_PRUNE = 1, ERR_LAND_PARTS = 2 };
// The conserved (positivity-fixed) carried fields (model.ts:1319-1327): rain, snow, graupel, hail, q, cloud, ice.
constexpr int N_CONSERVED = 7;

// Geometry of a size class (src/core/grid.ts:99-121).
struct Geo {
  int nx, ny, nz, layer, n;
  int mw, nm;            // kept Hermitian columns mw = nx/2 + 1 and modes nm = mw * ny
  double dx, dy, dz, height;
};

// Device arrays of a pool (all [slot][...]) and the class tables. R: arithmetic type of the intermediate arrays.
template <class R> struct Bufs {
  float *fields;   // [slot][N_FIELDS][n]
  float *stx;      // [slot][N_CARRIED][n] RK3 stage buffer X (stage 0 output) / mixed fields
  float *sty;      // [slot][N_CARRIED][n] RK3 stage buffer Y (stage 1 output)
  double *mass;    // [slot][3][n]  face mass fluxes mx, my, mz (float64 in both variants)
  double *phi;     // [slot][n]     potential of the face projection (float64, the spectral solves' output)
  double *cells;   // [slot][n]     right-hand side of a spectral solve (cells or interior layers) (float64)
  double *press;   // [slot][n]     pressure of the projection (model.ts:358), cells (float64)
  double *spec;    // [slot][nz][nm][2] Hermitian half spectra, re/im interleaved (float64: see spectral.h)
  double *ehat;    // [slot][nz][nm][2] the discharge's spectra of the electric right-hand side (kept between its solves)
  double *mean0;   // [slot][nz]    float64 layer sums of the right-hand side (the m = n = 0 mode)
  double *emean0;  // [slot][nz]    the electric solve's layer sums (kept for the discharge, as ehat)
  double *dspec;   // [slot][nz][nm][2] the discharge's own spectra (so that it can run beside the dynamics' solves)
  R *km, *kh;      // [slot][n]     eddy viscosities
  double *ephi;    // [slot][n]     electric potential (electric.ts:36-60) (float64)
  float *cols;     // [slot][N_COLUMNS][layer]
  double *lev;     // [slot][N_LEVELS][nz]
  double *par;     // [slot][N_PARAMS]
  double *scal;    // [slot][N_SCALARS]
  const double *zs, *dzs, *hz;   // class levels (nz, nz - 1, nz)
  const double *sym;  // [4][nm] per kept mode: face-projection lambda, pressure alpha, beta, electric lambda
  const double *twx;  // [nx][2] exp(-2 pi i k / nx)
  const double *twy;  // [ny][2] exp(-2 pi i k / ny)
  const int *runs; // active slots of this launch
  // Lightning state of each slot (lightning.ts:135-170): node charge density, the channel's per-node marks and front.
  R *lrho;         // [slot][n]     charge density of the last solve, C/m3
  // Per node: mark, failed, seen (flash stamps) and the node's place in the front; per front entry (the front of a
  // flash is a list in the reference's order): node, link (channel index) and flags.
  int *lmark;      // [slot][7][n]
  double *lnode;   // [slot][4][n]  per front entry: scale (link length times threshold), potential, weight; per node: scale
  const double *efac; // [3][nz - 2][nm] the electric modes' elimination: pivots, upper factors, lower coefficients
  int *chan;       // [slot][2][MAX_CHANNEL] channel nodes and their parents
  double *cdep;    // [slot][MAX_CHANNEL] charge deposited on each channel node
  double *flog;    // [slot][FLASH_LOG][FLASH_FIELDS]
  // Sums kept in parts so that they add in a fixed order, whatever order the blocks run in:
  double *fillp;   // [slot][N_CONSERVED][nz][ftiles][3] the positivity-fix sums of each level (and strip tile)
  int ftiles;      // strip tiles of a level (stageTile)
  double *gpart;   // [slot][2][gmax] a launch's block parts of the charge carried to the ground (falls, conduction)
  int gmax;
  unsigned char *zflag;  // [slot][N_CARRIED][nz] a level of a carried field holds a non-zero node (this step's transport)
  float *landed;   // [slot][4][LAND_PARTS][layer] each water fall's ground-node values at the start of its sub-steps

  __device__ float *f(int slot, int k, const Geo &g) const { return fields + (size_t(slot) * N_FIELDS + k) * g.n; }
  __device__ float *x(int slot, int k, const Geo &g) const { return stx + (size_t(slot) * N_CARRIED + k) * g.n; }
  __device__ float *y(int slot, int k, const Geo &g) const { return sty + (size_t(slot) * N_CARRIED + k) * g.n; }
  // Buffer set: 0 the state, 1 stage buffer X, 2 stage buffer Y.
  __device__ float *set(int which, int slot, int k, const Geo &g) const { return which == 0 ? f(slot, k, g) : which == 1 ? x(slot, k, g) : y(slot, k, g); }
  __device__ double *m(int slot, int k, const Geo &g) const { return mass + (size_t(slot) * 3 + k) * g.n; }
  template <class T> __device__ T *arr(T *base, int slot, const Geo &g) const { return base + size_t(slot) * g.n; }
  __device__ double *sp(int slot, const Geo &g) const { return spec + size_t(slot) * g.nz * g.nm * 2; }
  __device__ double *eh(int slot, const Geo &g) const { return ehat + size_t(slot) * g.nz * g.nm * 2; }
  __device__ double *m0(int slot, const Geo &g) const { return mean0 + size_t(slot) * g.nz; }
  __device__ double *em0(int slot, const Geo &g) const { return emean0 + size_t(slot) * g.nz; }
  __device__ double *dsp(int slot, const Geo &g) const { return dspec + size_t(slot) * g.nz * g.nm * 2; }
  __device__ float *col(int slot, int k, const Geo &g) const { return cols + (size_t(slot) * N_COLUMNS + k) * g.layer; }
  __device__ double *lv(int slot, int k, const Geo &g) const { return lev + (size_t(slot) * N_LEVELS + k) * g.nz; }
  __device__ double *pr(int slot) const { return par + size_t(slot) * N_PARAMS; }
  __device__ double *sc(int slot) const { return scal + size_t(slot) * N_SCALARS; }
  __device__ double *fillPart(int slot, int cf, int z, int tile, const Geo &g) const {
    return fillp + (((size_t(slot) * N_CONSERVED + cf) * g.nz + z) * ftiles + tile) * 3;
  }
  __device__ unsigned char *zf(int slot, int f, const Geo &g) const { return zflag + (size_t(slot) * N_CARRIED + f) * g.nz; }
  __device__ float *land(int slot, int k, int part, const Geo &g) const { return landed + ((size_t(slot) * 4 + k) * LAND_PARTS + part) * g.layer; }
  __device__ int *lm(int slot, int k, const Geo &g) const { return lmark + (size_t(slot) * N_LMARK + k) * g.n; }
  __device__ double *ln(int slot, int k, const Geo &g) const { return lnode + (size_t(slot) * N_LNODE + k) * g.n; }
};

// Quotient of 0 <= a < 2^22 by a divisor 1 <= d < 2^22 fixed for a loop: a float product with the divisor's reciprocal
// is within one of the quotient, and one correction step makes it exact (integer division by a variable is a long
// instruction sequence on the GPU).
struct Div {
  int d; float inv;
  __device__ explicit Div(int d_) : d(d_), inv(1.0f / float(d_)) {}
  __device__ int q(int a) const { int v = int(float(a) * inv), r = a - int(__umul24(unsigned(v), unsigned(d))); return v + (r >= d) - (r < 0); }
};
// Product of two non-negative ints below 2^24 (a full-rate 24-bit multiply; a 32-bit one is a quarter-rate instruction).
__device__ inline int mul24(int a, int b) { return int(__umul24(unsigned(a), unsigned(b))); }

// ---- wave64 / block reductions --------------------------------------------------------------------------------
template <class T> __device__ inline T waveSum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
template <class T> __device__ inline T waveMax(T v) {
  for (int o = 32; o > 0; o >>= 1) { T w = __shfl_xor(v, o, 64); v = w > v ? w : v; }
  return v;
}
// Sum over a block of up to 1024 threads (blockDim.x a multiple of 64); every thread gets the total.
template <class T> __device__ inline T blockSum(T v, T *shared) {
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
  v = waveSum(v);
  __syncthreads();
  if (lane == 0) shared[wave] = v;
  __syncthreads();
  T t = 0;
  for (int w = 0; w < waves; w++) t += shared[w];
  return t;
}
// Whether the run held no charge anywhere at this step's electric solve (k_charge_rhs): its potential is zero.
template <class B> __device__ inline bool unchargedRun(const B &b, int slot) {
  return reinterpret_cast<const unsigned long long *>(b.sc(slot))[S_CHARGED] == 0;
}
template <class T> __device__ inline T blockMax(T v, T *shared) {
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
  v = waveMax(v);
  __syncthreads();
  if (lane == 0) shared[wave] = v;
  __syncthreads();
  T t = shared[0];
  for (int w = 1; w < waves; w++) t = shared[w] > t ? shared[w] : t;
  return t;
}

// A launch's block part of a run's sum (the run's blocks are blockIdx.x of a grid whose y is the run), kept so that a
// later kernel (k_ground_total) adds the parts in block order: the sum does not depend on the order the blocks run in.
template <class B> __device__ inline void storePart(const B &b, int slot, int which, double t) {
  if (threadIdx.x == 0) b.gpart[(size_t(slot) * 2 + which) * b.gmax + blockIdx.x] = t;
}

}  // namespace ml
