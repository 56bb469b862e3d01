// Data-parallel kernels of the StormLab step, batched over runs: blockIdx.y (or .z) indexes the active-slot list,
// blockIdx.x tiles a run. R is the arithmetic type (float: production; double: the float64 variant of the same
// kernels). Fields are float32 in both. Each kernel names the reference function it computes.
#pragma once
#include "common.h"
#include "physics.h"

namespace ml {

constexpr int TPB = 256;  // threads per block of the per-node kernels (four wave64)

__device__ inline int idiv(int a, int b) { return a / b; }

// ---- step start: clear the per-step scalars of each run ----------------------------------------------------------
template <class R> __global__ void k_begin(Geo g, Bufs<R> b) {
  int slot = b.runs[blockIdx.x];
  double *s = b.sc(slot);
  for (int t = threadIdx.x; t < S_STEP_END; t += blockDim.x) s[t] = 0;
}

// ---- base state (src/core/environment.ts:45-73, profiles.ts:61-73) in float64, one block per run -----------------
__device__ inline double profTheta(const double *P, double z) {
  return z <= P[P_ZTR] ? P[P_THETA0] + P[P_DTHETA] * pow(z / P[P_ZTR], 1.25) : (P[P_THETA0] + P[P_DTHETA]) * exp(G * (z - P[P_ZTR]) / (CP * P[P_TTR]));
}
__device__ inline double profRh(const double *P, double z) { return z <= P[P_ZTR] ? 1 - P[P_RH_DROP] * pow(z / P[P_ZTR], 1.25) : P[P_RH_TOP]; }
__device__ inline double tempAt(const double *P, double z, double p) { return profTheta(P, z) * pow(p / 100000, KAPPA) - 273.15; }
__device__ inline double mixRatio(const double *P, double z, double t, double p) { double r = profRh(P, z) * qsatP(t, p); return r < P[P_QVMAX] ? r : P[P_QVMAX]; }
__device__ inline double virtualT(const double *P, double z, double p) { double t = tempAt(P, z, p); return (t + 273.15) * (1 + .61 * mixRatio(P, z, t, p)); }
// The hydrostatic integration every 10 m (environment.ts:64-73) is a serial recurrence: thread 0 of the block runs it
// into shared memory, then the levels are read off in parallel.
template <class R> __global__ void k_base_state(Geo g, Bufs<R> b) {
  int slot = b.runs[blockIdx.x];
  const double *P = b.pr(slot);
  extern __shared__ double lnP[];
  double top = g.height + 2000; int count = int(ceil(top / 10)) + 1;
  if (threadIdx.x == 0) {
    lnP[0] = log(101325.0);
    for (int k = 0; k + 1 < count; k++) {
      double z = k * 10.0, p = exp(lnP[k]);
      lnP[k + 1] = lnP[k] - G * 10 / (RD * .5 * (virtualT(P, z, p) + virtualT(P, z + 10, p)));
    }
  }
  __syncthreads();
  for (int z = threadIdx.x; z < g.nz; z += blockDim.x) {
    double alt = b.zs[z], x = (alt > 0 ? alt : 0) / 10; int k = int(floor(x)); if (k > count - 2) k = count - 2;
    double p = exp(lnP[k] + (lnP[k + 1] - lnP[k]) * (x - k)), exner = pow(p / 100000, KAPPA);
    double tEnv = tempAt(P, alt, p), theta = (tEnv + 273.15) * pow(100000 / p, KAPPA), q = mixRatio(P, alt, tEnv, p);
    double u = P[P_SHEAR] * tanh(alt / P[P_SHEAR_DEPTH]) + windError(P + P_WIND_U, alt), v = windError(P + P_WIND_V, alt);
    double rho = p / (RD * theta * exner * (1 + .61 * q));
    b.lv(slot, L_P, g)[z] = p; b.lv(slot, L_EXNER, g)[z] = exner; b.lv(slot, L_THETA, g)[z] = theta; b.lv(slot, L_Q, g)[z] = q;
    b.lv(slot, L_U, g)[z] = u; b.lv(slot, L_V, g)[z] = v; b.lv(slot, L_RHO, g)[z] = rho;
    b.lv(slot, L_UENV, g)[z] = u - P[P_FRAME_U]; b.lv(slot, L_VENV, g)[z] = v - P[P_FRAME_V];
    b.lv(slot, L_LW, g)[z] = b.hz[z] / g.dz * rho;                       // levelWeight (model.ts:362)
    b.lv(slot, L_EINIT, g)[z] = 2.84e5 * rho / 1.225;                     // lightning.ts:172
    double L = cbrt(g.dx * g.dy * b.hz[z]);
    b.lv(slot, L_CAPACITANCE, g)[z] = .25 * 2 * PI * EPSILON0 * L / log(L / 1.0);  // lightning.ts:175
  }
  __syncthreads();
  // The level coefficients of the face projection's operator (advection.ts:222-241, nz levels) and the pressure's
  // (pressure.ts:104-140, nz - 1 levels), with the divisions done once per run.
  const double *rho = b.lv(slot, L_RHO, g), *dzs = b.dzs, *hz = b.hz;
  const int L0 = g.nz, L1 = g.nz - 1;
  for (int k = threadIdx.x; k < g.nz; k += blockDim.x) {
    auto face = [&](int j) { return j < 0 || j >= L0 - 1 ? 0.0 : (rho[j] + rho[j + 1]) / 2; };
    b.lv(slot, L_F0LO, g)[k] = k > 0 ? face(k - 1) / dzs[k - 1] : 0.0;
    b.lv(slot, L_F0UP, g)[k] = k < L0 - 1 ? face(k) / dzs[k] : 0.0;
    b.lv(slot, L_F0RH, g)[k] = rho[k] * hz[k];
    if (k < L1) {
      b.lv(slot, L_P1A, g)[k] = rho[k] * dzs[k] / (2 * hz[k]) + rho[k + 1] * dzs[k] / (2 * hz[k + 1]);
      b.lv(slot, L_P1B, g)[k] = k > 0 ? rho[k] * dzs[k - 1] / (2 * hz[k]) : 0.0;
      b.lv(slot, L_P1C, g)[k] = k > 0 ? rho[k] / (hz[k] * dzs[k]) : 0.0;
      b.lv(slot, L_P1D, g)[k] = k < L1 - 1 ? rho[k + 1] * dzs[k + 1] / (2 * hz[k + 1]) : 0.0;
      b.lv(slot, L_P1E, g)[k] = k < L1 - 1 ? rho[k + 1] / (hz[k + 1] * dzs[k]) : 0.0;
    }
  }
}

// ---- initial state (model.ts:399-451 in single-storm mode), the run's thermal and its theta noise -----------------
template <class R> __global__ void k_init(Geo g, Bufs<R> b) {
  int slot = b.runs[blockIdx.y];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n) return;
  int z = i / g.layer, y = (i - z * g.layer) / g.nx, x = i - z * g.layer - y * g.nx;
  const double *P = b.pr(slot);
  for (int f = 0; f < N_FIELDS; f++) b.f(slot, f, g)[i] = 0;
  b.f(slot, F_U, g)[i] = float(b.lv(slot, L_U, g)[z] - P[P_FRAME_U]);
  b.f(slot, F_V, g)[i] = float(b.lv(slot, L_V, g)[z] - P[P_FRAME_V]);
  float th = float(b.lv(slot, L_THETA, g)[z]);
  double ax = (x * g.dx - P[P_BUBBLE_X]) / P[P_BUBBLE_RH], ay = (y * g.dy - P[P_BUBBLE_Y]) / P[P_BUBBLE_RH], az = (b.zs[z] - P[P_BUBBLE_Z]) / P[P_BUBBLE_RZ];
  double beta = sqrt(ax * ax + ay * ay + az * az);
  if (beta < 1) { double c = cos(PI * beta / 2); th = float(th + P[P_BUBBLE_AMP] * (c * c)); }
  // Draw k of mulberry32(noise seed) goes to the k-th node below the noise top in node order: the nodes below it are
  // the first levels, so that is node i (draw i + 1).
  if (b.zs[z] < P[P_NOISE_TOP]) th = float(th + (2 * mulberryAt(uint32_t(int(P[P_NOISE_SEED])), uint32_t(i + 1)) - 1) * P[P_NOISE_AMP]);
  b.f(slot, F_THETA, g)[i] = th;
  b.f(slot, F_Q, g)[i] = float(b.lv(slot, L_Q, g)[z]);
  if (i < g.layer) for (int k = 0; k < N_COLUMNS; k++) b.col(slot, k, g)[i] = 0;
}

// ---- fall speeds and Courant numbers (model.ts:1220-1235 'fall', 575-580 fallParts) ---------------------------------
// chargeCourant: also the charge falls' Courant numbers (0 when k_charge_courant takes them after the lightning of the
// step before has changed the charges).
template <class R> __global__ __launch_bounds__(TPB) void k_fall(Geo g, Bufs<R> b, double dt, int electrified, int chargeCourant) {
  int slot = b.runs[blockIdx.y];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double cour[N_FALLS] = {0, 0, 0, 0, 0, 0, 0};
  if (i < g.n) {
    int z = i / g.layer;
    const double *rho = b.lv(slot, L_RHO, g);
    R rz = R(rho[z]), r0 = R(rho[0]);
    float rain = b.f(slot, F_RAIN, g)[i], snow = b.f(slot, F_SNOW, g)[i], graupel = b.f(slot, F_GRAUPEL, g)[i], hail = b.f(slot, F_HAIL, g)[i];
    float vr = float(fallSpeedRain(R(rain), rz, r0)), vs = float(snowFallSpeed(R(snow), rz, r0));
    float vg = float(graupelFallSpeed(R(graupel), rz)), vh = float(hailFallSpeed(R(hail), rz));
    b.f(slot, F_FALL_RAIN, g)[i] = vr; b.f(slot, F_FALL_SNOW, g)[i] = vs; b.f(slot, F_FALL_GRAUPEL, g)[i] = vg; b.f(slot, F_FALL_HAIL, g)[i] = vh;
    // fallParts: the Courant number c vt of each fall where its field is non-zero (float64, as it decides the sub-steps).
    double c = dt * rho[z] / (b.lv(slot, L_LW, g)[z] * g.dz);
    cour[0] = rain != 0 ? c * vr : 0; cour[1] = snow != 0 ? c * vs : 0; cour[2] = graupel != 0 ? c * vg : 0; cour[3] = hail != 0 ? c * vh : 0;
    if (electrified) {
      float ice = b.f(slot, F_ICE, g)[i];
      R is = R(ice) + R(snow), gh = R(graupel) + R(hail);
      float vci = float(is > 0 ? R(snow) * R(vs) / is : R(0));
      float vcg = float(gh > 0 ? (R(graupel) * R(vg) + R(hail) * R(vh)) / gh : R(0));
      b.f(slot, F_FALL_CI, g)[i] = vci; b.f(slot, F_FALL_CG, g)[i] = vcg;
      if (chargeCourant) {
        cour[4] = b.f(slot, F_CHARGE_ICE, g)[i] != 0 ? c * vci : 0;
        cour[5] = b.f(slot, F_CHARGE_GRAUPEL, g)[i] != 0 ? c * vcg : 0;
        cour[6] = b.f(slot, F_CHARGE_RAIN, g)[i] != 0 ? c * vr : 0;
      }
    }
  }
  __shared__ double part[N_FALLS][TPB / 64];
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int k = 0; k < N_FALLS; k++) { double m = waveMax(cour[k]); if (lane == 0) part[k][wave] = m; }
  __syncthreads();
  if (threadIdx.x < N_FALLS) {
    double m = 0;
    for (int w = 0; w < TPB / 64; w++) m = part[threadIdx.x][w] > m ? part[threadIdx.x][w] : m;
    // Non-negative doubles order as their bit patterns.
    if (m > 0) atomicMax(reinterpret_cast<unsigned long long *>(b.sc(slot) + S_COURANT + threadIdx.x), (unsigned long long)__double_as_longlong(m));
  }
}

// The charge falls' Courant numbers from the fall speeds k_fall stored (the same values as k_fall's own).
template <class R> __global__ __launch_bounds__(TPB) void k_charge_courant(Geo g, Bufs<R> b, double dt) {
  int slot = b.runs[blockIdx.y];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double cour[3] = {0, 0, 0};
  if (i < g.n) {
    int z = i / g.layer;
    const double c = dt * b.lv(slot, L_RHO, g)[z] / (b.lv(slot, L_LW, g)[z] * g.dz);
    cour[0] = b.f(slot, F_CHARGE_ICE, g)[i] != 0 ? c * b.f(slot, F_FALL_CI, g)[i] : 0;
    cour[1] = b.f(slot, F_CHARGE_GRAUPEL, g)[i] != 0 ? c * b.f(slot, F_FALL_CG, g)[i] : 0;
    cour[2] = b.f(slot, F_CHARGE_RAIN, g)[i] != 0 ? c * b.f(slot, F_FALL_RAIN, g)[i] : 0;
  }
  __shared__ double part[3][TPB / 64];
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int k = 0; k < 3; k++) { double m = waveMax(cour[k]); if (lane == 0) part[k][wave] = m; }
  __syncthreads();
  if (threadIdx.x < 3) {
    double m = 0;
    for (int w = 0; w < TPB / 64; w++) m = part[threadIdx.x][w] > m ? part[threadIdx.x][w] : m;
    if (m > 0) atomicMax(reinterpret_cast<unsigned long long *>(b.sc(slot) + S_COURANT + 4 + threadIdx.x), (unsigned long long)__double_as_longlong(m));
  }
}

// ---- sedimentation: flux-form upwind fall per column with sub-steps (model.ts:565-598, sedimentStep 928-932) -------
// One thread per column; the column of the falling field and its speed are staged in LDS, so the bottom-up sweep of
// each sub-step reads no memory.
constexpr int SED_COLS = 64;
// Falls k0 .. k1 - 1 of the step's order (the water falls 0-3 and the charge falls 4-6 are independent of each other).
template <class R> __global__ __launch_bounds__(SED_COLS) void k_sediment(Geo g, Bufs<R> b, double dt, int k0, int k1) {
  int slot = b.runs[blockIdx.y];
  int col = blockIdx.x * SED_COLS + threadIdx.x;
  extern __shared__ float sed[];
  float *qa = sed, *va = sed + SED_COLS * g.nz;  // [z][thread]
  const double *rho = b.lv(slot, L_RHO, g), *lw = b.lv(slot, L_LW, g);
  const int fields[N_FALLS] = {F_RAIN, F_SNOW, F_GRAUPEL, F_HAIL, F_CHARGE_ICE, F_CHARGE_GRAUPEL, F_CHARGE_RAIN};
  const int falls[N_FALLS] = {F_FALL_RAIN, F_FALL_SNOW, F_FALL_GRAUPEL, F_FALL_HAIL, F_FALL_CI, F_FALL_CG, F_FALL_RAIN};
  double charge = 0;
  bool live = col < g.layer;
  for (int k = k0; k < k1; k++) {
    double courant = __longlong_as_double((long long)reinterpret_cast<unsigned long long *>(b.sc(slot))[S_COURANT + k]);
    // No node moves anything (every non-zero value has zero speed): the fall changes nothing.
    if (!(courant > 0)) continue;
    int parts = int(ceil(courant / .8)); if (parts < 1) parts = 1;
    R pdt = R(dt / parts);
    float *a = b.f(slot, fields[k], g) + col; const float *vt = b.f(slot, falls[k], g) + col;
    if (live) for (int z = 0; z < g.nz; z++) { qa[z * SED_COLS + threadIdx.x] = a[size_t(z) * g.layer]; va[z * SED_COLS + threadIdx.x] = vt[size_t(z) * g.layer]; }
    // A water fall keeps its ground-node values of each sub-step; k_ground puts them into the gauge.
    const bool deposit = k < 4;
    if (deposit && parts > LAND_PARTS && live && col == 0)
      atomicOr(reinterpret_cast<unsigned long long *>(b.sc(slot) + S_ERROR), (unsigned long long)ERR_LAND_PARTS);
    if (live) {
      for (int p = 0; p < parts; p++) {
        float a0 = qa[threadIdx.x], v0 = va[threadIdx.x];
        if (deposit) { if (p < LAND_PARTS) b.land(slot, k, p, g)[col] = a0; }
        else charge += double(a0) * v0 * rho[0] * double(pdt) * g.dx * g.dy;
        float here = a0, hv = v0;
        for (int z = 0; z < g.nz; z++) {
          R inflow = 0; float above = 0, av = 0;
          if (z < g.nz - 1) { above = qa[(z + 1) * SED_COLS + threadIdx.x]; av = va[(z + 1) * SED_COLS + threadIdx.x]; inflow = R(rho[z + 1]) * (R(above) * R(av)); }
          qa[z * SED_COLS + threadIdx.x] = float(R(here) + pdt * (inflow - R(rho[z]) * R(here) * R(hv)) / R(lw[z] * g.dz));
          here = above; hv = av;
        }
      }
      for (int z = 0; z < g.nz; z++) a[size_t(z) * g.layer] = qa[z * SED_COLS + threadIdx.x];
    }
  }
  if (k1 > 4) {
    __shared__ double red[SED_COLS / 64];
    storePart(b, slot, 0, blockSum(charge, red));
  }
}

// ---- the ground under the moving domain (model.ts:637-657 frameOffset, updateGround; 691-706 deposit, collectHail) ----
// The domain moves over the ground at the frame velocity, so at time t domain column (x, y) lies over the four ground
// columns from (x + x0, y + y0) with bilinear weights, (x0 + fx, y0 + fy) being the distance moved in columns, periodic
// (the identity without frame motion). Here each ground column gathers from the domain columns over it (up to four),
// in the order the domain columns deposit theirs (by column, then by corner), so every ground value is the same chain
// of float32 sums as when each domain column adds to its four ground columns in turn.
struct GroundSources {
  int n;
  int col[4];    // the domain columns, in deposit order
  double w[4];   // the weight of this ground column in each one's footprint (> 0)
};
__device__ inline GroundSources groundSources(const Geo &g, const double *P, int ground) {
  const double nx = g.nx, ny = g.ny, ox = P[P_FRAME_U] * P[P_TIME], oy = P[P_FRAME_V] * P[P_TIME];
  const double sx = fmod(fmod(ox / g.dx, nx) + nx, nx), sy = fmod(fmod(oy / g.dy, ny) + ny, ny);
  const int x0 = int(floor(sx)), y0 = int(floor(sy));
  const double fx = sx - x0, fy = sy - y0;
  const double w[4] = {(1 - fx) * (1 - fy), fx * (1 - fy), (1 - fx) * fy, fx * fy};
  const int gx = ground % g.nx, gy = ground / g.nx;
  GroundSources s;
  s.n = 0;
  int key[4];
  for (int j = 0; j < 4; j++) {
    if (!(w[j] > 0)) continue;
    // Corner j of domain column (x, y) is ground column (x + x0 + (j & 1), y + y0 + (j >> 1)); 0 <= x0 < nx.
    const int x = (gx - x0 - (j & 1) + g.nx) % g.nx, y = (gy - y0 - (j >> 1) + g.ny) % g.ny;
    int k = s.n++;
    const int kk = (x + g.nx * y) * 4 + j;
    while (k > 0 && key[k - 1] > kk) { key[k] = key[k - 1]; s.w[k] = s.w[k - 1]; k--; }
    key[k] = kk; s.w[k] = w[j];
  }
  for (int k = 0; k < s.n; k++) s.col[k] = key[k] >> 2;
  return s;
}
// One thread per ground column. what = 0: the hail swath (collectHail), between the fall speeds and the sedimentation
// that moves the hail; what = 1: the water falls' sub-steps into the rain gauge (sedimentPart's deposit, water falls in
// order, each sub-step over the columns in order), from the ground-node values k_sediment kept.
template <class R> __global__ __launch_bounds__(TPB) void k_ground(Geo g, Bufs<R> b, double dt, int what) {
  const int slot = b.runs[blockIdx.y];
  const int ground = blockIdx.x * blockDim.x + threadIdx.x;
  if (ground >= g.layer) return;
  const GroundSources s = groundSources(g, b.pr(slot), ground);
  const double *rho = b.lv(slot, L_RHO, g);
  const R r0 = R(rho[0]);
  if (what == 0) {
    const float *hail = b.f(slot, F_HAIL, g), *vh = b.f(slot, F_FALL_HAIL, g);
    float *hg = b.col(slot, C_HAIL_GROUND, g), *hs = b.col(slot, C_HAIL_SIZE, g);
    float total = hg[ground], largest = hs[ground];
    for (int k = 0; k < s.n; k++) {
      const float h = hail[s.col[k]];
      if (!(double(h) > 1e-9)) continue;
      const R mm = R(r0 * R(h) * R(vh[s.col[k]]) * R(dt));
      total = float(double(total) + __dmul_rn(s.w[k], double(mm)));
      largest = float(mmax(R(largest), hailMaxDiameter(R(h), r0) * 1000));
    }
    hg[ground] = total; hs[ground] = largest;
    return;
  }
  const int falls[4] = {F_FALL_RAIN, F_FALL_SNOW, F_FALL_GRAUPEL, F_FALL_HAIL};
  float *prec = b.col(slot, C_PRECIPITATION, g);
  float total = prec[ground];
  for (int k = 0; k < 4; k++) {
    const double courant = __longlong_as_double((long long)reinterpret_cast<const unsigned long long *>(b.sc(slot))[S_COURANT + k]);
    if (!(courant > 0)) continue;  // as k_sediment: nothing of this fall moves
    int parts = int(ceil(courant / .8)); if (parts < 1) parts = 1;
    const R pdt = R(dt / parts);
    const float *vt = b.f(slot, falls[k], g);
    for (int p = 0; p < parts && p < LAND_PARTS; p++) {
      const float *a = b.land(slot, k, p, g);
      for (int q = 0; q < s.n; q++) {
        const R mm = r0 * R(a[s.col[q]]) * R(vt[s.col[q]]) * pdt;
        total = float(double(total) + __dmul_rn(s.w[q], double(mm)));
      }
    }
  }
  prec[ground] = total;
}

// ---- face mass fluxes and their divergence (advection.ts:190-211 faces + divergence) --------------------------------
// float64 from the float32 winds.
__device__ inline void faceFluxes(const Geo &g, const double *rho, const float *u, const float *v, const float *w, int x, int y, int z, double &mx, double &my, double &mz) {
  int i = x + g.nx * (y + g.ny * z), l = z * g.layer;
  int xp = x + 1 == g.nx ? 0 : x + 1, yp = y + 1 == g.ny ? 0 : y + 1;
  double r = rho[z];
  mx = r * (double(u[i]) + double(u[xp + y * g.nx + l])) / 2;
  my = r * (double(v[i]) + double(v[x + yp * g.nx + l])) / 2;
  mz = z < g.nz - 1 ? (rho[z] + rho[z + 1]) / 2 * (double(w[i]) + double(w[i + g.layer])) / 2 : 0.0;
}
// Face potential correction (advection.ts:244-255 correctFaces) of the faces of node (x, y, z).
template <class R> __device__ inline void correctedFaces(const Geo &g, const Bufs<R> &b, const double *rho, const float *u, const float *v, const float *w, const double *phi, int x, int y, int z, double &mx, double &my, double &mz) {
  faceFluxes(g, rho, u, v, w, x, y, z, mx, my, mz);
  int i = x + g.nx * (y + g.ny * z), l = z * g.layer;
  int xp = x + 1 == g.nx ? 0 : x + 1, yp = y + 1 == g.ny ? 0 : y + 1;
  double r = rho[z];
  mx -= r * (phi[xp + y * g.nx + l] - phi[i]) / g.dx;
  my -= r * (phi[x + yp * g.nx + l] - phi[i]) / g.dy;
  if (z < g.nz - 1) mz -= (rho[z] + rho[z + 1]) / 2 * (phi[i + g.layer] - phi[i]) / b.dzs[z];
}
// corrected = 0: faces and divergence of the node velocities (the right-hand side of the face projection);
// corrected = 1: the projected faces (faces minus the potential's gradient), whose divergence the transport's
// node-relative flux form takes as zero (it is the divergence the projection removed).
template <class R> __global__ __launch_bounds__(TPB) void k_faces(Geo g, Bufs<R> b, int corrected) {
  int slot = b.runs[blockIdx.y];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n) return;
  int z = i / g.layer, y = (i - z * g.layer) / g.nx, x = i - z * g.layer - y * g.nx;
  const double *rho = b.lv(slot, L_RHO, g);
  const float *u = b.f(slot, F_U, g), *v = b.f(slot, F_V, g), *w = b.f(slot, F_W, g);
  const double *phi = b.arr(b.phi, slot, g);
  int xm = x == 0 ? g.nx - 1 : x - 1, ym = y == 0 ? g.ny - 1 : y - 1, xp = x + 1 == g.nx ? 0 : x + 1, yp = y + 1 == g.ny ? 0 : y + 1;
  int l = z * g.layer;
  double mx, my, mz;
  if (corrected) correctedFaces(g, b, rho, u, v, w, phi, x, y, z, mx, my, mz);
  else faceFluxes(g, rho, u, v, w, x, y, z, mx, my, mz);
  b.m(slot, 0, g)[i] = mx; b.m(slot, 1, g)[i] = my; b.m(slot, 2, g)[i] = mz;
  if (corrected) return;
  // The right-hand side of the face projection: the divergence of the faces in difference form,
  // mx_i - mx_{i-1} = rho (u_{i+1} - u_{i-1}) / 2, so no sum of two ~25 m/s winds is ever differenced (the
  // projection's low modes amplify its rounding); modes multiplies it by h.
  double rzd = rho[z];
  double dv = rzd * (double(u[xp + y * g.nx + l]) - double(u[xm + y * g.nx + l])) / (2 * g.dx) + rzd * (double(v[x + yp * g.nx + l]) - double(v[x + ym * g.nx + l])) / (2 * g.dy);
  double dup = z < g.nz - 1 ? (rho[z] + rho[z + 1]) / 2 * (double(w[i]) + double(w[i + g.layer])) : 0.0;
  double ddn = z > 0 ? (rho[z - 1] + rho[z]) / 2 * (double(w[i - g.layer]) + double(w[i])) : 0.0;
  dv += (dup - ddn) / (2 * b.hz[z]);
  b.arr(b.cells, slot, g)[i] = dv;
}

// ---- transport: flux-form WENO5 / linear upwind tendency and the RK3 stage update --------------------------------
// (advection.ts:294-399 tendencyWeno/tendencyLinear/weno/upwind5, model.ts:1237-1251 the stage update, 1204-1214 the
// buoyancy in the w tendency, 1045 the rigid walls of the stages)
//
// WENO5 weights without underflow in float: the weights are scale-free but for the reference's 1e-40 in epsilon, so
// the stencil is divided by its scale s (the sum of its magnitudes, at least 1e-19, so that 1 / s^2 stays a float),
// where that term is (1e-20 / s)^2 <= 0.01. Relative to s, epsilon lies in [4e-8 (or 0.01 at the floor), ~1] and the
// indicators in [epsilon, ~30], so the reference's one-division form (both sums times (b0 b1 b2)^2) stays within float
// range: a stencil below 1e-20 takes the linear weights, as the reference's 1e-40 makes them. Over stencils of every scale from 1e-30 to 1e8 (zeros, near-constant and mixed signs among them) the face
// value is within 2e-7 of the stencil's variation of the reference's float64 value.
// Reciprocal: the hardware's (1 ulp) in float, a division in double.
__device__ inline float recip(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline double recip(double x) { return 1.0 / x; }
__device__ inline float fmx(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ inline double fmx(double a, double b) { return __builtin_fmax(a, b); }

__device__ inline double upwind5d(double a, double b_, double c, double d, double e) { return (2 * a - 13 * b_ + 47 * c + 27 * d - 3 * e) / 60; }
// The face values relative to the stencil's centre c (the upwind node), from the differences of the other nodes to it:
// face - c, in the reference's formulas rewritten on da = a - c, db = b - c, dd = d - c, de = e - c (the weights of
// the five coefficients sum to one). Differences of nearby float values are exact, so the result's rounding is relative
// to the local variation instead of to the field's magnitude: a float32 flux of m (face - c) keeps the digits a float32
// flux of m face loses to the cancellation of neighbouring faces.
template <class R> __device__ inline R upwindRel(R da, R db, R dd, R de) { return (2 * da - 13 * db + 27 * dd - 3 * de) * R(1.0 / 60); }
// WENO5 (advection.ts:388-399) relative to c: the candidates (times six, the 1/6 joining the final normalisation) and
// the indicators from the differences to c, relative to s. Every indicator and epsilon is taken times 4 / s^2 and every
// weight times 10 / (b0 b1 b2)^2 of the reference's (a factor common to all three, which the normalisation removes).
template <class R> __device__ inline R wenoRel(R a, R b_, R c, R d, R e) {
  // No branch for the reference's all-zero shortcut (advection.ts:390): a zero stencil gives zero through the same
  // arithmetic, and the branch costs more than the weights it skips (measured).
  const R s = fmx(mabs(a) + mabs(b_) + mabs(c) + mabs(d) + mabs(e), R(1e-19));
  const R is = recip(s), tiny = R(2e-20) * is;
  const R da = (a - c) * is, db = (b_ - c) * is, dd = (d - c) * is, de = (e - c) * is;
  const R p0 = 2 * da - 7 * db, p1 = -db + 2 * dd, p2 = 5 * dd - de;
  const R s0 = da - 2 * db, t0 = da - 4 * db, s1 = db + dd, t1 = db - dd, s2 = de - 2 * dd, t2 = de - 4 * dd;
  const R eps = R(8e-7) * (is * is) * (a * a + b_ * b_ + c * c + d * d + e * e) + tiny * tiny;
  const R b0 = eps + (R(13.0 / 3) * s0 * s0 + t0 * t0), b1 = eps + (R(13.0 / 3) * s1 * s1 + t1 * t1), b2 = eps + (R(13.0 / 3) * s2 * s2 + t2 * t2);
  const R c0 = R(3) * (b0 * b0), c1 = b1 * b1, c2 = b2 * b2;
  // w0 = c1 c2, w1 = 6 c0 c2, w2 = 3 c0 c1 (with c0 here 3 b0^2).
  const R w0 = c1 * c2, c2x2 = c2 + c2;
  const R num = w0 * p0 + c0 * (c2x2 * p1 + c1 * p2), den = w0 + c0 * (c2x2 + c1);
  return s * (num * recip(6 * den));
}
// The flux m (face - q0) through the face between q0 and qp1, relative to q0 (the face's western, southern or lower
// node), WENO5 from the stencil chosen by the flow's direction.
template <class R> __device__ inline R relWeno(R m, R qm2, R qm1, R q0, R qp1, R qp2, R qp3) {
  const bool up = m >= 0;
  const R a = up ? qm2 : qp3, b_ = up ? qm1 : qp2, c = up ? q0 : qp1, d = up ? qp1 : q0, e = up ? qp2 : qm1;
  return m * (wenoRel(a, b_, c, d, e) + (c - q0));
}
// The WENO5 face value of each of n stencils (a, b, c, d, e, float32, flow from a towards e), relative to c, in the
// arithmetic type R: the transport kernels' own reconstruction, for its numerics check (ml_check_weno).
template <class R> __global__ void k_check_weno(const float *st, int n, double *out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float *v = st + size_t(k) * 5;
  out[k] = double(wenoRel(R(v[0]), R(v[1]), R(v[2]), R(v[3]), R(v[4])));
}
__device__ inline bool monotoneField(int f) { return !(f == F_U || f == F_V || f == F_W || f == F_THETA); }
// Index of a carried field among the conserved ones (rain, snow, graupel, hail, q, cloud, ice), -1 if it is not one.
// The conserved (positivity-fixed) carried fields by their index (model.ts:1319-1327).
__device__ inline int conservedField(int k) { const int c[N_CONSERVED] = {F_RAIN, F_SNOW, F_GRAUPEL, F_HAIL, F_Q, F_CLOUD, F_ICE}; return c[k]; }
__device__ inline int conservedIndex(int f) { return f <= F_HAIL ? f : f == F_Q ? 4 : f == F_CLOUD ? 5 : f == F_ICE ? 6 : -1; }
// Which levels of the monotone fields fbase .. fbase + nf - 1 hold a non-zero node, before their transport: a block of
// the transport whose field is zero everywhere its stage's result can reach from (3 levels a stage) has a zero
// tendency, and its result is its start (zero). One block per (level, field) and run; the other fields are flagged
// non-zero, and with `all` every field (the autotuner times the whole work).
template <class R> __global__ __launch_bounds__(256) void k_zero_flags(Geo g, Bufs<R> b, int fbase, int all) {
  const int slot = b.runs[blockIdx.y], f = fbase + int(blockIdx.x) / g.nz, z = int(blockIdx.x) % g.nz;
  int any = all || !monotoneField(f);
  if (!any) {
    const float *q = b.f(slot, f, g) + size_t(z) * g.layer;
    for (int c = threadIdx.x; c < g.layer && !any; c += blockDim.x) any = q[c] != 0;
  }
  any = __syncthreads_or(any);
  if (threadIdx.x == 0) b.zf(slot, f, g)[z] = any ? 1 : 0;
}
// Whether field f was zero at levels lo .. hi (clamped to the grid) by this step's flags.
template <class R> __device__ inline bool zeroLevels(const Bufs<R> &b, const Geo &g, int slot, int f, int lo, int hi) {
  if (!monotoneField(f)) return false;
  lo = lo < 0 ? 0 : lo; hi = hi > g.nz - 1 ? g.nz - 1 : hi;
  const unsigned char *fl = b.zf(slot, f, g);
  for (int z = lo; z <= hi; z++) if (fl[z]) return false;
  return true;
}
// Stage `stage` of levels z0 .. z1 - 1 reads its input at 3 levels around them, and that input is zero where the
// state was zero within 3 more levels for each stage before it.
__device__ inline int stageReach(int stage) { return 3 + 3 * stage; }

// The positivity-fix sums (removed, kept, negative; level-weighted) of one level (or strip tile) of conserved field cf,
// reduced over the block in a fixed order into their part of the run's sums; k_fill_total adds the parts in level
// order, so the sums do not depend on the order the blocks run in. Every thread of the block calls; fill is cleared.
template <class R> __device__ inline void storeFill(const Bufs<R> &b, const Geo &g, int slot, int cf, int z, int tile, double fill[3], double (*red)[3]) {
  for (int j = 0; j < 3; j++) fill[j] = waveSum(fill[j]);
  if ((threadIdx.x & 63) == 0) for (int j = 0; j < 3; j++) red[threadIdx.x >> 6][j] = fill[j];
  __syncthreads();
  if (threadIdx.x < 3) {
    double s = 0;
    for (int w = 0; w < int(blockDim.x) / 64; w++) s += red[w][threadIdx.x];
    b.fillPart(slot, cf, z, tile, g)[threadIdx.x] = s;
  }
  __syncthreads();
  fill[0] = fill[1] = fill[2] = 0;
}
// The charge carried to the ground by a launch (which: 0 the charge falls, 1 the conduction) from its `count` block
// parts, added in a fixed order.
template <class R> __global__ void k_ground_total(Geo g, Bufs<R> b, int which, int count) {
  const int slot = b.runs[blockIdx.x];
  const double *part = b.gpart + (size_t(slot) * 2 + which) * b.gmax;
  double s = 0;
  for (int k = threadIdx.x; k < count; k += blockDim.x) s += part[k];
  __shared__ double red[16];
  s = blockSum(s, red);
  if (threadIdx.x == 0 && s != 0) b.sc(slot)[S_CHARGE_GROUND] += s;
}
// The result of a stage of levels z0 .. z1 - 1 of a field that is zero within the stage's reach: stages 0 and 1 write
// their zero output (the start is zero); stage 2's output is the state, which keeps its zeros. The positivity-fix parts
// of a conserved field are zero. Every thread of the block calls (tile: the strip tile of a level, 0 otherwise; the
// strip layout passes the tile's nodes as z0 = z1 - 1 with its own writes).
template <class R> __device__ void zeroStage(const Bufs<R> &b, const Geo &g, int slot, int f, int z0, int z1, int stage, int tile, int cf, double (*red)[3]) {
  if (stage < 2) {
    float *out = b.set(stage + 1, slot, f, g);
    for (size_t i = size_t(z0) * g.layer + threadIdx.x; i < size_t(z1) * g.layer; i += blockDim.x) out[i] = 0.f;
  } else if (cf >= 0) {
    double fill[3] = {0, 0, 0};
    for (int z = z0; z < z1; z++) storeFill(b, g, slot, cf, z, tile, fill, red);
  }
}
// The run's positivity-fix sums from the parts the last transport stage stored (tiles: parts per level).
template <class R> __global__ void k_fill_total(Geo g, Bufs<R> b, int tiles) {
  const int slot = b.runs[blockIdx.x], t = threadIdx.x;
  if (t >= 3 * N_CONSERVED) return;
  double s = 0;
  if (tiles == 1) {
    // The parts in level order, 8 levels' loads in flight at a time (the adds stay in order).
    constexpr int ZB = 8;
    for (int z0 = 0; z0 < g.nz; z0 += ZB) {
      double v[ZB];
#pragma unroll
      for (int q = 0; q < ZB; q++) v[q] = z0 + q < g.nz ? b.fillPart(slot, t / 3, z0 + q, 0, g)[t % 3] : 0.0;
#pragma unroll
      for (int q = 0; q < ZB; q++) if (z0 + q < g.nz) s += v[q];
    }
  } else {
    for (int z = 0; z < g.nz; z++)
      for (int k = 0; k < tiles; k++) s += b.fillPart(slot, t / 3, z, k, g)[t % 3];
  }
  b.sc(slot)[S_FILL + t] = s;
}
constexpr double TINY_WATER = 1e-10;

// Transport blocks: K levels of one field of one run. stage 0: state -> X, stage 1: X -> Y, stage 2: Y -> state (each
// node reads its own start value first); stage 2 also sums the positivity fix of the conserved fields (model.ts:606-617
// fillNegative) for the physics kernel.
// Block -> (run index in the launch list, block within the run). Workgroups go to the eight XCDs in turn; with `xcd`
// set and a batch that divides into eight, each run's blocks are all given to one XCD, in order.
__device__ inline void blockRun(int nr, int perRun, int xcd, int &runIdx, int &b) {
  int p = blockIdx.x;
  if (xcd && nr % 8 == 0) { int x = p & 7, q = p >> 3; runIdx = x + 8 * (q / perRun); b = q % perRun; }
  else { runIdx = p / perRun; b = p % perRun; }
}

// ---- the node-relative flux form of the tendency ---------------------------------------------------------------------
// tend rho0 = -sum_faces (m (q_face - q)) / spacing for a node of value q: the reference's flux form
// -(sum (m q_face) / spacing - q div) (advection.ts:294-378) with div the divergence of the same face mass fluxes
// (advection.ts:255), rewritten on each face value's departure from the node's value, which the float rounding of
// the neighbouring faces' fluxes then no longer cancels out of.
// Linear upwind5 (u, v, w, theta) along one axis, in float64: D[k] = q(k - 3) - q(0), k = 0..6 (D[3] = 0); mp, mm the
// face mass fluxes at +1/2 and -1/2. Returns mp (face+ - q) - mm (face- - q).
__device__ inline double linearAxis(double mp, double mm, const double *D) {
  double fp, fm;
  if (mp >= 0) fp = upwindRel<double>(D[1], D[2], D[4], D[5]);                                      // stencil -2 .. 2, centre 0
  else { const double c = D[4]; fp = upwindRel<double>(D[6] - c, D[5] - c, -c, D[2] - c) + c; }     // stencil 3 .. -1, centre 1
  if (mm >= 0) { const double c = D[2]; fm = upwindRel<double>(D[0] - c, D[1] - c, -c, D[4] - c) + c; }  // -3 .. 1, centre -1
  else fm = upwindRel<double>(D[5], D[4], D[2], D[1]);                                              // stencil 2 .. -2, centre 0
  return mp * fp - mm * fm;
}
// The base profile's own tendency of a departure-transported field (u, v, theta), exact for the linear scheme: its
// vertical face values relative to the level's value, times the vertical face fluxes. The face values of a level are
// the same at every node of it: BaseFaces holds them for either flow direction.
struct BaseFaces { double topUp, topDown, bottomUp, bottomDown; };
__device__ inline BaseFaces baseFaces(const double *Bc) {
  const double Bz = Bc[3];
  return {upwind5d(Bc[1], Bc[2], Bc[3], Bc[4], Bc[5]) - Bz, upwind5d(Bc[6], Bc[5], Bc[4], Bc[3], Bc[2]) - Bz,
          upwind5d(Bc[0], Bc[1], Bc[2], Bc[3], Bc[4]) - Bz, upwind5d(Bc[5], Bc[4], Bc[3], Bc[2], Bc[1]) - Bz};
}
__device__ inline double baseTendency(const BaseFaces &f, double mzt, double mzb, double ihzrz) {
  const double pt = mzt >= 0 ? f.topUp : f.topDown, pb = mzb >= 0 ? f.bottomUp : f.bottomDown;
  return (mzt * pt - mzb * pb) * ihzrz;
}
// The buoyancy of w's tendency (model.ts:1204-1214) at node i of the stage's input set, in float64.
template <class R> __device__ inline double buoyancy(const Bufs<R> &b, const Geo &g, int set, int slot, size_t i, double thvEnv, double ithv) {
  const float th = b.set(set, slot, F_THETA, g)[i], qv = b.set(set, slot, F_Q, g)[i];
  const double loaded = double(b.set(set, slot, F_CLOUD, g)[i]) + double(b.set(set, slot, F_ICE, g)[i])
    + (double(b.set(set, slot, F_RAIN, g)[i]) + double(b.set(set, slot, F_SNOW, g)[i]) + double(b.set(set, slot, F_GRAUPEL, g)[i]) + double(b.set(set, slot, F_HAIL, g)[i]));
  const double dthv = (double(th) * (1 + .61 * double(qv)) - thvEnv) * ithv;
  return G * (dthv - loaded);
}
// The level's (base) values a vertical stencil reads, z - 3 .. z + 3 clamped to the column.
__device__ inline void columnLevels(const Geo &g, int z, int *zc) { for (int k = 0; k < 7; k++) { int zz = z - 3 + k; zc[k] = zz < 0 ? 0 : zz > g.nz - 1 ? g.nz - 1 : zz; } }

// The level transport layout: one block of 256 threads per (K levels, field, run), each level in turn whole in LDS
// (periodic in x and y, so the level is its own halo), each thread a few nodes. A WENO field's face fluxes are taken
// once a face, relative to the face's western / southern / lower node, and kept in LDS (the top face of a level is the
// bottom face of the next); a linear field's node takes its six faces itself, in float64. The stage update is float64
// from the float32 start, rounded once.
__host__ __device__ inline size_t stageLevelLdsBytes(const Geo &g, size_t rb) { return ((g.layer * 4 + 15) / 16) * 16 + 3 * size_t(g.layer) * rb; }
template <class R> __global__ __launch_bounds__(TPB) void k_stage_level(Geo g, Bufs<R> b, int stage, double dt, int nr, int nc, int fbase, int xcd, int K) {
  const int zBlocks = (g.nz + K - 1) / K;
  int runIdx, blk;
  blockRun(nr, zBlocks * nc, xcd, runIdx, blk);
  const int slot = b.runs[runIdx], f = fbase + blk % nc, z0 = (blk / nc) * K, z1 = z0 + K < g.nz ? z0 + K : g.nz;
  const int inSet = stage, outSet = stage == 2 ? 0 : stage + 1;
  const float *q = b.set(inSet, slot, f, g);
  float *out = b.set(outSet, slot, f, g);
  const float *start = b.f(slot, f, g);
  const int layer = g.layer, nx = g.nx, ny = g.ny, nz = g.nz;
  extern __shared__ char smem[];
  float *ql = reinterpret_cast<float *>(smem);
  R *fxl = reinterpret_cast<R *>(smem + ((layer * sizeof(float) + 15) / 16) * 16), *fyl = fxl + layer, *fzl = fyl + layer;
  __shared__ double red[TPB / 64][3];
  // u, v and theta: departures from the level's base profile (their vertical differences; the profile's own tendency
  // added exactly).
  const int lvl = f == F_U ? L_UENV : f == F_V ? L_VENV : f == F_THETA ? L_THETA : -1;
  const double *B = lvl >= 0 ? b.lv(slot, lvl, g) : nullptr;
  const bool mono = monotoneField(f), isW = f == F_W;
  const int cf = conservedIndex(f);
  if (zeroLevels(b, g, slot, f, z0 - stageReach(stage), z1 - 1 + stageReach(stage))) { zeroStage(b, g, slot, f, z0, z1, stage, 0, cf, red); return; }
  const Div dNx(nx);
  const double fracd = (stage == 0 ? 1.0 / 3 : stage == 1 ? 1.0 / 2 : 1.0) * dt;
  const double *rho = b.lv(slot, L_RHO, g);
  const double *mz = b.m(slot, 2, g);
  double fill[3] = {0, 0, 0};
  for (int z = z0; z < z1; z++) {
    const float *qz = q + size_t(z) * layer;
    for (int c = threadIdx.x; c < layer; c += TPB) ql[c] = qz[c];
    __syncthreads();
    const double *mx = b.m(slot, 0, g) + size_t(z) * layer, *my = b.m(slot, 1, g) + size_t(z) * layer;
    if (mono) {
      for (int c = threadIdx.x; c < layer; c += TPB) {
        const int y = dNx.q(c), x = c - y * nx, row = y * nx;
        auto X = [&](int o) { int xx = x + o; xx += xx < 0 ? nx : 0; xx -= xx >= nx ? nx : 0; return R(ql[row + xx]); };
        auto Y = [&](int o) { int yy = y + o; yy += yy < 0 ? ny : 0; yy -= yy >= ny ? ny : 0; return R(ql[yy * nx + x]); };
        fxl[c] = relWeno(R(mx[c]), X(-2), X(-1), X(0), X(1), X(2), X(3));
        fyl[c] = relWeno(R(my[c]), Y(-2), Y(-1), Y(0), Y(1), Y(2), Y(3));
      }
      __syncthreads();
    }
    const double ihzd = 1.0 / b.hz[z], irzd = 1.0 / rho[z], ihzrz = 1.0 / (b.hz[z] * rho[z]);
    const R idx = recip(R(g.dx)), idy = recip(R(g.dy)), ihz = R(ihzd), irz = R(irzd);
    double thvEnv = 0, ithv = 0;
    if (isW) { const double te = b.lv(slot, L_THETA, g)[z], qe = b.lv(slot, L_Q, g)[z]; thvEnv = te * (1 + .61 * qe); ithv = 1.0 / thvEnv; }
    int zc[7];
    columnLevels(g, z, zc);
    double Bc[7], Bd[7];
    for (int k = 0; k < 7; k++) { Bc[k] = B ? B[zc[k]] : 0; Bd[k] = B ? B[zc[k]] - B[z] : 0; }
    const double wz = b.lv(slot, L_LW, g)[z];
    for (int c = threadIdx.x; c < layer; c += TPB) {
      const int y = dNx.q(c), x = c - y * nx, row = y * nx;
      const size_t i = size_t(z) * layer + c;
      const int cw = row + (x == 0 ? nx - 1 : x - 1), cs = (y == 0 ? ny - 1 : y - 1) * nx + x;
      const double mzt = z == nz - 1 ? 0.0 : mz[i], mzb = z == 0 ? 0.0 : mz[i - layer];
      float v[7];
      for (int k = 0; k < 7; k++) v[k] = k == 3 ? ql[c] : q[size_t(zc[k]) * layer + c];
      double tendd;
      if (mono) {
        const R fzt = z == nz - 1 ? R(0) : relWeno(R(mzt), R(v[1]), R(v[2]), R(v[3]), R(v[4]), R(v[5]), R(v[6]));
        const R fzb = z == 0 ? R(0) : z > z0 ? fzl[c] : relWeno(R(mzb), R(v[0]), R(v[1]), R(v[2]), R(v[3]), R(v[4]), R(v[5]));
        fzl[c] = fzt;
        const R q0 = v[3];
        const R fluxX = fxl[c] - (fxl[cw] - R(mx[cw]) * (q0 - R(ql[cw])));
        const R fluxY = fyl[c] - (fyl[cs] - R(my[cs]) * (q0 - R(ql[cs])));
        const R fluxZ = fzt - (fzb - R(mzb) * (q0 - R(v[2])));
        tendd = double(-(fluxX * idx + fluxY * idy + fluxZ * ihz) * irz);
      } else {
        double Dx[7], Dy[7], Dz[7];
        const double q0 = v[3];
        for (int k = 0; k < 7; k++) {
          int xx = x + k - 3; xx += xx < 0 ? nx : 0; xx -= xx >= nx ? nx : 0;
          int yy = y + k - 3; yy += yy < 0 ? ny : 0; yy -= yy >= ny ? ny : 0;
          Dx[k] = double(ql[row + xx]) - q0; Dy[k] = double(ql[yy * nx + x]) - q0; Dz[k] = (double(v[k]) - q0) - Bd[k];
        }
        // Per node, as written: computing the base faces once a level, or 1 / dx once a launch, gives wrong u, v and w
        // in this kernel (measured; the plane kernel below does both and keeps parity).
        tendd = -(linearAxis(mx[c], mx[cw], Dx) / g.dx + linearAxis(my[c], my[cs], Dy) / g.dy + linearAxis(mzt, mzb, Dz) * ihzd) * irzd;
        if (B) tendd -= baseTendency(baseFaces(Bc), mzt, mzb, ihzrz);
        if (isW) tendd += buoyancy(b, g, inSet, slot, i, thvEnv, ithv);
      }
      float val = float(double(start[i]) + fracd * tendd);
      if (isW && stage < 2 && (z == 0 || z == nz - 1)) val = 0;
      out[i] = val;
      if (stage == 2) {
        const double xv = val;
        if (xv < TINY_WATER) { if (xv != 0) { fill[0] += wz * xv; if (xv < 0) fill[2] -= wz * xv; } }
        else fill[1] += wz * xv;
      }
    }
    // The level's part of the run's positivity-fix sums.
    if (stage == 2 && cf >= 0) storeFill(b, g, slot, cf, z, 0, fill, red);
    __syncthreads();
  }
}

// The third transport layout: K levels of one field per block as in k_stage_level, each level staged in LDS as a plane
// padded by three periodic rows and columns on every side, so a node's six-point stencils are LDS reads at fixed
// distances from its own place. The x and y face fluxes go to LDS planes whose west and south edges also hold the
// periodic faces, so the tendency takes its faces without wrap-around arithmetic; a thread steps through its nodes
// by (x, y) increments. The node's other inputs are read at a level offset held in scalar registers. The scheme
// (WENO5 or upwind5) and the base-profile departure are compile-time paths. The arithmetic and its order are those
// of k_stage_level, so both give the same values.
__host__ __device__ inline int align16i(int bytes) { return (bytes + 15) & ~15; }
__host__ __device__ inline size_t stagePlaneLdsBytes(const Geo &g, size_t rb) {
  return size_t(align16i((g.nx + 6) * (g.ny + 6) * 4)) + align16i(int((g.nx + 1) * g.ny * rb)) + align16i(int(g.nx * (g.ny + 1) * rb)) +
         size_t(g.layer) * rb;
}
// The plane layout's halo copies are one periodic image per side, so a level has at least six columns and rows.
__host__ __device__ inline bool planeFits(const Geo &g, size_t rb, size_t ldsMax) { return g.nx >= 6 && g.ny >= 6 && stagePlaneLdsBytes(g, rb) <= ldsMax; }
// Element i of an array whose base is uniform over the block: a 32-bit offset from a base in scalar registers.
template <class T> __device__ inline T at(const T *base, int i) { return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + unsigned(i) * unsigned(sizeof(T))); }
template <class T> __device__ inline void put(T *base, int i, T v) { *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + unsigned(i) * unsigned(sizeof(T))) = v; }
// Levels z0 .. z1-1 of field f of a run. BASE: u, v or theta, transported as departures from the base profile.
// cf: the conserved index of the field (-1: none); each level's positivity-fix sums go to their part of the run's.
template <class R, bool BASE, bool MONO> __device__ void stagePlane(const Geo &g, const Bufs<R> &b, int slot, int f, int z0, int z1, int stage, double dt, char *smem,
                                                                     int cf, double (*red)[3]) {
  double fill[3] = {0, 0, 0};
  const int nx = g.nx, ny = g.ny, nz = g.nz, layer = g.layer, W = nx + 6, X1 = nx + 1, T = TPB, planeN = W * (ny + 6);
  const int inSet = stage, outSet = stage == 2 ? 0 : stage + 1;
  const float *q = b.set(inSet, slot, f, g), *start = b.f(slot, f, g);
  float *out = b.set(outSet, slot, f, g);
  float *ql = reinterpret_cast<float *>(smem);
  R *fxl = reinterpret_cast<R *>(smem + align16i(planeN * 4));
  R *fyl = reinterpret_cast<R *>(reinterpret_cast<char *>(fxl) + align16i(int(X1 * ny * sizeof(R))));
  R *fzl = reinterpret_cast<R *>(reinterpret_cast<char *>(fyl) + align16i(int(nx * (ny + 1) * sizeof(R))));
  const double *B = BASE ? b.lv(slot, f == F_U ? L_UENV : f == F_V ? L_VENV : L_THETA, g) : nullptr;
  const bool isW = !BASE && !MONO && f == F_W;
  const double fracd = (stage == 0 ? 1.0 / 3 : stage == 1 ? 1.0 / 2 : 1.0) * dt;
  const double *rho = b.lv(slot, L_RHO, g);
  // This thread's first node and place in the plane, and the (x, y) steps between its nodes and places. A node's place
  // in the plane (pq) and its western face's place in the x-face plane (pf) step with it: by dq, df per step, and by
  // W - nx = 6 and X1 - nx = 1 more when the step wraps to the next row.
  const Div dNx(nx), dW(W), d6(6);
  const int y00 = dNx.q(threadIdx.x), x00 = threadIdx.x - y00 * nx, sy = T / nx, sx = T - sy * nx;
  const int pq00 = (y00 + 3) * W + x00 + 3, pf00 = y00 * X1 + x00, dq = sy * W + sx, df = sy * X1 + sx;
  for (int z = z0; z < z1; z++) {
    const size_t lz = size_t(z) * layer;
    const float *qz = q + lz;
    // Place (px, py) of the plane holds node ((px - 3) mod nx, (py - 3) mod ny): the values themselves (a base
    // profile is constant over a level, so the horizontal differences of departures are those of the values). The
    // level's nodes go to the interior, each at its own place pq; then the rim's periodic images: its top and bottom
    // three rows (W places each), then the three places at each end of the interior's rows.
    {
      int x = x00, pq = pq00;
      for (int c = threadIdx.x; c < layer; c += T) {
        ql[pq] = at(qz, c);
        x += sx; pq += dq;
        if (x >= nx) { x -= nx; pq += W - nx; }
      }
      for (int k = threadIdx.x; k < 6 * W; k += T) {
        const int r = dW.q(k), px = k - mul24(r, W), py = r < 3 ? r : ny + r, ys = r < 3 ? ny - 3 + r : r - 3;
        int xs = px - 3;
        xs += xs < 0 ? nx : 0; xs -= xs >= nx ? nx : 0;
        ql[mul24(py, W) + px] = at(qz, mul24(ys, nx) + xs);
      }
      for (int k = threadIdx.x; k < 6 * ny; k += T) {
        const int ys = d6.q(k), j = k - mul24(ys, 6), px = j < 3 ? j : nx + j, xs = j < 3 ? nx - 3 + j : j - 3;
        ql[mul24(ys + 3, W) + px] = at(qz, mul24(ys, nx) + xs);
      }
    }
    __syncthreads();
    const double *mxz = b.m(slot, 0, g) + lz, *myz = b.m(slot, 1, g) + lz;
    if (MONO) {
      // Each thread's node's eastern and northern faces, relative to the face's western / southern node.
      int x = x00, y = y00, pq = pq00, pf = pf00;
      for (int c = threadIdx.x; c < layer; c += T) {
        const float *s = ql + pq;
        const R fx = relWeno(R(at(mxz, c)), R(s[-2]), R(s[-1]), R(s[0]), R(s[1]), R(s[2]), R(s[3]));
        const R fy = relWeno(R(at(myz, c)), R(s[-2 * W]), R(s[-W]), R(s[0]), R(s[W]), R(s[2 * W]), R(s[3 * W]));
        R *ex = fxl + pf + 1;          // face x + 1/2; the row's first slot is the face west of x = 0
        ex[0] = fx;
        if (x == nx - 1) ex[-nx] = fx;
        fyl[c + nx] = fy;              // face y + 1/2; the first row is the face south of y = 0
        if (y == ny - 1) fyl[x] = fy;
        x += sx; y += sy; pq += dq; pf += df;
        if (x >= nx) { x -= nx; y++; pq += W - nx; pf += X1 - nx; }
      }
      __syncthreads();
    }
    const double ihzd = 1.0 / b.hz[z], irzd = 1.0 / rho[z], ihzrz = 1.0 / (b.hz[z] * rho[z]);
    const R idx = recip(R(g.dx)), idy = recip(R(g.dy)), ihz = R(ihzd), irz = R(irzd);
    double thvEnv = 0, ithv = 0;
    if (isW) { const double te = b.lv(slot, L_THETA, g)[z], qe = b.lv(slot, L_Q, g)[z]; thvEnv = te * (1 + .61 * qe); ithv = 1.0 / thvEnv; }
    int zc[7];
    columnLevels(g, z, zc);
    const float *qc[7];
    double Bc[7], Bd[7];
    for (int k = 0; k < 7; k++) { qc[k] = q + size_t(zc[k]) * layer; Bc[k] = BASE ? B[zc[k]] : 0; Bd[k] = BASE ? B[zc[k]] - B[z] : 0; }
    const BaseFaces bf = baseFaces(Bc);
    const double idxd = 1.0 / g.dx, idyd = 1.0 / g.dy;
    const double wz = b.lv(slot, L_LW, g)[z];
    const double *mzz = b.m(slot, 2, g) + lz;
    const double *mzb_ = z > 0 ? mzz - layer : mzz;
    const float *startz = start + lz;
    float *outz = out + lz;
    int x = x00, y = y00, pq = pq00, pf = pf00;
    for (int c = threadIdx.x; c < layer; c += T) {
      const float *s = ql + pq;
      float v[7];
      for (int k = 0; k < 7; k++) v[k] = k == 3 ? s[0] : at(qc[k], c);
      const double mzt = z == nz - 1 ? 0.0 : at(mzz, c), mzb = z == 0 ? 0.0 : at(mzb_, c);
      const int cw = x == 0 ? c + nx - 1 : c - 1, cs = y == 0 ? c + layer - nx : c - nx;
      double tendd;
      if (MONO) {
        const R fzt = z == nz - 1 ? R(0) : relWeno(R(mzt), R(v[1]), R(v[2]), R(v[3]), R(v[4]), R(v[5]), R(v[6]));
        R fzb;  // the bottom face relative to the node below
        if (z == 0) fzb = 0;
        else if (z > z0) fzb = fzl[c];
        else fzb = relWeno(R(mzb), R(v[0]), R(v[1]), R(v[2]), R(v[3]), R(v[4]), R(v[5]));
        fzl[c] = fzt;
        // The western, southern and bottom faces relative to this node: minus m (q - q_neighbour).
        const R q0 = v[3];
        const R *fxw = fxl + pf, *fys = fyl + c;
        const R fluxX = fxw[1] - (fxw[0] - R(at(mxz, cw)) * (q0 - R(s[-1])));
        const R fluxY = fys[nx] - (fys[0] - R(at(myz, cs)) * (q0 - R(s[-W])));
        const R fluxZ = fzt - (fzb - R(mzb) * (q0 - R(v[2])));
        tendd = double(-(fluxX * idx + fluxY * idy + fluxZ * ihz) * irz);
      } else {
        // A linear field's node takes its six faces itself, in float64, relative to its own value.
        double Dx[7], Dy[7], Dz[7];
        const double q0 = v[3];
        for (int k = 0; k < 7; k++) {
          Dx[k] = double(s[k - 3]) - q0; Dy[k] = double(s[(k - 3) * W]) - q0;
          Dz[k] = BASE ? (double(v[k]) - q0) - Bd[k] : double(v[k]) - q0;
        }
        tendd = -(linearAxis(at(mxz, c), at(mxz, cw), Dx) * idxd + linearAxis(at(myz, c), at(myz, cs), Dy) * idyd + linearAxis(mzt, mzb, Dz) * ihzd) * irzd;
        if (BASE) tendd -= baseTendency(bf, mzt, mzb, ihzrz);
        if (isW) tendd += buoyancy(b, g, inSet, slot, lz + c, thvEnv, ithv);
      }
      // The stage update in float64 from the float32 start, rounded once.
      float val = float(double(at(startz, c)) + fracd * tendd);
      if (isW && stage < 2 && (z == 0 || z == nz - 1)) val = 0;
      put(outz, c, val);
      if (MONO && stage == 2) {
        const double xv = val;
        if (xv < TINY_WATER) { if (xv != 0) { fill[0] += wz * xv; if (xv < 0) fill[2] -= wz * xv; } }
        else fill[1] += wz * xv;
      }
      x += sx; y += sy; pq += dq; pf += df;
      if (x >= nx) { x -= nx; y++; pq += W - nx; pf += X1 - nx; }
    }
    if (MONO && stage == 2 && cf >= 0) storeFill(b, g, slot, cf, z, 0, fill, red);
    __syncthreads();
  }
}
template <class R> __global__ __launch_bounds__(TPB) void k_stage_plane(Geo g, Bufs<R> b, int stage, double dt, int nr, int nc, int fbase, int xcd, int K) {
  const int zBlocks = (g.nz + K - 1) / K;
  int runIdx, blk;
  blockRun(nr, zBlocks * nc, xcd, runIdx, blk);
  const int slot = b.runs[runIdx], f = fbase + blk % nc, z0 = (blk / nc) * K, z1 = z0 + K < g.nz ? z0 + K : g.nz;
  extern __shared__ char smem[];
  __shared__ double red[TPB / 64][3];
  const int cf = conservedIndex(f);
  if (zeroLevels(b, g, slot, f, z0 - stageReach(stage), z1 - 1 + stageReach(stage))) { zeroStage(b, g, slot, f, z0, z1, stage, 0, cf, red); return; }
  // One kernel for both kinds of field: a launch runs the linear fields' blocks (float64 faces, memory latency) beside
  // the WENO fields' (arithmetic), which keeps the CUs busier than a kernel of each kind in turn (measured).
  if (f == F_U || f == F_V || f == F_THETA) stagePlane<R, true, false>(g, b, slot, f, z0, z1, stage, dt, smem, cf, red);
  else if (f == F_W) stagePlane<R, false, false>(g, b, slot, f, z0, z1, stage, dt, smem, cf, red);
  else stagePlane<R, false, true>(g, b, slot, f, z0, z1, stage, dt, smem, cf, red);
}

// ---- pressure projection: dual-cell divergence and the node-velocity correction (pressure.ts:50-92) --------------------
// set: which buffer set holds u, v, w (0 state, 1 X, 2 Y).
template <class R> __global__ __launch_bounds__(TPB) void k_pdiv(Geo g, Bufs<R> b, int set, double dt) {
  int slot = b.runs[blockIdx.y];
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  int cells = g.layer * (g.nz - 1);
  if (c >= cells) return;
  int k = c / g.layer, j = (c - k * g.layer) / g.nx, i = c - k * g.layer - j * g.nx;
  const float *u = b.set(set, slot, F_U, g), *v = b.set(set, slot, F_V, g), *w = b.set(set, slot, F_W, g);
  const double *rho = b.lv(slot, L_RHO, g);
  int l0 = k * g.layer, l1 = l0 + g.layer, y0 = j * g.nx, y1 = (j + 1 == g.ny ? 0 : j + 1) * g.nx, i1 = i + 1 == g.nx ? 0 : i + 1;
  int a = i + y0 + l0, bb = i1 + y0 + l0, e = i + y1 + l0, f = i1 + y1 + l0, gg = i + y0 + l1, h = i1 + y0 + l1, s = i + y1 + l1, t = i1 + y1 + l1;
  // float64 from the float32 winds (the sums of four winds are exact in float64), as the reference computes it.
  double r0 = rho[k] / 4, r1 = rho[k + 1] / 4;
  double ddx = (r0 * (double(u[bb]) + u[f] - u[a] - u[e]) + r1 * (double(u[h]) + u[t] - u[gg] - u[s])) / g.dx;
  double ddy = (r0 * (double(v[e]) + v[f] - v[a] - v[bb]) + r1 * (double(v[s]) + v[t] - v[gg] - v[h])) / g.dy;
  double ddz = (r1 * (double(w[gg]) + w[h] + w[s] + w[t]) - r0 * (double(w[a]) + w[bb] + w[e] + w[f])) / b.dzs[k];
  b.arr(b.cells, slot, g)[c] = (ddx + ddy + ddz) / dt;
}
// final = 0: the stage projection, in place on set `set`. final = 1: the end-of-step projection of the mixed wind in X
// into the state, with the safety limits (model.ts:1586-1595 clipWind), and the other mixed fields copied from X.
template <class R> __global__ __launch_bounds__(TPB) void k_pcorrect(Geo g, Bufs<R> b, int set, double dt, int final) {
  int slot = b.runs[blockIdx.y];
  int node = blockIdx.x * blockDim.x + threadIdx.x;
  bool live = node < g.n;
  unsigned long long clipped = 0;
  if (live) {
    int k = node / g.layer, j = (node - k * g.layer) / g.nx, i = node - k * g.layer - j * g.nx;
    const double *p = b.arr(b.press, slot, g);
    int cl = g.layer, nz = g.nz;
    int lo = (k > 0 ? k - 1 : 0) * cl, hi = (k < nz - 2 ? k : nz - 2) * cl;
    // The gradient and the correction in float64 from the float32 winds, rounded once (the correction of w is about as
    // large as w itself, so its float32 product would add an ulp a projection).
    double wLo = k > 0 ? b.dzs[k - 1] / b.hz[k] : 0.0, wHi = k < nz - 1 ? b.dzs[k] / b.hz[k] : 0.0;
    int jm = (j == 0 ? g.ny - 1 : j - 1) * g.nx, j0 = j * g.nx, im = i == 0 ? g.nx - 1 : i - 1;
    auto P = [&](int ii, int jj, int base) { return p[base + ii + jj]; };  // float64: differences before rounding
    // Differences of neighbouring pressures first (as for the divergence).
    double gx = (wLo * ((P(i, jm, lo) - P(im, jm, lo)) + (P(i, j0, lo) - P(im, j0, lo))) + wHi * ((P(i, jm, hi) - P(im, jm, hi)) + (P(i, j0, hi) - P(im, j0, hi)))) / (4 * g.dx);
    double gy = (wLo * ((P(im, j0, lo) - P(im, jm, lo)) + (P(i, j0, lo) - P(i, jm, lo))) + wHi * ((P(im, j0, hi) - P(im, jm, hi)) + (P(i, j0, hi) - P(i, jm, hi)))) / (4 * g.dy);
    float *u = b.set(set, slot, F_U, g), *v = b.set(set, slot, F_V, g), *w = b.set(set, slot, F_W, g);
    float un = float(double(u[node]) - dt * gx), vn = float(double(v[node]) - dt * gy), wn = w[node];
    if (k > 0 && k < nz - 1) {
      int below = (k - 1) * cl, above = k * cl;
      double gz = (((P(im, jm, above) - P(im, jm, below)) + (P(i, jm, above) - P(i, jm, below))) + ((P(im, j0, above) - P(im, j0, below)) + (P(i, j0, above) - P(i, j0, below)))) / (4 * b.hz[k]);
      wn = float(double(wn) - dt * gz);
    }
    if (final) {
      if (un < -120.f || un > 120.f) { un = un < -120.f ? -120.f : 120.f; clipped++; }
      if (vn < -120.f || vn > 120.f) { vn = vn < -120.f ? -120.f : 120.f; clipped++; }
      if (wn < -100.f || wn > 100.f) { wn = wn < -100.f ? -100.f : 100.f; clipped++; }
      b.f(slot, F_U, g)[node] = un; b.f(slot, F_V, g)[node] = vn; b.f(slot, F_W, g)[node] = wn;
      const int others[5] = {F_THETA, F_Q, F_CLOUD, F_ICE, F_ICENUMBER};
      for (int o = 0; o < 5; o++) b.f(slot, others[o], g)[node] = b.x(slot, others[o], g)[node];
    } else {
      u[node] = un; v[node] = vn; w[node] = wn;
    }
  }
  if (final) {
    unsigned long long any = __ballot(clipped != 0);
    if (any) {
      __shared__ unsigned long long red[TPB / 64];
      unsigned long long t = blockSum(clipped, red);
      if (threadIdx.x == 0) atomicAdd(reinterpret_cast<unsigned long long *>(b.sc(slot) + S_CLIPPED), t);
    }
  }
}

// ---- column physics of every node (model.ts:1116-1127 'finish' + physics(z) 1347-1531) ------------------------------
// The positivity fix of the conserved fields and the decay of the cold-pool indicator come first (the 'finish'
// phase), then the node's microphysics, charging, Coriolis, wind relaxation and the sponge.
// R here is the arithmetic of the node's physics, BR the pools' type: the physics runs in float64 in both variants
// (measured: float32 microphysics left the rain's charge 1.4e-5 from the float64 step after one step of a mature
// storm, as the charge follows water increments far below a float32 mass's rounding); fields stay float32.
template <class BR, class R> __global__ __launch_bounds__(TPB) void k_physics(Geo g, Bufs<BR> b, double dt, int electrified) {
  int slot = b.runs[blockIdx.y];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n) return;
  int z = i / g.layer;
  double *S = b.sc(slot);
  float F[N_CARRIED];
  int nc = electrified ? N_CARRIED : F_COLD + 1;
  for (int k = 0; k < nc; k++) F[k] = b.f(slot, k, g)[i];
  for (int k = nc; k < N_CARRIED; k++) F[k] = 0;
  // 'finish': fillNegative with the domain sums, then cold *= .9992^dt.
  for (int k = 0; k < N_CONSERVED; k++) {
    double removed = S[S_FILL + 3 * k], kept = S[S_FILL + 3 * k + 1], negative = S[S_FILL + 3 * k + 2];
    if (removed == 0 && negative == 0) continue;
    double scale = kept > 0 ? fmax(0.0, 1 + removed / kept) : 0;
    float &a = F[conservedField(k)];
    a = a < float(TINY_WATER) ? 0.f : float(double(a) * scale);
  }
  F[F_COLD] = float(double(F[F_COLD]) * pow(.9992, dt));
  const double *P = b.pr(slot);
  const double alt = b.zs[z], H = g.height;
  const R p = R(b.lv(slot, L_P, g)[z]), rhoZ = R(b.lv(slot, L_RHO, g)[z]), exner = R(b.lv(slot, L_EXNER, g)[z]), rhoGround = R(b.lv(slot, L_RHO, g)[0]);
  const double ue = b.lv(slot, L_UENV, g)[z], ve = b.lv(slot, L_VENV, g)[z], thEnv = b.lv(slot, L_THETA, g)[z];
  const double spongeStart = P[P_SPONGE_START], keep = exp(-1e-4 * dt), fc = P[P_CORIOLIS];
  const R Rdt = R(dt);
  float &u = F[F_U], &v = F[F_V], &w = F[F_W], &theta = F[F_THETA], &q = F[F_Q], &cloud = F[F_CLOUD], &rain = F[F_RAIN], &cold = F[F_COLD];
  float &ice = F[F_ICE], &iceNumber = F[F_ICENUMBER], &snow = F[F_SNOW], &graupel = F[F_GRAUPEL], &hail = F[F_HAIL];
  float &cI = F[F_CHARGE_ICE], &cG = F[F_CHARGE_GRAUPEL], &cR = F[F_CHARGE_RAIN], &cA = F[F_CHARGE_AIR];
  // The water of the four charge carriers before and after the microphysics, and the charges, in float64: the charge
  // follows mass changes that can be far below the float32 rounding of a sum such as ice + snow.
  double before[4], after[4], cNow[4];
  if (electrified) { before[0] = double(ice) + snow; before[1] = double(graupel) + hail; before[2] = rain; before[3] = cloud; }
  R evap = rainProcesses(q, cloud, rain, R(theta) * exner, p, rhoZ, Rdt);
  if (evap > 0) { R cool = R(LV / CP) / exner * evap; theta = float(R(theta) - cool); cold = float(R(cold) + cool); }
  freezeRain(theta, rain, graupel, exner);
  settleCrystals<R>(ice, iceNumber);
  cold = float(R(cold) + iceProcesses(theta, q, ice, snow, iceNumber, exner, p, rhoZ, Rdt));
  R iceBefore = ice, cloudBefore = cloud;
  cold = float(R(cold) + snowProcesses(theta, q, cloud, ice, snow, rain, exner, p, rhoZ, rhoGround, Rdt));
  R rainBefore = rain;
  cold = float(R(cold) + graupelProcesses(theta, q, cloud, ice, snow, rain, graupel, exner, p, rhoZ, Rdt));
  R rainFrozen = R(theta) * exner < R(273.15) ? mmax(R(0), rainBefore - R(rain)) : R(0);
  cold = float(R(cold) + hailProcesses(theta, q, cloud, ice, snow, rain, graupel, hail, exner, p, rhoZ, Rdt));
  if (iceBefore > 0) iceNumber = float(R(iceNumber) * (R(ice) / iceBefore));
  R splinters = splintering(R(theta) * exner - R(273.15)) * R(SPLINTERS) * mmax(R(0), cloudBefore - R(cloud))
    + (rainFrozen > 0 ? fragmentsPerKg(mpow(R(PI * RHO_W * N0R) / (rhoZ * rainBefore), R(.25))) * rainFrozen : R(0));
  if (splinters > 0) {
    R fromG = mmin(R(graupel), splinters * R(CRYSTAL_MASS)), fromS = mmin(R(snow), splinters * R(CRYSTAL_MASS) - fromG);
    graupel = float(R(graupel) - fromG); snow = float(R(snow) - fromS); ice = float(R(ice) + (fromG + fromS)); iceNumber = float(R(iceNumber) + (fromG + fromS) / R(CRYSTAL_MASS));
  }
  R cond = saturationAdjustMixed(theta, q, cloud, ice, exner, p, iceNumber);
  settleCrystals<R>(ice, iceNumber);
  if (electrified) {
    after[0] = double(ice) + snow; after[1] = double(graupel) + hail; after[2] = rain; after[3] = cloud;
    cNow[0] = cI; cNow[1] = cG; cNow[2] = cR; cNow[3] = cA;
    if (cNow[0] != 0 || cNow[1] != 0 || cNow[2] != 0 || cNow[3] != 0) transferCharge<double>(cNow, before, after);
    R tc = R(theta) * exner - R(273.15);
    bool wetG = wetGrowing(R(graupel), R(N0G), R(RHO_G), tc, R(q), p, rhoZ, R(cloud), R(rain), R(ice), R(snow));
    bool wetH = wetGrowing(R(hail), R(N0H), R(RHO_H), tc, R(q), p, rhoZ, R(cloud), R(rain), R(ice), R(snow));
    R separated = chargingRate(tc, rhoZ, R(cloud), R(ice), R(snow), R(graupel), R(hail), R(b.f(slot, F_FALL_SNOW, g)[i]), wetG, wetH, R(iceNumber) * rhoZ) * Rdt / rhoZ;
    R induced = inductiveRate(tc, rhoZ, R(cloud), R(graupel), R(hail), R(b.f(slot, F_FIELDZ, g)[i]), R(cNow[1] + double(separated)) * rhoZ, wetG, wetH) * Rdt / rhoZ;
    cI = float(cNow[0] - double(separated)); cG = float(cNow[1] + double(separated) + double(induced)); cR = float(cNow[2]); cA = float(cNow[3] - double(induced));
  }
  if (cond < 0) cold = float(R(cold) - R(LV / CP) / exner * cond * R(.14));
  // Coriolis and the relaxation of the wind's departure from the environment (float64: small departures of ~25 m/s).
  double du = (double(u) - ue) * keep, dv = (double(v) - ve) * keep;
  u = float(ue + du + fc * dv * dt); v = float(ve + dv - fc * du * dt);
  cold = cold < 1e-3f ? 0.f : (cold < 15.f ? cold : 15.f);
  if (alt > spongeStart) {
    double s = fmin(1.0, fmax(0.0, (alt - spongeStart) / (H - spongeStart))) * .06 * dt, sc = fmin(1.0, fmax(0.0, s));
    w = float(double(w) * (1 - s));
    u = float(double(u) + (ue - double(u)) * sc); v = float(double(v) + (ve - double(v)) * sc); theta = float(double(theta) + (thEnv - double(theta)) * sc);
  }
  for (int k = 0; k < nc; k++) b.f(slot, k, g)[i] = F[k];
  b.f(slot, F_CONDENSATE, g)[i] = float(R(cloud) + R(ice));
}

// ---- Smagorinsky K (turbulence.ts:87-137 viscosity: anisotropic widths, cube-root vertical length, theta_e stability)
template <class R> __global__ __launch_bounds__(TPB) void k_viscosity(Geo g, Bufs<R> b, double dt) {
  int slot = b.runs[blockIdx.y];
  const double cs = b.pr(slot)[P_CS];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n) return;
  const int nx = g.nx, ny = g.ny, nz = g.nz, layer = g.layer;
  int z = i / layer, y = (i - z * layer) / nx, x = i - z * layer - y * nx;
  const float *u = b.f(slot, F_U, g), *v = b.f(slot, F_V, g), *w = b.f(slot, F_W, g), *theta = b.f(slot, F_THETA, g), *q = b.f(slot, F_Q, g);
  const double *exner = b.lv(slot, L_EXNER, g);
  double lenH2 = cs * sqrt(g.dx * g.dy); lenH2 *= lenH2;
  double capH = .05 * fmin(g.dx, g.dy) * fmin(g.dx, g.dy) / dt;
  double local = z == 0 ? b.dzs[0] : z == nz - 1 ? b.dzs[nz - 2] : b.hz[z], len2 = cs * cbrt(g.dx * g.dy * local); len2 *= len2;
  double mn = fmin(fmin(g.dx, g.dy), fmin(z > 0 ? b.dzs[z - 1] : 1e300, z < nz - 1 ? b.dzs[z] : 1e300)), cap = .05 * mn * mn / dt;
  int up = z < nz - 1 ? layer : 0, down = z > 0 ? -layer : 0;
  // float64 arithmetic from the float32 fields (the node's K is a small difference of large terms where the air is
  // near neutral), rounded to the pool's type once.
  const double span = b.zs[z + (up ? 1 : 0)] - b.zs[z - (down ? 1 : 0)];
  int row = y * nx + z * layer;
  int e = (x + 1 == nx ? 0 : x + 1) + row, wst = (x == 0 ? nx - 1 : x - 1) + row;
  int n = x + (y + 1 == ny ? 0 : y + 1) * nx + z * layer, s = x + (y == 0 ? ny - 1 : y - 1) * nx + z * layer;
  const double hx = 2 * g.dx, hy = 2 * g.dy;
  double dudx = (double(u[e]) - double(u[wst])) / hx, dvdy = (double(v[n]) - double(v[s])) / hy, dwdz = (double(w[i + up]) - double(w[i + down])) / span;
  double dudy = (double(u[n]) - double(u[s])) / hy, dvdx = (double(v[e]) - double(v[wst])) / hx;
  double dudz = (double(u[i + up]) - double(u[i + down])) / span, dwdx = (double(w[e]) - double(w[wst])) / hx;
  double dvdz = (double(v[i + up]) - double(v[i + down])) / span, dwdy = (double(w[n]) - double(w[s])) / hy;
  double s2 = 2 * (dudx * dudx + dvdy * dvdy + dwdz * dwdz) + (dudy + dvdx) * (dudy + dvdx) + (dudz + dwdx) * (dudz + dwdx) + (dvdz + dwdy) * (dvdz + dwdy);
  double n2 = G / b.lv(slot, L_THETA, g)[z] * (double(theta[i + up]) - double(theta[i + down])) / span;
  if (b.f(slot, F_CONDENSATE, g)[i] > 1e-5f) {
    auto te = [&](int k, int level) { return double(theta[k]) * exp(LV * double(q[k]) / (CP * double(theta[k]) * exner[level])); };
    n2 = G / te(i, z) * (te(i + up, z + (up ? 1 : 0)) - te(i + down, z - (down ? 1 : 0))) / span;
  }
  const double rate = sqrt(fmax(0.0, s2 - n2 / (1.0 / 3)));  // turbulent Prandtl number 1/3
  b.arr(b.km, slot, g)[i] = R(fmin(cap, len2 * rate));
  b.arr(b.kh, slot, g)[i] = R(fmin(capH, lenH2 * rate));
}

// ---- subgrid mixing of the 8 mixed fields, surface drag and rigid walls (turbulence.ts:143-185 mix, model.ts:1161-1195)
// Gather form of the reference's face loop: each node sums the fluxes through its six faces. The new values go to X;
// the end-of-step projection moves them into the state.
template <class R> __global__ __launch_bounds__(TPB) void k_mix(Geo g, Bufs<R> b, double dt) {
  int slot = b.runs[blockIdx.y];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n) return;
  const int nx = g.nx, ny = g.ny, nz = g.nz, layer = g.layer;
  int z = i / layer, y = (i - z * layer) / nx, x = i - z * layer - y * nx;
  int row = y * nx + z * layer;
  int e = (x + 1 == nx ? 0 : x + 1) + row, wst = (x == 0 ? nx - 1 : x - 1) + row;
  int n = x + (y + 1 == ny ? 0 : y + 1) * nx + z * layer, s = x + (y == 0 ? ny - 1 : y - 1) * nx + z * layer;
  const R *km = b.arr(b.km, slot, g), *kh = b.arr(b.kh, slot, g);
  const double *rho = b.lv(slot, L_RHO, g);
  // float64 arithmetic from the float32 fields and the pool's K, each new value rounded once.
  const double dx = g.dx, dy = g.dy;
  const double khi = kh[i], kmi = km[i];
  const double kx0 = (khi + double(kh[e])) / 2, kx1 = (double(kh[wst]) + khi) / 2, ky0 = (khi + double(kh[n])) / 2, ky1 = (double(kh[s]) + khi) / 2;
  double kzu = 0, kzd = 0, cu = 0, cd = 0;
  if (z < nz - 1) { kzu = (kmi + double(km[i + layer])) / 2 * ((rho[z] + rho[z + 1]) / 2) / b.dzs[z]; cu = 1 / (rho[z] * b.hz[z]); }
  if (z > 0) { kzd = (double(km[i - layer]) + kmi) / 2 * ((rho[z - 1] + rho[z]) / 2) / b.dzs[z - 1]; cd = 1 / (rho[z] * b.hz[z]); }
  const int mixed[8] = {F_U, F_V, F_W, F_THETA, F_Q, F_CLOUD, F_ICE, F_ICENUMBER};
  const double factors[8] = {1, 1, 1, 3, 3, 3, 3, 3};
  float outv[8];
  for (int k = 0; k < 8; k++) {
    const float *a = b.f(slot, mixed[k], g);
    // Departures from the base state of the level (u, v: the environment's wind in the frame; theta, q: the base state).
    double base = 0, bu = 0, bd = 0;
    int lvl = mixed[k] == F_U ? L_UENV : mixed[k] == F_V ? L_VENV : mixed[k] == F_THETA ? L_THETA : mixed[k] == F_Q ? L_Q : -1;
    if (lvl >= 0) { const double *B = b.lv(slot, lvl, g); base = B[z]; bu = z < nz - 1 ? B[z + 1] : 0; bd = z > 0 ? B[z - 1] : 0; }
    const double here = double(a[i]) - base;
    double t = (kx0 * ((double(a[e]) - base) - here) - kx1 * (here - (double(a[wst]) - base))) / (dx * dx)
             + (ky0 * ((double(a[n]) - base) - here) - ky1 * (here - (double(a[s]) - base))) / (dy * dy);
    if (z < nz - 1) t += kzu * ((double(a[i + layer]) - bu) - here) * cu;
    if (z > 0) t -= kzd * (here - (double(a[i - layer]) - bd)) * cd;
    outv[k] = float(double(a[i]) + dt * factors[k] * t);
  }
  if (z == 0) {
    // Surface drag on the ground node: dV/dt = -C_D (|V| V - |V_env| V_env) / h0, on the wind over the ground.
    const double *P = b.pr(slot);
    double ue0 = b.lv(slot, L_U, g)[0], ve0 = b.lv(slot, L_V, g)[0], envSpeed = hypot(ue0, ve0), drag = P[P_DRAG] * dt / b.hz[0];
    double ug = double(outv[0]) + P[P_FRAME_U], vg = double(outv[1]) + P[P_FRAME_V], speed = hypot(ug, vg);
    outv[0] = float(outv[0] - drag * (speed * ug - envSpeed * ue0)); outv[1] = float(outv[1] - drag * (speed * vg - envSpeed * ve0));
  }
  if (z == 0 || z == nz - 1) outv[2] = 0;
  for (int k = 0; k < 8; k++) b.x(slot, mixed[k], g)[i] = outv[k];
}

// ---- end of step: advance the run's clock -----------------------------------------------------------------------
template <class R> __global__ void k_end(Geo g, Bufs<R> b, double dt) {
  int slot = b.runs[blockIdx.x];
  if (threadIdx.x == 0) {
    double *S = b.sc(slot), negative = 0;
    for (int k = 0; k < N_CONSERVED; k++) negative += S[S_FILL + 3 * k + 2];
    S[S_NEGATIVE] += negative;
    b.pr(slot)[P_TIME] += dt;
  }
}

// ---- diagnostics of a run (model.ts:1672-1696 and the harness's metrics): one block per run --------------------------
enum : int { D_MAXW = 0, D_MINW, D_MAXTHETAP, D_RAINTOTAL, D_WATER, D_CHARGE, N_DIAG_DEV };
template <class R> __global__ void k_diag(Geo g, Bufs<R> b, double *out) {
  int slot = b.runs[blockIdx.x];
  double mx = -1e300, mn = 1e300, tp = -1e300, rt = 0, water = 0, charge = 0;
  const float *w = b.f(slot, F_W, g), *th = b.f(slot, F_THETA, g);
  const double *te = b.lv(slot, L_THETA, g), *rho = b.lv(slot, L_RHO, g);
  for (int i = threadIdx.x; i < g.n; i += blockDim.x) {
    int z = i / g.layer;
    double wv = w[i]; mx = fmax(mx, wv); mn = fmin(mn, wv); tp = fmax(tp, double(th[i]) - te[z]);
    double lvl = double(b.f(slot, F_Q, g)[i]) + b.f(slot, F_CLOUD, g)[i] + b.f(slot, F_ICE, g)[i] + b.f(slot, F_RAIN, g)[i] + b.f(slot, F_SNOW, g)[i] + b.f(slot, F_GRAUPEL, g)[i] + b.f(slot, F_HAIL, g)[i];
    water += b.hz[z] * rho[z] * lvl / g.layer;
    double cq = double(b.f(slot, F_CHARGE_ICE, g)[i]) + b.f(slot, F_CHARGE_GRAUPEL, g)[i] + b.f(slot, F_CHARGE_RAIN, g)[i] + b.f(slot, F_CHARGE_AIR, g)[i];
    charge += cq * rho[z] * g.dx * g.dy * b.hz[z];
  }
  const float *prec = b.col(slot, C_PRECIPITATION, g);
  for (int c = threadIdx.x; c < g.layer; c += blockDim.x) { rt = fmax(rt, double(prec[c])); water += double(prec[c]) / g.layer; }
  __shared__ double red[16];
  mx = blockMax(mx, red); mn = -blockMax(-mn, red); tp = blockMax(tp, red); rt = blockMax(rt, red);
  water = blockSum(water, red); charge = blockSum(charge, red);
  if (threadIdx.x == 0) {
    double *o = out + size_t(blockIdx.x) * N_DIAG_DEV;
    o[D_MAXW] = mx; o[D_MINW] = mn; o[D_MAXTHETAP] = tp; o[D_RAINTOTAL] = rt; o[D_WATER] = water; o[D_CHARGE] = charge;
  }
}

}  // namespace ml
