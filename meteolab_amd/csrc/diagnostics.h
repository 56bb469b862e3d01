// The model's diagnostics (model.ts:1672-1696 AtmosphereModel.diagnostics, ModelDiagnostics :161-180) without terrain:
// domain maxima of the up- and downdraught, cloud, cloud top, thermal top and low cold pool; at the ground nodes the
// heaviest precipitation flux (surfaceflux.ts:22-37), rain total, gust and largest hail; the cloudy updraught cores
// (countCores, model.ts:1650-1670); the environment's 0-6 km shear; the outflow's virtual-temperature deficit against
// the ground air's 75th percentile (rkw.ts:72-90) and its largest value near the low-level rotation.
#pragma once
#include "common.h"
#include "physics.h"
#include "rotation.h"

namespace ml {

constexpr int MDIAG_THREADS = 1024;
constexpr int DIAG_ROW = 16;  // doubles a run's row of diagnostics takes in the pool's diagnostics buffer
enum : int {
  MD_UPDRAFT = 0, MD_DOWNDRAFT, MD_RAINRATE, MD_RAINTOTAL, MD_CLOUDTOP, MD_THERMALTOP, MD_MAXCLOUD, MD_COLDMAX, MD_CORES,
  MD_SHEAR06, MD_OUTFLOW, MD_OUTFLOWLOW, MD_GUST, MD_HAIL, MD_CLIPPED, N_MD
};
static_assert(N_MD <= DIAG_ROW, "a row holds the diagnostics");

// Levels the diagnostics look at: cores between the first level at or above 1.9 km and the last at or below 6.6 km
// (at most nz - 3); the cold pool below the first level at or above 1.5 km.
struct DiagLevels { int coreLo, coreHi, coldTop; };

// Per column: the ground air's virtual potential temperature then its deficit (float32), the core mask and the flood's
// stack; one bit per column of the core mask.
__host__ __device__ inline size_t modelDiagLdsBytes(const Geo &g) {
  return size_t(g.layer) * (4 + 4) + ((size_t(g.layer) + 7) / 8 * 8) + size_t((g.layer + 63) / 64) * 8;
}

// One block per run; out gets a row of DIAG_ROW doubles per block, in the order of MD_*.
template <class R> __global__ __launch_bounds__(MDIAG_THREADS) void k_model_diag(Geo g, Bufs<R> b, DiagLevels dl, double *out) {
  const int slot = b.runs[blockIdx.x], layer = g.layer, nx = g.nx, ny = g.ny;
  extern __shared__ char smem[];
  float *thv = reinterpret_cast<float *>(smem);
  int *stack = reinterpret_cast<int *>(thv + layer);
  unsigned char *mask = reinterpret_cast<unsigned char *>(stack + layer);
  unsigned long long *bits = reinterpret_cast<unsigned long long *>(mask + (layer + 7) / 8 * 8);
  __shared__ double red[MDIAG_THREADS / 64];
  __shared__ float warm;
  const float *w = b.f(slot, F_W, g), *cloud = b.f(slot, F_CLOUD, g), *ice = b.f(slot, F_ICE, g), *cold = b.f(slot, F_COLD, g);
  for (int c = threadIdx.x; c < layer; c += blockDim.x) mask[c] = 0;
  if (threadIdx.x == 0) warm = 0;
  __syncthreads();
  // Every node: the maxima start at 0, as the reference's.
  double up = 0, down = 0, maxCloud = 0, top = 0, thermalTop = 0, coldMax = 0;
  for (int i = threadIdx.x; i < g.n; i += blockDim.x) {
    const int z = i / layer, c = i - z * layer;
    const double wv = w[i], ci = double(cloud[i]) + double(ice[i]), alt = b.zs[z];
    up = fmax(up, wv); down = fmax(down, -wv); maxCloud = fmax(maxCloud, ci);
    if (ci > .00007) top = fmax(top, alt / 1000);
    if (wv > .6) thermalTop = fmax(thermalTop, alt / 1000);
    if (z < dl.coldTop) coldMax = fmax(coldMax, double(cold[i]));
    if (z >= dl.coreLo && z <= dl.coreHi && wv > 2 && ci > .00007) mask[c] = 1;
  }
  // The ground nodes (level 0).
  const double *P = b.pr(slot), rho0 = b.lv(slot, L_RHO, g)[0];
  const float *u = b.f(slot, F_U, g), *v = b.f(slot, F_V, g), *theta = b.f(slot, F_THETA, g), *q = b.f(slot, F_Q, g);
  const float *rain = b.f(slot, F_RAIN, g), *snow = b.f(slot, F_SNOW, g), *graupel = b.f(slot, F_GRAUPEL, g), *hail = b.f(slot, F_HAIL, g);
  const float *prec = b.col(slot, C_PRECIPITATION, g);
  double gust = 0, rainRate = 0, rainTotal = 0, hailMax = 0;
  for (int c = threadIdx.x; c < layer; c += blockDim.x) {
    gust = fmax(gust, hypot(double(u[c]) + P[P_FRAME_U], double(v[c]) + P[P_FRAME_V]));
    const double r = rain[c], s = snow[c], gr = graupel[c], h = hail[c];
    const double vg = gr > 0 ? graupelFallSpeed<double>(gr, rho0) : 0, vh = h > 0 ? hailFallSpeed<double>(h, rho0) : 0;
    const double flux = rho0 * r * fallSpeedRain<double>(r, rho0, rho0) + rho0 * (s * snowFallSpeed<double>(s, rho0, rho0) + gr * vg + h * vh);
    rainRate = fmax(rainRate, flux * 3600); rainTotal = fmax(rainTotal, double(prec[c]));
    hailMax = fmax(hailMax, hailMaxDiameter<double>(h, rho0) * 1000);
    thv[c] = float(double(theta[c]) * (1 + .61 * double(q[c])));
  }
  __syncthreads();
  // The warm air: the element of rank floor((layer - 1) 0.75) of the ground thv in ascending order (ties by column).
  const int k = int(floor((layer - 1) * .75));
  for (int c = threadIdx.x; c < layer; c += blockDim.x) {
    const float t = thv[c];
    int rank = 0;
    for (int j = 0; j < layer; j++) { const float o = thv[j]; rank += (o < t) || (o == t && j < c); }
    if (rank == k) warm = t;
  }
  __syncthreads();
  double outflow = 0;
  for (int c = threadIdx.x; c < layer; c += blockDim.x) {
    const float d = float(fmax(0.0, double(warm) - double(thv[c])));
    outflow = fmax(outflow, double(d));
    const unsigned long long m = __ballot(mask[c] != 0);
    if ((threadIdx.x & 63) == 0) bits[c >> 6] = m;
    thv[c] = d;  // each thread rewrites only its own columns, read by others after the barrier
  }
  __syncthreads();
  // deficitNear (rkw.ts:82-90): within 2 km of the low-level rotation's column, across the periodic edges.
  const int low = int(b.sc(slot)[S_ROT + ROT_LOW]), x0 = low % nx, y0 = low / nx;
  const int rx = int(floor(2000 / g.dx)), ry = int(floor(2000 / g.dy)), wx = 2 * rx + 1;
  double outflowLow = 0;
  for (int t = threadIdx.x; t < wx * (2 * ry + 1); t += blockDim.x) {
    const int j = t / wx - ry, i = t % wx - rx;
    const double ex = i * g.dx, ey = j * g.dy;
    if (ex * ex + ey * ey > 2000.0 * 2000.0) continue;
    const int xx = ((x0 + i) % nx + nx) % nx, yy = ((y0 + j) % ny + ny) % ny;
    outflowLow = fmax(outflowLow, double(thv[xx + nx * yy]));
  }
  up = blockMax(up, red); down = blockMax(down, red); maxCloud = blockMax(maxCloud, red); top = blockMax(top, red);
  thermalTop = blockMax(thermalTop, red); coldMax = blockMax(coldMax, red); gust = blockMax(gust, red);
  rainRate = blockMax(rainRate, red); rainTotal = blockMax(rainTotal, red); hailMax = blockMax(hailMax, red);
  outflow = blockMax(outflow, red); outflowLow = blockMax(outflowLow, red);
  if (threadIdx.x != 0) return;
  // The cores: 4-connected marked columns; the count does not depend on the order of the fill.
  int cores = 0;
  for (int word = 0; word < (layer + 63) / 64; word++) {
    for (unsigned long long m = bits[word]; m; m &= m - 1) {
      const int start = word * 64 + __builtin_ctzll(m);
      if (mask[start] != 1) continue;
      cores++;
      int sp = 0;
      mask[start] = 2; stack[sp++] = start;
      while (sp > 0) {
        const int a = stack[--sp], x = a % nx, y = a / nx;
        const int nb[4] = {(x + 1 == nx ? 0 : x + 1) + nx * y, (x == 0 ? nx - 1 : x - 1) + nx * y, x + nx * (y + 1 == ny ? 0 : y + 1), x + nx * (y == 0 ? ny - 1 : y - 1)};
        for (int e = 0; e < 4; e++) if (mask[nb[e]] == 1) { mask[nb[e]] = 2; stack[sp++] = nb[e]; }
      }
    }
  }
  long long clipped;
  __builtin_memcpy(&clipped, b.sc(slot) + S_CLIPPED, 8);
  double *o = out + size_t(blockIdx.x) * DIAG_ROW;
  o[MD_UPDRAFT] = up; o[MD_DOWNDRAFT] = down; o[MD_RAINRATE] = rainRate; o[MD_RAINTOTAL] = rainTotal; o[MD_CLOUDTOP] = top;
  o[MD_THERMALTOP] = thermalTop; o[MD_MAXCLOUD] = maxCloud; o[MD_COLDMAX] = coldMax; o[MD_CORES] = cores;
  o[MD_SHEAR06] = hypot(P[P_SHEAR] * tanh(6000 / P[P_SHEAR_DEPTH]) + windError(P + P_WIND_U, 6000) - (P[P_SHEAR] * tanh(0 / P[P_SHEAR_DEPTH]) + windError(P + P_WIND_U, 0)),
                        windError(P + P_WIND_V, 6000) - windError(P + P_WIND_V, 0));
  o[MD_OUTFLOW] = outflow; o[MD_OUTFLOWLOW] = outflowLow; o[MD_GUST] = gust; o[MD_HAIL] = hailMax; o[MD_CLIPPED] = double(clipped);
}

// probabilityMap (ensemble/combine.ts:70-75): per ground column the share of the launch's `count` runs whose column
// `column` reached `threshold` there (a float32 count, then the quotient rounded to float32).
template <class R> __global__ __launch_bounds__(256) void k_probability(Geo g, Bufs<R> b, int count, int column, double threshold, float *out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.layer) return;
  float n = 0;
  for (int k = 0; k < count; k++) if (double(b.col(b.runs[k], column, g)[c]) >= threshold) n += 1;
  out[c] = float(double(n) / count);
}

}  // namespace ml
