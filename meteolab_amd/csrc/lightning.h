// Lightning of the StormLab step, the data-parallel parts: the charge density of the four carriers and the right-hand
// side of the potential's Poisson equation (lightning.ts:228-232 solve, electric.ts:36-40), the node where |E|/E_init is
// largest (lightning.ts:235-248 strongest), the ion conduction with the ground's corona (lightning.ts:190-225 conduct)
// and the vertical field the inductive charging reads (model.ts:1549-1551).
#pragma once
#include "common.h"
#include "physics.h"
#include "spectral.h"

namespace ml {

constexpr double E_INIT = 2.84e5, RHO_STANDARD = 1.225, RHO_MIN = 1e-10;
constexpr double SIGMA_GROUND = 2e-14, SIGMA_SCALE = 6000, CLOUD_SIGMA = .1, CLOUDY = 1e-5, CORONA_FIELD = 5e3, CORONA_RATE = 5e-13;

// The charge density at node i, C/m3: the sum of the carriers (C/kg) times the air density `air` of its level.
template <class R> __device__ inline double chargeDensity(const Geo &g, const Bufs<R> &b, int slot, int i, double air) {
  double s = 0;
  s += b.f(slot, F_CHARGE_ICE, g)[i]; s += b.f(slot, F_CHARGE_GRAUPEL, g)[i]; s += b.f(slot, F_CHARGE_RAIN, g)[i]; s += b.f(slot, F_CHARGE_AIR, g)[i];
  return s * air;
}
// rho at node i into lrho; at the interior levels -rho/eps0 into `cells` (level k at layer k - 1). The cloudiness the
// lightning reads (cloud + ice + snow, float32 as lightning.ts:320-322 adds them) goes with it into F_CLOUDY, so that
// the discharge and the conduction of this step see this step's clouds even while the next step's dynamics runs.
template <class R> __device__ inline double chargeAt(const Geo &g, const Bufs<R> &b, int slot, int i) {
  int z = i / g.layer;
  double rho = chargeDensity(g, b, slot, i, b.lv(slot, L_RHO, g)[z]);
  b.arr(b.lrho, slot, g)[i] = R(rho);
  if (z > 0 && z < g.nz - 1) b.arr(b.cells, slot, g)[i - g.layer] = -rho / EPSILON0;
  b.f(slot, F_CLOUDY, g)[i] = (b.f(slot, F_CLOUD, g)[i] + b.f(slot, F_ICE, g)[i]) + b.f(slot, F_SNOW, g)[i];
  return rho;
}
// Also clears the run's strongest-field key, which k_strongest then raises after the solve, and marks the run charged
// when any of its nodes holds charge (an uncharged run's electric solve is zero, and skipped).
template <class R> __global__ __launch_bounds__(256) void k_charge_rhs(Geo g, Bufs<R> b) {
  int slot = b.runs[blockIdx.y];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) reinterpret_cast<unsigned long long *>(b.sc(slot))[S_STRONGEST] = 0;
  int charged = 0;
  if (i < g.n) charged = chargeAt(g, b, slot, i) != 0;
  if (__syncthreads_or(charged) && threadIdx.x == 0) atomicOr(reinterpret_cast<unsigned long long *>(b.sc(slot)) + S_CHARGED, 1ull);
}

// The largest (|E| / E_init)^2 over the interior levels, packed with its node (the first node in node order wins a tie,
// as the reference's strict comparison does) into the run's S_STRONGEST key.
template <class R> __global__ __launch_bounds__(256) void k_strongest(Geo g, Bufs<R> b) {
  int slot = b.runs[blockIdx.y];
  if (unchargedRun(b, slot)) return;  // no field: the key stays 0
  int i = g.layer + blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long key = 0;
  if (i < g.n - g.layer) {
    int z = i / g.layer, y = (i - z * g.layer) / g.nx, x = i - z * g.layer - y * g.nx, row = y * g.nx + z * g.layer;
    const double *phi = b.arr(b.ephi, slot, g);
    int xp = x + 1 == g.nx ? 0 : x + 1, xm = x == 0 ? g.nx - 1 : x - 1, yp = y + 1 == g.ny ? 0 : y + 1, ym = y == 0 ? g.ny - 1 : y - 1;
    R ex = R((phi[xp + row] - phi[xm + row]) / (2 * g.dx)), ey = R((phi[x + yp * g.nx + z * g.layer] - phi[x + ym * g.nx + z * g.layer]) / (2 * g.dy));
    R ez = R((phi[i + g.layer] - phi[i - g.layer]) / (b.zs[z + 1] - b.zs[z - 1]));
    R lim = R(b.lv(slot, L_EINIT, g)[z]);
    float r = float((ex * ex + ey * ey + ez * ez) / (lim * lim));
    if (r > 0) key = (static_cast<unsigned long long>(__float_as_uint(r)) << 32) | static_cast<unsigned int>(0xFFFFFFFFu - unsigned(i));
  }
  __shared__ unsigned long long red[4];
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int o = 32; o > 0; o >>= 1) { unsigned long long w = __shfl_xor(key, o, 64); key = w > key ? w : key; }
  if (lane == 0) red[wave] = key;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = red[0];
    for (int w = 1; w < int(blockDim.x) / 64; w++) m = red[w] > m ? red[w] : m;
    if (m) atomicMax(reinterpret_cast<unsigned long long *>(b.sc(slot) + S_STRONGEST), m);
  }
}

// Conductivity of a node: the clear air's, growing with height, a tenth of it in cloud (cloud + ice + snow > 1e-5).
template <class R> __device__ inline R sigmaAt(const Bufs<R> &b, const Geo &g, int slot, int i, int z) {
  R clear = R(SIGMA_GROUND * exp(b.zs[z] / SIGMA_SCALE));
  return b.f(slot, F_CLOUDY, g)[i] > float(CLOUDY) ? clear * R(CLOUD_SIGMA) : clear;
}
template <class R> __device__ inline R series(R a, R c) { return 2 * a * c / (a + c); }

// Gather form of the conduction's face loop at the interior levels: the charge each node gets through its six faces
// (the ground and the top are conductors at phi = 0) and, at level 1, the corona; then the vertical field.
template <class R> __global__ __launch_bounds__(256) void k_conduct(Geo g, Bufs<R> b, double dt) {
  int slot = b.runs[blockIdx.y];
  int i = g.layer + blockIdx.x * blockDim.x + threadIdx.x;
  double lost = 0;
  if (i < g.n - g.layer) {
    const int nx = g.nx, ny = g.ny, nz = g.nz, layer = g.layer;
    int z = i / layer, y = (i - z * layer) / nx, x = i - z * layer - y * nx, row = y * nx + z * layer;
    const double *phi = b.arr(b.ephi, slot, g);
    const double dx = g.dx, dy = g.dy;
    R across = R(dy * b.hz[z] * dt / dx), along = R(dx * b.hz[z] * dt / dy);
    int xp = (x + 1 == nx ? 0 : x + 1) + row, xm = (x == 0 ? nx - 1 : x - 1) + row;
    int yp = x + (y + 1 == ny ? 0 : y + 1) * nx + z * layer, ym = x + (y == 0 ? ny - 1 : y - 1) * nx + z * layer;
    R si = sigmaAt(b, g, slot, i, z);
    double pi = phi[i];
    R change = 0;
    change -= series(si, sigmaAt(b, g, slot, xp, z)) * R(pi - phi[xp]) * across;
    change -= series(si, sigmaAt(b, g, slot, yp, z)) * R(pi - phi[yp]) * along;
    change += series(sigmaAt(b, g, slot, xm, z), si) * R(phi[xm] - pi) * across;
    change += series(sigmaAt(b, g, slot, ym, z), si) * R(phi[ym] - pi) * along;
    // The face below (z - 1 to z) brings charge in, the face above (z to z + 1) takes it out.
    R upBelow = R(dx * dy * dt / b.dzs[z - 1]), upAbove = R(dx * dy * dt / b.dzs[z]);
    R qb = series(sigmaAt(b, g, slot, i - layer, z - 1), si) * R(phi[i - layer] - pi) * upBelow;
    R qa = series(si, sigmaAt(b, g, slot, i + layer, z + 1)) * R(pi - phi[i + layer]) * upAbove;
    change += qb; change -= qa;
    if (z == 1) lost -= double(qb);        // out of the ground into level 1
    if (z == nz - 2) lost += double(qa);   // out at the top
    if (z == 1) {
      // Corona: the field at the ground E0 = -phi_1 / dz0 beyond CORONA_FIELD drives ions of its sign up into level 1.
      double e0 = -double(pi) / b.dzs[0], over = fabs(e0) - CORONA_FIELD;
      if (over > 0) { double q = (e0 > 0 ? 1 : e0 < 0 ? -1 : 0) * CORONA_RATE * over * dx * dy * dt; change += R(q); lost -= q; }
    }
    if (change != 0) {
      double perKg = 1 / (b.lv(slot, L_RHO, g)[z] * dx * dy * b.hz[z]);
      float *air = b.f(slot, F_CHARGE_AIR, g);
      air[i] = float(double(air[i]) + double(change) * perKg);
    }
    b.f(slot, F_FIELDZ, g)[i] = float(-(phi[i + layer] - phi[i - layer]) / (b.zs[z + 1] - b.zs[z - 1]));
  }
  __shared__ double red[4];
  storePart(b, slot, 1, blockSum(lost, red));
}

}  // namespace ml
