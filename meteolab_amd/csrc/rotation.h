// The rotation diagnostics of a step (model.ts:1609-1648 updateRotation, run in afterStep before the lightning): per
// column the updraught helicity w zeta hz summed over the 2-5 km levels, and over 0-1 km and 0-3 km; the updraught
// cores (connected columns, 8 neighbours, periodic, whose 2-5 km nodes above CORE_ROTATION_W carry w hz) and the
// most cyclonic core's mean vorticity; how long the mesocyclone criterion has held.
#pragma once
#include "common.h"

namespace ml {

constexpr double CORE_ROTATION_W = 10;  // m/s (model.ts:112)
// Levels of a grid the diagnostics sum over (model.ts:389 uhLevels, :1612 the 1 and 3 km tops) and the mesocyclone
// threshold of the core's vorticity on its spacing (model.ts:129 mesoCoreZeta).
struct RotLevels { int z0, z1, top1, top3; double meso; };

// Vertical vorticity dv/dx - du/dy at a node, float64 from the float32 winds (model.ts:1598-1601).
__device__ inline double zetaAt(const Geo &g, const float *u, const float *v, int x, int y, int z) {
  const int nx = g.nx, l = z * g.layer, xp = x + 1 == nx ? 0 : x + 1, xm = x == 0 ? nx - 1 : x - 1;
  const int yp = (y + 1 == g.ny ? 0 : y + 1) * nx, ym = (y == 0 ? g.ny - 1 : y - 1) * nx;
  return (double(v[xp + nx * y + l]) - double(v[xm + nx * y + l])) / (2 * g.dx) - (double(u[x + yp + l]) - double(u[x + ym + l])) / (2 * g.dy);
}

constexpr int ROTATION_THREADS = 1024;
// Per column: coreW, coreWZ (float64), the column's helicity (float32), the flood fill's stack (int) and label (byte);
// then one bit per column of the columns that belong to a core.
__host__ __device__ inline size_t rotationLdsBytes(const Geo &g) {
  return size_t(g.layer) * (8 + 8 + 4 + 4) + ((size_t(g.layer) + 7) / 8 * 8) + size_t((g.layer + 63) / 64) * 8;
}

// The larger of two (value, column) pairs; equal values: the first column (the reference's strict > in column order).
__device__ inline void keepMax(double &v, int &c, double v2, int c2) { if (v2 > v || (v2 == v && c2 < c)) { v = v2; c = c2; } }
__device__ inline void blockMaxAt(double &v, int &c, double *redV, int *redC) {
  for (int o = 32; o > 0; o >>= 1) { const double v2 = __shfl_xor(v, o, 64); const int c2 = __shfl_xor(c, o, 64); keepMax(v, c, v2, c2); }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) { redV[wave] = v; redC[wave] = c; }
  __syncthreads();
  v = redV[0]; c = redC[0];
  for (int k = 1; k < int(blockDim.x) / 64; k++) keepMax(v, c, redV[k], redC[k]);
  __syncthreads();
}

// One block per run. A thread per column sums its levels in the reference's order (products rounded on their own, as
// the reference's separate multiply and add), the columns of a core marked one bit each; then thread 0 floods the cores
// from those bits in column order, depth first as the reference (so a core's float64 sums add in the same order).
// keep != 0 leaves how long the rotation has held as it is (refreshRotation, model.ts:1607: the diagnostics of fields
// put in from outside).
template <class R> __global__ __launch_bounds__(ROTATION_THREADS) void k_rotation(Geo g, Bufs<R> b, double dt, RotLevels rl, int keep) {
  const int slot = b.runs[blockIdx.x], nx = g.nx, ny = g.ny, layer = g.layer;
  extern __shared__ char smem[];
  double *coreW = reinterpret_cast<double *>(smem), *coreWZ = coreW + layer;
  float *col = reinterpret_cast<float *>(coreWZ + layer);
  int *stack = reinterpret_cast<int *>(col + layer);
  unsigned char *label = reinterpret_cast<unsigned char *>(stack + layer);
  unsigned long long *cores = reinterpret_cast<unsigned long long *>(label + (layer + 7) / 8 * 8);
  __shared__ double redV[ROTATION_THREADS / 64];
  __shared__ int redC[ROTATION_THREADS / 64];
  const float *u = b.f(slot, F_U, g), *v = b.f(slot, F_V, g), *w = b.f(slot, F_W, g);
  const double *hz = b.hz;
  const int top = max(rl.top3, rl.z1);
  // Maxima start at 0 (a column counts when its sum exceeds it), the minimum at 0.
  double vmax = 0, vmin = 0, v01 = 0, v03 = 0;
  int cmax = 0, c01 = 0, cmin = 0, c03 = 0;
  for (int c = threadIdx.x; c < layer; c += blockDim.x) {
    const int y = c / nx, x = c - y * nx;
    double low = 0, uh = 0, sw = 0, swz = 0;
    // One pass over the levels of both sums: each adds its own levels in increasing height, as the reference's two loops.
#pragma unroll 4
    for (int z = 1; z <= top; z++) {
      const double wn = w[c + z * layer], wz = __dmul_rn(__dmul_rn(wn, zetaAt(g, u, v, x, y, z)), hz[z]);
      if (z <= rl.top3) low += wz;
      if (z == rl.top1) keepMax(v01, c01, low, c);
      if (z >= rl.z0 && z <= rl.z1) {
        uh += wz;
        if (wn > CORE_ROTATION_W) { sw += __dmul_rn(wn, hz[z]); swz += wz; }
      }
    }
    keepMax(v03, c03, low, c);
    col[c] = float(uh); coreW[c] = sw; coreWZ[c] = swz; label[c] = 0;
    const unsigned long long bits = __ballot(sw > 0);
    if ((threadIdx.x & 63) == 0) cores[c >> 6] = bits;
    keepMax(vmax, cmax, uh, c);
    keepMax(vmin, cmin, -uh, c);
  }
  // addSwaths (ensemble/member.ts:85-100) after the step: the ground under the domain is the domain shifted by its
  // motion so far, in whole columns (the nearest), periodic; the strongest surface wind and helicity over each.
  const double *P = b.pr(slot);
  if (!keep && P[P_SWATHS] != 0) {
    const double t = P[P_TIME] + dt, fu = P[P_FRAME_U], fv = P[P_FRAME_V];
    const int sx = int(floor(fu * t / g.dx + .5)), sy = int(floor(fv * t / g.dy + .5));
    float *gust = b.col(slot, C_GUST_MAX, g), *uhm = b.col(slot, C_UH_MAX, g);
    for (int c = threadIdx.x; c < layer; c += blockDim.x) {
      const int y = c / nx, x = c - y * nx, gx = ((x + sx) % nx + nx) % nx, gy = ((y + sy) % ny + ny) % ny, gc = gx + nx * gy;
      const double wind = hypot(double(u[c]) + fu, double(v[c]) + fv);
      if (wind > gust[gc]) gust[gc] = float(wind);
      if (col[c] > uhm[gc]) uhm[gc] = col[c];
    }
  }
  blockMaxAt(vmax, cmax, redV, redC);
  blockMaxAt(vmin, cmin, redV, redC);
  blockMaxAt(v01, c01, redV, redC);
  blockMaxAt(v03, c03, redV, redC);
  if (threadIdx.x != 0) return;
  double coreZeta = 0, coreUH = 0, coreAnti = 0;
  for (int word = 0; word < (layer + 63) / 64; word++) {
    for (unsigned long long m = cores[word]; m; m &= m - 1) {
      const int start = word * 64 + __builtin_ctzll(m);
      if (label[start]) continue;
      double sw = 0, swz = 0, peak = 0;
      int sp = 0;
      label[start] = 1; stack[sp++] = start;
      while (sp > 0) {
        const int a = stack[--sp], x = a % nx, y = a / nx;
        sw += coreW[a]; swz += coreWZ[a]; peak = fmax(peak, double(col[a]));
        for (int oy = -1; oy <= 1; oy++)
          for (int ox = -1; ox <= 1; ox++) {
            const int bb = (x + ox + nx) % nx + nx * ((y + oy + ny) % ny);
            if (coreW[bb] > 0 && !label[bb]) { label[bb] = 1; stack[sp++] = bb; }
          }
      }
      if (swz / sw > coreZeta) { coreZeta = swz / sw; coreUH = peak; }
      coreAnti = fmax(coreAnti, -swz / sw);
    }
  }
  double *r = b.sc(slot) + S_ROT;
  r[ROT_UH] = vmax; r[ROT_X] = cmax % nx; r[ROT_Y] = cmax / nx; r[ROT_ANTI] = vmin;
  r[ROT_UH01] = v01; r[ROT_UH03] = v03; r[ROT_LOW] = c01;
  r[ROT_CORE_ZETA] = coreZeta; r[ROT_CORE_UH] = coreUH; r[ROT_CORE_ANTI] = coreAnti;
  if (!keep) r[ROT_PERSISTED] = coreZeta >= rl.meso ? r[ROT_PERSISTED] + dt : 0;
}

}  // namespace ml
