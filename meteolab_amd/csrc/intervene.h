// Interventions in a run going on (intervene.ts:38-68): a smooth patch, round on the map (an e-folding radius) and
// fading with height, added to the fields as they are, with the heat and vapour put in counted (AtmosphereModel.added,
// model.ts:227-228). One block of a run: a patch covers the lowest levels only (a few e-folding depths), so the work
// is a few thousand nodes, and one block sums the heat and vapour in a fixed order.
#pragma once
#include "common.h"
#include "physics.h"

namespace ml {

// A patch as the host resolves it from INTERVENTIONS (intervene.ts:22-30): e-folding radius and depth (m), per unit
// strength the potential temperature (K), vapour (kg/kg), vertical wind (m/s) and vortex speed (m/s) at the centre,
// the strength, the centre (m from the domain's corner), the domain's extent (m) and the levels below the patch's top.
struct Patch { double radius, depth, theta, vapour, w, spin, a0, x, y, width, depth_y; int levels; };
constexpr int INTERVENE_THREADS = 1024;

template <class R> __global__ __launch_bounds__(INTERVENE_THREADS) void k_intervene(Geo g, Bufs<R> b, int slot, Patch k) {
  // The reference's association order throughout (no fused multiply-adds), so the float32 results are its own.
#pragma clang fp contract(off)
  const double CP = 1004, MOIST_CAP = .95;
  const double *rho = b.lv(slot, L_RHO, g), *exner = b.lv(slot, L_EXNER, g), *pl = b.lv(slot, L_P, g);
  float *theta = b.f(slot, F_THETA, g), *q = b.f(slot, F_Q, g), *w = b.f(slot, F_W, g), *u = b.f(slot, F_U, g), *v = b.f(slot, F_V, g);
  double heat = 0, vapour = 0;
  const int nodes = k.levels * g.layer;
  for (int i = threadIdx.x; i < nodes; i += blockDim.x) {
    const int z = i / g.layer, c = i - z * g.layer, x = c % g.nx, y = c / g.nx;
    // The shortest way round the periodic domain.
    double ddx = x * g.dx - k.x, ddy = y * g.dy - k.y;
    if (ddx > k.width / 2) ddx -= k.width;
    if (ddx < -k.width / 2) ddx += k.width;
    if (ddy > k.depth_y / 2) ddy -= k.depth_y;
    if (ddy < -k.depth_y / 2) ddy += k.depth_y;
    const double mass = rho[z] * g.dx * g.dy * b.hz[z], zd = b.zs[z] / k.depth;
    const double rx = ddx / k.radius, ry = ddy / k.radius, r2 = rx * rx + ry * ry;
    const double a = exp(-r2 - zd * zd) * k.a0;
    if (k.theta != 0) { theta[i] = float(double(theta[i]) + k.theta * a); heat += CP * exner[z] * k.theta * a * mass; }
    if (k.w != 0) w[i] = float(double(w[i]) + k.w * a);
    if (k.vapour != 0) {
      const double before = q[i], cap = MOIST_CAP * qsatP<double>(double(theta[i]) * exner[z] - 273.15, pl[z]);
      const float after = float(k.vapour > 0 ? fmax(before, fmin(cap, before + k.vapour * a)) : fmax(0.0, before + k.vapour * a));
      q[i] = after;
      vapour += (double(after) - before) * mass;
    }
    if (k.spin != 0) {
      // A cyclonic (anticlockwise) vortex, non-divergent: the speed r/R exp((1 - r^2/R^2)/2) times spin, round the centre.
      const double r = sqrt(r2), speed = k.spin * k.a0 * r * exp((1 - r2) / 2) * exp(-(zd * zd)), d = hypot(ddx, ddy);
      if (d > 0) { u[i] = float(double(u[i]) - speed * ddy / d); v[i] = float(double(v[i]) + speed * ddx / d); }
    }
  }
  __shared__ double red[INTERVENE_THREADS / 64];
  heat = blockSum(heat, red);
  vapour = blockSum(vapour, red);
  if (threadIdx.x == 0) {
    double *sc = b.sc(slot);
    sc[S_ADDED] += heat; sc[S_ADDED + 1] += vapour; sc[S_ADDED + 2] += 1;
    sc[S_LAST_ADDED] = heat; sc[S_LAST_ADDED + 1] = vapour;
  }
}

}  // namespace ml
