// Column physics of the StormLab step on the device: warm rain, ice, snow, graupel, hail, charging.
//
// Each function works on one node. Field values are float (the model's float32 storage) and are passed by
// reference, so every store rounds to float as the reference's Float32Array stores do; the arithmetic type R is
// float (the production path) or double (the float64 variant of the same kernels). Reference file:line on each.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace ml {

// src/core/constants.ts:1
constexpr double G = 9.81, CP = 1004, LV = 2.5e6, KAPPA = .286, RD = 287.05, RV = 461.5, EPS = .622, OMEGA = 7.2921e-5;
constexpr double PI = 3.141592653589793;

// Math of the arithmetic type: float calls the single-precision device functions, double the double ones.
__device__ inline float mexp(float x) { return expf(x); }
__device__ inline double mexp(double x) { return exp(x); }
__device__ inline float mpow(float x, float y) { return powf(x, y); }
__device__ inline double mpow(double x, double y) { return pow(x, y); }
__device__ inline float msqrt(float x) { return sqrtf(x); }
__device__ inline double msqrt(double x) { return sqrt(x); }
__device__ inline float mcbrt(float x) { return cbrtf(x); }
__device__ inline double mcbrt(double x) { return cbrt(x); }
__device__ inline float mlog(float x) { return logf(x); }
__device__ inline double mlog(double x) { return log(x); }
__device__ inline float mabs(float x) { return fabsf(x); }
__device__ inline double mabs(double x) { return fabs(x); }
template <class R> __device__ inline R mmax(R a, R b) { return a > b ? a : b; }
template <class R> __device__ inline R mmin(R a, R b) { return a < b ? a : b; }
template <class R> __device__ inline R mmin3(R a, R b, R c) { return mmin(mmin(a, b), c); }
template <class R> __device__ inline R clampv(R x, R a, R b) { return mmax(a, mmin(b, x)); }
template <class R> __device__ inline R lerpv(R a, R b, R t) { return a + (b - a) * clampv(t, R(0), R(1)); }

// ---- microphysics.ts:5, 37-77 ---------------------------------------------------------------------------------
template <class R> __device__ inline R qsatP(R t, R p) {
  R es = R(611.2) * mexp(R(17.67) * t / (t + R(243.5)));
  return clampv(R(EPS) * es / mmax(R(1000), p - es), R(0), R(.045));
}
template <class R> __device__ inline R fallSpeedRain(R rain, R rho, R rhoGround) {
  return rain > 0 ? R(36.34) * mpow(R(.001) * rho * rain, R(.1364)) * msqrt(rhoGround / rho) : R(0);
}
template <class R> __device__ R rainProcesses(float &q, float &cloud, float &rain, R tk, R p, R rho, R dt) {
  R qc = cloud, qr = rain;
  R rate = R(.001) * mmax(R(0), qc - R(.001)) + (qr > 0 ? R(2.2) * qc * mpow(qr, R(.875)) : R(0));
  R production = mmin(qc, rate * dt);
  cloud = float(cloud - production); rain = float(rain + production);
  R qs = qsatP(tk - R(273.15), p);
  if (rain <= 0 || R(q) >= qs) return 0;
  R rq = R(.001) * rho * rain, ventilation = R(1.6) + R(124.9) * mpow(rq, R(.2046));
  R er = (1 - q / qs) * ventilation * mpow(rq, R(.525)) / (R(.001) * rho * (R(5.4e5) + R(2.55e6) / (p / 100 * qs)));
  R evaporation = mmin3(R(rain), er * dt, (qs - q) / (1 + R(LV) * R(LV) * qs / (R(CP) * R(RV) * tk * tk)));
  rain = float(rain - evaporation); q = float(q + evaporation);
  return evaporation;
}
template <class R> __device__ R saturationAdjust(float &theta, float &q, float &cloud, R exner, R p) {
  R total = 0;
  for (int k = 0; k < 2; k++) {
    R tk = theta * exner, qs = qsatP(tk - R(273.15), p);
    R d = (q - qs) / (1 + R(LV) * R(LV) * qs / (R(CP) * R(RV) * tk * tk));
    if (d < 0) d = mmax(d, -R(cloud));
    if (d == 0) break;
    q = float(q - d); cloud = float(cloud + d); theta = float(theta + R(LV) / (R(CP) * exner) * d); total += d;
  }
  return total;
}

// ---- ice.ts:91-303 -------------------------------------------------------------------------------------------
constexpr double LF = 3.34e5, LS = LV + LF, T0 = 273.15, T_HOMOGENEOUS = 235.15, CW = 4187;
constexpr double CRYSTAL_MASS = 1e-12, MAX_CRYSTAL_MASS = 9.4e-10, RHO_I = 900;
constexpr double DROPLET_MASS = PI / 6 * 1000 * 8.000000000000002e-15;  // 20 um droplets (ice.ts:102)
constexpr double SPLINTERS = 3.5e8;
constexpr double KA = 2.43e-2, PSI = 2.26e-5, NU = 1.51e-5, SC = .6, N0S = 3e6, RHO_S = 100;
constexpr double C_S = 4.836071225695503;      // 152.93 * 0.01^0.75 (ice.ts:122)
constexpr double SC_CBRT = 0.8434326653017492;  // 0.6^(1/3)
constexpr double D_S = .25, GAMMA_3D = 2.5493, GAMMA_4D = 8.2850, GAMMA_5D2 = 1.4569;

template <class R> __device__ inline R splintering(R tc) { return tc >= -3 || tc <= -8 ? R(0) : tc > -5 ? (R(-3) - tc) / 2 : (tc + 8) / 3; }
template <class R> __device__ inline R fragmentsPerKg(R lambda) { return R(2.5e13 * 7) / (lambda * 1000 * R(PI) / 6); }
template <class R> __device__ inline R qsatIce(R t, R p) {
  if (t >= 0) return qsatP(t, p);
  R es = R(611.21) * mexp(R(22.587) * t / (t + R(273.86)));
  return clampv(R(EPS) * es / mmax(R(1000), p - es), R(0), R(.045));
}
template <class R> __device__ inline R cooperCrystals(R tc) { return tc >= -5 ? R(0) : mmin(R(250e3), 5 * mexp(R(.304) * -tc)); }
template <class R> __device__ inline R snowSlope(R snow, R rho) { return mpow(R(PI * RHO_S * N0S) / (rho * snow), R(.25)); }
template <class R> __device__ inline R snowFallSpeed(R snow, R rho, R rhoGround) {
  return snow > R(1e-9) ? R(C_S * GAMMA_4D) / 6 * mpow(snowSlope(snow, rho), R(-D_S)) * msqrt(rhoGround / rho) : R(0);
}
template <class R> __device__ R saturationAdjustMixed(float &theta, float &q, float &cloud, float &ice, R exner, R p, float &number) {
  R heat = R(LF) / (R(CP) * exner);
  if (theta * exner >= R(T_HOMOGENEOUS)) {
    if (ice > 0 && theta * exner > R(T0)) {
      R melt = mmin(R(ice), (theta * exner - R(T0)) * R(CP) / R(LF));
      number = float(number * (1 - melt / ice));
      ice = float(ice - melt); cloud = float(cloud + melt); theta = float(theta - heat * melt);
    }
    return saturationAdjust(theta, q, cloud, exner, p);
  }
  if (cloud > 0) { number = float(number + cloud / R(DROPLET_MASS)); ice = float(ice + R(cloud)); theta = float(theta + heat * cloud); cloud = 0; }
  R total = 0;
  for (int k = 0; k < 2; k++) {
    R tk = theta * exner, qs = qsatIce(tk - R(T0), p);
    R d = (q - qs) / (1 + R(LS) * R(LS) * qs / (R(CP) * R(RV) * tk * tk));
    if (d < 0) d = mmax(d, -R(ice));
    if (d == 0) break;
    q = float(q - d); ice = float(ice + d); theta = float(theta + R(LS) / (R(CP) * exner) * d); total += d;
  }
  return total;
}
template <class R> __device__ inline void freezeRain(float &theta, float &rain, float &graupel, R exner) {
  R r = rain;
  if (r <= 0 || theta * exner >= R(T_HOMOGENEOUS)) return;
  rain = 0; graupel = float(graupel + r); theta = float(theta + R(LF) / (R(CP) * exner) * r);
}
__constant__ const float KOENIG_A1F[31] = {.7939e-7f, .7841e-6f, .3369e-5f, .4336e-5f, .5285e-5f, .3728e-5f, .1852e-5f, .2991e-6f, .4248e-6f, .7434e-6f, .1812e-5f, .4394e-5f, .9145e-5f, .1725e-4f, .3348e-4f,
  .1725e-4f, .9175e-5f, .4412e-5f, .2252e-5f, .9115e-6f, .4876e-6f, .3473e-6f, .4758e-6f, .6306e-6f, .8573e-6f, .7868e-6f, .7192e-6f, .6513e-6f, .5956e-6f, .5333e-6f, .4834e-6f};
__constant__ const double KOENIG_A1D[31] = {.7939e-7, .7841e-6, .3369e-5, .4336e-5, .5285e-5, .3728e-5, .1852e-5, .2991e-6, .4248e-6, .7434e-6, .1812e-5, .4394e-5, .9145e-5, .1725e-4, .3348e-4,
  .1725e-4, .9175e-5, .4412e-5, .2252e-5, .9115e-6, .4876e-6, .3473e-6, .4758e-6, .6306e-6, .8573e-6, .7868e-6, .7192e-6, .6513e-6, .5956e-6, .5333e-6, .4834e-6};
__constant__ const double KOENIG_A2D[31] = {.4006, .4831, .5320, .5307, .5319, .5249, .4888, .3894, .4047, .4318, .4771, .5183, .5463, .5651, .5813,
  .5655, .5478, .5203, .4906, .4447, .4126, .3960, .4149, .4320, .4506, .4483, .4460, .4433, .4413, .4382, .4361};
__constant__ const float KOENIG_A2F[31] = {.4006f, .4831f, .5320f, .5307f, .5319f, .5249f, .4888f, .3894f, .4047f, .4318f, .4771f, .5183f, .5463f, .5651f, .5813f,
  .5655f, .5478f, .5203f, .4906f, .4447f, .4126f, .3960f, .4149f, .4320f, .4506f, .4483f, .4460f, .4433f, .4413f, .4382f, .4361f};
__device__ inline float koenigA1(float, int k) { return KOENIG_A1F[k]; }
__device__ inline double koenigA1(double, int k) { return KOENIG_A1D[k]; }
__device__ inline float koenigA2(float, int k) { return KOENIG_A2F[k]; }
__device__ inline double koenigA2(double, int k) { return KOENIG_A2D[k]; }
template <class R> __device__ inline R koenigGrowth(R tc, R m) {
  R x = clampv(-tc, R(1), R(31)) - 1; int k = int(floor(x)); if (k > 29) k = 29; R f = x - k, g = m * 1000;
  return R(1e-3) * ((1 - f) * koenigA1(R(0), k) * mpow(g, koenigA2(R(0), k)) + f * koenigA1(R(0), k + 1) * mpow(g, koenigA2(R(0), k + 1)));
}
template <class R> __device__ R iceProcesses(float &theta, float &q, float &ice, float &snow, float &number, R exner, R p, R rho, R dt) {
  R tk = theta * exner, tc = tk - R(T0);
  if (tc >= 0 || tk < R(T_HOMOGENEOUS)) return 0;
  R qsi = qsatIce(tc, p), limit = (q - qsi) / (1 + R(LS) * R(LS) * qsi / (R(CP) * R(RV) * tk * tk)), heat = R(LS) / (R(CP) * exner), lacking = cooperCrystals(tc) / rho - number;
  R change = 0;
  if (limit > 0 && lacking > 0) { change = mmin(limit, lacking * R(CRYSTAL_MASS)); number = float(number + change / R(CRYSTAL_MASS)); }
  R qi = ice + change, n = number;
  if (qi > R(1e-12) && n > 0) {
    if (R(q) > qsi) change += mmin(n * koenigGrowth(tc, qi / n) * (q - qsi) / (qsatP(tc, p) - qsi) * dt, limit - change);
    else {
      R a = R(LS) * R(LS) / (R(KA) * R(RV) * tk * tk), b = 1 / (rho * qsi * R(PSI));
      R lost = mmin(qi, n * 2 * R(PI) * mcbrt(6 * rho * qi / (R(PI) * R(RHO_I) * (n * rho))) * (1 - q / qsi) / (a + b) * dt);
      change -= lost; number = float(number * (1 - lost / qi));
    }
  }
  ice = float(ice + change); q = float(q - change); theta = float(theta + heat * change);
  R excess = ice - number * R(MAX_CRYSTAL_MASS);
  if (excess > 0) { ice = float(ice - excess); snow = float(snow + excess); }
  return change < 0 ? -heat * change : R(0);
}
template <class R> __device__ inline void settleCrystals(float &ice, float &number) {
  R qi = ice;
  number = float(qi > R(1e-12) ? mmin(qi / R(CRYSTAL_MASS), mmax(qi / R(MAX_CRYSTAL_MASS), R(number))) : R(0));
}
template <class R> __device__ R snowProcesses(float &theta, float &q, float &cloud, float &ice, float &snow, float &rain, R exner, R p, R rho, R rhoGround, R dt) {
  R tk = theta * exner, tc = tk - R(T0), qs = snow;
  if (tc < 0 && ice > 0) {
    R aut = R(1e-3) * mexp(R(.025) * tc) * mmax(R(0), ice - R(.001));
    R moved = mmin(R(ice), aut * dt); ice = float(ice - moved); snow = float(snow + moved);
  }
  if (qs <= R(1e-9)) return 0;
  R lambda = snowSlope(qs, rho), fall = msqrt(rhoGround / rho);
  R sweep = R(PI * N0S * C_S * GAMMA_3D) / (4 * mpow(lambda, R(3 + D_S))) * fall;
  R vent = R(N0S) * (R(.78) / (lambda * lambda) + R(.31 * SC_CBRT * GAMMA_5D2) * msqrt(R(C_S)) * mpow(fall, R(.5)) / msqrt(R(NU)) * mpow(lambda, R(-(D_S + 5) / 2)));
  R cooling = 0;
  if (tc < 0) {
    R saci = mmin(R(ice), sweep * mexp(R(.025) * tc) * ice * dt); ice = float(ice - saci); snow = float(snow + saci);
    R sacw = mmin(R(cloud), sweep * cloud * dt); cloud = float(cloud - sacw); snow = float(snow + sacw);
    theta = float(theta + R(LF) / (R(CP) * exner) * sacw);
    R qsi = qsatIce(tc, p), a = R(LS) * R(LS) / (R(KA) * R(RV) * tk * tk), b = 1 / (rho * qsi * R(PSI));
    R rate = 2 * R(PI) * (q / qsi - 1) * vent / (rho * (a + b));
    R dep = rate > 0 ? mmin(rate * dt, (q - qsi) / (1 + R(LS) * R(LS) * qsi / (R(CP) * R(RV) * tk * tk))) : -mmin(R(snow), -rate * dt);
    snow = float(snow + dep); q = float(q - dep); theta = float(theta + R(LS) / (R(CP) * exner) * dep);
    if (dep < 0) cooling -= R(LS) / (R(CP) * exner) * dep;
  } else {
    R sacw = mmin(R(cloud), sweep * cloud * dt); cloud = float(cloud - sacw); rain = float(rain + sacw);
    R rate = 2 * R(PI) * R(KA) * tc * vent / (rho * R(LF)) + R(CW) * tc / R(LF) * sweep * cloud;
    R melt = mmin3(R(snow), rate * dt, R(CP) * tc / R(LF));
    snow = float(snow - melt); rain = float(rain + melt); theta = float(theta - R(LF) / (R(CP) * exner) * melt); cooling += R(LF) / (R(CP) * exner) * melt;
  }
  return cooling;
}

// ---- graupel.ts:6-72, hail.ts:84-175 ---------------------------------------------------------------------------
constexpr double N0G = 4e4, RHO_G = 400, CD = .6, N0R = 8e6, RHO_W = 1000, BIGG_B = 100, BIGG_A = .66;
constexpr double GAMMA_45 = 11.6317, GAMMA_35 = 3.3234, GAMMA_275 = 1.6084, SNOW_TO_GRAUPEL = 6e-4;
constexpr double N0H = 4e4, RHO_H = 900, CI = 2106, HAIL_CONVERSION_TIME = 60, HAIL_SAMPLE_VOLUME = 10;
template <class R> __device__ inline R slopeOf(R q, R rho, R n0, R rhoX) { return mpow(R(PI) * rhoX * n0 / (rho * q), R(.25)); }
template <class R> __device__ inline R fallCoefficientOf(R rho, R rhoX) { return msqrt(4 * R(G) * rhoX / (3 * R(CD) * rho)); }
template <class R> __device__ inline R sweepOf(R lambda, R a, R n0) { return R(PI) * n0 * a * R(GAMMA_35) / (4 * mpow(lambda, R(3.5))); }
template <class R> __device__ inline R ventilationOf(R lambda, R a, R n0) { return n0 * (R(.78) / (lambda * lambda) + R(.31 * SC_CBRT * GAMMA_275) * msqrt(a / R(NU)) * mpow(lambda, R(-2.75))); }
template <class R> __device__ inline R graupelFallSpeed(R g, R rho) { return g > R(1e-9) ? fallCoefficientOf(rho, R(RHO_G)) * R(GAMMA_45) / 6 * mpow(slopeOf(g, rho, R(N0G), R(RHO_G)), R(-.5)) : R(0); }
template <class R> __device__ inline R hailFallSpeed(R h, R rho) { return h > R(1e-9) ? fallCoefficientOf(rho, R(RHO_H)) * R(GAMMA_45) / 6 * mpow(slopeOf(h, rho, R(N0H), R(RHO_H)), R(-.5)) : R(0); }
template <class R> __device__ inline R hailMaxDiameter(R h, R rho) {
  if (!(h > R(1e-9))) return 0;
  R lambda = slopeOf(h, rho, R(N0H), R(RHO_H));
  return mmax(R(0), mlog(R(N0H * HAIL_SAMPLE_VOLUME) / lambda) / lambda);
}
template <class R> __device__ R graupelProcesses(float &theta, float &q, float &cloud, float &ice, float &snow, float &rain, float &graupel, R exner, R p, R rho, R dt) {
  R tk = theta * exner, tc = tk - R(T0), heat = 1 / (R(CP) * exner);
  R cooling = 0;
  if (tc < 0) {
    if (rain > 1e-9f) {
      R lambdaR = mpow(R(PI * RHO_W * N0R) / (rho * rain), R(.25));
      R rate = R(20 * PI * PI * BIGG_B * N0R * RHO_W) / rho * (mexp(R(-BIGG_A) * tc) - 1) / mpow(lambdaR, R(7));
      R frozen = mmin(R(rain), rate * dt); rain = float(rain - frozen); graupel = float(graupel + frozen); theta = float(theta + R(LF) * heat * frozen);
    }
    R aut = mmin(R(snow), R(1e-3) * mexp(R(.09) * tc) * mmax(R(0), snow - R(SNOW_TO_GRAUPEL)) * dt);
    snow = float(snow - aut); graupel = float(graupel + aut);
  }
  R qg = graupel;
  if (qg <= R(1e-9)) return 0;
  R lambda = slopeOf(qg, rho, R(N0G), R(RHO_G)), a = fallCoefficientOf(rho, R(RHO_G));
  R sweep = sweepOf(lambda, a, R(N0G)), vent = ventilationOf(lambda, a, R(N0G));
  if (tc < 0) {
    R gacw = mmin(R(cloud), sweep * cloud * dt), gacr = mmin(R(rain), sweep * rain * dt);
    R gaci = mmin(R(ice), sweep * R(.1) * ice * dt), gacs = mmin(R(snow), sweep * mexp(R(.09) * tc) * snow * dt);
    cloud = float(cloud - gacw); rain = float(rain - gacr); ice = float(ice - gaci); snow = float(snow - gacs); graupel = float(graupel + (gacw + gacr + gaci + gacs));
    theta = float(theta + R(LF) * heat * (gacw + gacr));
    R qsi = qsatIce(tc, p);
    if (R(q) < qsi) {
      R aa = R(LS) * R(LS) / (R(KA) * R(RV) * tk * tk), b = 1 / (rho * qsi * R(PSI));
      R sub = mmin3(R(graupel), 2 * R(PI) * (1 - q / qsi) * vent / (rho * (aa + b)) * dt, (qsi - q) / (1 + R(LS) * R(LS) * qsi / (R(CP) * R(RV) * tk * tk)));
      graupel = float(graupel - sub); q = float(q + sub); theta = float(theta - R(LS) * heat * sub); cooling += R(LS) * heat * sub;
    }
  } else {
    R gacw = mmin(R(cloud), sweep * cloud * dt); cloud = float(cloud - gacw); rain = float(rain + gacw);
    R rate = 2 * R(PI) * R(KA) * tc * vent / (rho * R(LF)) + R(CW) * tc / R(LF) * sweep * (R(cloud) + R(rain));
    R melt = mmin3(R(graupel), rate * dt, R(CP) * tc / R(LF));
    graupel = float(graupel - melt); rain = float(rain + melt); theta = float(theta - R(LF) * heat * melt); cooling += R(LF) * heat * melt;
  }
  return cooling;
}
template <class R> __device__ inline R wetGrowth(R lambda, R a, R n0, R tc, R qv, R p, R rho, R iceRate) {
  R heat = R(KA) * -tc + R(LV) * R(PSI) * rho * (qsatP(R(0), p) - qv);
  return mmax(R(0), 2 * R(PI) * ventilationOf(lambda, a, n0) * heat / (rho * (R(LF) + R(CW) * tc))) + iceRate * (1 - R(CI) * tc / (R(LF) + R(CW) * tc));
}
template <class R> __device__ inline bool wetGrowing(R q, R n0, R rhoX, R tc, R qv, R p, R rho, R cloud, R rain, R ice, R snow) {
  if (tc >= 0 || q <= R(1e-9)) return false;
  R lambda = slopeOf(q, rho, n0, rhoX), a = fallCoefficientOf(rho, rhoX), k = sweepOf(lambda, a, n0);
  return k * (cloud + rain + R(.1) * ice + mexp(R(.09) * tc) * snow) > wetGrowth(lambda, a, n0, tc, qv, p, rho, k * (ice + snow));
}
template <class R> __device__ R hailProcesses(float &theta, float &q, float &cloud, float &ice, float &snow, float &rain, float &graupel, float &hail, R exner, R p, R rho, R dt) {
  R tk = theta * exner, tc = tk - R(T0), heat = 1 / (R(CP) * exner);
  R cooling = 0;
  if (wetGrowing(R(graupel), R(N0G), R(RHO_G), tc, R(q), p, rho, R(cloud), R(rain), R(ice), R(snow))) {
    R turned = graupel * (1 - mexp(-dt / R(HAIL_CONVERSION_TIME))); graupel = float(graupel - turned); hail = float(hail + turned);
  }
  R qh = hail;
  if (qh <= R(1e-9)) return 0;
  R lambda = slopeOf(qh, rho, R(N0H), R(RHO_H)), a = fallCoefficientOf(rho, R(RHO_H)), k = sweepOf(lambda, a, R(N0H)), vent = ventilationOf(lambda, a, R(N0H));
  if (tc < 0) {
    R hacw = mmin(R(cloud), k * cloud * dt), hacr = mmin(R(rain), k * rain * dt), liquid = hacw + hacr;
    R haci = mmin(R(ice), k * R(.1) * ice * dt), hacs = mmin(R(snow), k * mexp(R(.09) * tc) * snow * dt), frozen = liquid;
    R wet = wetGrowth(lambda, a, R(N0H), tc, R(q), p, rho, k * (R(ice) + R(snow))) * dt;
    if (liquid + haci + hacs > wet) {
      haci = mmin(R(ice), k * ice * dt); hacs = mmin(R(snow), k * snow * dt);
      frozen = mmin(liquid, mmax(R(0), wet - haci - hacs));
    }
    cloud = float(cloud - hacw); rain = float(rain + (liquid - frozen - hacr)); ice = float(ice - haci); snow = float(snow - hacs); hail = float(hail + (frozen + haci + hacs));
    theta = float(theta + R(LF) * heat * frozen);
    R qsi = qsatIce(tc, p);
    if (R(q) < qsi) {
      R aa = R(LS) * R(LS) / (R(KA) * R(RV) * tk * tk), b = 1 / (rho * qsi * R(PSI));
      R sub = mmin3(R(hail), 2 * R(PI) * (1 - q / qsi) * vent / (rho * (aa + b)) * dt, (qsi - q) / (1 + R(LS) * R(LS) * qsi / (R(CP) * R(RV) * tk * tk)));
      hail = float(hail - sub); q = float(q + sub); theta = float(theta - R(LS) * heat * sub); cooling += R(LS) * heat * sub;
    }
  } else {
    R hacw = mmin(R(cloud), k * cloud * dt); cloud = float(cloud - hacw); rain = float(rain + hacw);
    R rate = 2 * R(PI) * R(KA) * tc * vent / (rho * R(LF)) + R(CW) * tc / R(LF) * k * (R(cloud) + R(rain));
    R melt = mmin3(R(hail), rate * dt, R(CP) * tc / R(LF));
    hail = float(hail - melt); rain = float(rain + melt); theta = float(theta - R(LF) * heat * melt); cooling += R(LF) * heat * melt;
  }
  return cooling;
}

// ---- takahashi.ts:44-54, charge.ts:98-226 -------------------------------------------------------------------------
#include "takahashi_table.inc"
__constant__ const double TAKAHASHI_LWC_D[30] = {0.01, 0.02, 0.03, 0.04, 0.05, 0.06, 0.07, 0.08, 0.09, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 20, 30};
template <class R> __device__ inline R takahashiCharge(R tc, R lwc) {
  if (tc >= 0 || tc <= -40 || lwc < R(.01)) return 0;
  R w = mmin(R(30), lwc);
  int row = 0;
  while (row < 28 && w >= R(TAKAHASHI_LWC_D[row + 1])) row++;
  R t = mmin(R(30), -mmax(R(-30), tc)); int col = int(floor(t)); if (col > 29) col = 29; R f = t - col;
  R a = R(TAKAHASHI_TABLE[row][col]) + f * (R(TAKAHASHI_TABLE[row][col + 1]) - R(TAKAHASHI_TABLE[row][col]));
  R b = R(TAKAHASHI_TABLE[row + 1][col]) + f * (R(TAKAHASHI_TABLE[row + 1][col + 1]) - R(TAKAHASHI_TABLE[row + 1][col]));
  R gl = mmin(R(1), (w - R(TAKAHASHI_LWC_D[row])) / (R(TAKAHASHI_LWC_D[row + 1]) - R(TAKAHASHI_LWC_D[row])));
  R cold = tc < -30 ? 1 - ((tc + 30) / 10) * ((tc + 30) / 10) : R(1);
  return cold * (a + gl * (b - a));
}
template <class R> __device__ inline R takahashiScaled(R table, R d, R dv) {
  R ds = d / R(100e-6);
  R q = table * mmin(5 * (ds * ds) * mabs(dv) / 8, R(10)) * R(1e-15);
  return mmax(R(-500e-15), mmin(R(500e-15), q));
}
constexpr double EPSILON0 = 8.8541878128e-12;
template <class R> struct Moments { R n, d1, d2, fall, d2v; };
template <class R> __device__ inline Moments<R> exponentialMoments(R n0, R density, R q, R rho) {
  R lambda = mpow(R(PI) * density * n0 / (rho * q), R(.25)), a = msqrt(4 * R(G) * density / (3 * R(CD) * rho));
  return {n0 / lambda, n0 / (lambda * lambda), 2 * n0 / mpow(lambda, R(3)), a * R(GAMMA_45) / 6 / msqrt(lambda), a * n0 * R(GAMMA_35) / mpow(lambda, R(3.5))};
}
template <class R> __device__ inline R rimerRate(const Moments<R> &m, R rebound, R ni, R di, R ns, R ds, R vs, R table) {
  R fall = m.fall, k = R(PI) / 4 * rebound;
  R rate = 0;
  if (ni > 0) { R sw = m.d2 + 2 * di * m.d1 + di * di * m.n; rate += clampv(ni * k * sw * fall * takahashiScaled(table, di, fall), R(-3e-9), R(3e-9)); }
  if (ns > 0) { R v = mabs(fall - vs), sw = m.d2 + 2 * ds * m.d1 + ds * ds * m.n; rate += clampv(ns * k * sw * v * takahashiScaled(table, ds, v), R(-3e-9), R(3e-9)); }
  return rate;
}
template <class R> __device__ inline R crystalDiameter(R qi, R rho, R tc) {
  if (!(qi > 0)) return 0;
  R cooper = mmin(R(150e3), 5 * mexp(R(.304) * mmax(R(0), -tc))), fromMass = rho * qi / R(PI / 6 * 900 * 1.0000000000000002e-12);
  R number = mmin(R(150e3), mmax(cooper, fromMass));
  return mmin(R(300e-6), mmax(R(20e-6), mcbrt(6 * rho * qi / (R(PI) * 900 * number))));
}
template <class R> __device__ R chargingRate(R tc, R rho, R cloud, R ice, R snow, R graupel, R hail, R snowFall, bool wetG, bool wetH, R crystals) {
  if (tc >= 0 || tc < -40 || cloud <= 0 || ice + snow <= 0 || graupel + hail <= 0) return 0;
  R rebound = R(.5) * (1 - R(.1) * mexp(R(.1) * tc)), water = rho * cloud * 1000;
  R ni = ice <= 0 ? R(0) : crystals;
  R di = ice <= 0 ? R(0) : crystals != 0 ? mmin(R(300e-6), mmax(R(20e-6), R(.1871) * mpow(rho * ice / crystals, R(.3429)))) : crystalDiameter(ice, rho, tc);
  R ss = snow > 0 ? mpow(R(PI * RHO_S * N0S) / (rho * snow), R(.25)) : R(0), ns = snow > 0 ? R(N0S) / ss : R(0), ds = snow > 0 ? R(1.8171) / ss : R(0);
  R rate = 0, table = takahashiCharge(tc, water);
  if (graupel > 0 && !wetG) rate += rimerRate(exponentialMoments(R(N0G), R(RHO_G), graupel, rho), rebound, ni, di, ns, ds, snowFall, table);
  if (hail > 0 && !wetH) rate += rimerRate(exponentialMoments(R(N0H), R(RHO_H), hail, rho), rebound, ni, di, ns, ds, snowFall, table);
  return rate;
}
template <class R> __device__ R inductiveRate(R tc, R rho, R cloud, R graupel, R hail, R ez, R rimerCharge, bool wetG, bool wetH) {
  if (tc >= 0 || cloud <= 0) return 0;
  const R d = R(20e-6);
  R droplets = rho * cloud / R(PI / 6 * 1000 * 8.000000000000002e-15), k = R(.6 * .01) * droplets;
  R polarize = 0, mass = 0, massFall = 0;
  if (!(graupel <= R(1e-9) || wetG)) {
    Moments<R> m = exponentialMoments(R(N0G), R(RHO_G), graupel, rho);
    polarize += R(PI * PI * PI * PI / 8 * EPSILON0) * ez * R(.5) * d * d * m.d2v; mass += graupel; massFall += graupel * m.fall;
  }
  if (!(hail <= R(1e-9) || wetH)) {
    Moments<R> m = exponentialMoments(R(N0H), R(RHO_H), hail, rho);
    polarize += R(PI * PI * PI * PI / 8 * EPSILON0) * ez * R(.5) * d * d * m.d2v; mass += hail; massFall += hail * m.fall;
  }
  if (mass == 0) return 0;
  return k * (polarize - R(PI * PI * PI / 24) * d * d * (massFall / mass) * rimerCharge);
}
template <class R> __device__ void transferCharge(R *charge, const R *before, const R *after) {
  R pool = 0, lost = 0, gained = 0;
  for (int k = 0; k < 4; k++) {
    if (after[k] < before[k]) { R share = before[k] > 0 ? (before[k] - after[k]) / before[k] : R(1), out = charge[k] * share; charge[k] -= out; pool += out; lost += before[k] - after[k]; }
    else gained += after[k] - before[k];
    if (k != 3 && after[k] <= 0 && charge[k] != 0) { pool += charge[k]; charge[k] = 0; }
  }
  if (pool == 0) return;
  R moved = gained > 0 ? mmin(R(1), lost > 0 ? gained / lost : R(1)) : R(0);
  if (moved > 0) for (int k = 0; k < 4; k++) if (after[k] > before[k]) charge[k] += pool * moved * (after[k] - before[k]) / gained;
  charge[3] += pool * (1 - moved);
}

// mulberry32 (src/core/math.ts:7-14): the k-th draw (k from 1) of a generator seeded with `seed`. The generator's
// state advances by a constant per draw, so any draw is computed on its own.
__host__ __device__ inline double mulberryAt(uint32_t seed, uint32_t k) {
  uint32_t s = seed + k * 0x6D2B79F5u;
  uint32_t t = (s ^ (s >> 15)) * (1u | s);
  t = (t + ((t ^ (t >> 7)) * (61u | t))) ^ t;
  return double(t ^ (t >> 14)) / 4294967296.0;
}

}  // namespace ml
