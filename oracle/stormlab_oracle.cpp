// StormLab step oracle: a CPU restatement of the reference's model step (float64 arithmetic on float32 fields).
//
// TEST INFRASTRUCTURE ONLY: the checker of the HIP step. Loaded by tests/, __graft_entry__.smoke() and bench.py's
// CPU-baseline leg; the product library never links or calls it.
//
// Every function names the reference file:line it restates (paths relative to the reference repository). Arithmetic
// follows the reference's association order (JS evaluates a * b * c as (a * b) * c, ** as Math.pow) so that, with the
// same libm, results agree to the last bit; fields are float32 arrays and every store rounds to float32 as a
// Float32Array store does. What no reference-held test pins (the composite trajectory, V8's pow/exp last bits) is
// stated as parity-unpinned where the tests document it.
#include "oracle.h"

// ORC_MUTANT (0 or undefined in the oracle itself) builds the oracle with one planted mistake, for the test that shows
// the checks of tests/test_oracle_physics.py catch each: 1 the Koenig growth table's -15 degC peak halved, 2 the charge
// that evaporates lost instead of left to the ions, 3 wet growth freezing all it collects, 4 sublimation growing the
// ice, 5 the stability's sign in the eddy viscosity, 6 no screening of the conduction in cloud, 7 the Takahashi table's
// sign, 8 the latent heat of vaporisation heating the deposition below -38 degC, 9 the sedimentation's inflow density from the
// level itself instead of the one above.
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

namespace {

// ---- src/core/constants.ts:1, src/core/math.ts:1-14 -------------------------------------------------------------
constexpr double G = 9.81, CP = 1004, LV = 2.5e6, KAPPA = .286, RD = 287.05, RV = 461.5, EPS = .622, OMEGA = 7.2921e-5;
constexpr double PI = 3.141592653589793;
inline double clamp01(double x, double a = 0, double b = 1) { return std::max(a, std::min(b, x)); }
inline double lerp(double a, double b, double t) { return a + (b - a) * clamp01(t); }
inline double fmax3(double a, double b, double c) { return std::max(std::max(a, b), c); }
inline double fmin3(double a, double b, double c) { return std::min(std::min(a, b), c); }

// mulberry32 (src/core/math.ts:7-14) on uint32 (JS int32 ops are modulo 2^32).
inline double mulberry(uint32_t &s) {
  s += 0x6D2B79F5u;
  uint32_t t = (s ^ (s >> 15)) * (1u | s);
  t = (t + ((t ^ (t >> 7)) * (61u | t))) ^ t;
  return double(t ^ (t >> 14)) / 4294967296.0;
}

typedef std::vector<float> F32;
typedef std::vector<double> F64;

// ---- grid (src/core/grid.ts:91-121) -------------------------------------------------------------------------------
double stretchRatio(double bottom, double height, int intervals) {
  if (bottom * intervals >= height) return 1;
  double lo = 1, hi = 2;
  while (bottom * (std::pow(hi, intervals) - 1) / (hi - 1) < height) hi *= 2;
  for (int k = 0; k < 100; k++) {
    double mid = (lo + hi) / 2;
    if (bottom * (std::pow(mid, intervals) - 1) / (mid - 1) < height) lo = mid; else hi = mid;
  }
  return (lo + hi) / 2;
}

struct Grid {
  int nx, ny, nz, n, layer;
  double width, depth, height, dx, dy, dz;
  bool uniform;
  F64 zs, dzs, hz;
  std::vector<int> xp, xm, yp, ym;  // yp/ym are row offsets (y * nx)
  explicit Grid(const orc_grid &o) {
    nx = o.nx; ny = o.ny; nz = o.nz; width = o.width; depth = o.depth; height = o.height;
    n = nx * ny * nz; layer = nx * ny;
    dz = height / (nz - 1);
    zs.assign(nz, 0.0);
    uniform = !(o.bottom > 0);
    if (uniform) for (int z = 0; z < nz; z++) zs[z] = z * dz;
    else {
      double r = stretchRatio(o.bottom, height, nz - 1);
      for (int z = 1; z < nz; z++) zs[z] = zs[z - 1] + o.bottom * std::pow(r, z - 1);
      // The top lands on `height`; the factor is re-read each time as the reference's loop does (grid.ts:107).
      for (int z = 1; z < nz; z++) zs[z] *= height / zs[nz - 1];
    }
    dzs.assign(nz - 1, 0.0);
    for (int k = 0; k < nz - 1; k++) dzs[k] = uniform ? dz : zs[k + 1] - zs[k];
    hz.assign(nz, 0.0);
    for (int k = 0; k < nz; k++) hz[k] = ((k > 0 ? dzs[k - 1] : 0) + (k < nz - 1 ? dzs[k] : 0)) / 2;
    dx = width / nx; dy = depth / ny;
    xp.resize(nx); xm.resize(nx); yp.resize(ny); ym.resize(ny);
    for (int x = 0; x < nx; x++) { xp[x] = (x + 1) % nx; xm[x] = (x + nx - 1) % nx; }
    for (int y = 0; y < ny; y++) { yp[y] = (y + 1) % ny * nx; ym[y] = (y + ny - 1) % ny * nx; }
  }
};

// ---- microphysics.ts:5, 37-39, 48-77 --------------------------------------------------------------------------
double qsatP(double t, double p) { double es = 611.2 * std::exp(17.67 * t / (t + 243.5)); return clamp01(EPS * es / std::max(1000.0, p - es), 0, .045); }
constexpr double N0_MP = 8e6;
double fallSpeedRain(double rain, double rho, double rhoGround) {
  return rain > 0 ? 36.34 * std::pow(.001 * rho * rain, .1364) * std::sqrt(rhoGround / rho) * 1 : 0;
}
double rainProcesses(float &q, float &cloud, float &rain, double tk, double p, double rho, double dt) {
  double qc = cloud, qr = rain;
  double rate = .001 * std::max(0.0, qc - .001) + (qr > 0 ? 2.2 * qc * std::pow(qr, .875) * 1 : 0);
  double production = std::min(qc, rate * dt);
  cloud = float(cloud - production); rain = float(rain + production);
  double qs = qsatP(tk - 273.15, p);
  if (rain <= 0 || q >= qs) return 0;
  double rq = .001 * rho * rain, ventilation = 1.6 + 124.9 * std::pow(rq, .2046);
  double er = (1 - q / qs) * ventilation * std::pow(rq, .525) / (.001 * rho * (5.4e5 + 2.55e6 / (p / 100 * qs)));
  double evaporation = fmin3(rain, er * dt, (qs - q) / (1 + LV * LV * qs / (CP * RV * tk * tk)));
  rain = float(rain - evaporation); q = float(q + evaporation);
  return evaporation;
}
double saturationAdjust(float &theta, float &q, float &cloud, double exner, double p) {
  double total = 0;
  for (int k = 0; k < 2; k++) {
    double tk = theta * exner, qs = qsatP(tk - 273.15, p);
    double d = (q - qs) / (1 + LV * LV * qs / (CP * RV * tk * tk));
    if (d < 0) d = std::max(d, -double(cloud));
    if (d == 0) break;
    q = float(q - d); cloud = float(cloud + d); theta = float(theta + LV / (CP * exner) * d); total += d;
  }
  return total;
}

// ---- ice.ts:91-303 -----------------------------------------------------------------------------------------------
constexpr double LF = 3.34e5, LS = LV + LF, T0 = 273.15, T_HOMOGENEOUS = 235.15, CW = 4187;
constexpr double NUCLEATION_START = -5, MAX_NUCLEATED = 250e3, CRYSTAL_MASS = 1e-12, MAX_CRYSTAL_MASS = 9.4e-10, RHO_I = 900;
const double DROPLET_MASS = PI / 6 * 1000 * std::pow(20e-6, 3);
constexpr double SPLINTERS = 3.5e8;
double splintering(double tc) { return tc >= -3 || tc <= -8 ? 0 : tc > -5 ? (-3 - tc) / 2 : (tc + 8) / 3; }
double fragmentsPerKg(double lambda) { return 2.5e13 * 7 / (lambda * 1000 * PI / 6); }
constexpr double KA = 2.43e-2, PSI = 2.26e-5, NU = 1.51e-5, SC = .6;
constexpr double N0S = 3e6, RHO_S = 100;
const double C_S = 152.93 * std::pow(.01, .75);
constexpr double D_S = .25, GAMMA_3D = 2.5493, GAMMA_4D = 8.2850, GAMMA_5D2 = 1.4569, ICE_AUTOCONVERSION_THRESHOLD = .001;

double qsatIce(double t, double p) {
  if (t >= 0) return qsatP(t, p);
  double es = 611.21 * std::exp(22.587 * t / (t + 273.86));
  return clamp01(EPS * es / std::max(1000.0, p - es), 0, .045);
}
double cooperCrystals(double tc) { return tc >= NUCLEATION_START ? 0 : std::min(MAX_NUCLEATED, 5 * std::exp(.304 * -tc)); }
double snowSlope(double snow, double rho) { return std::pow(PI * RHO_S * N0S / (rho * snow), .25); }
double snowFallSpeed(double snow, double rho, double rhoGround) {
  return snow > 1e-9 ? C_S * GAMMA_4D / 6 * std::pow(snowSlope(snow, rho), -D_S) * std::sqrt(rhoGround / rho) : 0;
}
double saturationAdjustMixed(float &theta, float &q, float &cloud, float &ice, double exner, double p, float &number) {
  double heat = LF / (CP * exner);
  if (theta * exner >= T_HOMOGENEOUS) {
    if (ice > 0 && theta * exner > T0) {
      double melt = std::min(double(ice), (theta * exner - T0) * CP / LF);
      number = float(number * (1 - melt / ice));
      ice = float(ice - melt); cloud = float(cloud + melt); theta = float(theta - heat * melt);
    }
    return saturationAdjust(theta, q, cloud, exner, p);
  }
  if (cloud > 0) { number = float(number + cloud / DROPLET_MASS); ice = float(ice + double(cloud)); theta = float(theta + heat * cloud); cloud = 0; }
  double total = 0;
  for (int k = 0; k < 2; k++) {
    double tk = theta * exner, qs = qsatIce(tk - T0, p);
    double d = (q - qs) / (1 + LS * LS * qs / (CP * RV * tk * tk));
    if (d < 0) d = std::max(d, -double(ice));
    if (d == 0) break;
#if ORC_MUTANT == 8
    q = float(q - d); ice = float(ice + d); theta = float(theta + LV / (CP * exner) * d); total += d;
#else
    q = float(q - d); ice = float(ice + d); theta = float(theta + LS / (CP * exner) * d); total += d;
#endif
  }
  return total;
}
void freezeRain(float &theta, float &rain, float &graupel, double exner) {
  double r = rain;
  if (r <= 0 || theta * exner >= T_HOMOGENEOUS) return;
  rain = 0; graupel = float(graupel + r); theta = float(theta + LF / (CP * exner) * r);
}
double crystalSize(double qi, double n, double rho) { return std::cbrt(6 * rho * qi / (PI * RHO_I * n)); }
const double KOENIG_A1[31] = {.7939e-7, .7841e-6, .3369e-5, .4336e-5, .5285e-5, .3728e-5, .1852e-5, .2991e-6, .4248e-6, .7434e-6, .1812e-5, .4394e-5, .9145e-5, .1725e-4, .3348e-4,
  .1725e-4, .9175e-5, .4412e-5, .2252e-5, .9115e-6, .4876e-6, .3473e-6, .4758e-6, .6306e-6, .8573e-6, .7868e-6, .7192e-6, .6513e-6, .5956e-6, .5333e-6, .4834e-6};
const double KOENIG_A2[31] = {.4006, .4831, .5320, .5307, .5319, .5249, .4888, .3894, .4047, .4318, .4771, .5183, .5463, .5651, .5813,
  .5655, .5478, .5203, .4906, .4447, .4126, .3960, .4149, .4320, .4506, .4483, .4460, .4433, .4413, .4382, .4361};
double koenigGrowth(double tc, double m) {
  double x = clamp01(-tc, 1, 31) - 1; int k = std::min(29, int(std::floor(x))); double f = x - k, g = m * 1000;
  double a1 = KOENIG_A1[k], a2 = KOENIG_A1[k + 1];
#if ORC_MUTANT == 1
  if (k == 14) a1 /= 2;
  if (k + 1 == 14) a2 /= 2;
#endif
  return 1e-3 * ((1 - f) * a1 * std::pow(g, KOENIG_A2[k]) + f * a2 * std::pow(g, KOENIG_A2[k + 1]));
}
double iceProcesses(float &theta, float &q, float &ice, float &snow, float &number, double exner, double p, double rho, double dt) {
  double tk = theta * exner, tc = tk - T0;
  if (tc >= 0 || tk < T_HOMOGENEOUS) return 0;
  double qsi = qsatIce(tc, p), limit = (q - qsi) / (1 + LS * LS * qsi / (CP * RV * tk * tk)), heat = LS / (CP * exner), lacking = cooperCrystals(tc) / rho - number;
  double change = 0;
  if (limit > 0 && lacking > 0) { change = std::min(limit, lacking * CRYSTAL_MASS); number = float(number + change / CRYSTAL_MASS); }
  double qi = ice + change, n = number;
  if (qi > 1e-12 && n > 0) {
    if (q > qsi) change += std::min(n * koenigGrowth(tc, qi / n) * (q - qsi) / (qsatP(tc, p) - qsi) * dt, limit - change);
    else {
      double a = LS * LS / (KA * RV * tk * tk), b = 1 / (rho * qsi * PSI);
      double lost = std::min(qi, n * 2 * PI * crystalSize(qi, n * rho, rho) * (1 - q / qsi) / (a + b) * dt);
#if ORC_MUTANT == 4
      change += lost;
#else
      change -= lost;
#endif
      number = float(number * (1 - lost / qi));
    }
  }
  ice = float(ice + change); q = float(q - change); theta = float(theta + heat * change);
  double excess = ice - number * MAX_CRYSTAL_MASS;
  if (excess > 0) { ice = float(ice - excess); snow = float(snow + excess); }
  return change < 0 ? -heat * change : 0;
}
void settleCrystals(float &ice, float &number) {
  double qi = ice;
  number = float(qi > 1e-12 ? std::min(qi / CRYSTAL_MASS, std::max(qi / MAX_CRYSTAL_MASS, double(number))) : 0);
}
double snowProcesses(float &theta, float &q, float &cloud, float &ice, float &snow, float &rain,
                     double exner, double p, double rho, double rhoGround, double dt) {
  double tk = theta * exner, tc = tk - T0, qs = snow;
  if (tc < 0 && ice > 0) {
    double aut = 1e-3 * std::exp(.025 * tc) * std::max(0.0, ice - ICE_AUTOCONVERSION_THRESHOLD);
    double moved = std::min(double(ice), aut * dt); ice = float(ice - moved); snow = float(snow + moved);
  }
  if (qs <= 1e-9) return 0;
  double lambda = snowSlope(qs, rho), fall = std::sqrt(rhoGround / rho);
  double sweep = PI * N0S * C_S * GAMMA_3D / (4 * std::pow(lambda, 3 + D_S)) * fall;
  double vent = N0S * (.78 / (lambda * lambda) + .31 * std::pow(SC, 1.0 / 3) * GAMMA_5D2 * std::sqrt(C_S) * std::pow(fall, .5) / std::sqrt(NU) * std::pow(lambda, -(D_S + 5) / 2));
  double cooling = 0;
  if (tc < 0) {
    double saci = std::min(double(ice), sweep * std::exp(.025 * tc) * ice * dt); ice = float(ice - saci); snow = float(snow + saci);
    double sacw = std::min(double(cloud), sweep * cloud * dt); cloud = float(cloud - sacw); snow = float(snow + sacw);
    theta = float(theta + LF / (CP * exner) * sacw);
    double qsi = qsatIce(tc, p), a = LS * LS / (KA * RV * tk * tk), b = 1 / (rho * qsi * PSI);
    double rate = 2 * PI * (q / qsi - 1) * vent / (rho * (a + b));
    double dep = rate > 0 ? std::min(rate * dt, (q - qsi) / (1 + LS * LS * qsi / (CP * RV * tk * tk))) : -std::min(double(snow), -rate * dt);
    snow = float(snow + dep); q = float(q - dep); theta = float(theta + LS / (CP * exner) * dep);
    if (dep < 0) cooling -= LS / (CP * exner) * dep;
  } else {
    double sacw = std::min(double(cloud), sweep * cloud * dt); cloud = float(cloud - sacw); rain = float(rain + sacw);
    double rate = 2 * PI * KA * tc * vent / (rho * LF) + CW * tc / LF * sweep * cloud;
    double melt = fmin3(snow, rate * dt, CP * tc / LF);
    snow = float(snow - melt); rain = float(rain + melt); theta = float(theta - LF / (CP * exner) * melt); cooling += LF / (CP * exner) * melt;
  }
  return cooling;
}

// ---- graupel.ts:6-72 -----------------------------------------------------------------------------------------------
constexpr double N0G = 4e4, RHO_G = 400, CDG = .6, N0R = 8e6, RHO_W = 1000, BIGG_B = 100, BIGG_A = .66;
constexpr double GAMMA_45 = 11.6317, GAMMA_35 = 3.3234, GAMMA_275 = 1.6084, SNOW_TO_GRAUPEL_THRESHOLD = 6e-4;
double graupelSlope(double g, double rho) { return std::pow(PI * RHO_G * N0G / (rho * g), .25); }
double graupelFallCoefficient(double rho) { return std::sqrt(4 * G * RHO_G / (3 * CDG * rho)); }
double graupelFallSpeed(double g, double rho) { return g > 1e-9 ? graupelFallCoefficient(rho) * GAMMA_45 / 6 * std::pow(graupelSlope(g, rho), -.5) : 0; }
double graupelProcesses(float &theta, float &q, float &cloud, float &ice, float &snow, float &rain, float &graupel,
                        double exner, double p, double rho, double dt) {
  double tk = theta * exner, tc = tk - T0, heat = 1 / (CP * exner);
  double cooling = 0;
  if (tc < 0) {
    if (rain > 1e-9) {
      double lambdaR = std::pow(PI * RHO_W * N0R / (rho * rain), .25);
      double rate = 20 * PI * PI * BIGG_B * N0R * RHO_W / rho * (std::exp(-BIGG_A * tc) - 1) / std::pow(lambdaR, 7);
      double frozen = std::min(double(rain), rate * dt); rain = float(rain - frozen); graupel = float(graupel + frozen); theta = float(theta + LF * heat * frozen);
    }
    double aut = std::min(double(snow), 1e-3 * std::exp(.09 * tc) * std::max(0.0, snow - SNOW_TO_GRAUPEL_THRESHOLD) * dt);
    snow = float(snow - aut); graupel = float(graupel + aut);
  }
  double qg = graupel;
  if (qg <= 1e-9) return 0;
  double lambda = graupelSlope(qg, rho), a = graupelFallCoefficient(rho);
  double sweep = PI * N0G * a * GAMMA_35 / (4 * std::pow(lambda, 3.5));
  double vent = N0G * (.78 / (lambda * lambda) + .31 * std::pow(SC, 1.0 / 3) * GAMMA_275 * std::sqrt(a / NU) * std::pow(lambda, -2.75));
  if (tc < 0) {
    double gacw = std::min(double(cloud), sweep * cloud * dt), gacr = std::min(double(rain), sweep * rain * dt);
    double gaci = std::min(double(ice), sweep * .1 * ice * dt), gacs = std::min(double(snow), sweep * std::exp(.09 * tc) * snow * dt);
    cloud = float(cloud - gacw); rain = float(rain - gacr); ice = float(ice - gaci); snow = float(snow - gacs); graupel = float(graupel + (gacw + gacr + gaci + gacs));
    theta = float(theta + LF * heat * (gacw + gacr));
    double qsi = qsatIce(tc, p);
    if (q < qsi) {
      double aa = LS * LS / (KA * RV * tk * tk), b = 1 / (rho * qsi * PSI);
      double sub = fmin3(graupel, 2 * PI * (1 - q / qsi) * vent / (rho * (aa + b)) * dt, (qsi - q) / (1 + LS * LS * qsi / (CP * RV * tk * tk)));
      graupel = float(graupel - sub); q = float(q + sub); theta = float(theta - LS * heat * sub); cooling += LS * heat * sub;
    }
  } else {
    double gacw = std::min(double(cloud), sweep * cloud * dt); cloud = float(cloud - gacw); rain = float(rain + gacw);
    double rate = 2 * PI * KA * tc * vent / (rho * LF) + CW * tc / LF * sweep * (double(cloud) + rain);
    double melt = fmin3(graupel, rate * dt, CP * tc / LF);
    graupel = float(graupel - melt); rain = float(rain + melt); theta = float(theta - LF * heat * melt); cooling += LF * heat * melt;
  }
  return cooling;
}

// ---- hail.ts:84-175 ------------------------------------------------------------------------------------------------
constexpr double N0H = 4e4, RHO_H = 900, CDH = .6, CI = 2106, HAIL_CONVERSION_TIME = 60, HAIL_SAMPLE_VOLUME = 10;
double slopeOf(double q, double rho, double n0, double rhoX) { return std::pow(PI * rhoX * n0 / (rho * q), .25); }
double fallCoefficientOf(double rho, double rhoX) { return std::sqrt(4 * G * rhoX / (3 * CDH * rho)); }
double sweepOf(double lambda, double a, double n0) { return PI * n0 * a * GAMMA_35 / (4 * std::pow(lambda, 3.5)); }
double ventilationOf(double lambda, double a, double n0) { return n0 * (.78 / (lambda * lambda) + .31 * std::pow(SC, 1.0 / 3) * GAMMA_275 * std::sqrt(a / NU) * std::pow(lambda, -2.75)); }
double hailFallSpeed(double h, double rho) { return h > 1e-9 ? fallCoefficientOf(rho, RHO_H) * GAMMA_45 / 6 * std::pow(slopeOf(h, rho, N0H, RHO_H), -.5) : 0; }
double hailMaxDiameter(double h, double rho) {
  if (!(h > 1e-9)) return 0;
  double lambda = slopeOf(h, rho, N0H, RHO_H);
  return std::max(0.0, std::log(N0H * HAIL_SAMPLE_VOLUME / lambda) / lambda);
}
double wetGrowth(double lambda, double a, double n0, double tc, double qv, double p, double rho, double iceRate) {
  double heat = KA * -tc + LV * PSI * rho * (qsatP(0, p) - qv);
  return std::max(0.0, 2 * PI * ventilationOf(lambda, a, n0) * heat / (rho * (LF + CW * tc))) + iceRate * (1 - CI * tc / (LF + CW * tc));
}
bool wetGrowing(double q, double n0, double rhoX, double tc, double qv, double p, double rho, double cloud, double rain, double ice, double snow) {
  if (tc >= 0 || q <= 1e-9) return false;
  double lambda = slopeOf(q, rho, n0, rhoX), a = fallCoefficientOf(rho, rhoX), k = sweepOf(lambda, a, n0);
  return k * (cloud + rain + .1 * ice + std::exp(.09 * tc) * snow) > wetGrowth(lambda, a, n0, tc, qv, p, rho, k * (ice + snow));
}
double hailProcesses(float &theta, float &q, float &cloud, float &ice, float &snow, float &rain, float &graupel, float &hail,
                     double exner, double p, double rho, double dt) {
  double tk = theta * exner, tc = tk - T0, heat = 1 / (CP * exner);
  double cooling = 0;
  if (wetGrowing(graupel, N0G, RHO_G, tc, q, p, rho, cloud, rain, ice, snow)) {
    double turned = graupel * (1 - std::exp(-dt / HAIL_CONVERSION_TIME)); graupel = float(graupel - turned); hail = float(hail + turned);
  }
  double qh = hail;
  if (qh <= 1e-9) return 0;
  double lambda = slopeOf(qh, rho, N0H, RHO_H), a = fallCoefficientOf(rho, RHO_H), k = sweepOf(lambda, a, N0H), vent = ventilationOf(lambda, a, N0H);
  if (tc < 0) {
    double hacw = std::min(double(cloud), k * cloud * dt), hacr = std::min(double(rain), k * rain * dt), liquid = hacw + hacr;
    double haci = std::min(double(ice), k * .1 * ice * dt), hacs = std::min(double(snow), k * std::exp(.09 * tc) * snow * dt), frozen = liquid;
    double wet = wetGrowth(lambda, a, N0H, tc, q, p, rho, k * (double(ice) + snow)) * dt;
    if (liquid + haci + hacs > wet) {
      haci = std::min(double(ice), k * ice * dt); hacs = std::min(double(snow), k * snow * dt);
#if ORC_MUTANT == 3
      frozen = liquid;
#else
      frozen = std::min(liquid, std::max(0.0, wet - haci - hacs));
#endif
    }
    cloud = float(cloud - hacw); rain = float(rain + (liquid - frozen - hacr)); ice = float(ice - haci); snow = float(snow - hacs); hail = float(hail + (frozen + haci + hacs));
    theta = float(theta + LF * heat * frozen);
    double qsi = qsatIce(tc, p);
    if (q < qsi) {
      double aa = LS * LS / (KA * RV * tk * tk), b = 1 / (rho * qsi * PSI);
      double sub = fmin3(hail, 2 * PI * (1 - q / qsi) * vent / (rho * (aa + b)) * dt, (qsi - q) / (1 + LS * LS * qsi / (CP * RV * tk * tk)));
      hail = float(hail - sub); q = float(q + sub); theta = float(theta - LS * heat * sub); cooling += LS * heat * sub;
    }
  } else {
    double hacw = std::min(double(cloud), k * cloud * dt); cloud = float(cloud - hacw); rain = float(rain + hacw);
    double rate = 2 * PI * KA * tc * vent / (rho * LF) + CW * tc / LF * k * (double(cloud) + rain);
    double melt = fmin3(hail, rate * dt, CP * tc / LF);
    hail = float(hail - melt); rain = float(rain + melt); theta = float(theta - LF * heat * melt); cooling += LF * heat * melt;
  }
  return cooling;
}

// ---- takahashi.ts:5-54, charge.ts:98-234, electric.ts:6 -------------------------------------------------------------
const double TAKAHASHI_LWC[30] = {0.01, 0.02, 0.03, 0.04, 0.05, 0.06, 0.07, 0.08, 0.09, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 20, 30};
#include "takahashi_table.inc"
double takahashiCharge(double tc, double lwc) {
  if (tc >= 0 || tc <= -40 || lwc < .01) return 0;
  double w = std::min(30.0, lwc);
  int row = 0;
  while (row < 30 - 2 && w >= TAKAHASHI_LWC[row + 1]) row++;
  double t = std::min(30.0, -std::max(-30.0, tc)); int col = std::min(29, int(std::floor(t))); double f = t - col;
  auto at = [&](const double *r) { return r[col] + f * (r[col + 1] - r[col]); };
  double a = at(TAKAHASHI_TABLE[row]), b = at(TAKAHASHI_TABLE[row + 1]), g = std::min(1.0, (w - TAKAHASHI_LWC[row]) / (TAKAHASHI_LWC[row + 1] - TAKAHASHI_LWC[row]));
  double cold = tc < -30 ? 1 - std::pow((tc + 30) / 10, 2) : 1;
#if ORC_MUTANT == 7
  return -cold * (a + g * (b - a));
#else
  return cold * (a + g * (b - a));
#endif
}
double takahashiScaled(double table, double d, double dv) {
  double q = table * std::min(5 * std::pow(d / 100e-6, 2) * std::fabs(dv) / 8, 10.0) * 1e-15;
  return std::max(-500e-15, std::min(500e-15, q));
}
constexpr double EPSILON0 = 8.8541878128e-12;
constexpr double MAX_CRYSTALS = 150e3, CRYSTAL_SIZE = 100e-6, CRYSTAL_MIN = 20e-6, CRYSTAL_MAX = 300e-6, TAK_RATE_MAX = 3e-9;
constexpr double COLLECTION_EFFICIENCY = .6, CHARGING_SPEED = 6, DROPLET_REBOUND = .01, DROPLET_SIZE = 20e-6, CONTACT_COS = .5;
double crystalNumber(double tc, double qi, double rho) {
  double cooper = std::min(MAX_CRYSTALS, 5 * std::exp(.304 * std::max(0.0, -tc))), fromMass = rho * qi / (PI / 6 * 900 * std::pow(CRYSTAL_SIZE, 3));
  return std::min(MAX_CRYSTALS, std::max(cooper, fromMass));
}
double crystalDiameter(double qi, double rho, double tc) {
  return qi > 0 ? std::min(CRYSTAL_MAX, std::max(CRYSTAL_MIN, std::cbrt(6 * rho * qi / (PI * 900 * crystalNumber(tc, qi, rho))))) : 0;
}
struct Moments { double n, d1, d2, fall, d2v; };
Moments exponentialMoments(double n0, double density, double q, double rho) {
  double lambda = std::pow(PI * density * n0 / (rho * q), .25), a = std::sqrt(4 * G * density / (3 * CDG * rho));
  return {n0 / lambda, n0 / (lambda * lambda), 2 * n0 / std::pow(lambda, 3), a * GAMMA_45 / 6 / std::sqrt(lambda), a * n0 * GAMMA_35 / std::pow(lambda, 3.5)};
}
double rimerRate(const Moments &m, double rebound, double ni, double di, double ns, double ds, double vs, double table) {
  double fall = m.fall, k = PI / 4 * rebound;
  auto swept = [&](double d) { return m.d2 + 2 * d * m.d1 + d * d * m.n; };
  auto pair = [](double r) { return std::max(-TAK_RATE_MAX, std::min(TAK_RATE_MAX, r)); };
  double rate = 0;
  if (ni > 0) rate += pair(ni * k * swept(di) * fall * takahashiScaled(table, di, fall));
  if (ns > 0) { double v = std::fabs(fall - vs); rate += pair(ns * k * swept(ds) * v * takahashiScaled(table, ds, v)); }
  return rate;
}
// chargingRate with the Takahashi scheme (charge.ts:138-158) and the model's own crystals (`crystals`, per m3); without
// them (`haveCrystals` false, the reference's `crystals` undefined) the number and size estimated from the mass.
double chargingRate(double tc, double rho, double cloud, double ice, double snow, double graupel, double hail, double snowFall, bool wetG, bool wetH, double crystals, bool haveCrystals = true) {
  if (tc >= 0 || tc < -40 || cloud <= 0 || ice + snow <= 0 || graupel + hail <= 0) return 0;
  double rebound = .5 * (1 - .1 * std::exp(.1 * tc)), water = rho * cloud * 1000;
  double ni = ice <= 0 ? 0 : haveCrystals ? crystals : crystalNumber(tc, ice, rho);
  auto size = [](double m) { return .1871 * std::pow(m, .3429); };
  double di = ice <= 0 ? 0 : haveCrystals && crystals != 0 ? std::min(CRYSTAL_MAX, std::max(CRYSTAL_MIN, size(rho * ice / crystals))) : crystalDiameter(ice, rho, tc);
  double ss = snow > 0 ? std::pow(PI * RHO_S * N0S / (rho * snow), .25) : 0, ns = snow > 0 ? N0S / ss : 0, ds = snow > 0 ? 1.8171 / ss : 0;
  double rate = 0, table = takahashiCharge(tc, water);
  if (graupel > 0 && !wetG) rate += rimerRate(exponentialMoments(N0G, RHO_G, graupel, rho), rebound, ni, di, ns, ds, snowFall, table);
  if (hail > 0 && !wetH) rate += rimerRate(exponentialMoments(N0H, RHO_H, hail, rho), rebound, ni, di, ns, ds, snowFall, table);
  return rate;
}
double inductiveRate(double tc, double rho, double cloud, double graupel, double hail, double ez, double rimerCharge, bool wetG, bool wetH) {
  if (tc >= 0 || cloud <= 0) return 0;
  double d = DROPLET_SIZE, droplets = rho * cloud / (PI / 6 * 1000 * std::pow(d, 3)), k = COLLECTION_EFFICIENCY * DROPLET_REBOUND * droplets;
  double polarize = 0, mass = 0, massFall = 0;
  const double n0s[2] = {N0G, N0H}, dens[2] = {RHO_G, RHO_H}, qs[2] = {graupel, hail}; const bool wets[2] = {wetG, wetH};
  for (int j = 0; j < 2; j++) {
    if (qs[j] <= 1e-9 || wets[j]) continue;
    Moments m = exponentialMoments(n0s[j], dens[j], qs[j], rho);
    polarize += std::pow(PI, 4) / 8 * EPSILON0 * ez * CONTACT_COS * d * d * m.d2v;
    mass += qs[j]; massFall += qs[j] * m.fall;
  }
  if (mass == 0) return 0;
  return k * (polarize - std::pow(PI, 3) / 24 * d * d * (massFall / mass) * rimerCharge);
}
void transferCharge(double *charge, const double *before, const double *after) {
  double pool = 0, lost = 0, gained = 0;
  for (int k = 0; k < 4; k++) {
    if (after[k] < before[k]) { double share = before[k] > 0 ? (before[k] - after[k]) / before[k] : 1, out = charge[k] * share; charge[k] -= out; pool += out; lost += before[k] - after[k]; }
    else gained += after[k] - before[k];
    if (k != 3 && after[k] <= 0 && charge[k] != 0) { pool += charge[k]; charge[k] = 0; }
  }
  if (pool == 0) return;
  double moved = gained > 0 ? std::min(1.0, lost > 0 ? gained / lost : 1) : 0;
  if (moved > 0) for (int k = 0; k < 4; k++) if (after[k] > before[k]) charge[k] += pool * moved * (after[k] - before[k]) / gained;
#if ORC_MUTANT != 2
  charge[3] += pool * (1 - moved);
#endif
}

// ---- turbulence.ts:87-185 ------------------------------------------------------------------------------------------
// Turbulence.viscosity (turbulence.ts:87-137) as the baseline configures it (model.ts:778-792): anisotropic (or, with
// `anisotropic` false, the class's default: kh = km), the cube vertical length, explicit vertical mixing, the theta_e
// stability of cloudy nodes (condensate above 1e-5; a null condensate: dry air everywhere). km, kh: n float64.
void viscosityOf(const Grid &g, const float *u, const float *v, const float *w, const float *theta, const double *thetaEnv, double cs, double dt,
                 const float *q, const float *condensate, const double *exner, double *km, double *kh, bool anisotropic = true) {
  int nx = g.nx, ny = g.ny, nz = g.nz, layer = g.layer;
  double lenH2 = std::pow(cs * std::sqrt(g.dx * g.dy), 2), capH = .05 * std::pow(std::min(g.dx, g.dy), 2) / dt;
  const double inf = std::numeric_limits<double>::infinity();
  for (int z = 0, i = 0; z < nz; z++) {
    double local = z == 0 ? g.dzs[0] : z == nz - 1 ? g.dzs[nz - 2] : g.hz[z], len2 = std::pow(cs * std::cbrt(g.dx * g.dy * local), 2);
    double cap = .05 * std::pow(std::min(std::min(g.dx, g.dy), std::min(z > 0 ? g.dzs[z - 1] : inf, z < nz - 1 ? g.dzs[z] : inf)), 2) / dt;
    int up = z < nz - 1 ? layer : 0, below = z > 0 ? -layer : 0; double span = g.zs[z + (up ? 1 : 0)] - g.zs[z - (below ? 1 : 0)];
    for (int y = 0; y < ny; y++) {
      int row = y * nx + z * layer;
      for (int x = 0; x < nx; x++, i++) {
        int e = g.xp[x] + row, wst = g.xm[x] + row, n = x + g.yp[y] + z * layer, s = x + g.ym[y] + z * layer; double hx = 2 * g.dx, hy = 2 * g.dy; int down = below;
        double dudx = (double(u[e]) - u[wst]) / hx, dvdy = (double(v[n]) - v[s]) / hy, dwdz = (double(w[i + up]) - w[i + down]) / span;
        double dudy = (double(u[n]) - u[s]) / hy, dvdx = (double(v[e]) - v[wst]) / hx;
        double dudz = (double(u[i + up]) - u[i + down]) / span, dwdx = (double(w[e]) - w[wst]) / hx;
        double dvdz = (double(v[i + up]) - v[i + down]) / span, dwdy = (double(w[n]) - w[s]) / hy;
        double s2 = 2 * (dudx * dudx + dvdy * dvdy + dwdz * dwdz) + std::pow(dudy + dvdx, 2) + std::pow(dudz + dwdx, 2) + std::pow(dvdz + dwdy, 2);
        double n2 = G / thetaEnv[z] * (double(theta[i + up]) - theta[i + down]) / span;
        if (condensate && condensate[i] > 1e-5) {
          auto te = [&](int k, int level) { return theta[k] * std::exp(LV * q[k] / (CP * theta[k] * exner[level])); };
          n2 = G / te(i, z) * (te(i + up, z + (up ? 1 : 0)) - te(i + down, z - (down ? 1 : 0))) / span;
        }
#if ORC_MUTANT == 5
        double rate = std::sqrt(std::max(0.0, s2 + n2 / (1.0 / 3)));
#else
        double rate = std::sqrt(std::max(0.0, s2 - n2 / (1.0 / 3)));
#endif
        km[i] = std::min(cap, len2 * rate); kh[i] = anisotropic ? std::min(capH, lenH2 * rate) : km[i];
      }
    }
  }
}
// Turbulence.mix (turbulence.ts:143-185), explicit, flat ground: the departure of `a` from `base` (per level; null: none)
// diffused with K times `factor`; t: n float32 scratch.
void mixOf(const Grid &g, const double *rho, const double *km, const double *kh, float *a, const double *base, double factor, double dt, float *t) {
  int nx = g.nx, ny = g.ny, nz = g.nz, layer = g.layer; const F64 &h = g.hz;
  std::fill(t, t + g.n, 0.0f);
  auto dev = [&](int i, int z) { return a[i] - (base ? base[z] : 0); };
  for (int z = 0, i = 0; z < nz; z++) {
    double rUp = z < nz - 1 ? (rho[z] + rho[z + 1]) / 2 : 0;
    for (int y = 0; y < ny; y++) {
      int row = y * nx + z * layer;
      for (int x = 0; x < nx; x++, i++) {
        double here = dev(i, z); int e = g.xp[x] + row, n = x + g.yp[y] + z * layer;
        { double fx = factor * (kh[i] + kh[e]) / 2 * (dev(e, z) - here) / g.dx; t[i] = float(t[i] + fx / g.dx); t[e] = float(t[e] - fx / g.dx); }
        { double fy = factor * (kh[i] + kh[n]) / 2 * (dev(n, z) - here) / g.dy; t[i] = float(t[i] + fy / g.dy); t[n] = float(t[n] - fy / g.dy); }
        if (z < nz - 1) {
          double fz = factor * rUp * (km[i] + km[i + layer]) / 2 * (dev(i + layer, z + 1) - here) / g.dzs[z];
          t[i] = float(t[i] + fz / (rho[z] * h[z])); t[i + layer] = float(t[i + layer] - fz / (rho[z + 1] * h[z + 1]));
        }
      }
    }
  }
  for (int i = 0; i < g.n; i++) a[i] = float(a[i] + dt * t[i]);
}

// ---- spectral.ts:8-151 ---------------------------------------------------------------------------------------------
struct FFT {
  int n; F64 twRe, twIm; std::vector<int> factors; F64 outRe, outIm, scrRe, scrIm;
  explicit FFT(int len) : n(len) {
    twRe.resize(n); twIm.resize(n);
    for (int k = 0; k < n; k++) { twRe[k] = std::cos(2 * PI * k / n); twIm[k] = -std::sin(2 * PI * k / n); }
    int rest = n, p = 4, largest = 1;
    while (rest > 1) {
      while (rest % p) { p = p == 4 ? 2 : p == 2 ? 3 : p + 2; if (p * p > rest) p = rest; }
      rest /= p; factors.push_back(p); factors.push_back(rest); largest = std::max(largest, p);
    }
    outRe.assign(n, 0); outIm.assign(n, 0); scrRe.assign(largest, 0); scrIm.assign(largest, 0);
  }
  void transform(double *re, double *im, int sign) {
    if (n == 1) return;
    if (sign > 0) for (int k = 0; k < n; k++) im[k] = -im[k];
    work(0, re, im, 0, 1, 0);
    for (int k = 0; k < n; k++) { re[k] = outRe[k]; im[k] = sign > 0 ? -outIm[k] : outIm[k]; }
  }
  void work(int out, const double *re, const double *im, int src, int stride, int f) {
    int p = factors[f], m = factors[f + 1];
    double *oRe = outRe.data(), *oIm = outIm.data();
    if (m == 1) for (int k = 0; k < p; k++) { oRe[out + k] = re[src + k * stride]; oIm[out + k] = im[src + k * stride]; }
    else for (int k = 0; k < p; k++) work(out + k * m, re, im, src + k * stride, stride * p, f + 2);
    const double *tr = twRe.data(), *ti = twIm.data();
    if (p == 2) {
      for (int u = 0, t = 0; u < m; u++, t += stride) {
        int a = out + u, b = a + m; double c = tr[t], s = ti[t], xr = oRe[b] * c - oIm[b] * s, xi = oRe[b] * s + oIm[b] * c;
        oRe[b] = oRe[a] - xr; oIm[b] = oIm[a] - xi; oRe[a] += xr; oIm[a] += xi;
      }
      return;
    }
    if (p == 4) {
      for (int u = 0, t = 0; u < m; u++, t += stride) {
        int a = out + u, b = a + m, c = b + m, d = c + m, t2 = 2 * t, t3 = 3 * t;
        double r0 = oRe[b] * tr[t] - oIm[b] * ti[t], i0 = oRe[b] * ti[t] + oIm[b] * tr[t];
        double r1 = oRe[c] * tr[t2] - oIm[c] * ti[t2], i1 = oRe[c] * ti[t2] + oIm[c] * tr[t2];
        double r2 = oRe[d] * tr[t3] - oIm[d] * ti[t3], i2 = oRe[d] * ti[t3] + oIm[d] * tr[t3];
        double r5 = oRe[a] - r1, i5 = oIm[a] - i1, ra = oRe[a] + r1, ia = oIm[a] + i1, r3 = r0 + r2, i3 = i0 + i2, r4 = r0 - r2, i4 = i0 - i2;
        oRe[c] = ra - r3; oIm[c] = ia - i3; oRe[a] = ra + r3; oIm[a] = ia + i3;
        oRe[b] = r5 + i4; oIm[b] = i5 - r4; oRe[d] = r5 - i4; oIm[d] = i5 + r4;
      }
      return;
    }
    double *sRe = scrRe.data(), *sIm = scrIm.data();
    for (int u = 0; u < m; u++) {
      for (int q = 0; q < p; q++) { sRe[q] = oRe[out + u + q * m]; sIm[q] = oIm[out + u + q * m]; }
      for (int q1 = 0, k = u; q1 < p; q1++, k += m) {
        double a = sRe[0], b = sIm[0]; int t = 0; int step = stride * k;
        for (int q = 1; q < p; q++) { t += step; if (t >= n) t -= n; double c = tr[t], s = ti[t]; a += sRe[q] * c - sIm[q] * s; b += sRe[q] * s + sIm[q] * c; }
        oRe[out + k] = a; oIm[out + k] = b;
      }
    }
  }
};

struct HorizontalDFT {
  int nx, ny, mh; F64 re, im; FFT fx, fy; F64 rowRe, rowIm;
  HorizontalDFT(int nx_, int ny_, int layers) : nx(nx_), ny(ny_), mh(nx_ / 2), re(size_t(nx_) * ny_ * layers), im(size_t(nx_) * ny_ * layers), fx(nx_), fy(ny_),
    rowRe(std::max(nx_, ny_)), rowIm(std::max(nx_, ny_)) {}
  template <class Src> void forwardLayers(const Src &src, int k0, int k1) {
    int cl = nx * ny;
    for (int k = k0; k < k1; k++) {
      int base = k * cl;
      for (int j = 0; j < ny; j += 2) {
        int row = base + j * nx; bool pair = j + 1 < ny; int row2 = row + nx;
        for (int i = 0; i < nx; i++) { rowRe[i] = src(row + i); rowIm[i] = pair ? src(row2 + i) : 0; }
        fx.transform(rowRe.data(), rowIm.data(), -1);
        for (int m = 0; m <= mh; m++) {
          int c = m ? nx - m : 0; double zr = rowRe[m], zi = rowIm[m], cr = rowRe[c], ci = rowIm[c];
          re[row + m] = (zr + cr) / 2; im[row + m] = (zi - ci) / 2;
          if (pair) { re[row2 + m] = (zi + ci) / 2; im[row2 + m] = (cr - zr) / 2; }
        }
      }
      for (int m = 0; m <= mh; m++) {
        for (int j = 0; j < ny; j++) { rowRe[j] = re[base + j * nx + m]; rowIm[j] = im[base + j * nx + m]; }
        fy.transform(rowRe.data(), rowIm.data(), -1);
        for (int n = 0; n < ny; n++) { re[base + n * nx + m] = rowRe[n]; im[base + n * nx + m] = rowIm[n]; }
      }
    }
  }
  template <class Dst> void inverseLayers(Dst &&dst, int k0, int k1) {
    int cl = nx * ny;
    for (int k = k0; k < k1; k++) {
      int base = k * cl;
      for (int m = 0; m <= mh; m++) {
        for (int n = 0; n < ny; n++) { rowRe[n] = re[base + n * nx + m]; rowIm[n] = im[base + n * nx + m]; }
        fy.transform(rowRe.data(), rowIm.data(), 1);
        for (int j = 0; j < ny; j++) { re[base + j * nx + m] = rowRe[j]; im[base + j * nx + m] = rowIm[j]; }
      }
      for (int j = 0; j < ny; j += 2) {
        int row = base + j * nx; bool pair = j + 1 < ny; int row2 = row + nx;
        for (int m = 0; m < nx; m++) {
          int c = m <= mh ? m : nx - m; double s = m <= mh ? 1 : -1; bool real = c == 0 || 2 * c == nx;
          double ar = re[row + c], ai = real ? 0 : s * im[row + c], br = pair ? re[row2 + c] : 0, bi = pair && !real ? s * im[row2 + c] : 0;
          rowRe[m] = ar - bi; rowIm[m] = ai + br;
        }
        fx.transform(rowRe.data(), rowIm.data(), 1);
        for (int i = 0; i < nx; i++) { dst(row + i, rowRe[i] / cl); if (pair) dst(row2 + i, rowIm[i] / cl); }
      }
    }
  }
};

// ---- pressure.ts:18-141 --------------------------------------------------------------------------------------------
struct PressureSolver {
  const Grid &g; int cells; HorizontalDFT dft; F64 alpha, beta, cp, dRe, dIm, rho;
  PressureSolver(const Grid &grid, const F64 &density) : g(grid), cells(grid.nx * grid.ny * (grid.nz - 1)), dft(grid.nx, grid.ny, grid.nz - 1),
    alpha(grid.nx * grid.ny), beta(grid.nx * grid.ny), cp(grid.nz), dRe(grid.nz), dIm(grid.nz), rho(density) {
    auto exact = [](double x) { return x < 1e-12 ? 0 : x > 1 - 1e-12 ? 1 : x; };
    for (int n = 0; n < g.ny; n++) for (int m = 0; m < g.nx; m++) {
      double sx = exact(std::pow(std::sin(PI * m / g.nx), 2)), cx = exact(std::pow(std::cos(PI * m / g.nx), 2)), sy = exact(std::pow(std::sin(PI * n / g.ny), 2)), cy = exact(std::pow(std::cos(PI * n / g.ny), 2));
      alpha[m + g.nx * n] = 4 * (sx * cy / (g.dx * g.dx) + cx * sy / (g.dy * g.dy));
      beta[m + g.nx * n] = cx * cy;
    }
  }
  void divergence(const float *u, const float *v, const float *w, double dt, double *out) const {
    int nx = g.nx, ny = g.ny, layer = g.layer;
    for (int k = 0, c = 0; k < g.nz - 1; k++) {
      int l0 = k * layer, l1 = l0 + layer; double r0 = rho[k] / 4, r1 = rho[k + 1] / 4, dz = g.dzs[k];
      for (int j = 0; j < ny; j++) {
        int y0 = j * nx, y1 = ((j + 1) % ny) * nx;
        for (int i = 0; i < nx; i++, c++) {
          int i1 = (i + 1) % nx;
          int a = i + y0 + l0, b = i1 + y0 + l0, e = i + y1 + l0, f = i1 + y1 + l0, gg = i + y0 + l1, h = i1 + y0 + l1, s = i + y1 + l1, t = i1 + y1 + l1;
          double ddx = (r0 * (double(u[b]) + u[f] - u[a] - u[e]) + r1 * (double(u[h]) + u[t] - u[gg] - u[s])) / g.dx;
          double ddy = (r0 * (double(v[e]) + v[f] - v[a] - v[b]) + r1 * (double(v[s]) + v[t] - v[gg] - v[h])) / g.dy;
          double ddz = (r1 * (double(w[gg]) + w[h] + w[s] + w[t]) - r0 * (double(w[a]) + w[b] + w[e] + w[f])) / dz;
          out[c] = (ddx + ddy + ddz) / dt;
        }
      }
    }
  }
  void correct(const double *p, float *u, float *v, float *w, double dt) const {
    int nx = g.nx, ny = g.ny, nz = g.nz, cl = nx * ny, layer = g.layer;
    for (int k = 0; k < nz; k++) {
      int lo = std::max(0, k - 1) * cl, hi = std::min(nz - 2, k) * cl;
      double wLo = k > 0 ? g.dzs[k - 1] / g.hz[k] : 0, wHi = k < nz - 1 ? g.dzs[k] / g.hz[k] : 0;
      for (int j = 0; j < ny; j++) {
        int jm = ((j + ny - 1) % ny) * nx, j0 = j * nx;
        for (int i = 0; i < nx; i++) {
          int im = (i + nx - 1) % nx, node = i + j * nx + k * layer;
          auto P = [&](int ii, int jj, int base) { return p[base + ii + jj]; };
          double gx = (wLo * (P(i, jm, lo) + P(i, j0, lo) - P(im, jm, lo) - P(im, j0, lo)) + wHi * (P(i, jm, hi) + P(i, j0, hi) - P(im, jm, hi) - P(im, j0, hi))) / (4 * g.dx);
          double gy = (wLo * (P(im, j0, lo) + P(i, j0, lo) - P(im, jm, lo) - P(i, jm, lo)) + wHi * (P(im, j0, hi) + P(i, j0, hi) - P(im, jm, hi) - P(i, jm, hi))) / (4 * g.dy);
          u[node] = float(u[node] - dt * gx); v[node] = float(v[node] - dt * gy);
          if (k > 0 && k < nz - 1) {
            int below = (k - 1) * cl, above = k * cl;
            double gz = (P(im, jm, above) + P(i, jm, above) + P(im, j0, above) + P(i, j0, above) - P(im, jm, below) - P(i, jm, below) - P(im, j0, below) - P(i, j0, below)) / (4 * g.hz[k]);
            w[node] = float(w[node] - dt * gz);
          }
        }
      }
    }
  }
  void solve(const double *rhs, double *p) {
    int nc = g.nz - 1, cl = g.nx * g.ny;
    dft.forwardLayers([&](int i) { return rhs[i]; }, 0, nc);
    for (int n = 0; n < g.ny; n++) for (int m = 0; m <= dft.mh; m++) {
      int mode = m + g.nx * n; double a = alpha[mode], b = beta[mode];
      if ((a == 0 && b == 0) || (nc == 1 && a == 0)) { for (int k = 0; k < nc; k++) dft.re[mode + k * cl] = dft.im[mode + k * cl] = 0; continue; }
      tridiagonal(mode, a, b, nc, cl);
    }
    dft.inverseLayers([&](int i, double v) { p[i] = v; }, 0, nc);
  }
  void tridiagonal(int mode, double a, double b, int nc, int cl) {
    F64 &re = dft.re, &im = dft.im; const F64 &dzs = g.dzs, &hz = g.hz;
    bool pinned = a == 0;
    for (int k = 0; k < nc; k++) {
      double di = -a / 2 * (rho[k] * dzs[k] / (2 * hz[k]) + rho[k + 1] * dzs[k] / (2 * hz[k + 1])), lo = 0, up = 0;
      if (k > 0) { lo = -a / 2 * rho[k] * dzs[k - 1] / (2 * hz[k]); double v = b * rho[k] / (hz[k] * dzs[k]); lo += v; di -= v; }
      if (k < nc - 1) { up = -a / 2 * rho[k + 1] * dzs[k + 1] / (2 * hz[k + 1]); double v = b * rho[k + 1] / (hz[k + 1] * dzs[k]); up += v; di -= v; }
      double r = re[mode + k * cl], s = im[mode + k * cl];
      if (pinned && k == 0) { lo = 0; di = 1; up = 0; r = 0; s = 0; }
      if (pinned && k == 1) lo = 0;
      double m = k == 0 ? di : di - lo * cp[k - 1];
      cp[k] = up / m;
      dRe[k] = (r - lo * (k == 0 ? 0 : dRe[k - 1])) / m;
      dIm[k] = (s - lo * (k == 0 ? 0 : dIm[k - 1])) / m;
    }
    for (int k = nc - 1; k >= 0; k--) {
      if (k < nc - 1) { dRe[k] -= cp[k] * dRe[k + 1]; dIm[k] -= cp[k] * dIm[k + 1]; }
      re[mode + k * cl] = dRe[k]; im[mode + k * cl] = dIm[k];
    }
  }
};

// ---- advection.ts:17-399 -------------------------------------------------------------------------------------------
double upwind5(double a, double b, double c, double d, double e) { return (2 * a - 13 * b + 47 * c + 27 * d - 3 * e) / 60; }
double weno(double a, double b, double c, double d, double e) {
  if (a == 0 && b == 0 && c == 0 && d == 0 && e == 0) return 0;
  double p0 = (2 * a - 7 * b + 11 * c) / 6, p1 = (-b + 5 * c + 2 * d) / 6, p2 = (2 * c + 5 * d - e) / 6;
  double s0 = a - 2 * b + c, t0 = a - 4 * b + 3 * c, s1 = b - 2 * c + d, t1 = b - d, s2 = c - 2 * d + e, t2 = 3 * c - 4 * d + e;
  double eps = 2e-7 * (a * a + b * b + c * c + d * d + e * e) + 1e-40;
  double b0 = eps + 13.0 / 12 * s0 * s0 + .25 * t0 * t0, b1 = eps + 13.0 / 12 * s1 * s1 + .25 * t1 * t1, b2 = eps + 13.0 / 12 * s2 * s2 + .25 * t2 * t2;
  double c0 = b0 * b0, c1 = b1 * b1, c2 = b2 * b2, w0 = .1 * c1 * c2, w1 = .6 * c0 * c2, w2 = .3 * c0 * c1;
  return (w0 * p0 + w1 * p1 + w2 * p2) / (w0 + w1 + w2);
}
const double RK3[3] = {1.0 / 3, 1.0 / 2, 1};

struct FluxTransport {
  const Grid &g; F64 rho, mx, my, mz, div, fx, fy, fz, phi, cp, dRe, dIm, lambda; HorizontalDFT dft;
  std::vector<int> xo[6], yo[6];
  FluxTransport(const Grid &grid, const F64 &density) : g(grid), rho(density), mx(grid.n), my(grid.n), mz(grid.n), div(grid.n), fx(grid.n), fy(grid.n), fz(grid.n),
    phi(grid.n), cp(grid.nz), dRe(grid.nz), dIm(grid.nz), lambda(grid.nx * grid.ny), dft(grid.nx, grid.ny, grid.nz) {
    for (int o = -2; o <= 3; o++) {
      xo[o + 2].resize(g.nx); yo[o + 2].resize(g.ny);
      for (int x = 0; x < g.nx; x++) xo[o + 2][x] = (x + o + 2 * g.nx) % g.nx;
      for (int y = 0; y < g.ny; y++) yo[o + 2][y] = ((y + o + 2 * g.ny) % g.ny) * g.nx;
    }
    for (int mode = 0; mode < g.nx * g.ny; mode++) {
      int m = mode % g.nx, n = mode / g.nx;
      lambda[mode] = m == 0 && n == 0 ? 0 : 4 * std::pow(std::sin(PI * m / g.nx), 2) / (g.dx * g.dx) + 4 * std::pow(std::sin(PI * n / g.ny), 2) / (g.dy * g.dy);
    }
  }
  void setVelocity(const float *u, const float *v, const float *w) { faces(u, v, w); divergence(); forward(); modes(); inverse(); correctFaces(); divergence(); }
  void faces(const float *u, const float *v, const float *w) {
    int nx = g.nx, ny = g.ny, nz = g.nz, layer = g.layer;
    for (int z = 0, i = 0; z < nz; z++) {
      int l = z * layer; double r = rho[z], rUp = z < nz - 1 ? (rho[z] + rho[z + 1]) / 2 : 0;
      for (int y = 0; y < ny; y++) for (int x = 0; x < nx; x++, i++) {
        mx[i] = r * (double(u[i]) + u[xo[3][x] + y * nx + l]) / 2;
        my[i] = r * (double(v[i]) + v[x + yo[3][y] + l]) / 2;
        mz[i] = z < nz - 1 ? rUp * (double(w[i]) + w[i + layer]) / 2 : 0;
      }
    }
  }
  void divergence() {
    int nx = g.nx, ny = g.ny, layer = g.layer;
    for (int z = 0, i = 0; z < g.nz; z++) {
      int l = z * layer;
      for (int y = 0; y < ny; y++) for (int x = 0; x < nx; x++, i++)
        div[i] = (mx[i] - mx[xo[1][x] + y * nx + l]) / g.dx + (my[i] - my[x + yo[1][y] + l]) / g.dy + (mz[i] - (z > 0 ? mz[i - layer] : 0)) / g.hz[z];
    }
  }
  void forward() { dft.forwardLayers([&](int i) { return div[i]; }, 0, g.nz); }
  void inverse() { dft.inverseLayers([&](int i, double v) { phi[i] = v; }, 0, g.nz); }
  void modes() {
    int nx = g.nx, nz = g.nz, layer = g.layer; const F64 &h = g.hz, &dzs = g.dzs;
    F64 &re = dft.re, &im = dft.im;
    auto face = [&](int k) { return k < 0 || k >= nz - 1 ? 0 : (rho[k] + rho[k + 1]) / 2; };
    for (int n = 0; n < g.ny; n++) for (int m = 0; m <= dft.mh; m++) {
      int mode = m + nx * n; double lam = lambda[mode]; bool pinned = lam == 0;
      for (int k = 0; k < nz; k++) {
        double lo = k > 0 ? face(k - 1) / dzs[k - 1] : 0, up = k < nz - 1 ? face(k) / dzs[k] : 0, di = -rho[k] * h[k] * lam - lo - up, rr = h[k] * re[mode + k * layer], ri = h[k] * im[mode + k * layer];
        if (pinned && k == 0) { lo = 0; di = 1; up = 0; rr = 0; ri = 0; }
        if (pinned && k == 1) lo = 0;
        double piv = k == 0 ? di : di - lo * cp[k - 1];
        cp[k] = up / piv;
        dRe[k] = (rr - (k == 0 ? 0 : lo * dRe[k - 1])) / piv;
        dIm[k] = (ri - (k == 0 ? 0 : lo * dIm[k - 1])) / piv;
      }
      for (int k = nz - 1; k >= 0; k--) {
        if (k < nz - 1) { dRe[k] -= cp[k] * dRe[k + 1]; dIm[k] -= cp[k] * dIm[k + 1]; }
        re[mode + k * layer] = dRe[k]; im[mode + k * layer] = dIm[k];
      }
    }
  }
  void correctFaces() {
    int nx = g.nx, ny = g.ny, nz = g.nz, layer = g.layer;
    auto face = [&](int k) { return k < 0 || k >= nz - 1 ? 0 : (rho[k] + rho[k + 1]) / 2; };
    for (int z = 0, i = 0; z < nz; z++) {
      int l = z * layer; double r = rho[z], rUp = face(z);
      for (int y = 0; y < ny; y++) for (int x = 0; x < nx; x++, i++) {
        mx[i] -= r * (phi[xo[3][x] + y * nx + l] - phi[i]) / g.dx;
        my[i] -= r * (phi[x + yo[3][y] + l] - phi[i]) / g.dy;
        if (z < nz - 1) mz[i] -= rUp * (phi[i + layer] - phi[i]) / g.dzs[z];
      }
    }
  }
  // tendencyWeno / tendencyLinear (advection.ts:294-378) over all levels. The active-row skipping is left out: it
  // writes exactly what the full computation gives (the reference's own test, advection.test.ts:48-65).
  void tendency(const float *q, double *tend, bool monotone) {
    int nx = g.nx, ny = g.ny, nz = g.nz, layer = g.layer;
    const std::vector<int> &xm2 = xo[0], &xm1 = xo[1], &xp1 = xo[3], &xp2 = xo[4], &xp3 = xo[5];
    const std::vector<int> &ym2 = yo[0], &ym1 = yo[1], &yp1 = yo[3], &yp2 = yo[4], &yp3 = yo[5];
    auto rec = monotone ? weno : upwind5;
    for (int z = 0, i = 0; z < nz; z++) {
      int l = z * layer; bool top = z == nz - 1;
      int zm2 = std::max(0, z - 2) * layer, zm1 = std::max(0, z - 1) * layer, zp1 = std::min(nz - 1, z + 1) * layer, zp2 = std::min(nz - 1, z + 2) * layer, zp3 = std::min(nz - 1, z + 3) * layer;
      for (int y = 0; y < ny; y++) {
        int row = y * nx + l, c0 = y * nx;
        for (int x = 0; x < nx; x++, i++) {
          double m1 = mx[i];
          fx[i] = m1 >= 0 ? m1 * rec(q[xm2[x] + row], q[xm1[x] + row], q[i], q[xp1[x] + row], q[xp2[x] + row]) : m1 * rec(q[xp3[x] + row], q[xp2[x] + row], q[xp1[x] + row], q[i], q[xm1[x] + row]);
          double m2 = my[i];
          fy[i] = m2 >= 0 ? m2 * rec(q[x + ym2[y] + l], q[x + ym1[y] + l], q[i], q[x + yp1[y] + l], q[x + yp2[y] + l]) : m2 * rec(q[x + yp3[y] + l], q[x + yp2[y] + l], q[x + yp1[y] + l], q[i], q[x + ym1[y] + l]);
          double m3 = mz[i]; int c = x + c0;
          fz[i] = top ? 0 : m3 >= 0 ? m3 * rec(q[c + zm2], q[c + zm1], q[i], q[c + zp1], q[c + zp2]) : m3 * rec(q[c + zp3], q[c + zp2], q[c + zp1], q[i], q[c + zm1]);
        }
      }
    }
    for (int z = 0, i = 0; z < nz; z++) {
      int l = z * layer; double r = rho[z], hz = g.hz[z];
      for (int y = 0; y < ny; y++) for (int x = 0; x < nx; x++, i++) {
        double flux = (fx[i] - fx[xm1[x] + y * nx + l]) / g.dx + (fy[i] - fy[x + ym1[y] + l]) / g.dy + (fz[i] - (z > 0 ? fz[i - layer] : 0)) / hz;
        tend[i] = -(flux - q[i] * div[i]) / r;
      }
    }
  }
};

// ---- electric.ts:16-61 ----------------------------------------------------------------------------------------------
struct ElectricSolver {
  const Grid &g; HorizontalDFT dft; F64 lambda, interior, cp, dRe, dIm;
  explicit ElectricSolver(const Grid &grid) : g(grid), dft(grid.nx, grid.ny, grid.nz - 2), lambda(grid.nx * grid.ny),
    interior(size_t(grid.nx) * grid.ny * (grid.nz - 2)), cp(grid.nz), dRe(grid.nz), dIm(grid.nz) {
    for (int n = 0; n < g.ny; n++) for (int m = 0; m < g.nx; m++)
      lambda[m + g.nx * n] = 4 * std::pow(std::sin(PI * m / g.nx), 2) / (g.dx * g.dx) + 4 * std::pow(std::sin(PI * n / g.ny), 2) / (g.dy * g.dy);
  }
  void solve(const double *rho, double *phi) {
    int nz = g.nz, layer = g.layer, ni = nz - 2, cl = layer;
    for (int i = 0; i < cl * ni; i++) interior[i] = -rho[i + layer] / EPSILON0;
    dft.forwardLayers([&](int i) { return interior[i]; }, 0, ni);
    F64 &re = dft.re, &im = dft.im;
    for (int n = 0; n < g.ny; n++) for (int m = 0; m <= dft.mh; m++) {
      int mode = m + g.nx * n; double lam = lambda[mode];
      for (int j = 0; j < ni; j++) {
        int k = j + 1; double lo = 1 / (g.hz[k] * g.dzs[k - 1]), up = 1 / (g.hz[k] * g.dzs[k]), di = -lo - up - lam;
        double a = j > 0 ? lo : 0, c = j < ni - 1 ? up : 0, r = re[mode + j * cl], s = im[mode + j * cl];
        double piv = j == 0 ? di : di - a * cp[j - 1];
        cp[j] = c / piv;
        dRe[j] = (r - (j == 0 ? 0 : a * dRe[j - 1])) / piv;
        dIm[j] = (s - (j == 0 ? 0 : a * dIm[j - 1])) / piv;
      }
      for (int j = ni - 1; j >= 0; j--) {
        if (j < ni - 1) { dRe[j] -= cp[j] * dRe[j + 1]; dIm[j] -= cp[j] * dIm[j + 1]; }
        re[mode + j * cl] = dRe[j]; im[mode + j * cl] = dIm[j];
      }
    }
    dft.inverseLayers([&](int i, double v) { interior[i] = v; }, 0, ni);
    for (int i = 0; i < layer; i++) { phi[i] = 0; phi[i + (nz - 1) * layer] = 0; }
    for (int i = 0; i < cl * ni; i++) phi[i + layer] = interior[i];
  }
};

// ---- lightning.ts:14-420 --------------------------------------------------------------------------------------------
constexpr double E_INIT = 2.84e5, RHO_STANDARD = 1.225, F_CRIT = .43, RHO_MIN = 1e-10, CHANNEL_RADIUS = 1, SHIELDING = .25;
constexpr int MAX_CHANNEL = 3000, STROKES = 4, MAX_FLASHES = 10;
constexpr double FLASH_WINDOW = 3, SIGMA_GROUND = 2e-14, SIGMA_SCALE = 6000, CLOUD_SIGMA = ORC_MUTANT == 6 ? 1 : .1, CLOUDY = 1e-5, CORONA_FIELD = 5e3, CORONA_RATE = 5e-13;
int flashLimit(double dt) { return std::max(1, int(std::floor(MAX_FLASHES * dt / FLASH_WINDOW + .5))); }

// A flash (lightning.ts Flash, :57-72): the channel's nodes in growth order, each one's parent (-1 for the first) and
// polarity (the sign of the charge its deposit takes), and the ground node a CG reached (-1 for an IC).
struct Flash { double time; bool cg; int start; int nodes; double potential, charge, moment; int ground; std::vector<int> channel, parent, polarity; };

struct Lightning {
  const Grid &g; ElectricSolver solver; F64 rho, phi, change, sigma, air, eInit, capacitance; F32 cloudy; bool hasCloud = false;
  double fromGround = 0, moment = 0, conducted = 0; bool corona = true;
  std::vector<int> mark, failed; int flashId = 0; uint32_t rng = 0; uint32_t seed;
  std::vector<int> frontier; F64 weight; int size = 0; std::vector<int> seen; F64 scale; std::vector<int> link;
  Lightning(const Grid &grid, const F64 &airDensity, int sd) : g(grid), solver(grid), rho(grid.n), phi(grid.n), change(grid.n), sigma(grid.n), air(airDensity),
    eInit(grid.nz), capacitance(grid.nz), mark(grid.n), failed(grid.n), seed(uint32_t(sd)), frontier(grid.n), weight(grid.n), seen(grid.n), scale(grid.n), link(grid.n) {
    for (int z = 0; z < g.nz; z++) eInit[z] = E_INIT * air[z] / RHO_STANDARD;
    for (int z = 0; z < g.nz; z++) { double L = std::cbrt(g.dx * g.dy * g.hz[z]); capacitance[z] = SHIELDING * 2 * PI * EPSILON0 * L / std::log(L / CHANNEL_RADIUS); }
    rng = seed;
  }
  double conduct(float *airCharge, double dt) {
    int nx = g.nx, ny = g.ny, nz = g.nz, layer = g.layer;
    for (int z = 0, i = 0; z < nz; z++) { double clear = SIGMA_GROUND * std::exp(g.zs[z] / SIGMA_SCALE); for (int e = i + layer; i < e; i++) sigma[i] = hasCloud && cloudy[i] > CLOUDY ? clear * CLOUD_SIGMA : clear; }
    std::fill(change.begin(), change.end(), 0.0);
    auto series = [](double a, double b) { return 2 * a * b / (a + b); };
    double lost = 0;
    for (int z = 0; z < nz - 1; z++) {
      double across = g.dy * g.hz[z] * dt / g.dx, along = g.dx * g.hz[z] * dt / g.dy, up = g.dx * g.dy * dt / g.dzs[z];
      for (int y = 0; y < ny; y++) for (int x = 0; x < nx; x++) {
        int i = x + y * nx + z * layer;
        if (z > 0) {
          int ix = g.xp[x] + y * nx + z * layer, iy = x + g.yp[y] + z * layer;
          double qx = series(sigma[i], sigma[ix]) * (phi[i] - phi[ix]) * across, qy = series(sigma[i], sigma[iy]) * (phi[i] - phi[iy]) * along;
          change[i] -= qx + qy; change[ix] += qx; change[iy] += qy;
        }
        int j = i + layer; double q = series(sigma[i], sigma[j]) * (phi[i] - phi[j]) * up;
        if (z == 0) { change[j] += q; lost -= q; } else if (z == nz - 2) { change[i] -= q; lost += q; } else { change[i] -= q; change[j] += q; }
      }
    }
    if (corona) for (int c = 0; c < layer; c++) {
      double e0 = -phi[c + layer] / g.dzs[0], over = std::fabs(e0) - CORONA_FIELD;
      if (over <= 0) continue;
      double q = (e0 > 0 ? 1 : e0 < 0 ? -1 : 0) * CORONA_RATE * over * g.dx * g.dy * dt;
      change[c + layer] += q; lost -= q;
    }
    for (int z = 1; z < nz - 1; z++) {
      double perKg = 1 / (air[z] * g.dx * g.dy * g.hz[z]);
      for (int i = z * layer, e = i + layer; i < e; i++) if (change[i] != 0) airCharge[i] = float(airCharge[i] + change[i] * perKg);
    }
    conducted += lost;
    return lost;
  }
  void solve(float *const *carriers) {
    for (int z = 0, i = 0; z < g.nz; z++) for (int e = i + g.layer; i < e; i++) { double s = 0; for (int c = 0; c < 4; c++) s += carriers[c][i]; rho[i] = s * air[z]; }
    solver.solve(rho.data(), phi.data());
  }
  std::pair<int, double> strongest(int stamp) {
    int nx = g.nx, ny = g.ny, nz = g.nz, layer = g.layer; int best = -1; double ratio = 0;
    for (int z = 1; z < nz - 1; z++) {
      double dz = g.zs[z + 1] - g.zs[z - 1], limit2 = std::pow(eInit[z], 2);
      for (int y = 0; y < ny; y++) for (int x = 0; x < nx; x++) {
        int i = x + y * nx + z * layer, row = y * nx + z * layer;
        double ex = (phi[g.xp[x] + row] - phi[g.xm[x] + row]) / (2 * g.dx), ey = (phi[x + g.yp[y] + z * layer] - phi[x + g.ym[y] + z * layer]) / (2 * g.dy), ez = (phi[i + layer] - phi[i - layer]) / dz;
        double r = (ex * ex + ey * ey + ez * ez) / limit2;
        if (r > ratio && failed[i] != stamp) { ratio = r; best = i; }
      }
    }
    return {best, std::sqrt(ratio)};
  }
  void setCloud(const float *cloud, const float *ice, const float *snow) {
    if (!hasCloud) { cloudy.assign(g.n, 0.0f); hasCloud = true; }
    for (int i = 0; i < g.n; i++) cloudy[i] = 0;
    for (int i = 0; i < g.n; i++) cloudy[i] = float(cloudy[i] + cloud[i]);
    for (int i = 0; i < g.n; i++) cloudy[i] = float(cloudy[i] + ice[i]);
    for (int i = 0; i < g.n; i++) cloudy[i] = float(cloudy[i] + snow[i]);
  }
  // cloud([cloudy]) (lightning.ts:179-186) with one array: the cloudiness itself.
  void setCloudy(const float *c) { cloudy.assign(c, c + g.n); hasCloud = true; }
  bool clear(int i) const { return std::fabs(rho[i]) < RHO_MIN || (hasCloud && cloudy[i] <= CLOUDY); }
  std::vector<Flash> discharge(float *const *carriers, double time, int limit) {
    std::vector<Flash> flashes; int stamp = ++flashId;
    int64_t tt = int64_t(std::floor(time * 10 + .5));
    rng = seed ^ (uint32_t(tt) * 0x9E3779B1u);
    for (int k = 0; k <= limit; k++) {
      solve(carriers);
      if (k == limit) break;
      auto sr = strongest(stamp);
      if (sr.second <= 1) break;
      Flash f = flash(sr.first, carriers, time);
      if (f.charge == 0) { failed[sr.first] = stamp; continue; }
      flashes.push_back(f);
    }
    return flashes;
  }
  Flash flash(int start, float *const *carriers, double time) {
    int layer = g.layer, id = ++flashId;
    std::vector<int> nodes, parent, polarity;
    double sum = 0, potential = 0, before = 0, neutralized = 0, lowered = 0; int ground = -1; bool touched = false, left = false; int strokes = 0;
    size = 0; moment = 0;
    auto add = [&](int j, int from) {
      mark[j] = id; nodes.push_back(j); parent.push_back(from);
      if (j < layer) { ground = j; before = potential; potential = 0; touched = true; return; }
      if (ground < 0) { sum += phi[j]; potential = sum / nodes.size(); }
      extend(j, int(nodes.size()) - 1, id);
    };
    add(start, -1);
    for (;;) {
      size_t grown = nodes.size(); bool event = false;
      while (nodes.size() < size_t(MAX_CHANNEL)) {
        double total = 0;
        for (int f = 0; f < size; f++) {
          int j = frontier[f];
          double w = (ground >= 0 && (j < layer || phi[j] * before <= 0)) || (clear(j) && (phi[j] - potential) * potential >= 0) ? 0 : std::fabs(potential - phi[j]) / scale[j] - 1;
          weight[f] = w > 0 ? w : 0; total += weight[f];
        }
        if (total == 0) break;
        double pick = mulberry(rng) * total; int n = 0;
        while (n < size - 1 && (pick -= weight[n]) > 0) n++;
        int j = frontier[n], from = link[j];
        if (!left && ground < 0 && clear(j)) { left = true; event = true; break; }
        frontier[n] = frontier[--size];
        if (clear(nodes[from])) for (int f = size - 1; f >= 0; f--) if (link[frontier[f]] == from) frontier[f] = frontier[--size];
        add(j, from);
        if (touched) { touched = false; event = true; break; }
      }
      if (ground >= 0) { double q = deposit(nodes, polarity, potential, true, carriers); lowered += q; }
      else neutralized += deposit(nodes, polarity, potential, false, carriers);
      if (ground >= 0 && (++strokes >= STROKES || (!event && nodes.size() == grown))) break;
      if (ground < 0 && !event) break;
      if (nodes.size() >= size_t(MAX_CHANNEL)) break;
      solve(carriers);
      if (ground < 0) { sum = 0; for (int i : nodes) sum += phi[i]; potential = sum / nodes.size(); }
    }
    Flash f; f.time = time; f.cg = ground >= 0; f.start = start; f.nodes = int(nodes.size()); f.potential = ground >= 0 ? before : potential;
    f.ground = ground; f.channel = nodes; f.parent = parent; f.polarity = polarity;
    f.charge = ground >= 0 ? lowered : neutralized; f.moment = ground >= 0 ? moment : 0;
    return f;
  }
  double deposit(const std::vector<int> &nodes, std::vector<int> &polarity, double V, bool grounded, float *const *carriers) {
    int layer = g.layer; F64 dep(nodes.size(), 0.0);
    double positive = 0, negative = 0;
    for (size_t c = 0; c < nodes.size(); c++) {
      int i = nodes[c], z = i / layer;
      if (z == 0) { if (c >= polarity.size()) polarity.push_back(0); continue; }
      double q = capacitance[z] * (V - phi[i]), here = rho[i] * g.dx * g.dy * g.hz[z];
      if (c >= polarity.size()) polarity.push_back(q > 0 ? 1 : q < 0 ? -1 : 0);
      if (q * here >= 0) continue;
      double d = (q > 0 ? 1 : -1) * std::min(std::fabs(q), std::fabs(here));
      dep[c] = d;
      if (d > 0) positive += d; else negative -= d;
    }
    double charge;
    if (!grounded) {
      double both = std::min(positive, negative);
      for (size_t c = 0; c < nodes.size(); c++) dep[c] *= dep[c] > 0 ? (positive > 0 ? both / positive : 0) : (negative > 0 ? both / negative : 0);
      charge = both;
    } else {
      charge = negative - positive;
      fromGround -= charge;
      for (size_t c = 0; c < nodes.size(); c++) if (dep[c] != 0) moment -= dep[c] * g.zs[nodes[c] / layer];
    }
    for (size_t c = 0; c < nodes.size(); c++) if (dep[c] != 0) neutralize(carriers, nodes[c], dep[c]);
    return charge;
  }
  void extend(int i, int c, int id) {
    int nx = g.nx, ny = g.ny, nz = g.nz, layer = g.layer, z = i / layer, y = (i - z * layer) / nx, x = i - z * layer - y * nx;
    for (int k = std::max(0, z - 1); k <= std::min(nz - 2, z + 1); k++) {
      double hz = g.zs[k] - g.zs[z], threshold = F_CRIT * .5 * (eInit[z] + eInit[k]);
      for (int dj = -1; dj <= 1; dj++) for (int di = -1; di <= 1; di++) {
        if (x + di < 0 || x + di >= nx || y + dj < 0 || y + dj >= ny) continue;
        int j = x + di + (y + dj) * nx + k * layer;
        if (mark[j] == id || (k >= z && std::fabs(rho[j]) < RHO_MIN)) continue;
        double s = std::sqrt(std::pow(di * g.dx, 2) + std::pow(dj * g.dy, 2) + hz * hz) * threshold;
        if (seen[j] != id) { seen[j] = id; frontier[size++] = j; }
        else if (s >= scale[j]) continue;
        scale[j] = s; link[j] = c;
      }
    }
  }
  void neutralize(float *const *carriers, int i, double q) {
    int z = i / g.layer; double perKg = q / (air[z] * g.dx * g.dy * g.hz[z]);
    double opposite = 0, sq = q > 0 ? 1 : q < 0 ? -1 : 0;
    for (int c = 0; c < 4; c++) if (carriers[c][i] * q < 0) opposite -= carriers[c][i] * sq;
    if (opposite <= 0) return;
    for (int c = 0; c < 4; c++) if (carriers[c][i] * q < 0) carriers[c][i] = float(carriers[c][i] + perKg * std::fabs(double(carriers[c][i])) / opposite);
  }
};

// The sounding's wind error at height z (perturbConfig's `at`, member.ts:56): d at 0, 3, 6 and 10 km, linear between those
// heights, held below and above them.
static double windError(const double *d, double z) {
  static const double at[4] = {0, 3000, 6000, 10000};
  int k = 0;
  while (k < 2 && z >= at[k + 1]) k++;
  return d[k] + (d[k + 1] - d[k]) * std::min(std::max((z - at[k]) / (at[k + 1] - at[k]), 0.0), 1.0);
}

// ---- environment.ts:104-131 with an analytic profile (profiles.ts:349-361) ------------------------------------------
struct Environment {
  F64 p, exner, theta, q, u, v, rho, lnP; orc_profile prof;
  double ptheta(double z) const { return z <= prof.ztr ? prof.theta0 + prof.dtheta * std::pow(z / prof.ztr, 1.25) : (prof.theta0 + prof.dtheta) * std::exp(G * (z - prof.ztr) / (CP * prof.ttr)); }
  double prh(double z) const { return z <= prof.ztr ? 1 - prof.rh_drop * std::pow(z / prof.ztr, 1.25) : prof.rh_top; }
  double temperatureAt(double z, double pp) const { return ptheta(z) * std::pow(pp / 100000, KAPPA) - 273.15; }
  double mixingRatio(double z, double t, double pp) const { return std::min(prh(z) * qsatP(t, pp), prof.qv_max); }
  double pressureAt(double z) const {
    double x = std::max(0.0, z) / 10; int k = std::min(int(std::floor(x)), int(lnP.size()) - 2);
    return std::exp(lnP[k] + (lnP[k + 1] - lnP[k]) * (x - k));
  }
  Environment(const Grid &g, const orc_profile &pr) : prof(pr) {
    double top = g.height + 2000; int n = int(std::ceil(top / 10)) + 1;
    lnP.assign(n, 0.0); lnP[0] = std::log(101325.0);
    auto tv = [&](double z, double pp) { double t = temperatureAt(z, pp), qq = mixingRatio(z, t, pp); return (t + 273.15) * (1 + .61 * qq); };
    for (int k = 0; k + 1 < n; k++) { double z = k * 10.0, pp = std::exp(lnP[k]); lnP[k + 1] = lnP[k] - G * 10 / (RD * .5 * (tv(z, pp) + tv(z + 10, pp))); }
    int nz = g.nz; p.assign(nz, 0); exner.assign(nz, 0); theta.assign(nz, 0); q.assign(nz, 0); u.assign(nz, 0); v.assign(nz, 0); rho.assign(nz, 0);
    for (int z = 0; z < nz; z++) {
      double alt = g.zs[z];
      p[z] = pressureAt(alt); exner[z] = std::pow(p[z] / 100000, KAPPA);
      double tEnv = temperatureAt(alt, pressureAt(alt));
      theta[z] = (tEnv + 273.15) * std::pow(100000 / pressureAt(alt), KAPPA);
      q[z] = mixingRatio(alt, tEnv - 0, pressureAt(alt));
      u[z] = prof.shear * std::tanh(alt / prof.shear_depth) + windError(prof.wind_u, alt); v[z] = windError(prof.wind_v, alt);
      rho[z] = p[z] / (RD * theta[z] * exner[z] * (1 + .61 * q[z]));
    }
  }
};

// ---- the model (model.ts:183-1697), baseline configuration -----------------------------------------------------------
struct Model {
  Grid g; Environment env; orc_run run;
  F32 f[ORC_FIELDS]; F32 col[ORC_COLUMNS]; F32 start[ORC_CARRIED]; F64 tend[ORC_CARRIED];
  F32 condensate, tmp;
  FluxTransport flux; PressureSolver solver; std::unique_ptr<Lightning> lightning;
  F64 pressure, divergence, km, kh, levelWeight, uEnv, vEnv;
  double time = 0, negativeFilled = 0, chargeGround = 0, firstFlash = -1; long clipped = 0; int flashesIC = 0, flashesCG = 0;
  double addedHeat = 0, addedVapour = 0, addedCount = 0;  // AtmosphereModel.added (model.ts:227-228)
  bool electrified = true;
  double cs, dragCd, f_coriolis;
  // RotationState (model.ts:142-158) in the order of ORC_ROT_*, and the levels updateRotation sums over.
  double rotation[ORC_ROT] = {0};
  std::vector<float> uhColumn;  // the columns' 2-5 km helicity of the last updateRotation (model.ts uhColumn)
  bool swaths = false;
  int uhZ0 = 0, uhZ1 = -1, top1 = -1, top3 = -1;

  Model(const orc_grid &go, const orc_profile &pr, const orc_run &r) : g(go), env(g, pr), run(r), flux(g, env.rho), solver(g, env.rho) {
    for (int k = 0; k < ORC_FIELDS; k++) f[k].assign(g.n, 0.0f);
    for (int k = 0; k < ORC_COLUMNS; k++) col[k].assign(g.layer, 0.0f);
    for (int k = 0; k < ORC_CARRIED; k++) { start[k].assign(g.n, 0.0f); tend[k].assign(g.n, 0.0); }
    condensate.assign(g.n, 0.0f); tmp.assign(g.n, 0.0f);
    pressure.assign(solver.cells, 0.0); divergence.assign(solver.cells, 0.0); km.assign(g.n, 0.0); kh.assign(g.n, 0.0);
    levelWeight.resize(g.nz); for (int z = 0; z < g.nz; z++) levelWeight[z] = g.hz[z] / g.dz * env.rho[z];
    uEnv.resize(g.nz); vEnv.resize(g.nz); for (int z = 0; z < g.nz; z++) { uEnv[z] = env.u[z] - run.frame_u; vEnv[z] = env.v[z] - run.frame_v; }
    cs = .18 * run.turbulence / .55;
    dragCd = std::pow(.4 / std::log(2 * g.hz[0] / 4 / run.roughness), 2);
    f_coriolis = 2 * OMEGA * std::sin(run.latitude * PI / 180);
    // model.ts:389 (uhLevels) and :1612 (the last levels at or below 1 and 3 km).
    uhZ0 = g.nz;
    for (int z = 0; z < g.nz; z++) {
      if (g.zs[z] >= 2000 && uhZ0 == g.nz) uhZ0 = z;
      if (g.zs[z] <= 5000) uhZ1 = z;
      if (g.zs[z] <= 1000) top1 = z;
      if (g.zs[z] <= 3000) top3 = z;
    }
    initialize();
  }
  float *F(int k) { return f[k].data(); }

  // initialize() (model.ts:399-451) in single-storm mode (no noise, bubble strength 0), then the harness's thermal and noise.
  void initialize() {
    int nx = g.nx, ny = g.ny, nz = g.nz;
    for (int z = 0, i = 0; z < nz; z++) for (int y = 0; y < ny; y++) for (int x = 0; x < nx; x++, i++) {
      f[ORC_U][i] = float(env.u[z] - run.frame_u); f[ORC_V][i] = float(env.v[z] - run.frame_v); f[ORC_W][i] = 0;
      f[ORC_THETA][i] = float(env.theta[z]); f[ORC_Q][i] = float(env.q[z]);
    }
    // The thermal: cos^2 of the normalised distance (model.ts:431-436, ELLIPSOID with amplitude and centre of the run).
    for (int z = 0; z < nz; z++) for (int y = 0; y < ny; y++) for (int x = 0; x < nx; x++) {
      double ax = (x * g.dx - run.bubble_x) / run.bubble_rh, ay = (y * g.dy - run.bubble_y) / run.bubble_rh, az = (g.zs[z] - run.bubble_z) / run.bubble_rz;
      double beta = std::sqrt(ax * ax + ay * ay + az * az);
      if (beta < 1) { int i = x + nx * (y + ny * z); double c = std::cos(PI * beta / 2); f[ORC_THETA][i] = float(f[ORC_THETA][i] + run.bubble_amp * (c * c)); }
    }
    uint32_t s = uint32_t(run.noise_seed);
    for (int z = 0, i = 0; z < nz; z++) for (int y = 0; y < ny; y++) for (int x = 0; x < nx; x++, i++) {
      if (!(g.zs[z] < run.noise_top)) continue;
      f[ORC_THETA][i] = float(f[ORC_THETA][i] + (2 * mulberry(s) - 1) * run.noise_amp);
    }
  }

  // runTask 'fall' (model.ts:1220-1235).
  void fallPhase() {
    const F64 &rho = env.rho;
    for (int k = 0; k < g.nz; k++) for (int i = k * g.layer, e = i + g.layer; i < e; i++) {
      f[ORC_FALL_RAIN][i] = float(fallSpeedRain(f[ORC_RAIN][i], rho[k], rho[0]));
      f[ORC_FALL_SNOW][i] = float(snowFallSpeed(f[ORC_SNOW][i], rho[k], rho[0]));
      f[ORC_FALL_GRAUPEL][i] = float(graupelFallSpeed(f[ORC_GRAUPEL][i], rho[k]));
      f[ORC_FALL_HAIL][i] = float(hailFallSpeed(f[ORC_HAIL][i], rho[k]));
      if (electrified) {
        double is = double(f[ORC_ICE][i]) + f[ORC_SNOW][i], gh = double(f[ORC_GRAUPEL][i]) + f[ORC_HAIL][i];
        f[ORC_FALL_CHARGE_ICE][i] = float(is > 0 ? f[ORC_SNOW][i] * double(f[ORC_FALL_SNOW][i]) / is : 0);
        f[ORC_FALL_CHARGE_GRAUPEL][i] = float(gh > 0 ? (f[ORC_GRAUPEL][i] * double(f[ORC_FALL_GRAUPEL][i]) + f[ORC_HAIL][i] * double(f[ORC_FALL_HAIL][i])) / gh : 0);
      }
    }
  }
  // updateGround (model.ts:644-657): the four ground columns under each domain column, with bilinear weights, from the
  // distance the domain has moved over the ground (frameOffset, model.ts:637). Without frame motion the identity.
  std::vector<int> groundCell; F64 groundWeight;
  void updateGround() {
    const int nx = g.nx, ny = g.ny;
    groundCell.resize(size_t(g.layer) * 4); groundWeight.resize(size_t(g.layer) * 4);
    auto mod = [](double x, double n) { return std::fmod(std::fmod(x, n) + n, n); };
    const double ox = run.frame_u * time, oy = run.frame_v * time, sx = mod(ox / g.dx, nx), sy = mod(oy / g.dy, ny);
    const int x0 = int(std::floor(sx)), y0 = int(std::floor(sy));
    const double fx = sx - x0, fy = sy - y0;
    for (int y = 0; y < ny; y++) for (int x = 0; x < nx; x++) {
      const int i = x + nx * y, ax = (x + x0) % nx, bx = (ax + 1) % nx, ay = (y + y0) % ny * nx, by = (ay / nx + 1) % ny * nx, k = i * 4;
      groundCell[k] = ax + ay; groundCell[k + 1] = bx + ay; groundCell[k + 2] = ax + by; groundCell[k + 3] = bx + by;
      groundWeight[k] = (1 - fx) * (1 - fy); groundWeight[k + 1] = fx * (1 - fy); groundWeight[k + 2] = (1 - fx) * fy; groundWeight[k + 3] = fx * fy;
    }
  }
  // deposit (model.ts:691-694): mm of water falling out of domain column i onto the ground columns under it.
  void deposit(int i, double mm, F32 &target) {
    const int k = i * 4;
    for (int j = 0; j < 4; j++) if (groundWeight[k + j] > 0) target[groundCell[k + j]] = float(target[groundCell[k + j]] + groundWeight[k + j] * mm);
  }
  // collectHail (model.ts:697-706).
  void collectHail(double dt) {
    const F64 &rho = env.rho;
    for (int b = 0; b < g.layer; b++) {
      double h = f[ORC_HAIL][b];
      if (!(h > 1e-9)) continue;
      deposit(b, rho[0] * h * f[ORC_FALL_HAIL][b] * dt, col[ORC_HAIL_GROUND]);
      double size = hailMaxDiameter(h, rho[0]) * 1000;
      for (int j = 0; j < 4; j++) {
        if (!(groundWeight[b * 4 + j] > 0)) continue;
        float &hs = col[ORC_HAIL_SIZE][groundCell[b * 4 + j]];
        hs = float(std::max(double(hs), size));
      }
    }
  }
  // fallParts, sedimentPart, sediment (model.ts:565-598).
  int fallParts(const float *rain, const float *vt, double dt) {
    double courant = 0;
    for (int z = 0; z < g.nz; z++) { double c = dt * env.rho[z] / (levelWeight[z] * g.dz); for (int i = z * g.layer; i < (z + 1) * g.layer; i++) if (rain[i] != 0) courant = std::max(courant, c * vt[i] / 1); }
    return std::max(1, int(std::ceil(courant / .8)));
  }
  void sedimentPart(float *rain, const float *vt, double dt, bool deposit) {
    int nz = g.nz, layer = g.layer; const F64 &rho = env.rho, &lw = levelWeight;
    if (deposit) for (int c = 0; c < layer; c++) this->deposit(c, rho[0] * rain[c] * vt[c] * dt, col[ORC_PRECIPITATION]);
    else {
      double out = 0;
      for (int i = 0; i < layer; i++) out += double(rain[i]) * vt[i];
      out *= rho[0];
      chargeGround += out * dt * g.dx * g.dy;
    }
    for (int z = 0; z < nz; z++) {
      int l = z * layer; double inflow = z < nz - 1 ? rho[ORC_MUTANT == 9 ? z : z + 1] : 0;
      for (int i = l; i < l + layer; i++)
        rain[i] = float(rain[i] + dt * (inflow * (z < nz - 1 ? double(rain[i + layer]) * vt[i + layer] : 0) - rho[z] * rain[i] * vt[i]) / (lw[z] * g.dz * 1));
    }
  }
  void sediment(int field, int fall, double dt, bool deposit) {
    int parts = fallParts(F(field), F(fall), dt);
    for (int p = 0; p < parts; p++) sedimentPart(F(field), F(fall), dt / parts, deposit);
  }
  void sedimentStep(double dt) {
    sediment(ORC_RAIN, ORC_FALL_RAIN, dt, true); sediment(ORC_SNOW, ORC_FALL_SNOW, dt, true);
    sediment(ORC_GRAUPEL, ORC_FALL_GRAUPEL, dt, true); sediment(ORC_HAIL, ORC_FALL_HAIL, dt, true);
    if (electrified) { sediment(ORC_CHARGE_ICE, ORC_FALL_CHARGE_ICE, dt, false); sediment(ORC_CHARGE_GRAUPEL, ORC_FALL_CHARGE_GRAUPEL, dt, false); sediment(ORC_CHARGE_RAIN, ORC_FALL_RAIN, dt, false); }
  }
  static bool monotone(int k) { return !(k == ORC_U || k == ORC_V || k == ORC_W || k == ORC_THETA); }
  static bool conserve(int k) { return k == ORC_RAIN || k == ORC_SNOW || k == ORC_GRAUPEL || k == ORC_HAIL || k == ORC_Q || k == ORC_CLOUD || k == ORC_ICE; }
  int carriedCount() const { return electrified ? ORC_CARRIED : ORC_COLD + 1; }

  // stageBuoyancy (model.ts:1204-1214).
  void stageBuoyancy(double *t) {
    for (int z = 0; z < g.nz; z++) {
      double thvEnv = env.theta[z] * (1 + .61 * env.q[z]), ref = thvEnv;
      for (int i = z * g.layer, c = 0; c < g.layer; i++, c++) {
        double loaded = double(f[ORC_CLOUD][i]) + f[ORC_ICE][i] + (double(f[ORC_RAIN][i]) + f[ORC_SNOW][i] + f[ORC_GRAUPEL][i] + f[ORC_HAIL][i]);
        t[i] += G * ((f[ORC_THETA][i] * (1 + .61 * f[ORC_Q][i]) - ref) / ref - loaded);
      }
    }
  }
  // dynamics (model.ts:1029-1052), WENO with the pressure in the stages.
  void dynamics(double dt) {
    flux.setVelocity(F(ORC_U), F(ORC_V), F(ORC_W));
    int nc = carriedCount(), top = (g.nz - 1) * g.layer;
    for (int stage = 0; stage < 3; stage++) {
      for (int k = 0; k < nc; k++) {
        if (stage == 0) start[k] = f[k];
        flux.tendency(F(k), tend[k].data(), monotone(k));
        if (k == ORC_W) stageBuoyancy(tend[k].data());
      }
      double fraction = RK3[stage] * dt;
      for (int k = 0; k < nc; k++) for (int i = 0; i < g.n; i++) f[k][i] = float(start[k][i] + fraction * tend[k][i]);
      if (stage < 2) { for (int b = 0; b < g.layer; b++) { f[ORC_W][b] = 0; f[ORC_W][b + top] = 0; } project(dt, false); }
    }
  }
  // fillNegative (model.ts:606-617).
  double fillNegative(float *a) {
    const double TINY_WATER = 1e-10;
    double removed = 0, kept = 0, negative = 0;
    for (int z = 0, i = 0; z < g.nz; z++) {
      double level = levelWeight[z];
      for (int e = i + g.layer; i < e; i++) { double x = a[i], wz = level; if (x < TINY_WATER) { if (x != 0) { removed += wz * x; if (x < 0) negative -= wz * x; } } else kept += wz * x; }
    }
    if (removed == 0 && negative == 0) return 0;
    double scale = kept > 0 ? std::max(0.0, 1 + removed / kept) : 0;
    for (int i = 0; i < g.n; i++) a[i] = a[i] < TINY_WATER ? 0 : float(a[i] * scale);
    return negative;
  }
  // physicsStep (model.ts:1116-1127): the 'finish' phase then physics() of every level.
  void physicsStep(double dt) {
    int nc = carriedCount();
    for (int k = 0; k < nc; k++) {
      if (conserve(k)) negativeFilled += fillNegative(F(k));
      if (k == ORC_COLD) { double keep = std::pow(.9992, dt); for (int i = 0; i < g.n; i++) f[k][i] = float(f[k][i] * keep); }
    }
    for (int z = 0; z < g.nz; z++) physics(z, dt);
  }
  // physics(z) (model.ts:1347-1531), baseline branches: ice, one-moment, electrified, staged buoyancy, no sun/zones/fronts.
  void physics(int z, double dt) {
    int layer = g.layer; const double alt = g.zs[z], p = env.p[z], rhoZ = env.rho[z], exner = env.exner[z], ue = uEnv[z], ve = vEnv[z];
    const double thEnv = env.theta[z], rhoGround = env.rho[0], H = g.height;
    double spongeStart = std::max(run.tropopause_km * 1000 + 2500, 14000.0), keep = std::exp(-1e-4 * dt), fc = f_coriolis;
    float *u = F(ORC_U), *v = F(ORC_V), *w = F(ORC_W), *theta = F(ORC_THETA), *q = F(ORC_Q), *cloud = F(ORC_CLOUD), *rain = F(ORC_RAIN), *cold = F(ORC_COLD);
    float *ice = F(ORC_ICE), *iceNumber = F(ORC_ICENUMBER), *snow = F(ORC_SNOW), *graupel = F(ORC_GRAUPEL), *hail = F(ORC_HAIL);
    float *cI = F(ORC_CHARGE_ICE), *cG = F(ORC_CHARGE_GRAUPEL), *cR = F(ORC_CHARGE_RAIN), *cA = F(ORC_CHARGE_AIR);
    for (int i = z * layer, e = i + layer; i < e; i++) {
      double before[4], after[4], cNow[4];
      if (electrified) { before[0] = double(ice[i]) + snow[i]; before[1] = double(graupel[i]) + hail[i]; before[2] = rain[i]; before[3] = cloud[i]; }
      double evap = rainProcesses(q[i], cloud[i], rain[i], theta[i] * exner, p, rhoZ, dt);
      if (evap > 0) { double cool = LV / CP / exner * evap; theta[i] = float(theta[i] - cool); cold[i] = float(cold[i] + cool); }
      freezeRain(theta[i], rain[i], graupel[i], exner);
      settleCrystals(ice[i], iceNumber[i]);
      cold[i] = float(cold[i] + iceProcesses(theta[i], q[i], ice[i], snow[i], iceNumber[i], exner, p, rhoZ, dt));
      double iceBefore = ice[i], cloudBefore = cloud[i];
      cold[i] = float(cold[i] + snowProcesses(theta[i], q[i], cloud[i], ice[i], snow[i], rain[i], exner, p, rhoZ, rhoGround, dt));
      double rainBefore = rain[i];
      cold[i] = float(cold[i] + graupelProcesses(theta[i], q[i], cloud[i], ice[i], snow[i], rain[i], graupel[i], exner, p, rhoZ, dt));
      double rainFrozen = theta[i] * exner < 273.15 ? std::max(0.0, rainBefore - rain[i]) : 0;
      cold[i] = float(cold[i] + hailProcesses(theta[i], q[i], cloud[i], ice[i], snow[i], rain[i], graupel[i], hail[i], exner, p, rhoZ, dt));
      if (iceBefore > 0) iceNumber[i] = float(iceNumber[i] * (ice[i] / iceBefore));
      double splinters = splintering(theta[i] * exner - 273.15) * SPLINTERS * std::max(0.0, cloudBefore - cloud[i])
        + (rainFrozen > 0 ? fragmentsPerKg(std::pow(PI * RHO_W * N0R / (rhoZ * rainBefore), .25)) * rainFrozen : 0);
      if (splinters > 0) {
        double fromG = std::min(double(graupel[i]), splinters * CRYSTAL_MASS), fromS = std::min(double(snow[i]), splinters * CRYSTAL_MASS - fromG);
        graupel[i] = float(graupel[i] - fromG); snow[i] = float(snow[i] - fromS); ice[i] = float(ice[i] + (fromG + fromS)); iceNumber[i] = float(iceNumber[i] + (fromG + fromS) / CRYSTAL_MASS);
      }
      double cond = saturationAdjustMixed(theta[i], q[i], cloud[i], ice[i], exner, p, iceNumber[i]);
      settleCrystals(ice[i], iceNumber[i]);
      if (electrified) {
        after[0] = double(ice[i]) + snow[i]; after[1] = double(graupel[i]) + hail[i]; after[2] = rain[i]; after[3] = cloud[i];
        cNow[0] = cI[i]; cNow[1] = cG[i]; cNow[2] = cR[i]; cNow[3] = cA[i];
        if (cNow[0] != 0 || cNow[1] != 0 || cNow[2] != 0 || cNow[3] != 0) transferCharge(cNow, before, after);
        double tc = theta[i] * exner - 273.15;
        bool wetG = wetGrowing(graupel[i], N0G, RHO_G, tc, q[i], p, rhoZ, cloud[i], rain[i], ice[i], snow[i]);
        bool wetH = wetGrowing(hail[i], N0H, RHO_H, tc, q[i], p, rhoZ, cloud[i], rain[i], ice[i], snow[i]);
        double separated = chargingRate(tc, rhoZ, cloud[i], ice[i], snow[i], graupel[i], hail[i], f[ORC_FALL_SNOW][i], wetG, wetH, iceNumber[i] * rhoZ) * dt / rhoZ;
        double induced = inductiveRate(tc, rhoZ, cloud[i], graupel[i], hail[i], f[ORC_FIELDZ][i], (cNow[1] + separated) * rhoZ, wetG, wetH) * dt / rhoZ;
        cI[i] = float(cNow[0] - separated); cG[i] = float(cNow[1] + separated + induced); cR[i] = float(cNow[2]); cA[i] = float(cNow[3] - induced);
      }
      if (cond < 0) cold[i] = float(cold[i] - LV / CP / exner * cond * .14);
      double du = (u[i] - ue) * keep, dv = (v[i] - ve) * keep;
      u[i] = float(ue + du + fc * dv * dt); v[i] = float(ve + dv - fc * du * dt);
      cold[i] = cold[i] < 1e-3 ? 0.0f : float(std::min(double(cold[i]), 15.0));
      if (alt > spongeStart) {
        double s = clamp01((alt - spongeStart) / (H - spongeStart)) * .06 * dt;
        w[i] = float(w[i] * (1 - s)); u[i] = float(lerp(u[i], ue, s)); v[i] = float(lerp(v[i], ve, s)); theta[i] = float(lerp(theta[i], thEnv, s));
      }
    }
    for (int i = z * layer, e = i + layer; i < e; i++) condensate[i] = float(double(cloud[i]) + ice[i]);
  }
  void viscosity(double dt) {
    viscosityOf(g, F(ORC_U), F(ORC_V), F(ORC_W), F(ORC_THETA), env.theta.data(), cs, dt, F(ORC_Q), condensate.data(), env.exner.data(), km.data(), kh.data());
  }
  void mix(float *a, const double *base, double factor, double dt) { mixOf(g, env.rho.data(), km.data(), kh.data(), a, base, factor, dt, tmp.data()); }
  // mixStep (model.ts:1161-1195), flat ground.
  void mixStep(double dt) {
    viscosity(dt);
    mix(F(ORC_U), uEnv.data(), 1, dt); mix(F(ORC_V), vEnv.data(), 1, dt); mix(F(ORC_W), nullptr, 1, dt);
    mix(F(ORC_THETA), env.theta.data(), 3, dt); mix(F(ORC_Q), env.q.data(), 3, dt); mix(F(ORC_CLOUD), nullptr, 3, dt);
    mix(F(ORC_ICE), nullptr, 3, dt); mix(F(ORC_ICENUMBER), nullptr, 3, dt);
    double cd = dragCd * dt, drag = cd / g.hz[0], ue0 = env.u[0], ve0 = env.v[0], envSpeed = std::hypot(ue0, ve0);
    float *u = F(ORC_U), *v = F(ORC_V), *w = F(ORC_W);
    for (int b = 0; b < g.layer; b++) {
      double ug = u[b] + run.frame_u, vg = v[b] + run.frame_v, speed = std::hypot(ug, vg);
      u[b] = float(u[b] - drag * (speed * ug - envSpeed * ue0)); v[b] = float(v[b] - drag * (speed * vg - envSpeed * ve0));
      w[b] = 0; w[b + (g.nz - 1) * g.layer] = 0;
    }
  }
  // project (model.ts:1575-1583) and clipWind (model.ts:1586-1595).
  void project(double dt, bool limit) {
    solver.divergence(F(ORC_U), F(ORC_V), F(ORC_W), dt, divergence.data());
    solver.solve(divergence.data(), pressure.data());
    solver.correct(pressure.data(), F(ORC_U), F(ORC_V), F(ORC_W), dt);
    if (limit) clipWind();
  }
  void clipWind() {
    float *u = F(ORC_U), *v = F(ORC_V), *w = F(ORC_W);
    for (int i = 0; i < g.n; i++) {
      if (u[i] < -120 || u[i] > 120) { u[i] = float(clamp01(u[i], -120, 120)); clipped++; }
      if (v[i] < -120 || v[i] > 120) { v[i] = float(clamp01(v[i], -120, 120)); clipped++; }
      if (w[i] < -100 || w[i] > 100) { w[i] = float(clamp01(w[i], -100, 100)); clipped++; }
    }
  }
  // discharge (model.ts:1537-1552).
  void discharge(double dt) {
    if (!lightning) lightning.reset(new Lightning(g, env.rho, run.seed + 1));
    lightning->corona = true;
    lightning->setCloud(F(ORC_CLOUD), F(ORC_ICE), F(ORC_SNOW));
    float *carriers[4] = {F(ORC_CHARGE_ICE), F(ORC_CHARGE_GRAUPEL), F(ORC_CHARGE_RAIN), F(ORC_CHARGE_AIR)};
    for (const Flash &fl : lightning->discharge(carriers, time + dt, flashLimit(dt))) {
      if (fl.cg) { chargeGround += fl.charge; flashesCG++; } else flashesIC++;
      if (firstFlash < 0) firstFlash = fl.time;
    }
    chargeGround += lightning->conduct(F(ORC_CHARGE_AIR), dt);
    const F64 &phi = lightning->phi; float *ez = F(ORC_FIELDZ);
    for (int z = 1; z < g.nz - 1; z++) { double h = g.zs[z + 1] - g.zs[z - 1]; for (int c = 0, i = z * g.layer; c < g.layer; c++, i++) ez[i] = float(-(phi[i + g.layer] - phi[i - g.layer]) / h); }
  }
  // step (model.ts:873-877) = beginStep, dynamics, endStep.
  // zeta (model.ts:1598-1601): vertical vorticity at a node from the periodic neighbours.
  double zeta(int x, int y, int z) const {
    const int nx = g.nx, l = z * g.layer, xp = (x + 1) % nx, xm = (x + nx - 1) % nx, yp = ((y + 1) % g.ny) * nx, ym = ((y + g.ny - 1) % g.ny) * nx;
    const float *u = f[ORC_U].data(), *v = f[ORC_V].data();
    return (double(v[xp + nx * y + l]) - double(v[xm + nx * y + l])) / (2 * g.dx) - (double(u[x + yp + l]) - double(u[x + ym + l])) / (2 * g.dy);
  }
  // updateRotation (model.ts:1609-1648): the columns' updraught helicity, the updraught cores by a depth-first fill,
  // and how long the core's rotation has exceeded mesoCoreZeta (model.ts:129).
  void updateRotation(double dt) {
    const int nx = g.nx, ny = g.ny, layer = g.layer;
    const float *w = f[ORC_W].data();
    std::vector<float> uhc(layer);
    std::vector<double> coreW(layer), coreWZ(layer);
    double max = 0, min = 0, max01 = 0, max03 = 0; int bx = 0, by = 0, low01 = 0;
    for (int y = 0; y < ny; y++) for (int x = 0; x < nx; x++) {
      double uh = 0, low = 0, sw = 0, swz = 0;
      for (int z = 1; z <= top3; z++) { low += double(w[x + nx * y + z * layer]) * zeta(x, y, z) * g.hz[z]; if (z == top1 && low > max01) { max01 = low; low01 = x + nx * y; } }
      max03 = std::max(max03, low);
      for (int z = uhZ0; z <= uhZ1; z++) {
        const double wn = w[x + nx * y + z * layer], wz = wn * zeta(x, y, z) * g.hz[z];
        uh += wz;
        if (wn > 10) { sw += wn * g.hz[z]; swz += wz; }
      }
      uhc[x + nx * y] = float(uh); coreW[x + nx * y] = sw; coreWZ[x + nx * y] = swz;
      if (uh > max) { max = uh; bx = x; by = y; }
      min = std::min(min, uh);
    }
    std::vector<unsigned char> label(layer, 0);
    std::vector<int> stack;
    double coreZeta = 0, coreUH = 0, coreAnti = 0;
    for (int start = 0; start < layer; start++) {
      if (coreW[start] <= 0 || label[start]) continue;
      double sw = 0, swz = 0, peak = 0;
      label[start] = 1; stack.push_back(start);
      while (!stack.empty()) {
        const int a = stack.back(); stack.pop_back();
        const int x = a % nx, y = (a - x) / nx;
        sw += coreW[a]; swz += coreWZ[a]; peak = std::max(peak, double(uhc[a]));
        for (int oy = -1; oy <= 1; oy++) for (int ox = -1; ox <= 1; ox++) {
          const int b = (x + ox + nx) % nx + nx * ((y + oy + ny) % ny);
          if (coreW[b] > 0 && !label[b]) { label[b] = 1; stack.push_back(b); }
        }
      }
      if (swz / sw > coreZeta) { coreZeta = swz / sw; coreUH = peak; }
      coreAnti = std::max(coreAnti, -swz / sw);
    }
    const double meso = 3e-3 * std::max(1.0, 1200 / g.dx);
    double *r = rotation;
    r[ORC_ROT_UH] = max; r[ORC_ROT_X] = bx; r[ORC_ROT_Y] = by; r[ORC_ROT_ANTI] = -min;
    r[ORC_ROT_PERSISTED] = coreZeta >= meso ? r[ORC_ROT_PERSISTED] + dt : 0;
    r[ORC_ROT_UH01] = max01; r[ORC_ROT_UH03] = max03; r[ORC_ROT_LOW] = low01;
    r[ORC_ROT_CORE_ZETA] = coreZeta; r[ORC_ROT_CORE_UH] = coreUH; r[ORC_ROT_CORE_ANTI] = coreAnti;
    uhColumn = uhc;
  }
  // addSwaths (ensemble/member.ts:85-100), periodic ground (groundRepeats of periodic edges): the ground under the
  // domain is the domain shifted by its motion so far, in whole columns (Math.round: halves up).
  void addSwaths() {
    const int nx = g.nx, ny = g.ny;
    const double fu = run.frame_u, fv = run.frame_v;
    const int sx = int(std::floor(fu * time / g.dx + .5)), sy = int(std::floor(fv * time / g.dy + .5));
    float *gust = col[ORC_GUST_MAX].data(), *uhm = col[ORC_UH_MAX].data();
    for (int y = 0; y < ny; y++) {
      const int gy = ((y + sy) % ny + ny) % ny;
      for (int x = 0; x < nx; x++) {
        const int gx = ((x + sx) % nx + nx) % nx, c = x + nx * y, gc = gx + nx * gy;
        const double wind = std::hypot(double(f[ORC_U][c]) + fu, double(f[ORC_V][c]) + fv);
        if (wind > gust[gc]) gust[gc] = float(wind);
        if (uhColumn[c] > uhm[gc]) uhm[gc] = uhColumn[c];
      }
    }
  }
  // countCores (model.ts:1650-1670): cloudy updraught columns between ~1.9 and 6.6 km, 4-connected across the
  // periodic edges.
  int countCores() const {
    const int nx = g.nx, ny = g.ny, nz = g.nz, layer = g.layer;
    int lo = -1, last = -1;
    for (int z = 0; z < nz; z++) { if (lo < 0 && g.zs[z] >= 1900) lo = z; if (g.zs[z] <= 6600) last = z; }
    const int hi = std::min(nz - 3, last);
    std::vector<unsigned char> mask(layer, 0);
    for (int c = 0; c < layer; c++)
      for (int z = std::max(lo, 0); lo >= 0 && z <= hi; z++) {
        const int i = c + z * layer;
        if (f[ORC_W][i] > 2 && double(f[ORC_CLOUD][i]) + double(f[ORC_ICE][i]) > .00007) { mask[c] = 1; break; }
      }
    int cores = 0; std::vector<int> stack;
    for (int i = 0; i < layer; i++) {
      if (mask[i] != 1) continue;
      cores++; mask[i] = 2; stack.push_back(i);
      while (!stack.empty()) {
        const int a = stack.back(); stack.pop_back();
        const int x = a % nx, y = a / nx;
        const int nb[4] = {(x + 1) % nx + nx * y, (x - 1 + nx) % nx + nx * y, x + nx * ((y + 1) % ny), x + nx * ((y - 1 + ny) % ny)};
        for (int b : nb) if (mask[b] == 1) { mask[b] = 2; stack.push_back(b); }
      }
    }
    return cores;
  }
  // outflowDeficit (rkw.ts:72-80): the ground air's thv as float32, the warm air its element floor((layer-1) 0.75) in
  // ascending order, the deficits as float32; returns the largest.
  double outflowDeficit(std::vector<float> &deficit, double &warm) const {
    const int layer = g.layer;
    std::vector<float> thv(layer), sorted;
    deficit.assign(layer, 0);
    for (int c = 0; c < layer; c++) thv[c] = float(double(f[ORC_THETA][c]) * (1 + .61 * double(f[ORC_Q][c])));
    sorted = thv; std::sort(sorted.begin(), sorted.end());
    warm = sorted[size_t(std::floor((layer - 1) * .75))];
    double max = 0;
    for (int c = 0; c < layer; c++) { deficit[c] = float(std::max(0.0, warm - thv[c])); max = std::max(max, double(deficit[c])); }
    return max;
  }
  // deficitNear (rkw.ts:82-90): the largest deficit within 2 km of column c, across the periodic edges.
  double deficitNear(const std::vector<float> &deficit, int c) const {
    const int nx = g.nx, ny = g.ny, x0 = c % nx, y0 = c / nx, rx = int(std::floor(2000 / g.dx)), ry = int(std::floor(2000 / g.dy));
    double max = 0;
    for (int j = -ry; j <= ry; j++) for (int i = -rx; i <= rx; i++) {
      if ((i * g.dx) * (i * g.dx) + (j * g.dy) * (j * g.dy) > 2000.0 * 2000.0) continue;
      max = std::max(max, double(deficit[((x0 + i) % nx + nx) % nx + nx * (((y0 + j) % ny + ny) % ny)]));
    }
    return max;
  }
  // diagnostics (model.ts:1672-1696) without terrain (the ground node of a column is its level 0), one-moment
  // precipitation (surfaceflux.ts:22-37), the outflow deficit (rkw.ts:72-90), in the order of ORC_MD_*.
  void diagnostics(double *out) const {
    const int nz = g.nz, layer = g.layer;
    const float *w = f[ORC_W].data(), *cloud = f[ORC_CLOUD].data(), *ice = f[ORC_ICE].data(), *cold = f[ORC_COLD].data();
    double up = 0, down = 0, top = 0, coldMax = 0, thermalTop = 0, maxCloud = 0;
    for (int z = 0, i = 0; z < nz; z++) for (int c = 0; c < layer; c++, i++) {
      const double alt = g.zs[z], ci = double(cloud[i]) + double(ice[i]);
      up = std::max(up, double(w[i])); down = std::max(down, -double(w[i])); maxCloud = std::max(maxCloud, ci);
      if (ci > .00007) top = std::max(top, alt / 1000);
      if (w[i] > .6) thermalTop = std::max(thermalTop, alt / 1000);
      if (alt < 1500) coldMax = std::max(coldMax, double(cold[i]));
    }
    const double u6 = env.prof.shear * std::tanh(6000 / env.prof.shear_depth) + windError(env.prof.wind_u, 6000), u0 = env.prof.shear * std::tanh(0 / env.prof.shear_depth) + windError(env.prof.wind_u, 0);
    const double v6 = windError(env.prof.wind_v, 6000), v0 = windError(env.prof.wind_v, 0);
    const double rho0 = env.rho[0];
    double rainRate = 0, rainTotal = 0, gust = 0, hail = 0;
    for (int c = 0; c < layer; c++) gust = std::max(gust, std::hypot(double(f[ORC_U][c]) + run.frame_u, double(f[ORC_V][c]) + run.frame_v));
    for (int c = 0; c < layer; c++) {
      const double r = f[ORC_RAIN][c], sn = f[ORC_SNOW][c], gr = f[ORC_GRAUPEL][c], h = f[ORC_HAIL][c];
      const double vg = gr > 0 ? graupelFallSpeed(gr, rho0) : 0, vh = h > 0 ? hailFallSpeed(h, rho0) : 0;
      const double flux = rho0 * r * fallSpeedRain(r, rho0, rho0) + rho0 * (sn * snowFallSpeed(sn, rho0, rho0) + gr * vg + h * vh);
      rainRate = std::max(rainRate, flux * 3600); rainTotal = std::max(rainTotal, double(col[ORC_PRECIPITATION][c]));
    }
    for (int c = 0; c < layer; c++) hail = std::max(hail, hailMaxDiameter(f[ORC_HAIL][c], rho0) * 1000);
    std::vector<float> deficit; double warm;
    const double outflow = outflowDeficit(deficit, warm), outflowLow = deficitNear(deficit, int(rotation[ORC_ROT_LOW]));
    out[ORC_MD_UPDRAFT] = up; out[ORC_MD_DOWNDRAFT] = down; out[ORC_MD_RAINRATE] = rainRate; out[ORC_MD_RAINTOTAL] = rainTotal;
    out[ORC_MD_CLOUDTOP] = top; out[ORC_MD_THERMALTOP] = thermalTop; out[ORC_MD_MAXCLOUD] = maxCloud; out[ORC_MD_COLDMAX] = coldMax;
    out[ORC_MD_CORES] = countCores(); out[ORC_MD_SHEAR06] = std::hypot(u6 - u0, v6 - v0); out[ORC_MD_OUTFLOW] = outflow;
    out[ORC_MD_OUTFLOWLOW] = outflowLow; out[ORC_MD_GUST] = gust; out[ORC_MD_HAIL] = hail; out[ORC_MD_CLIPPED] = double(clipped);
  }
  void step(double dt) {
    updateGround();
    fallPhase();
    collectHail(dt);
    sedimentStep(dt);
    dynamics(dt);
    physicsStep(dt);
    mixStep(dt);
    project(dt, true);
    updateRotation(dt);
    if (electrified) discharge(dt);
  }
  void diag(double *out) {
    double maxw = -1e300, minw = 1e300, maxtp = -1e300, rainTotal = 0, water = 0;
    for (int z = 0, i = 0; z < g.nz; z++) for (int c = 0; c < g.layer; c++, i++) {
      maxw = std::max(maxw, double(f[ORC_W][i])); minw = std::min(minw, double(f[ORC_W][i]));
      maxtp = std::max(maxtp, f[ORC_THETA][i] - env.theta[z]);
    }
    for (int c = 0; c < g.layer; c++) rainTotal = std::max(rainTotal, double(col[ORC_PRECIPITATION][c]));
    for (int z = 0; z < g.nz; z++) {
      double level = 0;
      for (int i = z * g.layer; i < (z + 1) * g.layer; i++) level += double(f[ORC_Q][i]) + f[ORC_CLOUD][i] + f[ORC_ICE][i] + f[ORC_RAIN][i] + f[ORC_SNOW][i] + f[ORC_GRAUPEL][i] + f[ORC_HAIL][i];
      water += g.hz[z] * env.rho[z] * level / g.layer;
    }
    double fallen = 0; for (int c = 0; c < g.layer; c++) fallen += col[ORC_PRECIPITATION][c];
    water += fallen / g.layer;
    out[ORC_D_TIME] = time; out[ORC_D_MAXW] = maxw; out[ORC_D_MAXTHETAP] = maxtp; out[ORC_D_RAINTOTAL] = rainTotal; out[ORC_D_FIRSTFLASH] = firstFlash;
    out[ORC_D_FLASHES_IC] = flashesIC; out[ORC_D_FLASHES_CG] = flashesCG; out[ORC_D_CLIPPED] = double(clipped); out[ORC_D_NEGFILLED] = negativeFilled;
    out[ORC_D_CHARGEGROUND] = chargeGround; out[ORC_D_MINW] = minw; out[ORC_D_WATER] = water;
  }
};

}  // namespace

extern "C" {

void *orc_create(const orc_grid *grid, const orc_profile *profile, const orc_run *run) { return new Model(*grid, *profile, *run); }
void orc_destroy(void *m) { delete static_cast<Model *>(m); }
void orc_step(void *m, double dt, int n) {
  Model *mm = static_cast<Model *>(m);
  for (int k = 0; k < n; k++) { mm->step(dt); mm->time += dt; if (mm->swaths) mm->addSwaths(); }
}
void orc_set_swaths(void *m, int on) { static_cast<Model *>(m)->swaths = on != 0; }
void orc_add_swaths(void *m) { Model *mm = static_cast<Model *>(m); if (!mm->uhColumn.empty()) mm->addSwaths(); }
int orc_uh_column(void *m, float *out) {
  Model *mm = static_cast<Model *>(m);
  if (mm->uhColumn.empty()) return -1;
  std::memcpy(out, mm->uhColumn.data(), mm->uhColumn.size() * 4);
  return 0;
}
// One part of a step, split where the reference's GPU step splits it (src/gpu/step.ts:12-31): 0 beginStep with the
// sedimentation and the transport (GpuDynamics.step), 1 physicsStep (on the CPU there), 2 mixStep and the projection
// with the wind clipping (GpuDynamics.mix, endGpuStep), 3 afterStep's lightning and time += dt. Parts 0-3 in order
// are orc_step.
int orc_step_part(void *m, int part, double dt) {
  Model *mm = static_cast<Model *>(m);
  switch (part) {
    case 0: mm->updateGround(); mm->fallPhase(); mm->collectHail(dt); mm->sedimentStep(dt); mm->dynamics(dt); return 0;
    case 1: mm->physicsStep(dt); return 0;
    case 2: mm->mixStep(dt); mm->project(dt, true); return 0;
    case 3: mm->updateRotation(dt); if (mm->electrified) mm->discharge(dt); mm->time += dt; if (mm->swaths) mm->addSwaths(); return 0;
    default: return -1;
  }
}
void orc_set_lightning(void *m, int on) { static_cast<Model *>(m)->electrified = on != 0; }
int orc_get_field(void *m, int field, float *out) {
  Model *mm = static_cast<Model *>(m); if (field < 0 || field >= ORC_FIELDS) return -1;
  std::memcpy(out, mm->f[field].data(), sizeof(float) * mm->g.n); return 0;
}
int orc_set_field(void *m, int field, const float *in) {
  Model *mm = static_cast<Model *>(m); if (field < 0 || field >= ORC_FIELDS) return -1;
  std::memcpy(mm->f[field].data(), in, sizeof(float) * mm->g.n); return 0;
}
int orc_get_column(void *m, int column, float *out) {
  Model *mm = static_cast<Model *>(m); if (column < 0 || column >= ORC_COLUMNS) return -1;
  std::memcpy(out, mm->col[column].data(), sizeof(float) * mm->g.layer); return 0;
}
int orc_set_column(void *m, int column, const float *in) {
  Model *mm = static_cast<Model *>(m); if (column < 0 || column >= ORC_COLUMNS) return -1;
  std::memcpy(mm->col[column].data(), in, sizeof(float) * mm->g.layer); return 0;
}
void orc_base_state(void *m, double *out) {
  Model *mm = static_cast<Model *>(m); int nz = mm->g.nz; const Environment &e = mm->env;
  for (int z = 0; z < nz; z++) { out[z] = e.p[z]; out[nz + z] = e.exner[z]; out[2 * nz + z] = e.theta[z]; out[3 * nz + z] = e.q[z]; out[4 * nz + z] = e.u[z]; out[5 * nz + z] = e.v[z]; out[6 * nz + z] = e.rho[z]; }
}
void orc_grid_levels(void *m, double *zs, double *hz) { Model *mm = static_cast<Model *>(m); for (int z = 0; z < mm->g.nz; z++) { zs[z] = mm->g.zs[z]; hz[z] = mm->g.hz[z]; } }
void orc_diag(void *m, double *out) { static_cast<Model *>(m)->diag(out); }
void orc_pressure(void *m, double *out) { Model *mm = static_cast<Model *>(m); std::memcpy(out, mm->pressure.data(), sizeof(double) * mm->pressure.size()); }
void orc_set_time(void *m, double t) { static_cast<Model *>(m)->time = t; }
// intervene (intervene.ts:38-68) with the patch table INTERVENTIONS (intervene.ts:22-30): a smooth patch round on the
// map (e-folding radius, the thermal's own bubbleRadius, default 4.2 km) fading with height, added to the fields as
// they are; the heat (J) and vapour (kg) put in are counted into AtmosphereModel.added. Flat ground (no solid nodes).
int orc_intervene(void *m, int kind, double x, double y, double strength, double thermal_radius, double *out) {
  struct Patch { double radius, depth, top, theta, vapour, w, spin; };
  static const Patch table[7] = {
      {-1, 1800, 2700, 3.2, 0, 1.1, 0},  // thermal
      {5000, 1000, -1, 2, 0, 0, 0},      // warm
      {5000, 1000, -1, -2, 0, 0, 0},     // cool
      {5000, 1200, -1, 0, .002, 0, 0},   // moisten
      {5000, 1200, -1, 0, -.002, 0, 0},  // dry
      {4000, 700, -1, -6, 0, 0, 0},      // coldPool
      {3000, 3000, -1, 0, 0, 0, 10},     // vortex
  };
  if (kind < 0 || kind > 6) return -1;
  Model *mm = static_cast<Model *>(m);
  const Grid &g = mm->g; const Environment &e = mm->env; const Patch &k = table[kind];
  const double MOIST_CAP = .95, REACH = 3;
  const double radius = k.radius > 0 ? k.radius : (thermal_radius > 0 ? thermal_radius : 4.2) * 1000, a0 = strength;
  const double top = k.top > 0 ? k.top : k.depth * REACH, W = g.width, D = g.depth;
  double heat = 0, vapour = 0;
  for (int z = 0; z < g.nz && g.zs[z] < top; z++) {
    const double mass = e.rho[z] * g.dx * g.dy * g.hz[z], exner = e.exner[z], p = e.p[z];
    for (int yy = 0; yy < g.ny; yy++) for (int xx = 0; xx < g.nx; xx++) {
      double ddx = xx * g.dx - x, ddy = yy * g.dy - y;
      if (ddx > W / 2) ddx -= W;
      if (ddx < -W / 2) ddx += W;
      if (ddy > D / 2) ddy -= D;
      if (ddy < -D / 2) ddy += D;
      const int i = xx + g.nx * (yy + g.ny * z);
      const double r2 = (ddx / radius) * (ddx / radius) + (ddy / radius) * (ddy / radius);
      const double a = std::exp(-r2 - (g.zs[z] / k.depth) * (g.zs[z] / k.depth)) * a0;
      float &theta = mm->f[ORC_THETA][i], &q = mm->f[ORC_Q][i];
      if (k.theta != 0) { theta = float(theta + k.theta * a); heat += CP * exner * k.theta * a * mass; }
      if (k.w != 0) mm->f[ORC_W][i] = float(mm->f[ORC_W][i] + k.w * a);
      if (k.vapour != 0) {
        const double before = q, cap = MOIST_CAP * qsatP(theta * exner - 273.15, p);
        q = float(k.vapour > 0 ? std::max(before, std::min(cap, before + k.vapour * a)) : std::max(0.0, before + k.vapour * a));
        vapour += (q - before) * mass;
      }
      if (k.spin != 0) {
        const double r = std::sqrt(r2), zd = g.zs[z] / k.depth;
        const double speed = k.spin * a0 * r * std::exp((1 - r2) / 2) * std::exp(-(zd * zd)), d = std::hypot(ddx, ddy);
        if (d > 0) { mm->f[ORC_U][i] = float(mm->f[ORC_U][i] - speed * ddy / d); mm->f[ORC_V][i] = float(mm->f[ORC_V][i] + speed * ddx / d); }
      }
    }
  }
  mm->addedHeat += heat; mm->addedVapour += vapour; mm->addedCount += 1;
  if (out) { out[0] = heat; out[1] = vapour; }
  return 0;
}
void orc_added(void *m, double *out) { Model *mm = static_cast<Model *>(m); out[0] = mm->addedHeat; out[1] = mm->addedVapour; out[2] = mm->addedCount; }
void orc_update_rotation(void *m, double dt, int keep) {
  Model *mm = static_cast<Model *>(m);
  const double held = mm->rotation[ORC_ROT_PERSISTED];
  mm->updateRotation(dt);
  if (keep) mm->rotation[ORC_ROT_PERSISTED] = held;
}
void orc_model_diagnostics(void *m, double *out) { static_cast<Model *>(m)->diagnostics(out); }
double orc_outflow_deficit(void *m, float *deficit, double *warm) {
  std::vector<float> d; double w;
  const double max = static_cast<Model *>(m)->outflowDeficit(d, w);
  if (deficit) std::memcpy(deficit, d.data(), d.size() * 4);
  if (warm) *warm = w;
  return max;
}
double orc_deficit_near(void *m, const float *deficit, int column) {
  Model *mm = static_cast<Model *>(m);
  return mm->deficitNear(std::vector<float>(deficit, deficit + mm->g.layer), column);
}
void orc_rotation(void *m, double *out) { std::memcpy(out, static_cast<Model *>(m)->rotation, sizeof(double) * ORC_ROT); }

double orc_weno(double a, double b, double c, double d, double e) { return weno(a, b, c, d, e); }
double orc_upwind5(double a, double b, double c, double d, double e) { return upwind5(a, b, c, d, e); }
void orc_fft(int n, double *re, double *im, int sign) { FFT(n).transform(re, im, sign); }
void orc_hdft_forward(int nx, int ny, int layers, const double *src, double *re, double *im) {
  HorizontalDFT d(nx, ny, layers); d.forwardLayers([&](int i) { return src[i]; }, 0, layers);
  std::memcpy(re, d.re.data(), sizeof(double) * d.re.size()); std::memcpy(im, d.im.data(), sizeof(double) * d.im.size());
}
void orc_hdft_inverse(int nx, int ny, int layers, double *re, double *im, double *dst) {
  HorizontalDFT d(nx, ny, layers); std::memcpy(d.re.data(), re, sizeof(double) * d.re.size()); std::memcpy(d.im.data(), im, sizeof(double) * d.im.size());
  d.inverseLayers([&](int i, double v) { dst[i] = v; }, 0, layers);
}
double orc_qsat(double tc, double p) { return qsatP(tc, p); }
double orc_qsat_ice(double tc, double p) { return qsatIce(tc, p); }
double orc_fall_speed(double rain, double rho, double rg) { return fallSpeedRain(rain, rho, rg); }
double orc_snow_fall_speed(double s, double rho, double rg) { return snowFallSpeed(s, rho, rg); }
double orc_graupel_fall_speed(double g, double rho) { return graupelFallSpeed(g, rho); }
double orc_hail_fall_speed(double h, double rho) { return hailFallSpeed(h, rho); }
double orc_rain_processes(float *q, float *cloud, float *rain, double tk, double p, double rho, double dt) { return rainProcesses(*q, *cloud, *rain, tk, p, rho, dt); }
double orc_saturation_adjust(float *theta, float *q, float *cloud, double exner, double p) { return saturationAdjust(*theta, *q, *cloud, exner, p); }
double orc_takahashi(double tc, double lwc) { return takahashiCharge(tc, lwc); }
double orc_takahashi_collision(double tc, double lwc, double d, double dv) { return takahashiScaled(takahashiCharge(tc, lwc), d, dv); }
int orc_flash_limit(double dt) { return flashLimit(dt); }
double orc_stretch_ratio(double b, double h, int m) { return stretchRatio(b, h, m); }
double orc_project(const orc_grid *grid, const double *rho, float *u, float *v, float *w, double dt, double *div_before) {
  Grid g(*grid); F64 r(rho, rho + g.nz); PressureSolver s(g, r);
  F64 div(s.cells), p(s.cells);
  s.divergence(u, v, w, dt, div.data());
  double before = 0; for (double d : div) before = std::max(before, std::fabs(d));
  if (div_before) *div_before = before;
  s.solve(div.data(), p.data()); s.correct(p.data(), u, v, w, dt); s.divergence(u, v, w, dt, div.data());
  double after = 0; for (double d : div) after = std::max(after, std::fabs(d));
  return after;
}
void orc_advect(const orc_grid *grid, const double *rho, const float *u, const float *v, const float *w, float *q, double dt, int steps, int monotone) {
  Grid g(*grid); F64 r(rho, rho + g.nz); FluxTransport t(g, r);
  t.setVelocity(u, v, w);
  F32 st(g.n); F64 tend(g.n);
  for (int s = 0; s < steps; s++) {
    std::copy(q, q + g.n, st.begin());
    for (int k = 0; k < 3; k++) { t.tendency(q, tend.data(), monotone != 0); for (int i = 0; i < g.n; i++) q[i] = float(st[i] + RK3[k] * dt * tend[i]); }
  }
}
// ---- the column physics' pieces on one node (ice.ts, graupel.ts, hail.ts, charge.ts) ---------------------------------
double orc_sat_adjust_mixed(float *theta, float *q, float *cloud, float *ice, double exner, double p, float *number) {
  float spare = 0;
  return saturationAdjustMixed(*theta, *q, *cloud, *ice, exner, p, number ? *number : spare);
}
void orc_freeze_rain(float *theta, float *rain, float *graupel, double exner) { freezeRain(*theta, *rain, *graupel, exner); }
double orc_ice_processes(float *theta, float *q, float *ice, float *snow, float *number, double exner, double p, double rho, double dt) {
  return iceProcesses(*theta, *q, *ice, *snow, *number, exner, p, rho, dt);
}
void orc_settle_crystals(float *ice, float *number) { settleCrystals(*ice, *number); }
double orc_snow_processes(float *theta, float *q, float *cloud, float *ice, float *snow, float *rain, double exner, double p, double rho, double rho_ground, double dt) {
  return snowProcesses(*theta, *q, *cloud, *ice, *snow, *rain, exner, p, rho, rho_ground, dt);
}
double orc_graupel_processes(float *theta, float *q, float *cloud, float *ice, float *snow, float *rain, float *graupel, double exner, double p, double rho, double dt) {
  return graupelProcesses(*theta, *q, *cloud, *ice, *snow, *rain, *graupel, exner, p, rho, dt);
}
double orc_hail_processes(float *theta, float *q, float *cloud, float *ice, float *snow, float *rain, float *graupel, float *hail, double exner, double p, double rho, double dt) {
  return hailProcesses(*theta, *q, *cloud, *ice, *snow, *rain, *graupel, *hail, exner, p, rho, dt);
}
double orc_hail_max_diameter(double h, double rho) { return hailMaxDiameter(h, rho); }
int orc_wet_growing(int rimer, double q, double tc, double qv, double p, double rho, double cloud, double rain, double ice, double snow) {
  return wetGrowing(q, rimer ? N0H : N0G, rimer ? RHO_H : RHO_G, tc, qv, p, rho, cloud, rain, ice, snow) ? 1 : 0;
}
double orc_cooper_crystals(double tc) { return cooperCrystals(tc); }
double orc_splintering(double tc) { return splintering(tc); }
double orc_fragments_per_kg(double lambda) { return fragmentsPerKg(lambda); }
double orc_koenig_growth(double tc, double m) { return koenigGrowth(tc, m); }
double orc_crystal_size(double qi, double n, double rho) { return crystalSize(qi, n, rho); }
double orc_constant(int which) { const double c[] = {LF, LS, CP, KAPPA, SPLINTERS, CRYSTAL_MASS, LV}; return which >= 0 && which < 7 ? c[which] : 0; }
double orc_charging_rate(double tc, double rho, double cloud, double ice, double snow, double graupel, double hail, double snow_fall, int wet_g, int wet_h, double crystals, int have_crystals) {
  return chargingRate(tc, rho, cloud, ice, snow, graupel, hail, snow_fall, wet_g != 0, wet_h != 0, crystals, have_crystals != 0);
}
double orc_inductive_rate(double tc, double rho, double cloud, double graupel, double hail, double ez, double rimer_charge, int wet_g, int wet_h) {
  return inductiveRate(tc, rho, cloud, graupel, hail, ez, rimer_charge, wet_g != 0, wet_h != 0);
}
void orc_transfer_charge(double *charge, const double *before, const double *after) { transferCharge(charge, before, after); }
double orc_crystal_number(double tc, double qi, double rho) { return crystalNumber(tc, qi, rho); }
double orc_crystal_diameter(double qi, double rho, double tc) { return crystalDiameter(qi, rho, tc); }

// ---- turbulence and sedimentation on their own ------------------------------------------------------------------------
void orc_viscosity(const orc_grid *grid, const float *u, const float *v, const float *w, const float *theta, const double *theta_env, double cs, double dt,
                   const float *q, const float *condensate, const double *exner, double *km, double *kh, int anisotropic) {
  Grid g(*grid); viscosityOf(g, u, v, w, theta, theta_env, cs, dt, q, condensate, exner, km, kh, anisotropic != 0);
}
void orc_mix(const orc_grid *grid, const double *rho, const double *km, const double *kh, float *a, const double *base, double factor, double dt) {
  Grid g(*grid); F32 t(g.n); mixOf(g, rho, km, kh, a, base, factor, dt, t.data());
}
int orc_fall_parts(void *m, int field, int fall, double dt) {
  Model *mm = static_cast<Model *>(m);
  if (field < 0 || field >= ORC_FIELDS || fall < 0 || fall >= ORC_FIELDS) return -1;
  return mm->fallParts(mm->F(field), mm->F(fall), dt);
}
int orc_sediment(void *m, int field, int fall, double dt, int deposit) {
  Model *mm = static_cast<Model *>(m);
  if (field < 0 || field >= ORC_FIELDS || fall < 0 || fall >= ORC_FIELDS) return -1;
  mm->updateGround();
  mm->sediment(field, fall, dt, deposit != 0);
  return 0;
}
double orc_charge_ground(void *m) { return static_cast<Model *>(m)->chargeGround; }

// ---- lightning on its own (lightning.ts Lightning) ---------------------------------------------------------------------
struct OrcLightning { Grid g; F64 air; Lightning l; std::vector<Flash> last; OrcLightning(const orc_grid &o, const double *a, int seed) : g(o), air(a, a + g.nz), l(g, air, seed) {} };
void *orc_lightning_create(const orc_grid *grid, const double *air, int seed) { return new OrcLightning(*grid, air, seed); }
void orc_lightning_destroy(void *l) { delete static_cast<OrcLightning *>(l); }
void orc_lightning_solve(void *l, float *ice, float *graupel, float *rain, float *air) {
  float *c[4] = {ice, graupel, rain, air}; static_cast<OrcLightning *>(l)->l.solve(c);
}
double orc_lightning_strongest(void *l, int *node) {
  auto r = static_cast<OrcLightning *>(l)->l.strongest(-1);
  if (node) *node = r.first;
  return r.second;
}
int orc_lightning_discharge(void *l, float *ice, float *graupel, float *rain, float *air, double time, int limit, double *records, int max) {
  OrcLightning *o = static_cast<OrcLightning *>(l);
  float *c[4] = {ice, graupel, rain, air};
  o->last = o->l.discharge(c, time, limit);
  int n = int(o->last.size());
  for (int k = 0; k < n && k < max; k++) {
    const Flash &f = o->last[k]; double *r = records + size_t(k) * 8;
    r[0] = f.time; r[1] = f.cg ? 1 : 0; r[2] = f.start; r[3] = f.nodes; r[4] = f.potential; r[5] = f.charge; r[6] = f.moment; r[7] = f.ground;
  }
  return n;
}
int orc_lightning_channel(void *l, int flash, int *nodes, int *parent, int *polarity, int max) {
  OrcLightning *o = static_cast<OrcLightning *>(l);
  if (flash < 0 || flash >= int(o->last.size())) return -1;
  const Flash &f = o->last[flash]; int n = int(f.channel.size());
  for (int k = 0; k < n && k < max; k++) { nodes[k] = f.channel[k]; parent[k] = f.parent[k]; polarity[k] = k < int(f.polarity.size()) ? f.polarity[k] : 0; }
  return n;
}
double orc_lightning_conduct(void *l, float *air_charge, double dt) { return static_cast<OrcLightning *>(l)->l.conduct(air_charge, dt); }
void orc_lightning_set_cloudy(void *l, const float *cloudy) { static_cast<OrcLightning *>(l)->l.setCloudy(cloudy); }
void orc_lightning_set_corona(void *l, int on) { static_cast<OrcLightning *>(l)->l.corona = on != 0; }
void orc_lightning_phi(void *l, double *out) { OrcLightning *o = static_cast<OrcLightning *>(l); std::memcpy(out, o->l.phi.data(), sizeof(double) * o->g.n); }
void orc_lightning_totals(void *l, double *from_ground, double *conducted) {
  OrcLightning *o = static_cast<OrcLightning *>(l);
  if (from_ground) *from_ground = o->l.fromGround;
  if (conducted) *conducted = o->l.conducted;
}

uint32_t orc_mulberry32(uint32_t *state) { double v = mulberry(*state); return uint32_t(v * 4294967296.0); }
void orc_electric_solve(const orc_grid *grid, const double *rho, double *phi) { Grid g(*grid); ElectricSolver s(g); s.solve(rho, phi); }

}  // extern "C"
