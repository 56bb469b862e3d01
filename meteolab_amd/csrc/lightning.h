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

// rho = (sum of the carriers) * air density at node i; at the interior levels -rho/eps0 into `cells` (level k at
// layer k - 1).
template <class R> __device__ inline void chargeAt(const Geo &g, const Bufs<R> &b, int slot, int i) {
  int z = i / g.layer;
  double s = 0;
  s += b.f(slot, F_CHARGE_ICE, g)[i]; s += b.f(slot, F_CHARGE_GRAUPEL, g)[i]; s += b.f(slot, F_CHARGE_RAIN, g)[i]; s += b.f(slot, F_CHARGE_AIR, g)[i];
  double rho = s * b.lv(slot, L_RHO, g)[z];
  b.arr(b.lrho, slot, g)[i] = R(rho);
  if (z > 0 && z < g.nz - 1) b.arr(b.cells, slot, g)[i - g.layer] = -rho / EPSILON0;
}
template <class R> __global__ __launch_bounds__(256) void k_charge_rhs(Geo g, Bufs<R> b) {
  int slot = b.runs[blockIdx.y];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < g.n) chargeAt(g, b, slot, i);
}

// The largest (|E| / E_init)^2 over the interior levels, packed with its node (the first node in node order wins a tie,
// as the reference's strict comparison does) into the run's S_STRONGEST key.
template <class R> __global__ __launch_bounds__(256) void k_strongest(Geo g, Bufs<R> b) {
  int slot = b.runs[blockIdx.y];
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
  float cloudy = (b.f(slot, F_CLOUD, g)[i] + b.f(slot, F_ICE, g)[i]) + b.f(slot, F_SNOW, g)[i];
  return cloudy > float(CLOUDY) ? clear * R(CLOUD_SIGMA) : clear;
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
  double t = blockSum(lost, red);
  if (threadIdx.x == 0 && t != 0) atomicAdd(b.sc(slot) + S_CHARGE_GROUND, t);
}

// ---- the discharge (lightning.ts:261-420 discharge, flash, deposit, extend, neutralize), one block per run ----------
// The step's solve and strongest node come from the data-parallel kernels above; a run whose field exceeds the
// initiation threshold nowhere leaves at once. Otherwise the block grows the flashes as the reference does, node by
// node with the run's own mulberry32 stream, and solves the potential again itself (all levels by the block) after
// each stage of a flash and each flash. The channel's serial choices (the pick, the front's bookkeeping, the links of
// a new node) are made by thread 0; the weights, sums, deposits and solves by the whole block.
constexpr int DISCHARGE_THREADS = 1024;
constexpr double F_CRIT = .43;

struct FlashState {
  int size, count, ground, touched, left, brk, event, id, pickN, best;
  uint32_t rng;
  double sum, potential, before, moment, pick, ratio;
};

__device__ inline uint32_t mulberryNext(uint32_t &s, double &out) {
  s += 0x6D2B79F5u;
  uint32_t t = (s ^ (s >> 15)) * (1u | s);
  t = (t + ((t ^ (t >> 7)) * (61u | t))) ^ t;
  out = double(t ^ (t >> 14)) / 4294967296.0;
  return s;
}

template <class R> struct Discharge {
  const Geo &g; const Bufs<R> &b; const Plan &pl; int slot; char *smem; double *red; FlashState &st;
  float *car[4];
  const double *air, *einit, *cap;
  double *phi;
  R *rho;
  int *mark, *failed, *seen, *link, *front, *nodes, *parent;
  double *scale, *weight, *dep;

  __device__ Discharge(const Geo &g_, const Bufs<R> &b_, const Plan &p_, int s_, char *m_, double *r_, FlashState &st_)
      : g(g_), b(b_), pl(p_), slot(s_), smem(m_), red(r_), st(st_) {
    car[0] = b.f(slot, F_CHARGE_ICE, g); car[1] = b.f(slot, F_CHARGE_GRAUPEL, g); car[2] = b.f(slot, F_CHARGE_RAIN, g); car[3] = b.f(slot, F_CHARGE_AIR, g);
    air = b.lv(slot, L_RHO, g); einit = b.lv(slot, L_EINIT, g); cap = b.lv(slot, L_CAPACITANCE, g);
    phi = b.arr(b.ephi, slot, g); rho = b.arr(b.lrho, slot, g);
    mark = b.lm(slot, 0, g); failed = b.lm(slot, 1, g); seen = b.lm(slot, 2, g); link = b.lm(slot, 3, g); front = b.lm(slot, 4, g);
    scale = b.ln(slot, 0, g); weight = b.ln(slot, 1, g);
    nodes = b.chan + size_t(slot) * 2 * MAX_CHANNEL; parent = nodes + MAX_CHANNEL; dep = b.cdep + size_t(slot) * MAX_CHANNEL;
  }
  // Clear air for the channel: no charge to speak of, or no cloud (cloud + ice + snow, lightning.ts:320-322).
  __device__ bool clear(int i) const {
    float cloudy = (b.f(slot, F_CLOUD, g)[i] + b.f(slot, F_ICE, g)[i]) + b.f(slot, F_SNOW, g)[i];
    return fabs(double(rho[i])) < RHO_MIN || cloudy <= float(CLOUDY);
  }
  // The potential of the carriers as they now stand (Lightning.solve, lightning.ts:228-232), by this block alone.
  __device__ void solve() {
    for (int i = threadIdx.x; i < g.n; i += blockDim.x) chargeAt(g, b, slot, i);
    __syncthreads();
    const int L = g.nz - 2;
    for (int lev = 0; lev < L; lev++) forwardLevel(g, b, pl, slot, lev, smem, red);
    double *cp = weight;  // the front's weights are free between picks: elimination factors [k][mode]
    for (int mode = threadIdx.x; mode < g.nm; mode += blockDim.x) modeSolve(g, b, slot, mode, 2, L, cp + mode, g.nm);
    __syncthreads();
    for (int lev = 0; lev < L; lev++) inverseLevel(g, b, pl, slot, lev, 2, smem);
  }
  // The interior node with the largest |E|/E_init that has not failed this step (lightning.ts:235-248).
  __device__ void strongest(int stamp) {
    unsigned long long key = 0;
    for (int i = g.layer + threadIdx.x; i < g.n - g.layer; i += blockDim.x) {
      if (failed[i] == stamp) continue;
      int z = i / g.layer, y = (i - z * g.layer) / g.nx, x = i - z * g.layer - y * g.nx, row = y * g.nx + z * g.layer;
      int xp = x + 1 == g.nx ? 0 : x + 1, xm = x == 0 ? g.nx - 1 : x - 1, yp = y + 1 == g.ny ? 0 : y + 1, ym = y == 0 ? g.ny - 1 : y - 1;
      R ex = R((phi[xp + row] - phi[xm + row]) / (2 * g.dx)), ey = R((phi[x + yp * g.nx + z * g.layer] - phi[x + ym * g.nx + z * g.layer]) / (2 * g.dy));
      R ez = R((phi[i + g.layer] - phi[i - g.layer]) / (b.zs[z + 1] - b.zs[z - 1]));
      R lim = R(einit[z]);
      float r = float((ex * ex + ey * ey + ez * ez) / (lim * lim));
      unsigned long long k = r > 0 ? (static_cast<unsigned long long>(__float_as_uint(r)) << 32) | static_cast<unsigned int>(0xFFFFFFFFu - unsigned(i)) : 0ull;
      key = k > key ? k : key;
    }
    key = blockMax(key, reinterpret_cast<unsigned long long *>(red));
    if (threadIdx.x == 0) {
      st.best = key ? int(0xFFFFFFFFu - unsigned(key & 0xFFFFFFFFull)) : -1;
      st.ratio = key ? sqrt(double(__uint_as_float(unsigned(key >> 32)))) : 0;
    }
    __syncthreads();
  }
  // Puts the nodes beside channel node i (channel index c) on the front of flash `id` (lightning.ts:384-400). Thread 0.
  __device__ void extend(int i, int c, int id) {
    int z = i / g.layer, y = (i - z * g.layer) / g.nx, x = i - z * g.layer - y * g.nx;
    for (int k = z > 0 ? z - 1 : 0; k <= (z + 1 < g.nz - 2 ? z + 1 : g.nz - 2); k++) {
      double hz = b.zs[k] - b.zs[z], threshold = F_CRIT * .5 * (einit[z] + einit[k]);
      for (int dj = -1; dj <= 1; dj++) for (int di = -1; di <= 1; di++) {
        if (x + di < 0 || x + di >= g.nx || y + dj < 0 || y + dj >= g.ny) continue;
        int j = x + di + (y + dj) * g.nx + k * g.layer;
        if (mark[j] == id || (k >= z && fabs(double(rho[j])) < RHO_MIN)) continue;
        double s = sqrt((di * g.dx) * (di * g.dx) + (dj * g.dy) * (dj * g.dy) + hz * hz) * threshold;
        if (seen[j] != id) { seen[j] = id; front[st.size++] = j; }
        else if (s >= scale[j]) continue;
        scale[j] = s; link[j] = c;
      }
    }
  }
  // Adds node j to the channel (lightning.ts:293-303). Thread 0.
  __device__ void add(int j, int from) {
    mark[j] = st.id; nodes[st.count] = j; parent[st.count] = from; st.count++;
    if (j < g.layer) { st.ground = j; st.before = st.potential; st.potential = 0; st.touched = 1; return; }
    if (st.ground < 0) { st.sum += double(phi[j]); st.potential = st.sum / st.count; }
    extend(j, st.count - 1, st.id);
  }
  // Adds charge q at node i, from the carriers of the opposite sign in proportion to their charge (lightning.ts:406-413).
  __device__ void neutralize(int i, double q) {
    int z = i / g.layer;
    double perKg = q / (air[z] * g.dx * g.dy * b.hz[z]), opposite = 0, sq = q > 0 ? 1 : q < 0 ? -1 : 0;
    for (int c = 0; c < 4; c++) if (double(car[c][i]) * q < 0) opposite -= double(car[c][i]) * sq;
    if (opposite <= 0) return;
    for (int c = 0; c < 4; c++) if (double(car[c][i]) * q < 0) car[c][i] = float(double(car[c][i]) + perKg * fabs(double(car[c][i])) / opposite);
  }
  // The charge of the channel at potential V onto its nodes (lightning.ts:330-368); returns the neutralized (IC) or
  // lowered (CG) charge. Whole block.
  __device__ double deposit(double V, bool grounded) {
    double pos = 0, neg = 0;
    for (int c = threadIdx.x; c < st.count; c += blockDim.x) {
      int i = nodes[c], z = i / g.layer;
      dep[c] = 0;
      if (z == 0) continue;
      double q = cap[z] * (V - double(phi[i])), here = double(rho[i]) * g.dx * g.dy * b.hz[z];
      if (q * here >= 0) continue;
      double d = (q > 0 ? 1 : -1) * fmin(fabs(q), fabs(here));
      dep[c] = d;
      if (d > 0) pos += d; else neg -= d;
    }
    pos = blockSum(pos, red); neg = blockSum(neg, red);
    double charge, moment = 0;
    if (!grounded) {
      double both = fmin(pos, neg), sp = pos > 0 ? both / pos : 0, sn = neg > 0 ? both / neg : 0;
      for (int c = threadIdx.x; c < st.count; c += blockDim.x) dep[c] *= dep[c] > 0 ? sp : sn;
      charge = both;
    } else {
      charge = neg - pos;
      for (int c = threadIdx.x; c < st.count; c += blockDim.x) if (dep[c] != 0) moment -= dep[c] * b.zs[nodes[c] / g.layer];
      moment = blockSum(moment, red);
      if (threadIdx.x == 0) { st.moment += moment; b.sc(slot)[S_FROM_GROUND] -= charge; }
    }
    for (int c = threadIdx.x; c < st.count; c += blockDim.x) if (dep[c] != 0) neutralize(nodes[c], dep[c]);
    __syncthreads();
    return charge;
  }
  // Chooses the next link of the front (lightning.ts:307-321): weights by the block, the pick by a prefix sum.
  __device__ bool pickLink() {
    const int size = st.size, ground = st.ground;
    const double potential = st.potential, before = st.before;
    const int per = (size + int(blockDim.x) - 1) / int(blockDim.x), f0 = threadIdx.x * per, f1 = min(size, f0 + per);
    double local = 0;
    for (int f = f0; f < f1; f++) {
      int j = front[f];
      double pj = phi[j];
      double w = (ground >= 0 && (j < g.layer || pj * before <= 0)) || (clear(j) && (pj - potential) * potential >= 0) ? 0 : fabs(potential - pj) / scale[j] - 1;
      weight[f] = w > 0 ? w : 0; local += weight[f];
    }
    // Exclusive prefix of the threads' sums (wave scans, then the waves' totals).
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double inc = local;
    for (int o = 1; o < 64; o <<= 1) { double v = __shfl_up(inc, o, 64); if (lane >= o) inc += v; }
    __shared__ double waveTot[DISCHARGE_THREADS / 64];
    if (lane == 63) waveTot[wave] = inc;
    __syncthreads();
    double base = 0, total = 0;
    for (int w = 0; w < int(blockDim.x) / 64; w++) { if (w < wave) base += waveTot[w]; total += waveTot[w]; }
    if (threadIdx.x == 0) { st.pickN = size - 1; if (total > 0) { double r; mulberryNext(st.rng, r); st.pick = r * total; } }
    __syncthreads();
    if (total == 0) return false;
    double run = base + inc - local, pick = st.pick;
    for (int f = f0; f < f1; f++) { run += weight[f]; if (run >= pick) { atomicMin(&st.pickN, f); break; } }
    __syncthreads();
    return true;
  }
  // One flash from node `start` (lightning.ts:287-328). Returns its charge (0: none) and records its kind.
  __device__ double flash(int start, bool &cg, int &count, double &potentialOut, double &momentOut) {
    double neutralized = 0, lowered = 0;
    int strokes = 0;
    if (threadIdx.x == 0) {
      double *S = b.sc(slot);
      long long id; memcpy(&id, S + S_FLASH_ID, 8); id++; memcpy(S + S_FLASH_ID, &id, 8);
      st.id = int(id); st.size = 0; st.count = 0; st.sum = 0; st.potential = 0; st.before = 0; st.ground = -1; st.touched = 0; st.left = 0; st.moment = 0;
      add(start, -1);
    }
    __syncthreads();
    for (;;) {
      const int grown = st.count;
      if (threadIdx.x == 0) st.event = 0;
      __syncthreads();
      while (st.count < MAX_CHANNEL) {
        if (!pickLink()) break;
        if (threadIdx.x == 0) {
          st.brk = 0;
          int n = st.pickN, j = front[n], from = link[j];
          if (!st.left && st.ground < 0 && clear(j)) { st.left = 1; st.event = 1; st.brk = 1; }
          else {
            front[n] = front[--st.size];
            if (clear(nodes[from])) for (int f = st.size - 1; f >= 0; f--) if (link[front[f]] == from) front[f] = front[--st.size];
            add(j, from);
            if (st.touched) { st.touched = 0; st.event = 1; st.brk = 1; }
          }
        }
        __syncthreads();
        if (st.brk) break;
      }
      __syncthreads();
      const bool grounded = st.ground >= 0;
      double q = deposit(st.potential, grounded);
      if (grounded) lowered += q; else neutralized += q;
      if (grounded && (++strokes >= STROKES || (!st.event && st.count == grown))) break;
      if (!grounded && !st.event) break;
      if (st.count >= MAX_CHANNEL) break;
      solve();
      if (!grounded) {
        double s = 0;
        for (int c = threadIdx.x; c < st.count; c += blockDim.x) s += double(phi[nodes[c]]);
        s = blockSum(s, red);
        if (threadIdx.x == 0) { st.sum = s; st.potential = s / st.count; }
        __syncthreads();
      }
    }
    cg = st.ground >= 0; count = st.count; potentialOut = cg ? st.before : st.potential; momentOut = cg ? st.moment : 0;
    return cg ? lowered : neutralized;
  }
};

template <class R> __global__ __launch_bounds__(DISCHARGE_THREADS) void k_discharge(Geo g, Bufs<R> b, Plan pl, double dt, int limit) {
  int slot = b.runs[blockIdx.x];
  double *S = b.sc(slot);
  unsigned long long key0 = reinterpret_cast<unsigned long long *>(S)[S_STRONGEST];
  // |E| / E_init <= 1 everywhere: no flash this step (the common case).
  if (!(__uint_as_float(unsigned(key0 >> 32)) > 1.f)) return;
  extern __shared__ char smem[];
  __shared__ double red[DISCHARGE_THREADS / 64];
  __shared__ FlashState st;
  Discharge<R> d(g, b, pl, slot, smem, red, st);
  const double *P = b.pr(slot);
  const double time = P[P_TIME] + dt;
  if (threadIdx.x == 0) {
    long long id; memcpy(&id, S + S_FLASH_ID, 8); id++; memcpy(S + S_FLASH_ID, &id, 8);
    st.id = int(id);
    // The stream of this step: the seed (SimConfig.seed + 1) xor the model time in tenths of a second times 0x9E3779B1.
    uint32_t seed = uint32_t(int(P[P_SEED]) + 1), tt = uint32_t((long long)floor(time * 10 + .5));
    st.rng = seed ^ (tt * 0x9E3779B1u);
    st.best = int(0xFFFFFFFFu - unsigned(key0 & 0xFFFFFFFFull));
    st.ratio = sqrt(double(__uint_as_float(unsigned(key0 >> 32))));
  }
  __syncthreads();
  const int stamp = st.id;
  for (int k = 0; k <= limit; k++) {
    if (k > 0) d.solve();
    if (k == limit) break;
    if (k > 0) d.strongest(stamp);
    if (!(st.ratio > 1)) break;
    int start = st.best;
    bool cg; int count; double potential, moment;
    double charge = d.flash(start, cg, count, potential, moment);
    if (charge == 0) { if (threadIdx.x == 0) d.failed[start] = stamp; __syncthreads(); continue; }
    if (threadIdx.x == 0) {
      long long c; memcpy(&c, S + (cg ? S_FLASH_CG : S_FLASH_IC), 8); c++; memcpy(S + (cg ? S_FLASH_CG : S_FLASH_IC), &c, 8);
      if (cg) S[S_CHARGE_GROUND] += charge;
      if (S[S_FIRST_FLASH] < 0) S[S_FIRST_FLASH] = time;
      long long nlog; memcpy(&nlog, S + S_FLASH_LOG, 8);
      double *e = b.flog + (size_t(slot) * FLASH_LOG + size_t(nlog % FLASH_LOG)) * FLASH_FIELDS;
      e[FL_TIME] = time; e[FL_CG] = cg; e[FL_START] = start; e[FL_NODES] = count; e[FL_CHARGE] = charge; e[FL_POTENTIAL] = potential; e[FL_MOMENT] = moment;
      nlog++; memcpy(S + S_FLASH_LOG, &nlog, 8);
    }
    __syncthreads();
  }
}

}  // namespace ml
