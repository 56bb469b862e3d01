// The lightning discharge of the StormLab step (lightning.ts:261-420: discharge, flash, deposit, extend, neutralize),
// one block of DISCHARGE_THREADS per run whose field exceeds the initiation threshold somewhere.
//
// The block grows the flashes as the reference does, link by link with the run's own mulberry32 stream, and solves
// the potential again itself after each stage of a flash and after each flash (solveBlock: every interior level, a
// chunk of levels at a time through LDS). The front of a flash is a list of entries in the reference's order; each
// entry keeps its node, its link (channel index), its scale, the node's potential of the last solve and its clear-air
// and ground flags, so the weights of a pick read the front alone; a node's place in the front is kept beside its
// marks. A pick's serial decisions are thread 0's; the links of a new channel node are made by the lanes of wave 0,
// one neighbour each, and appended in the reference's loop order.
#pragma once
#include "lightning.h"

namespace ml {

constexpr int DISCHARGE_THREADS = 512;
constexpr double F_CRIT = .43;
constexpr int FF_CLEAR = 1, FF_GROUND = 2;  // flags of a front entry: clear air, ground level
constexpr int PICK_REGS = 8;                // front entries a thread keeps in registers between the two passes of a pick
constexpr int PRUNE_MAX = 32;               // links from one channel node (at most its 26 neighbours)

// The block's barrier in this file. The waves of a block hand data to each other through global memory here (the
// front's entries past its LDS part, the potential, the spectra, charges, marks), and __syncthreads() alone does not
// order those accesses: its workgroup-scope fences emit no wait for the wave's outstanding vector memory operations on
// gfx950 (the block's waves share one CU), and measured runs showed a wave's global store issued before the barrier
// unseen by another wave's load after it now and then (flash counts and potentials that changed with the code's
// layout and timing, identical runs in one launch agreeing). So each wave first waits for its own loads and stores.
__device__ inline void blockSync() {
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
}

// ---- the block's electric solve (electric.ts:36-60) ---------------------------------------------------------------
// The modes' tridiagonal systems (electric.ts:42-52) depend on the grid alone, so the pool computes their elimination
// once, as modeSolve does it: per interior level k and mode, the pivot, the upper factor and the lower coefficient,
// [3][nz - 2][nm].
__global__ void k_electric_factors(Geo g, const double *sym, const double *dzs, const double *hz, double *fac) {
  const int mode = blockIdx.x * blockDim.x + threadIdx.x, L = g.nz - 2;
  if (mode >= g.nm) return;
  const double a = sym[3 * g.nm + mode];
  double upper = 0;
  for (int k = 0; k < L; k++) {
    double lo, di, up, rf;
    modeRow<double>(2, k, L, a, 0.0, nullptr, dzs, hz, lo, di, up, rf);
    const double piv = k == 0 ? di : di - lo * upper;
    upper = up / piv;
    fac[size_t(k) * g.nm + mode] = piv;
    fac[(size_t(L) + k) * g.nm + mode] = upper;
    fac[(size_t(2 * L) + k) * g.nm + mode] = lo;
  }
}

// Complex values a level takes in LDS (its row pairs, then its kept columns), and the LDS of a chunk of C levels: two
// buffers and the twiddles.
__host__ __device__ inline int solveStride(const Geo &g) { return g.layer / 2 > g.nm ? g.layer / 2 : g.nm; }
__host__ __device__ inline size_t solveLdsBytes(const Geo &g, int C) { return (size_t(2) * C * solveStride(g) + g.nx + g.ny) * 16; }
// Between solves the same LDS holds the front of a flash: potential and scale (float64), node (int), link (int16, a
// channel index below MAX_CHANNEL) and flags, 23 bytes an entry. Entries beyond the capacity live in the run's global
// front arrays.
constexpr int FRONT_ENTRY_BYTES = 23;
// LDS pointers by address space, so that an LDS access is a DS instruction of its own (never a FLAT access that a select
// between an LDS and a global place would otherwise make of it).
template <class T> using Lds = __attribute__((address_space(3))) T;
__host__ __device__ inline int frontCapacity(const Geo &g, int C) { return int(solveLdsBytes(g, C) / FRONT_ENTRY_BYTES); }
static_assert(MAX_CHANNEL <= 32767, "a front entry's link is an int16");

// The potential of the carriers as they now stand (Lightning.solve, lightning.ts:228-232) by one block: the charge
// density of every node into lrho, the interior levels' transforms C levels at a time (the transforms of
// forwardLevel / inverseLevel, several levels per Stockham pass), the modes' columns with the pool's elimination
// (float64, KB levels' loads in flight per mode). Ends with a barrier.
// The right-hand side's spectra stay in ehat between the solves of a discharge. With `incremental` they are not
// transformed again: a deposit changes the charge of the channel's nodes only, and the change of node (x, y) of
// interior level l adds d exp(-2 pi i (m x / nx + n y / ny)) to mode (m, n) of level l (the transform is linear);
// chan[c] and delta[c] (c < count) are the channel's nodes and their right-hand sides' changes (lrho already new).
template <class R> __device__ void solveBlock(const Geo &g, const Bufs<R> &b, const Plan &pl, int slot, char *smem, int C, const int *chan,
                                              const double *delta, int count, bool incremental) {
  typedef cx<double> Z;
  const int nx = g.nx, ny = g.ny, mw = g.mw, nm = g.nm, layer = g.layer, half = ny / 2, hl = layer / 2, L = g.nz - 2;
  const int ls = solveStride(g);
  Z *A = reinterpret_cast<Z *>(smem), *B = A + C * ls, *tx = B + C * ls, *ty = tx + nx;
  for (int k = threadIdx.x; k < nx; k += blockDim.x) tx[k] = {b.twx[2 * k], b.twx[2 * k + 1]};
  for (int k = threadIdx.x; k < ny; k += blockDim.x) ty[k] = {b.twy[2 * k], b.twy[2 * k + 1]};
  const double *air = b.lv(slot, L_RHO, g);
  R *rho = b.arr(b.lrho, slot, g);
  double *spec = b.dsp(slot, g), *hat = b.eh(slot, g), *m0 = b.em0(slot, g), *phi = b.arr(b.ephi, slot, g);
  const int lane = threadIdx.x & 63, waves = blockDim.x >> 6;
  const Div dHl(hl), dNx(nx), dHm(half * mw), dMw(mw), dNm(nm), dNy(ny), dLayer(layer);
  // The changed nodes as (level, y, x) keys and changes, compacted in channel order into the transform buffers.
  int *keys = reinterpret_cast<int *>(A);
  double *ds = reinterpret_cast<double *>(keys + MAX_CHANNEL);
  if (incremental && size_t(2) * C * ls * sizeof(Z) < size_t(MAX_CHANNEL) * (sizeof(int) + sizeof(double))) incremental = false;
  if (incremental) {
    __shared__ int waveCount[DISCHARGE_THREADS / 64];
    int total = 0;
    for (int base = 0; base < count; base += blockDim.x) {
      const int c = base + threadIdx.x;
      bool keep = false;
      int key = 0;
      double d = 0;
      if (c < count) {
        d = delta[c];
        if (d != 0) {
          const int i = chan[c], z = dLayer.q(i), r = i - z * layer, y = dNx.q(r), x = r - y * nx;
          key = ((z - 1) << 24) | (y << 12) | x;
          keep = true;
        }
      }
      const unsigned long long bal = __ballot(keep);
      if (lane == 0) waveCount[threadIdx.x >> 6] = __popcll(bal);
      blockSync();
      int at = total, all = 0;
      for (int v = 0; v < waves; v++) { if (v < int(threadIdx.x >> 6)) at += waveCount[v]; all += waveCount[v]; }
      if (keep) { at += __popcll(bal & ((1ull << lane) - 1)); keys[at] = key; ds[at] = d; }
      total += all;
      blockSync();
    }
    // Each thread its modes: the changes of a level summed in registers, then added to the level's spectrum. The
    // changes are taken EU at a time: their keys, the twiddles' LDS loads and products first, then the sums in
    // change order (the same additions in the same order as one change at a time).
    constexpr int EU = 8;
    for (int mode = threadIdx.x; mode < nm; mode += blockDim.x) {
      const int n = dMw.q(mode), m = mode - n * mw;
      int cur = -1;
      double ar = 0, ai = 0, s0 = 0;
      auto flush = [&]() {
        if (cur >= 0) {
          double *o = hat + (size_t(cur) * nm + mode) * 2;
          o[0] += ar; o[1] += ai;
          if (mode == 0) m0[cur] += s0;
        }
      };
      auto twiddle = [&](int key) {
        const int y = (key >> 12) & 0xFFF, x = key & 0xFFF;
        const int mx = m * x - dNx.q(m * x) * nx, nyy = n * y - dNy.q(n * y) * ny;
        return cmul(tx[mx], ty[nyy]);
      };
      int e = 0;
      for (; e + EU <= total; e += EU) {
        int kq[EU];
        double dq[EU];
        Z wq[EU];
#pragma unroll
        for (int q = 0; q < EU; q++) { kq[q] = keys[e + q]; dq[q] = ds[e + q]; }
#pragma unroll
        for (int q = 0; q < EU; q++) wq[q] = twiddle(kq[q]);
#pragma unroll
        for (int q = 0; q < EU; q++) {
          const int l = kq[q] >> 24;
          if (l != cur) { flush(); cur = l; ar = 0; ai = 0; s0 = 0; }
          ar += dq[q] * wq[q].re; ai += dq[q] * wq[q].im; s0 += dq[q];
        }
      }
      for (; e < total; e++) {
        const int key = keys[e], l = key >> 24;
        const double d = ds[e];
        if (l != cur) { flush(); cur = l; ar = 0; ai = 0; s0 = 0; }
        const Z w = twiddle(key);
        ar += d * w.re; ai += d * w.im; s0 += d;
      }
      flush();
    }
    blockSync();
  } else {
  // The bottom and top levels' charge density (the interior levels' comes with their transforms).
  for (int t = threadIdx.x; t < 2 * layer; t += blockDim.x) {
    const int i = t < layer ? t : (g.nz - 1) * layer + (t - layer);
    rho[i] = R(chargeDensity(g, b, slot, i, air[i / layer]));
  }
  for (int l0 = 0; l0 < L; l0 += C) {
    const int cn = L - l0 < C ? L - l0 : C;
    // Right-hand side -rho / eps0 of level l0 + c (node level l0 + c + 1), rows 2p and 2p + 1 as one complex row; the
    // densities wait in B (free until the first pass) so that this loop reads global memory only and its loads overlap.
#pragma unroll 4
    for (int t = threadIdx.x; t < cn * hl; t += blockDim.x) {
      const int c = dHl.q(t), r = t - c * hl, p = dNx.q(r), x = r - p * nx, z = l0 + c + 1;
      const int i0 = z * layer + 2 * p * nx + x, i1 = i0 + nx;
      const double r0 = chargeDensity(g, b, slot, i0, air[z]), r1 = chargeDensity(g, b, slot, i1, air[z]);
      A[c * ls + r] = {-r0 / EPSILON0, -r1 / EPSILON0};
      B[c * ls + r] = {r0, r1};
    }
    blockSync();
    for (int t = threadIdx.x; t < cn * hl; t += blockDim.x) {
      const int c = dHl.q(t), r = t - c * hl, p = dNx.q(r), x = r - p * nx, z = l0 + c + 1;
      const int i0 = z * layer + 2 * p * nx + x;
      rho[i0] = R(B[c * ls + r].re); rho[i0 + nx] = R(B[c * ls + r].im);
    }
    // The level sums (the m = n = 0 mode's right-hand side) in a fixed order: wave w sums levels w, w + waves, ...
    for (int c = threadIdx.x >> 6; c < cn; c += waves) {
      double s = 0;
      for (int r = lane; r < hl; r += 64) s += A[c * ls + r].re + A[c * ls + r].im;
      s = waveSum(s);
      if (lane == 0) m0[l0 + c] = s;
    }
    blockSync();
    Z *X = fftStages(A, B, nx, cn * half, nx, 1, tx, pl.rx, pl.nrx, -1, half, ls);
    Z *Y = X == A ? B : A;
    for (int t = threadIdx.x; t < cn * half * mw; t += blockDim.x) {
      const int c = dHm.q(t), r = t - c * (half * mw), p = dMw.q(r), m = r - p * mw;
      const Z z = X[c * ls + p * nx + m], w = X[c * ls + p * nx + (m ? nx - m : 0)];
      Y[c * ls + 2 * p * mw + m] = {(z.re + w.re) / 2, (z.im - w.im) / 2};
      Y[c * ls + (2 * p + 1) * mw + m] = {(z.im + w.im) / 2, (w.re - z.re) / 2};
    }
    blockSync();
    Z *S = fftStages(Y, X, ny, cn * mw, 1, mw, ty, pl.ry, pl.nry, -1, mw, ls);
    for (int t = threadIdx.x; t < cn * nm; t += blockDim.x) {
      const int c = dNm.q(t), m = t - c * nm;
      double *o = hat + (size_t(l0 + c) * nm + m) * 2;
      o[0] = S[c * ls + m].re; o[1] = S[c * ls + m].im;
    }
    blockSync();
  }
  }
  // The modes' columns: forward elimination from the right-hand side's spectra into the potential's, then back
  // substitution.
  constexpr int KB = 8;
  const double *fp = b.efac, *fu = b.efac + size_t(L) * nm, *fl = b.efac + size_t(2 * L) * nm;
  for (int mode = threadIdx.x; mode < nm; mode += blockDim.x) {
    double pr = 0, pi = 0;
    for (int k0 = 0; k0 < L; k0 += KB) {
      double vr[KB], vi[KB], pv[KB], lo[KB];
#pragma unroll
      for (int q = 0; q < KB; q++) {
        const int k = k0 + q < L ? k0 + q : L - 1;
        const double *s = hat + (size_t(k) * nm + mode) * 2;
        vr[q] = mode ? s[0] : m0[k]; vi[q] = mode ? s[1] : 0.0;
        pv[q] = fp[size_t(k) * nm + mode]; lo[q] = fl[size_t(k) * nm + mode];
      }
#pragma unroll
      for (int q = 0; q < KB; q++) {
        if (k0 + q < L) {
          pr = (vr[q] - lo[q] * pr) / pv[q]; pi = (vi[q] - lo[q] * pi) / pv[q];
          double *s = spec + (size_t(k0 + q) * nm + mode) * 2;
          s[0] = pr; s[1] = pi;
        }
      }
    }
    for (int k1 = L - 2; k1 >= 0; k1 -= KB) {
      double vr[KB], vi[KB], cu[KB];
#pragma unroll
      for (int q = 0; q < KB; q++) {
        const int k = k1 - q >= 0 ? k1 - q : 0;
        const double *s = spec + (size_t(k) * nm + mode) * 2;
        vr[q] = s[0]; vi[q] = s[1]; cu[q] = fu[size_t(k) * nm + mode];
      }
#pragma unroll
      for (int q = 0; q < KB; q++) {
        if (k1 - q >= 0) {
          pr = vr[q] - cu[q] * pr; pi = vi[q] - cu[q] * pi;
          double *s = spec + (size_t(k1 - q) * nm + mode) * 2;
          s[0] = pr; s[1] = pi;
        }
      }
    }
  }
  blockSync();
  // Inverse transforms into the potential's interior levels, normalised by nx ny.
  const double norm = 1.0 / double(layer);
  const int mh = nx / 2;
  for (int l0 = 0; l0 < L; l0 += C) {
    const int cn = L - l0 < C ? L - l0 : C;
#pragma unroll 4
    for (int t = threadIdx.x; t < cn * nm; t += blockDim.x) {
      const int c = dNm.q(t), m = t - c * nm;
      const double *s = spec + (size_t(l0 + c) * nm + m) * 2;
      A[c * ls + m] = {s[0], s[1]};
    }
    blockSync();
    Z *Yv = fftStages(A, B, ny, cn * mw, 1, mw, ty, pl.ry, pl.nry, +1, mw, ls);
    Z *Zr = Yv == A ? B : A;
    for (int t = threadIdx.x; t < cn * half * nx; t += blockDim.x) {
      const int c = dHl.q(t), r = t - c * hl, p = dNx.q(r), m = r - p * nx;
      const int cc = m <= mh ? m : nx - m;
      const double sg = m <= mh ? 1.0 : -1.0;
      const bool real = cc == 0 || 2 * cc == nx;
      const Z a = Yv[c * ls + 2 * p * mw + cc], bb = Yv[c * ls + (2 * p + 1) * mw + cc];
      const double ar = a.re, ai = real ? 0.0 : sg * a.im, br = bb.re, bi = real ? 0.0 : sg * bb.im;
      Zr[c * ls + r] = {ar - bi, ai + br};
    }
    blockSync();
    Z *Xv = fftStages(Zr, Yv, nx, cn * half, nx, 1, tx, pl.rx, pl.nrx, +1, half, ls);
    for (int t = threadIdx.x; t < cn * hl; t += blockDim.x) {
      const int c = dHl.q(t), r = t - c * hl, p = dNx.q(r), x = r - p * nx;
      double *dst = phi + size_t(l0 + c + 1) * layer;
      dst[2 * p * nx + x] = Xv[c * ls + r].re * norm;
      dst[(2 * p + 1) * nx + x] = Xv[c * ls + r].im * norm;
    }
    blockSync();
  }
}

// ---- the flashes -----------------------------------------------------------------------------------------------

// A front entry as a pick hands it on.
struct FrontEntry { int node, link, flag; double phi, scale; };

struct FlashState {
  int size, count, ground, touched, left, brk, event, id, pickN, best, prune;
  int hat;  // the right-hand side's spectra of the charge before the last deposit are in ehat (the next solve updates them)
  // The step's failure mark (the flash id at the launch's start). Written once before the first barrier and never
  // after: every thread reads it from here, not from `id`, which thread 0 moves on at each flash's start with no
  // barrier between (a wave still to read `id` would then take the next id and miss the marks of failed starts).
  int stamp;
  uint32_t rng;
  double sum, potential, before, moment, ratio;
  FrontEntry picked, last;  // the picked entry and the front's last one (which takes the picked one's place)
};

__device__ inline uint32_t mulberryNext(uint32_t &s, double &out) {
  s += 0x6D2B79F5u;
  uint32_t t = (s ^ (s >> 15)) * (1u | s);
  t = (t + ((t ^ (t >> 7)) * (61u | t))) ^ t;
  out = double(t ^ (t >> 14)) / 4294967296.0;
  return s;
}

template <class R> struct Discharge {
  const Geo &g; const Bufs<R> &b; const Plan &pl; int slot, chunk; char *smem; double *red; FlashState &st;
  // The front: entries f < qcap in LDS (qphi .. qflag, carved from smem while no solve runs), the rest in the global
  // arrays fnode .. fscale (which also keep the LDS entries' node, link and scale while a solve runs).
  int qcap;
  Lds<double> *qphi, *qscale;
  Lds<int> *qnode;
  Lds<short> *qlink;
  Lds<unsigned char> *qflag;
  unsigned char *cflag;  // LDS: channel node c is clear air (as of the last solve)
  const double *zsL, *eiL;  // LDS: the level heights and E_init of the levels
  Div dLay, dRow;           // node index -> level, level place -> row
  int *prunes;           // LDS: the front places a prune takes off
  float *car[4];
  const float *cloudy;
  const double *air, *cap;
  double *phi;
  R *rho;
  int *mark, *failed, *seen, *pos, *fnode, *flink, *fflag, *nodes, *parent;
  double *fscale, *fphi, *fweight, *nscale, *dep;

  __device__ Discharge(const Geo &g_, const Bufs<R> &b_, const Plan &p_, int s_, int chunk_, char *m_, double *r_, FlashState &st_, unsigned char *cf_, int *pr_, int cap_,
                        const double *zs_, const double *ei_)
      : g(g_), b(b_), pl(p_), slot(s_), chunk(chunk_), smem(m_), red(r_), st(st_), qcap(cap_), cflag(cf_), zsL(zs_), eiL(ei_), dLay(g_.layer), dRow(g_.nx), prunes(pr_) {
    qphi = (Lds<double> *)(smem); qscale = qphi + qcap; qnode = (Lds<int> *)(qscale + qcap);
    qlink = (Lds<short> *)(qnode + qcap); qflag = (Lds<unsigned char> *)(qlink + qcap);
    car[0] = b.f(slot, F_CHARGE_ICE, g); car[1] = b.f(slot, F_CHARGE_GRAUPEL, g); car[2] = b.f(slot, F_CHARGE_RAIN, g); car[3] = b.f(slot, F_CHARGE_AIR, g);
    cloudy = b.f(slot, F_CLOUDY, g);
    air = b.lv(slot, L_RHO, g); cap = b.lv(slot, L_CAPACITANCE, g);
    phi = b.arr(b.ephi, slot, g); rho = b.arr(b.lrho, slot, g);
    mark = b.lm(slot, LM_MARK, g); failed = b.lm(slot, LM_FAILED, g); seen = b.lm(slot, LM_SEEN, g); pos = b.lm(slot, LM_POS, g);
    fnode = b.lm(slot, LM_FNODE, g); flink = b.lm(slot, LM_FLINK, g); fflag = b.lm(slot, LM_FFLAG, g);
    fscale = b.ln(slot, LN_FSCALE, g); fphi = b.ln(slot, LN_FPHI, g); fweight = b.ln(slot, LN_FWEIGHT, g); nscale = b.ln(slot, LN_NSCALE, g);
    nodes = b.chan + size_t(slot) * 2 * MAX_CHANNEL; parent = nodes + MAX_CHANNEL; dep = b.cdep + size_t(slot) * MAX_CHANNEL;
  }
  // Front entry f, wherever it lives.
  __device__ int nodeAt(int f) const { if (f < qcap) return qnode[f]; return fnode[f]; }
  __device__ int linkAt(int f) const { if (f < qcap) return int(qlink[f]); return flink[f]; }
  __device__ int flagAt(int f) const { if (f < qcap) return int(qflag[f]); return fflag[f]; }
  __device__ double phiAt(int f) const { if (f < qcap) return qphi[f]; return fphi[f]; }
  __device__ double scaleAt(int f) const { if (f < qcap) return qscale[f]; return fscale[f]; }
  __device__ FrontEntry entryAt(int f) const { return {nodeAt(f), linkAt(f), flagAt(f), phiAt(f), scaleAt(f)}; }
  __device__ void put(int f, int node, int link, int flag, double p, double s) {
    if (f < qcap) { qnode[f] = node; qlink[f] = short(link); qflag[f] = (unsigned char)flag; qphi[f] = p; qscale[f] = s; }
    else { fnode[f] = node; flink[f] = link; fflag[f] = flag; fphi[f] = p; fscale[f] = s; }
  }
  __device__ void relink(int f, int link, double s) {
    if (f < qcap) { qlink[f] = short(link); qscale[f] = s; } else { flink[f] = link; fscale[f] = s; }
  }
  // Clear air for the channel: no charge to speak of, or no cloud (cloud + ice + snow, lightning.ts:320-322).
  __device__ bool clear(int i) const {
    return fabs(double(rho[i])) < RHO_MIN || cloudy[i] <= float(CLOUDY);
  }
  __device__ int flagsOf(int i) const { return (clear(i) ? FF_CLEAR : 0) | (i < g.layer ? FF_GROUND : 0); }
  // The potential of the carriers as they now stand, then the front's and the channel's potentials and flags from it.
  __device__ void solve() {
    // The LDS entries' node, link and scale wait in the global arrays while the solve takes the LDS; each comes back
    // with its node's new potential and flags (each thread handles the same entries both times).
    const int size = st.size, kept = min(size, qcap);
    for (int f = threadIdx.x; f < kept; f += blockDim.x) { fnode[f] = qnode[f]; flink[f] = qlink[f]; fscale[f] = qscale[f]; }
    blockSync();
    solveBlock(g, b, pl, slot, smem, chunk, nodes, dep, st.count, st.hat != 0);
    if (threadIdx.x == 0) st.hat = 1;
    for (int f = threadIdx.x; f < size; f += blockDim.x) {
      const int j = fnode[f];
      put(f, j, flink[f], flagsOf(j), phi[j], fscale[f]);
    }
    for (int c = threadIdx.x; c < st.count; c += blockDim.x) cflag[c] = clear(nodes[c]) ? 1 : 0;
    blockSync();
  }
  // The interior node with the largest |E|/E_init that has not failed this step (lightning.ts:235-248).
  __device__ void strongest(int stamp) {
    // SU nodes a thread per round, every load of the round issued before any of them is used: one memory round trip a
    // round instead of one a node.
    constexpr int SU = 6;
    unsigned long long key = 0;
    const int lo = g.layer, hi = g.n - g.layer, nx = g.nx, ny = g.ny, layer = g.layer;
    for (int base = lo + threadIdx.x; base < hi; base += SU * int(blockDim.x)) {
      double pxp[SU], pxm[SU], pyp[SU], pym[SU], pzp[SU], pzm[SU], dzz[SU], lim2[SU];
      int fl[SU];
#pragma unroll
      for (int u = 0; u < SU; u++) {
        const int i0 = base + u * int(blockDim.x), i = i0 < hi ? i0 : lo;
        const int z = dLay.q(i), r = i - z * layer, y = dRow.q(r), x = r - y * nx, row = y * nx + z * layer;
        const int xp = x + 1 == nx ? 0 : x + 1, xm = x == 0 ? nx - 1 : x - 1, yp = y + 1 == ny ? 0 : y + 1, ym = y == 0 ? ny - 1 : y - 1;
        pxp[u] = phi[xp + row]; pxm[u] = phi[xm + row]; pyp[u] = phi[x + yp * nx + z * layer]; pym[u] = phi[x + ym * nx + z * layer];
        pzp[u] = phi[i + layer]; pzm[u] = phi[i - layer]; dzz[u] = zsL[z + 1] - zsL[z - 1]; lim2[u] = eiL[z];
        fl[u] = i0 < hi ? failed[i] : stamp;
      }
#pragma unroll
      for (int u = 0; u < SU; u++) {
        if (fl[u] == stamp) continue;
        const int i = base + u * int(blockDim.x);
        R ex = R((pxp[u] - pxm[u]) / (2 * g.dx)), ey = R((pyp[u] - pym[u]) / (2 * g.dy));
        R ez = R((pzp[u] - pzm[u]) / dzz[u]);
        R lim = R(lim2[u]);
        float r = float((ex * ex + ey * ey + ez * ez) / (lim * lim));
        unsigned long long k = r > 0 ? (static_cast<unsigned long long>(__float_as_uint(r)) << 32) | static_cast<unsigned int>(0xFFFFFFFFu - unsigned(i)) : 0ull;
        key = k > key ? k : key;
      }
    }
    key = blockMax(key, reinterpret_cast<unsigned long long *>(red));
    if (threadIdx.x == 0) {
      st.best = key ? int(0xFFFFFFFFu - unsigned(key & 0xFFFFFFFFull)) : -1;
      st.ratio = key ? sqrt(double(__uint_as_float(unsigned(key >> 32)))) : 0;
    }
    blockSync();
  }
  // The links of new channel node i (channel index c) to the nodes beside it (lightning.ts:388-404); wave 0, lane l
  // the l-th neighbour of the reference's loop (level, row, column). `size` is the front's size before. Returns
  // whether an entry it wrote or relinked lies past the front's LDS part (in global memory, read next by every wave).
  __device__ bool extend(int i, int c, int id, int size) {
    const int lane = threadIdx.x;
    const int z = dLay.q(i), r = i - z * g.layer, y = dRow.q(r), x = r - y * g.nx;
    const int k = z - 1 + lane / 9, dj = (lane / 3) % 3 - 1, di = lane % 3 - 1;
    bool fresh = false, better = false;
    int j = 0, at = 0, fl = 0;
    double s = 0, pj = 0;
    if (lane < 27 && k >= 0 && k <= g.nz - 2 && x + di >= 0 && x + di < g.nx && y + dj >= 0 && y + dj < g.ny) {
      j = x + di + (y + dj) * g.nx + k * g.layer;
      // Everything the decision reads, in one round of loads.
      const int mk = mark[j], sn = seen[j], ps = pos[j];
      const double rj = double(rho[j]), sj = nscale[j], phj = phi[j];
      const float cj = cloudy[j];
      // The link's scale from the level tables in LDS, while the node's loads are in flight.
      const double hz = zsL[k] - zsL[z], threshold = F_CRIT * .5 * (eiL[z] + eiL[k]);
      const double sl = sqrt((di * g.dx) * (di * g.dx) + (dj * g.dy) * (dj * g.dy) + hz * hz) * threshold;
      // Node i itself is on the channel (its mark was just set).
      if (!(j == i || mk == id || (k >= z && fabs(rj) < RHO_MIN))) {
        s = sl;
        if (sn != id) {
          fresh = true; pj = phj;
          fl = (fabs(rj) < RHO_MIN || cj <= float(CLOUDY) ? FF_CLEAR : 0) | (j < g.layer ? FF_GROUND : 0);
        }
        // Seen before: a shorter link replaces the entry's (a node seen and taken off the front does not return).
        else if (ps >= 0 && s < sj) { better = true; at = ps; }
      }
    }
    const unsigned long long nb = __ballot(fresh);
    if (fresh) {
      const int f = size + __popcll(nb & ((1ull << lane) - 1));
      put(f, j, c, fl, pj, s); nscale[j] = s; seen[j] = id; pos[j] = f;
    }
    if (better) { relink(at, c, s); nscale[j] = s; }
    if (lane == 0) st.size = size + __popcll(nb);
    return size + __popcll(nb) > qcap || __ballot(better && at >= qcap) != 0;
  }
  // Adds node j (flags fl, potential pj) to the channel from channel node `from` (lightning.ts:293-303): thread 0 the
  // channel, wave 0 the new links. Ends with a barrier. Of wave 0's global stores only front entries past the LDS part
  // are read by other waves before a later barrier that waits for them, so wave 0 waits here only when it wrote one.
  __device__ void add(int j, int from, int fl, double pj) {
    if (threadIdx.x < 64) {
      const int c = st.count, size = st.size, id = st.id;
      const bool grounded = j < g.layer;
      if (threadIdx.x == 0) {
        mark[j] = id; nodes[c] = j; parent[c] = from; cflag[c] = (fl & FF_CLEAR) ? 1 : 0; st.count = c + 1;
        if (grounded) { st.ground = j; st.before = st.potential; st.potential = 0; st.touched = 1; }
        else if (st.ground < 0) { st.sum += pj; st.potential = st.sum / (c + 1); }
      }
      if (!grounded && extend(j, c, id, size)) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    }
    __syncthreads();
  }
  // Takes entry n (node j) off the front: the last entry moves into its place (lightning.ts:329). Thread 0.
  __device__ void removeAt(int n, int j) {
    const int last = --st.size;
    if (n != last) {
      const FrontEntry e = entryAt(last);
      put(n, e.node, e.link, e.flag, e.phi, e.scale);
      pos[e.node] = n;
    }
    pos[j] = -1;
  }
  // The same with the last entry already at hand (the pick hands it on). Thread 0.
  __device__ void removeAt(int n, int j, const FrontEntry &e) {
    const int last = --st.size;
    if (n != last) {
      put(n, e.node, e.link, e.flag, e.phi, e.scale);
      pos[e.node] = n;
    }
    pos[j] = -1;
  }
  // Takes every link from channel node `from` off the front as the reference's loop from the end does
  // (lightning.ts:331): the block finds the places in order, thread 0 takes them off from the highest down.
  __device__ void prune(int from) {
    const int size = st.size;
    const int per = (size + int(blockDim.x) - 1) / int(blockDim.x), f0 = threadIdx.x * per, f1 = min(size, f0 + per);
    int mine = 0;
    for (int f = f0; f < f1; f++) mine += linkAt(f) == from;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int inc = mine;
    for (int o = 1; o < 64; o <<= 1) { int v = __shfl_up(inc, o, 64); if (lane >= o) inc += v; }
    __shared__ int waveCnt[DISCHARGE_THREADS / 64];
    if (lane == 63) waveCnt[wave] = inc;
    blockSync();
    int at = 0, total = 0;
    for (int w = 0; w < int(blockDim.x) / 64; w++) { if (w < wave) at += waveCnt[w]; total += waveCnt[w]; }
    at += inc - mine;
    for (int f = f0; f < f1; f++) if (linkAt(f) == from) { if (at < PRUNE_MAX) prunes[at] = f; at++; }
    blockSync();
    if (threadIdx.x == 0) {
      if (total > PRUNE_MAX) atomicOr(reinterpret_cast<unsigned long long *>(b.sc(slot) + S_ERROR), (unsigned long long)ERR_PRUNE);
      for (int k = min(total, PRUNE_MAX) - 1; k >= 0; k--) removeAt(prunes[k], nodeAt(prunes[k]));
    }
    blockSync();
  }
  // The charge of the channel at potential V onto its nodes (lightning.ts:330-368); returns the neutralized (IC) or
  // lowered (CG) charge. Whole block.
  __device__ void neutralize(int i, double q) {
    int z = i / g.layer;
    double perKg = q / (air[z] * g.dx * g.dy * b.hz[z]), opposite = 0, sq = q > 0 ? 1 : q < 0 ? -1 : 0;
    for (int c = 0; c < 4; c++) if (double(car[c][i]) * q < 0) opposite -= double(car[c][i]) * sq;
    if (opposite <= 0) return;
    for (int c = 0; c < 4; c++) if (double(car[c][i]) * q < 0) car[c][i] = float(double(car[c][i]) + perKg * fabs(double(car[c][i])) / opposite);
  }
  __device__ double deposit(double V, bool grounded) {
    double pos_ = 0, neg = 0;
    for (int c = threadIdx.x; c < st.count; c += blockDim.x) {
      int i = nodes[c], z = i / g.layer;
      dep[c] = 0;
      if (z == 0) continue;
      double q = cap[z] * (V - double(phi[i])), here = double(rho[i]) * g.dx * g.dy * b.hz[z];
      if (q * here >= 0) continue;
      double d = (q > 0 ? 1 : -1) * fmin(fabs(q), fabs(here));
      dep[c] = d;
      if (d > 0) pos_ += d; else neg -= d;
    }
    pos_ = blockSum(pos_, red); neg = blockSum(neg, red);
    double charge, moment = 0;
    if (!grounded) {
      double both = fmin(pos_, neg), sp = pos_ > 0 ? both / pos_ : 0, sn = neg > 0 ? both / neg : 0;
      for (int c = threadIdx.x; c < st.count; c += blockDim.x) dep[c] *= dep[c] > 0 ? sp : sn;
      charge = both;
    } else {
      charge = neg - pos_;
      for (int c = threadIdx.x; c < st.count; c += blockDim.x) if (dep[c] != 0) moment -= dep[c] * b.zs[nodes[c] / g.layer];
      moment = blockSum(moment, red);
      if (threadIdx.x == 0) { st.moment += moment; b.sc(slot)[S_FROM_GROUND] -= charge; }
    }
    // The carriers' change, and for the next solve each node's new charge density (lrho) and the change of its
    // right-hand side -rho / eps0 (interior levels; 0 elsewhere) in place of its deposit.
    for (int c = threadIdx.x; c < st.count; c += blockDim.x) {
      if (dep[c] == 0) continue;
      const int i = nodes[c], z = i / g.layer;
      const double before = chargeDensity(g, b, slot, i, air[z]);
      neutralize(i, dep[c]);
      const double after = chargeDensity(g, b, slot, i, air[z]);
      rho[i] = R(after);
      dep[c] = z >= 1 && z <= g.nz - 2 ? -(after - before) / EPSILON0 : 0.0;
    }
    blockSync();
    return charge;
  }
  // The weight of front entry f (lightning.ts:317-321).
  __device__ double weightAt(int f, int ground, double potential, double before) const {
    const double pj = phiAt(f);
    const int fl = flagAt(f);
    const double w = (ground >= 0 && ((fl & FF_GROUND) || pj * before <= 0)) || ((fl & FF_CLEAR) && (pj - potential) * potential >= 0) ? 0 : fabs(potential - pj) / scaleAt(f) - 1;
    return w > 0 ? w : 0;
  }
  // Chooses the next link of the front (lightning.ts:307-324): the weights of each thread's run of entries (the
  // entries kept in registers), their prefix over the block, the draw by thread 0, and the first entry whose prefix
  // reaches it; the threads holding the picked entry and the last one hand them on in st.picked and st.last.
  __device__ bool pickLink() {
    const int size = st.size, ground = st.ground;
    const double potential = st.potential, before = st.before;
    // The run's stream as it stands: every thread reads it before the first barrier, thread 0 advances it after it.
    const uint32_t rng0 = st.rng;
    const int per = (size + int(blockDim.x) - 1) / int(blockDim.x), f0 = threadIdx.x * per, f1 = min(size, f0 + per);
    double w[PICK_REGS], local = 0;
#pragma unroll
    for (int q = 0; q < PICK_REGS; q++) {
      const int f = f0 + q;
      w[q] = f < f1 ? weightAt(f, ground, potential, before) : 0.0;
      local += w[q];
    }
    for (int f = f0 + PICK_REGS; f < f1; f++) { const double v = weightAt(f, ground, potential, before); fweight[f] = v; local += v; }
    // Exclusive prefix of the threads' sums (wave scans, then the waves' totals).
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double inc = local;
    for (int o = 1; o < 64; o <<= 1) { double v = __shfl_up(inc, o, 64); if (lane >= o) inc += v; }
    __shared__ double waveTot[DISCHARGE_THREADS / 64];
    if (lane == 63) waveTot[wave] = inc;
    if (threadIdx.x == 0) st.pickN = size - 1;
    blockSync();
    double base = 0, total = 0;
    for (int v = 0; v < int(blockDim.x) / 64; v++) { if (v < wave) base += waveTot[v]; total += waveTot[v]; }
    if (total == 0) return false;
    // The draw: every thread takes it from the same state and the same total, so all hold the same value.
    uint32_t rs = rng0;
    double r;
    mulberryNext(rs, r);
    const double pick = r * total;
    double run = base + inc - local;
    int mine = -1;
#pragma unroll
    for (int q = 0; q < PICK_REGS; q++) {
      if (mine < 0 && f0 + q < f1) { run += w[q]; if (run >= pick) mine = f0 + q; }
    }
    for (int f = f0 + PICK_REGS; mine < 0 && f < f1; f++) { run += fweight[f]; if (run >= pick) mine = f; }
    // The first entry whose prefix reaches the draw: a thread's entries come before the next thread's, so a wave's is
    // its lowest lane's that found one, and the block's the least of the waves'.
    const unsigned long long any = __ballot(mine >= 0);
    if (any && lane == __ffsll((long long)any) - 1) atomicMin(&st.pickN, mine);
    blockSync();
    // Thread 0 advances the stream and hands on the picked entry and the front's last one (which takes the picked
    // one's place).
    if (threadIdx.x == 0) { st.rng = rs; st.picked = entryAt(st.pickN); st.last = entryAt(size - 1); }
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
      st.picked = {start, -1, flagsOf(start), phi[start], 0.0};
    }
    blockSync();
    add(st.picked.node, -1, st.picked.flag, st.picked.phi);
    for (;;) {
      const int grown = st.count;
      if (threadIdx.x == 0) st.event = 0;
      blockSync();
      while (st.count < MAX_CHANNEL) {
        if (!pickLink()) break;
        if (threadIdx.x == 0) {
          st.brk = 0; st.prune = -1;
          const FrontEntry p = st.picked;
          // The first step out of the cloud ends the stage; the link stays on the front.
          if (!st.left && st.ground < 0 && (p.flag & FF_CLEAR)) { st.left = 1; st.event = 1; st.brk = 1; }
          else {
            // The last entry moves into the picked one's place: other waves read it next if that is in global memory.
            const bool glob = st.pickN >= qcap && st.pickN != st.size - 1;
            removeAt(st.pickN, p.node, st.last);
            if (cflag[p.link]) st.prune = p.link;
            if (glob) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
          }
        }
        __syncthreads();
        if (st.brk) break;
        if (st.prune >= 0) prune(st.prune);
        add(st.picked.node, st.picked.link, st.picked.flag, st.picked.phi);
        if (st.touched) {
          blockSync();
          if (threadIdx.x == 0) { st.touched = 0; st.event = 1; }
          blockSync();
          break;
        }
      }
      blockSync();
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
        blockSync();
      }
    }
    cg = st.ground >= 0; count = st.count; potentialOut = cg ? st.before : st.potential; momentOut = cg ? st.moment : 0;
    return cg ? lowered : neutralized;
  }
};

// The discharge of each run of the launch (lightning.ts:261-276): one block per run, `chunk` levels a solve pass.
// advanced: the run's clock has already moved on by dt (the discharge of a step run beside the next step's dynamics).
template <class R> __global__ __launch_bounds__(DISCHARGE_THREADS) void k_discharge(Geo g, Bufs<R> b, Plan pl, double dt, int limit, int chunk, int advanced, int cap) {
  int slot = b.runs[blockIdx.x];
  double *S = b.sc(slot);
  unsigned long long key0 = reinterpret_cast<unsigned long long *>(S)[S_STRONGEST];
  // |E| / E_init <= 1 everywhere: no flash this step (the common case).
  if (!(__uint_as_float(unsigned(key0 >> 32)) > 1.f)) return;
  extern __shared__ char smem[];
  __shared__ double red[DISCHARGE_THREADS / 64];
  __shared__ FlashState st;
  __shared__ unsigned char cflag[MAX_CHANNEL];
  __shared__ int prunes[PRUNE_MAX];
  __shared__ double zsL[64], eiL[64];  // nz <= 64 (pool creation checks it)
  for (int k = threadIdx.x; k < g.nz; k += blockDim.x) { zsL[k] = b.zs[k]; eiL[k] = b.lv(slot, L_EINIT, g)[k]; }
  Discharge<R> d(g, b, pl, slot, chunk, smem, red, st, cflag, prunes, cap, zsL, eiL);
  const double *P = b.pr(slot);
  const double time = advanced ? P[P_TIME] : P[P_TIME] + dt;
  if (threadIdx.x == 0) {
    long long id; memcpy(&id, S + S_FLASH_ID, 8); id++; memcpy(S + S_FLASH_ID, &id, 8);
    // The step's electric solve kept its right-hand side's spectra and level sums: every solve here is an update.
    st.id = int(id); st.stamp = int(id); st.size = 0; st.count = 0; st.hat = 1;
    // The stream of this step: the seed (SimConfig.seed + 1) xor the model time in tenths of a second times 0x9E3779B1.
    uint32_t seed = uint32_t(int(P[P_SEED]) + 1), tt = uint32_t((long long)floor(time * 10 + .5));
    st.rng = seed ^ (tt * 0x9E3779B1u);
    st.best = int(0xFFFFFFFFu - unsigned(key0 & 0xFFFFFFFFull));
    st.ratio = sqrt(double(__uint_as_float(unsigned(key0 >> 32))));
  }
  blockSync();
  const int stamp = st.stamp;
  for (int k = 0; k <= limit; k++) {
    if (k > 0) d.solve();
    if (k == limit) break;
    if (k > 0) d.strongest(stamp);
    if (!(st.ratio > 1)) break;
    int start = st.best;
    bool cg; int count; double potential, moment;
    double charge = d.flash(start, cg, count, potential, moment);
    if (charge == 0) { if (threadIdx.x == 0) d.failed[start] = stamp; blockSync(); continue; }
    if (threadIdx.x == 0) {
      long long c; memcpy(&c, S + (cg ? S_FLASH_CG : S_FLASH_IC), 8); c++; memcpy(S + (cg ? S_FLASH_CG : S_FLASH_IC), &c, 8);
      if (cg) S[S_CHARGE_GROUND] += charge;
      if (S[S_FIRST_FLASH] < 0) S[S_FIRST_FLASH] = time;
      long long nlog; memcpy(&nlog, S + S_FLASH_LOG, 8);
      double *e = b.flog + (size_t(slot) * FLASH_LOG + size_t(nlog % FLASH_LOG)) * FLASH_FIELDS;
      e[FL_TIME] = time; e[FL_CG] = cg; e[FL_START] = start; e[FL_NODES] = count; e[FL_CHARGE] = charge; e[FL_POTENTIAL] = potential; e[FL_MOMENT] = moment;
      nlog++; memcpy(S + S_FLASH_LOG, &nlog, 8);
    }
    blockSync();
  }
}

}  // namespace ml
