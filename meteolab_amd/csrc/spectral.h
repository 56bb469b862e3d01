// Spectral solves of the StormLab step: the horizontal 2-D real DFT of a stack of levels (spectral.ts:81-151
// HorizontalDFT, Hermitian half m <= nx/2, two real rows per complex transform), the vertical tridiagonal solve of
// every horizontal mode (advection.ts:222-241 face projection, pressure.ts:104-140 pressure, electric.ts:36-60
// electric potential), and the inverse transform.
//
// One block per (level, run) transforms a level in LDS: mixed-radix Stockham FFTs (radices 4, 2, 3, 5 from the
// factorisation of nx and ny, so 40 = 4·2·5 and 80 = 4·4·5), rows first, then the kept columns. One thread per
// (mode, run) solves the mode's column by the Thomas algorithm. The solves run in float64 whatever the kernels'
// arithmetic type: the low horizontal modes are ill-conditioned vertical systems that amplify float32 rounding of the
// transforms and the elimination (measured: float32 spectra left v and w 1.2e-5 from the float64 step after one step);
// the m = n = 0 mode (the level means) takes float64 level sums as its right-hand side.
#pragma once
#include "common.h"

namespace ml {

template <class R> struct cx { R re, im; };
template <class R> __device__ inline cx<R> cmul(cx<R> a, cx<R> b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }

// Radix plans of a size class: radices of nx then of ny.
struct Plan { int rx[8], nrx, ry[8], nry; };

constexpr int SPEC_THREADS = 256;

// One Stockham pass of radix P (a compile-time radix, so a butterfly's values stay in registers).
template <int P, class R> __device__ inline void fftPass(const cx<R> *src, cx<R> *dst, int N, int Ns, int count, int sStride, int kStride, const cx<R> *tw, int sign,
                                                         int per, int outer) {
  const int span = N / P, total = count * span, tstep = N / (Ns * P), wp = N / P;
  const Div dSpan(span), dNs(Ns), dPer(per < total ? per : total);
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    const int s = dSpan.q(t), j = t - s * span, jn = dNs.q(j), k = j - jn * Ns, lvl = per < total ? dPer.q(s) : 0;
    const int base = lvl * outer + (s - lvl * per) * sStride;
    cx<R> v[P], o[P];
#pragma unroll
    for (int r = 0; r < P; r++) {
      cx<R> x = src[base + (j + r * span) * kStride];
      cx<R> w = tw[k * r * tstep];
      if (sign > 0) w.im = -w.im;
      v[r] = r == 0 ? x : cmul(x, w);
    }
    if constexpr (P == 2) {
      o[0] = {v[0].re + v[1].re, v[0].im + v[1].im}; o[1] = {v[0].re - v[1].re, v[0].im - v[1].im};
    } else if constexpr (P == 4) {
      // W_4 = -i forward, +i inverse.
      cx<R> s02 = {v[0].re + v[2].re, v[0].im + v[2].im}, d02 = {v[0].re - v[2].re, v[0].im - v[2].im};
      cx<R> s13 = {v[1].re + v[3].re, v[1].im + v[3].im}, d13 = {v[1].re - v[3].re, v[1].im - v[3].im};
      cx<R> jd13 = sign < 0 ? cx<R>{d13.im, -d13.re} : cx<R>{-d13.im, d13.re};  // d13 times W_4
      o[0] = {s02.re + s13.re, s02.im + s13.im}; o[2] = {s02.re - s13.re, s02.im - s13.im};
      o[1] = {d02.re + jd13.re, d02.im + jd13.im}; o[3] = {d02.re - jd13.re, d02.im - jd13.im};
    } else if constexpr (P == 3) {
      // X1,2 = v0 - (v1 + v2)/2 -+ i sin(2 pi/3) (v1 - v2) forward (+- inverse).
      const R h = R(0.86602540378443864676);
      cx<R> t1 = {v[1].re + v[2].re, v[1].im + v[2].im}, t2 = {v[1].re - v[2].re, v[1].im - v[2].im};
      cx<R> a = {v[0].re - t1.re / 2, v[0].im - t1.im / 2}, bq = {h * t2.re, h * t2.im};
      cx<R> ib = sign < 0 ? cx<R>{bq.im, -bq.re} : cx<R>{-bq.im, bq.re};  // -i b forward, +i b inverse
      o[0] = {v[0].re + t1.re, v[0].im + t1.im};
      o[1] = {a.re + ib.re, a.im + ib.im}; o[2] = {a.re - ib.re, a.im - ib.im};
    } else if constexpr (P == 5) {
      // With t1 = v1 + v4, t2 = v2 + v3, t3 = v1 - v4, t4 = v2 - v3 and c_k = cos(2 pi k/5), s_k = sin(2 pi k/5):
      // X1,4 = v0 + c1 t1 + c2 t2 -+ i (s1 t3 + s2 t4), X2,3 = v0 + c2 t1 + c1 t2 -+ i (s2 t3 - s1 t4) forward.
      const R c1 = R(0.30901699437494742410), c2 = R(-0.80901699437494742410), s1 = R(0.95105651629515357212), s2 = R(0.58778525229247312917);
      cx<R> t1 = {v[1].re + v[4].re, v[1].im + v[4].im}, t2 = {v[2].re + v[3].re, v[2].im + v[3].im};
      cx<R> t3 = {v[1].re - v[4].re, v[1].im - v[4].im}, t4 = {v[2].re - v[3].re, v[2].im - v[3].im};
      cx<R> a1 = {v[0].re + c1 * t1.re + c2 * t2.re, v[0].im + c1 * t1.im + c2 * t2.im};
      cx<R> a2 = {v[0].re + c2 * t1.re + c1 * t2.re, v[0].im + c2 * t1.im + c1 * t2.im};
      cx<R> b1 = {s1 * t3.re + s2 * t4.re, s1 * t3.im + s2 * t4.im}, b2 = {s2 * t3.re - s1 * t4.re, s2 * t3.im - s1 * t4.im};
      cx<R> ib1 = sign < 0 ? cx<R>{b1.im, -b1.re} : cx<R>{-b1.im, b1.re}, ib2 = sign < 0 ? cx<R>{b2.im, -b2.re} : cx<R>{-b2.im, b2.re};
      o[0] = {v[0].re + t1.re + t2.re, v[0].im + t1.im + t2.im};
      o[1] = {a1.re + ib1.re, a1.im + ib1.im}; o[4] = {a1.re - ib1.re, a1.im - ib1.im};
      o[2] = {a2.re + ib2.re, a2.im + ib2.im}; o[3] = {a2.re - ib2.re, a2.im - ib2.im};
    } else {
#pragma unroll
      for (int q = 0; q < P; q++) {
        cx<R> acc = v[0];
#pragma unroll
        for (int r = 1; r < P; r++) {
          cx<R> w = tw[((r * q) % P) * wp];
          if (sign > 0) w.im = -w.im;
          cx<R> m = cmul(v[r], w);
          acc.re += m.re; acc.im += m.im;
        }
        o[q] = acc;
      }
    }
    const int d = jn * Ns * P + k;
#pragma unroll
    for (int q = 0; q < P; q++) dst[base + (d + q * Ns) * kStride] = o[q];
  }
  __syncthreads();
}

// `count` sequences of length N in LDS, element k of sequence s at base(s) + k * kStride, base(s) = (s / per) * outer +
// (s % per) * sStride (several levels' sequences at once: `per` sequences a level, levels `outer` apart);
// tw[k] = exp(-2 pi i k/N) (sign -1, forward) or its conjugate (sign +1). Stockham stages ping-pong between a and b;
// returns the buffer that holds the result. Ends with a barrier.
// Inlined into every caller, so that the buffers' address space (LDS) is known and the passes use DS instructions: an
// out-of-line copy takes generic pointers (FLAT loads and stores) and costs the discharge kernel its spill-free
// register budget.
template <class R> __device__ __attribute__((always_inline)) cx<R> *fftStages(cx<R> *a, cx<R> *b, int N, int count, int sStride, int kStride, const cx<R> *tw, const int *rad, int nrad, int sign,
                                               int per = 1 << 30, int outer = 0) {
  cx<R> *src = a, *dst = b;
  int Ns = 1;
  for (int st = 0; st < nrad; st++) {
    const int p = rad[st];
    switch (p) {
      case 2: fftPass<2>(src, dst, N, Ns, count, sStride, kStride, tw, sign, per, outer); break;
      case 3: fftPass<3>(src, dst, N, Ns, count, sStride, kStride, tw, sign, per, outer); break;
      case 4: fftPass<4>(src, dst, N, Ns, count, sStride, kStride, tw, sign, per, outer); break;
      default: fftPass<5>(src, dst, N, Ns, count, sStride, kStride, tw, sign, per, outer); break;
    }
    cx<R> *t = src; src = dst; dst = t;
    Ns *= p;
  }
  return src;
}

// LDS of a level transform: two complex buffers of max(nx ny / 2, mw ny) and the twiddles of x and y.
__host__ __device__ inline size_t specLdsBytes(const Geo &g, size_t realBytes) {
  size_t buf = size_t(g.layer / 2 > g.nm ? g.layer / 2 : g.nm);
  return (2 * buf + g.nx + g.ny) * 2 * realBytes;
}

// Forward transform of level `lev` of the right-hand side in `cells` into the spectra, and its float64 sum (the block
// works together; `smem` holds the LDS of specLdsBytes, `red` 16 doubles).
// keep: also write the spectrum to the discharge's right-hand-side spectra (the electric solve, whose right-hand side
// the discharge then updates node by node).
template <class R> __device__ void forwardLevel(const Geo &g, const Bufs<R> &b, const Plan &pl, int slot, int lev, char *smem, double *red, bool keep) {
  typedef double S;
  const int nx = g.nx, ny = g.ny, mw = g.mw, layer = g.layer, half = ny / 2;
  const int buf = layer / 2 > g.nm ? layer / 2 : g.nm;
  cx<S> *A = reinterpret_cast<cx<S> *>(smem), *B = A + buf, *tx = B + buf, *ty = tx + nx;
  for (int k = threadIdx.x; k < nx; k += blockDim.x) { tx[k] = {b.twx[2 * k], b.twx[2 * k + 1]}; }
  for (int k = threadIdx.x; k < ny; k += blockDim.x) { ty[k] = {b.twy[2 * k], b.twy[2 * k + 1]}; }
  const double *src = b.arr(b.cells, slot, g) + size_t(lev) * layer;
  double sum = 0;
  for (int t = threadIdx.x; t < layer / 2; t += blockDim.x) {
    int p = t / nx, x = t - p * nx;
    S v0 = src[2 * p * nx + x], v1 = src[(2 * p + 1) * nx + x];
    A[t] = {v0, v1};
    sum += v0 + v1;
  }
  sum = blockSum(sum, red);  // contains barriers
  if (threadIdx.x == 0) (keep ? b.em0(slot, g) : b.m0(slot, g))[lev] = sum;
  // Row pairs: z = row_2p + i row_2p+1, complex FFT of length nx.
  cx<S> *Z = fftStages(A, B, nx, half, nx, 1, tx, pl.rx, pl.nrx, -1);
  cx<S> *C = Z == A ? B : A;
  // Split into the two rows' Hermitian halves (m = 0 .. nx/2): X_2p = (Z_m + conj Z_-m)/2, X_2p+1 = (Z_m - conj Z_-m)/2i.
  for (int t = threadIdx.x; t < half * mw; t += blockDim.x) {
    int p = t / mw, m = t - p * mw;
    cx<S> z = Z[p * nx + m], c = Z[p * nx + (m ? nx - m : 0)];
    C[(2 * p) * mw + m] = {(z.re + c.re) / 2, (z.im - c.im) / 2};
    C[(2 * p + 1) * mw + m] = {(z.im + c.im) / 2, (c.re - z.re) / 2};
  }
  __syncthreads();
  // Columns: mw sequences of length ny, element n of column m at n * mw + m.
  cx<S> *Y = fftStages(C, Z, ny, mw, 1, mw, ty, pl.ry, pl.nry, -1);
  S *out = b.sp(slot, g) + size_t(lev) * g.nm * 2, *kept = b.eh(slot, g) + size_t(lev) * g.nm * 2;
  for (int t = threadIdx.x; t < g.nm; t += blockDim.x) {
    out[2 * t] = Y[t].re; out[2 * t + 1] = Y[t].im;
    if (keep) { kept[2 * t] = Y[t].re; kept[2 * t + 1] = Y[t].im; }
  }
  __syncthreads();
}

// Inverse transform of level `lev` of the spectra into dst (kind 0: the face potential phi, 1: the pressure, 2: the
// electric potential's interior levels 1 .. nz-2), normalised by nx ny.
template <class R> __device__ void inverseLevel(const Geo &g, const Bufs<R> &b, const Plan &pl, int slot, int lev, int kind, char *smem) {
  typedef double S;
  const int nx = g.nx, ny = g.ny, mw = g.mw, layer = g.layer, half = ny / 2, mh = nx / 2;
  const int buf = layer / 2 > g.nm ? layer / 2 : g.nm;
  cx<S> *A = reinterpret_cast<cx<S> *>(smem), *B = A + buf, *tx = B + buf, *ty = tx + nx;
  for (int k = threadIdx.x; k < nx; k += blockDim.x) { tx[k] = {b.twx[2 * k], b.twx[2 * k + 1]}; }
  for (int k = threadIdx.x; k < ny; k += blockDim.x) { ty[k] = {b.twy[2 * k], b.twy[2 * k + 1]}; }
  const S *in = b.sp(slot, g) + size_t(lev) * g.nm * 2;
  for (int t = threadIdx.x; t < g.nm; t += blockDim.x) A[t] = {in[2 * t], in[2 * t + 1]};
  __syncthreads();
  cx<S> *Y = fftStages(A, B, ny, mw, 1, mw, ty, pl.ry, pl.nry, +1);
  cx<S> *Z = Y == A ? B : A;
  // Rows 2p, 2p+1 from their Hermitian halves: z_m = A_m + i B_m with A_-m = conj A_m; the m = 0 and m = nx/2 terms
  // are taken real (spectral.ts inverseLayers).
  for (int t = threadIdx.x; t < half * nx; t += blockDim.x) {
    int p = t / nx, m = t - p * nx;
    int c = m <= mh ? m : nx - m; S s = m <= mh ? S(1) : S(-1); bool real = c == 0 || 2 * c == nx;
    cx<S> a = Y[(2 * p) * mw + c], bb = Y[(2 * p + 1) * mw + c];
    S ar = a.re, ai = real ? S(0) : s * a.im, br = bb.re, bi = real ? S(0) : s * bb.im;
    Z[p * nx + m] = {ar - bi, ai + br};
  }
  __syncthreads();
  cx<S> *X = fftStages(Z, Y, nx, half, nx, 1, tx, pl.rx, pl.nrx, +1);
  double *dst = kind == 0 ? b.arr(b.phi, slot, g) : kind == 1 ? b.arr(b.press, slot, g) : b.arr(b.ephi, slot, g) + layer;
  dst += size_t(lev) * layer;
  const S norm = S(1) / S(layer);
  for (int t = threadIdx.x; t < layer / 2; t += blockDim.x) {
    int p = t / nx, x = t - p * nx;
    dst[2 * p * nx + x] = X[t].re * norm;
    dst[(2 * p + 1) * nx + x] = X[t].im * norm;
  }
  __syncthreads();
}

// The electric solve (keep, kind 2) of a run without charge: its right-hand side's kept spectra and level sums are
// zero, the modes are skipped, and its potential is zero.
template <class R> __global__ __launch_bounds__(SPEC_THREADS) void k_spec_fwd(Geo g, Bufs<R> b, Plan pl, int L, int keep) {
  extern __shared__ char smem[];
  __shared__ double red[SPEC_THREADS / 64];
  const int slot = b.runs[blockIdx.y], lev = blockIdx.x;
  if (keep && unchargedRun(b, slot)) {
    double *kept = b.eh(slot, g) + size_t(lev) * g.nm * 2;
    for (int t = threadIdx.x; t < 2 * g.nm; t += blockDim.x) kept[t] = 0;
    if (threadIdx.x == 0) b.em0(slot, g)[lev] = 0;
    return;
  }
  forwardLevel(g, b, pl, slot, lev, smem, red, keep != 0);
}
template <class R> __global__ __launch_bounds__(SPEC_THREADS) void k_spec_inv(Geo g, Bufs<R> b, Plan pl, int L, int kind) {
  extern __shared__ char smem[];
  const int slot = b.runs[blockIdx.y], lev = blockIdx.x;
  if (kind == 2 && unchargedRun(b, slot)) {
    double *dst = b.arr(b.ephi, slot, g) + size_t(lev + 1) * g.layer;
    for (int t = threadIdx.x; t < g.layer; t += blockDim.x) dst[t] = 0;
    return;
  }
  inverseLevel(g, b, pl, slot, lev, kind, smem);
}

// Row k of the tridiagonal system of mode (symbol a, b) of kind 0 (face projection, advection.ts:222-241, L = nz),
// 1 (pressure, pressure.ts:119-140, L = nz - 1) or 2 (electric, electric.ts:42-52, L = nz - 2): lo, di, up and the
// factor on the right-hand side.
template <class T> __device__ inline void modeRow(int kind, int k, int L, T a, T bsym, const double *rho, const double *dzs, const double *hz, T &lo, T &di, T &up, T &rf) {
  rf = 1;
  if (kind == 0) {
    auto face = [&](int j) { return j < 0 || j >= L - 1 ? T(0) : T((rho[j] + rho[j + 1]) / 2); };
    lo = k > 0 ? face(k - 1) / T(dzs[k - 1]) : T(0);
    up = k < L - 1 ? face(k) / T(dzs[k]) : T(0);
    di = -T(rho[k]) * T(hz[k]) * a - lo - up;
    rf = T(hz[k]);
  } else if (kind == 1) {
    di = -a / 2 * (T(rho[k]) * T(dzs[k]) / (2 * T(hz[k])) + T(rho[k + 1]) * T(dzs[k]) / (2 * T(hz[k + 1]))); lo = 0; up = 0;
    if (k > 0) { lo = -a / 2 * T(rho[k]) * T(dzs[k - 1]) / (2 * T(hz[k])); T v = bsym * T(rho[k]) / (T(hz[k]) * T(dzs[k])); lo += v; di -= v; }
    if (k < L - 1) { up = -a / 2 * T(rho[k + 1]) * T(dzs[k + 1]) / (2 * T(hz[k + 1])); T v = bsym * T(rho[k + 1]) / (T(hz[k + 1]) * T(dzs[k])); up += v; di -= v; }
  } else {
    int kk = k + 1;
    T l0 = T(1 / (hz[kk] * dzs[kk - 1])), u0 = T(1 / (hz[kk] * dzs[kk]));
    di = -l0 - u0 - a;
    lo = k > 0 ? l0 : T(0); up = k < L - 1 ? u0 : T(0);
  }
}

constexpr int MODE_THREADS = 64;
// Solves the column of `mode` in place in the spectra (float64) by the Thomas algorithm; cp[k * cpStride] holds the
// elimination's factors. The rows come from the run's level coefficients (face projection, pressure) or the pool's
// elimination of the electric operator, so a level costs one reciprocal of its pivot.
template <class R> __device__ void modeSolve(const Geo &g, const Bufs<R> &b, int slot, int mode, int kind, int L, double *cpv, int cpStride) {
  double *spec = b.sp(slot, g);
  const int nm = g.nm;
  const size_t ls = size_t(nm) * 2;
  const double a = b.sym[(kind == 0 ? 0 : kind == 1 ? 1 : 3) * nm + mode], bsym = kind == 1 ? b.sym[2 * nm + mode] : 0.0;
  if (kind == 1 && ((a == 0 && bsym == 0) || (L == 1 && a == 0))) {
    for (int k = 0; k < L; k++) { spec[k * ls + 2 * mode] = 0; spec[k * ls + 2 * mode + 1] = 0; }
    return;
  }
  const bool pinned = kind != 2 && a == 0;
  const double *c0 = b.lv(slot, kind == 0 ? L_F0LO : L_P1A, g);  // the run's level coefficients: [5][nz] from here
  const double *ep = b.efac, *eu = b.efac + size_t(L) * nm, *el = b.efac + size_t(2 * L) * nm;
  const int nz = g.nz;
  const double h = -a / 2;
  // The m = n = 0 mode (the level means) takes its right-hand side from the float64 level sums (its imaginary part is 0).
  const double *m0 = mode == 0 ? (kind == 2 ? b.em0(slot, g) : b.m0(slot, g)) : nullptr;
  // KB levels' loads in flight at a time: the right-hand sides first, then their elimination.
  constexpr int KB = 8;
  double pr = 0, pi = 0, cprev = 0;
  for (int k0 = 0; k0 < L; k0 += KB) {
    double vr[KB], vi[KB];
#pragma unroll
    for (int q = 0; q < KB; q++) {
      const int k = k0 + q < L ? k0 + q : L - 1;
      vr[q] = m0 ? m0[k] : spec[k * ls + 2 * mode]; vi[q] = m0 ? 0.0 : spec[k * ls + 2 * mode + 1];
    }
#pragma unroll
    for (int q = 0; q < KB; q++) {
      const int k = k0 + q;
      if (k >= L) break;
      double lo, di, up, rf = 1;
      if (kind == 0) {
        lo = c0[k]; up = c0[nz + k]; di = -c0[2 * nz + k] * a - lo - up; rf = b.hz[k];
      } else if (kind == 1) {
        di = h * c0[k]; lo = 0; up = 0;
        if (k > 0) { const double v = bsym * c0[2 * nz + k]; lo = h * c0[nz + k] + v; di -= v; }
        if (k < L - 1) { const double v = bsym * c0[4 * nz + k]; up = h * c0[3 * nz + k] + v; di -= v; }
      } else {
        lo = el[size_t(k) * nm + mode]; up = 0; di = 0;
      }
      double r = rf * vr[q], s_ = rf * vi[q];
      if (pinned && k == 0) { lo = 0; di = 1; up = 0; r = 0; s_ = 0; }
      if (pinned && k == 1) lo = 0;
      double inv, cp;
      if (kind == 2) { inv = 1.0 / ep[size_t(k) * nm + mode]; cp = eu[size_t(k) * nm + mode]; }
      else { inv = 1.0 / (k == 0 ? di : di - lo * cprev); cp = up * inv; }
      cpv[k * cpStride] = cprev = cp;
      pr = (r - (k == 0 ? 0.0 : lo * pr)) * inv;
      pi = (s_ - (k == 0 ? 0.0 : lo * pi)) * inv;
      spec[k * ls + 2 * mode] = pr; spec[k * ls + 2 * mode + 1] = pi;
    }
  }
  for (int k1 = L - 2; k1 >= 0; k1 -= KB) {
    double vr[KB], vi[KB];
#pragma unroll
    for (int q = 0; q < KB; q++) {
      const int k = k1 - q >= 0 ? k1 - q : 0;
      vr[q] = spec[k * ls + 2 * mode]; vi[q] = spec[k * ls + 2 * mode + 1];
    }
#pragma unroll
    for (int q = 0; q < KB; q++) {
      const int k = k1 - q;
      if (k < 0) break;
      const double c = cpv[k * cpStride];
      pr = vr[q] - c * pr; pi = vi[q] - c * pi;
      spec[k * ls + 2 * mode] = pr; spec[k * ls + 2 * mode + 1] = pi;
    }
  }
}
// One thread per (mode, run); the elimination's factors stay in LDS, its right-hand sides overwrite the spectra.
template <class R> __global__ __launch_bounds__(MODE_THREADS) void k_modes(Geo g, Bufs<R> b, int kind, int L) {
  int slot = b.runs[blockIdx.y];
  int mode = blockIdx.x * MODE_THREADS + threadIdx.x;
  if (mode >= g.nm || (kind == 2 && unchargedRun(b, slot))) return;
  extern __shared__ char smem[];
  modeSolve(g, b, slot, mode, kind, L, reinterpret_cast<double *>(smem) + threadIdx.x, MODE_THREADS);
}

}  // namespace ml
