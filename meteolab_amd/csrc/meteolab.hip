// Host runtime of the StormLab step on MI355X and its C ABI (include/meteolab.h).
//
// A context owns one device and one stream. A pool holds the runs of one grid size in slots; each array kind is one
// allocation over the slots, so the kernels of a step cover a whole batch of runs in one launch each. The launch
// sequence of a batch (one model step) is captured into a hipGraph once and replayed: all data-dependent control of a
// step (sedimentation sub-steps, positivity-fix scales, lightning) is decided on the device, so a replay needs no
// host round trip.
#include "../../include/meteolab.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "kernels.h"
#include "discharge.h"
#include "lightning.h"
#include "spectral.h"
#include "rotation.h"
#include "diagnostics.h"
#include "intervene.h"

using namespace ml;

// The scratch of ml_model_diagnostics_noalloc, in the library's own device memory (each device holds its own copy until
// the process ends): the diagnostics rows of up to ML_NOALLOC_RUNS runs, then their run list. In host code `scratch`
// names a host-side stand-in; the device address of each device's copy comes from hipGetSymbolAddress at context
// creation (ml_ctx::scratch), and only that address is handed to kernels and copies.
constexpr size_t NOALLOC_ROWS_BYTES = size_t(ML_NOALLOC_RUNS) * DIAG_ROW * sizeof(double);
constexpr size_t NOALLOC_BYTES = NOALLOC_ROWS_BYTES + size_t(ML_NOALLOC_RUNS) * sizeof(int);
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {

constexpr int MAX_POOLS = 4096, SLOT_BITS = 16;

// The transport kernel's launch: layout 1 is k_stage_level (`levels` whole levels of one field per block), layout 2 is
// k_stage_plane (as layout 1, each level in a padded LDS plane); `xcd` keeps each run's blocks on one XCD. (Layout 0,
// strips of a level, is not built; `groups` is kept for the reported shape and is 1.)
struct Tuning { int layout = 1, groups = 1, xcd = 0, levels = 1; float ms = 0; };

struct Batch {
  std::vector<int> runs;
  int *dev = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // With lightning: the step without its discharge (first of a call), and the step beside the discharge of the step
  // before it (the rest); a call ends with the last step's discharge.
  hipGraph_t firstGraph = nullptr, nextGraph = nullptr;
  hipGraphExec_t first = nullptr, next = nullptr;
  double dt = 0;
  int overlap = -1;  // the context's overlap setting the graphs were captured with
  Tuning tune{};   // the transport kernel's layout, picked by the autotuner when the batch is first stepped
  bool tuned = false;
};

struct Pool {
  Geo g{};
  double width = 0, depth = 0;  // the domain's extent (m), as the grid gives it
  Plan plan{};
  int cap = 0, electrified = 1;
  size_t rb = 4;  // bytes of the arithmetic type
  void *fields = nullptr, *stx = nullptr, *sty = nullptr, *mass = nullptr, *phi = nullptr, *cells = nullptr, *press = nullptr,
       *spec = nullptr, *mean0 = nullptr, *km = nullptr, *kh = nullptr, *ephi = nullptr, *cols = nullptr, *lev = nullptr,
       *par = nullptr, *scal = nullptr, *zs = nullptr, *dzs = nullptr, *hz = nullptr, *sym = nullptr, *twx = nullptr, *twy = nullptr,
       *lrho = nullptr, *lmark = nullptr, *lnode = nullptr, *chan = nullptr, *cdep = nullptr, *flog = nullptr, *efac = nullptr,
       *ehat = nullptr, *emean0 = nullptr, *dspec = nullptr, *fillp = nullptr, *gpart = nullptr, *zflag = nullptr, *landed = nullptr;
  int ftiles = 1, gmax = 1;  // strip tiles of a level; blocks of a run in the launches that carry charge to the ground
  int chunk = 1;          // levels per pass of the discharge's own electric solve (as many as the block's LDS holds)
  int frontCap = 0;       // entries of a flash's front the discharge keeps in that LDS between its solves
  RotLevels rot{};        // the rotation diagnostics' levels and threshold
  int *single = nullptr;  // single[s] = s: the run list of one slot
  DiagLevels dlev{};      // the levels the model diagnostics look at
  double *diag = nullptr; // [cap][DIAG_ROW]: rows of diagnostics
  int *dlist = nullptr;   // [cap]: the run list of a diagnostics call
  void *pmap = nullptr;   // [layer] float: a probability map
  std::vector<double> zsH, dzsH, hzH;
  std::vector<char> used;
  // Captured steps by batch size: the graph of a size is reused for any runs of the pool, its device run list rewritten
  // when the runs change (the continuous-batching case: runs join and leave between calls).
  std::map<int, Batch *> batches;
};

}  // namespace

struct ml_ctx {
  int device = 0, precision = 32;
  hipStream_t stream = nullptr;
  // The lightning's stream: a step's discharge runs on it beside the next step's dynamics (default priority, as the
  // main stream), forked from and joined back into the main stream by the two events.
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  int overlap = 1;  // a step's discharge beside the next step's dynamics (ml_ctx_set_overlap)
  hipDeviceProp_t prop{};
  std::string err;
  void *scratch = nullptr;  // this device's copy of the static scratch array (hipGetSymbolAddress)
  std::vector<Pool *> pools;
};

namespace {

int fail(ml_ctx *c, int code, const std::string &msg) { if (c) c->err = msg; return code; }
#define HIPCHK(c, x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return fail(c, ML_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); } while (0)

// Geometric stretching ratio (grid.ts:91-97) and the levels of createGrid (grid.ts:99-121).
double stretchRatio(double bottom, double height, int intervals) {
  if (bottom * intervals >= height) return 1;
  double lo = 1, hi = 2;
  while (bottom * (std::pow(hi, intervals) - 1) / (hi - 1) < height) hi *= 2;
  for (int k = 0; k < 100; k++) { double mid = (lo + hi) / 2; if (bottom * (std::pow(mid, intervals) - 1) / (mid - 1) < height) lo = mid; else hi = mid; }
  return (lo + hi) / 2;
}
void levels(const ml_grid &o, std::vector<double> &zs, std::vector<double> &dzs, std::vector<double> &hz) {
  int nz = o.nz; double dz = o.height / (nz - 1);
  zs.assign(nz, 0.0);
  bool uniform = !(o.bottom > 0);
  if (uniform) for (int z = 0; z < nz; z++) zs[z] = z * dz;
  else {
    double r = stretchRatio(o.bottom, o.height, nz - 1);
    for (int z = 1; z < nz; z++) zs[z] = zs[z - 1] + o.bottom * std::pow(r, z - 1);
    for (int z = 1; z < nz; z++) zs[z] *= o.height / zs[nz - 1];
  }
  dzs.assign(nz - 1, 0.0);
  for (int k = 0; k < nz - 1; k++) dzs[k] = uniform ? dz : zs[k + 1] - zs[k];
  hz.assign(nz, 0.0);
  for (int k = 0; k < nz; k++) hz[k] = ((k > 0 ? dzs[k - 1] : 0) + (k < nz - 1 ? dzs[k] : 0)) / 2;
}
// Radices of an FFT length: fours first, then twos, threes, fives (spectral.ts:20-24 tries 4, 2, 3, 5, ...).
bool radices(int n, int *out, int &count) {
  count = 0;
  const int order[4] = {4, 2, 3, 5};
  for (int p : order) while (n % p == 0 && n > 1) { if (count == 8) return false; out[count++] = p; n /= p; }
  return n == 1;
}

template <class R> Bufs<R> bufs(const Pool &p, const int *runs) {
  Bufs<R> b;
  b.fields = (float *)p.fields; b.stx = (float *)p.stx; b.sty = (float *)p.sty; b.mass = (double *)p.mass; b.phi = (double *)p.phi;
  b.cells = (double *)p.cells; b.press = (double *)p.press; b.spec = (double *)p.spec; b.mean0 = (double *)p.mean0; b.km = (R *)p.km; b.kh = (R *)p.kh;
  b.ephi = (double *)p.ephi; b.cols = (float *)p.cols; b.lev = (double *)p.lev; b.par = (double *)p.par; b.scal = (double *)p.scal;
  b.zs = (const double *)p.zs; b.dzs = (const double *)p.dzs; b.hz = (const double *)p.hz; b.sym = (const double *)p.sym;
  b.twx = (const double *)p.twx; b.twy = (const double *)p.twy; b.runs = runs;
  b.efac = (const double *)p.efac; b.lrho = (R *)p.lrho; b.lmark = (int *)p.lmark; b.lnode = (double *)p.lnode; b.chan = (int *)p.chan; b.cdep = (double *)p.cdep; b.flog = (double *)p.flog;
  b.ehat = (double *)p.ehat; b.emean0 = (double *)p.emean0; b.dspec = (double *)p.dspec;
  b.fillp = (double *)p.fillp; b.ftiles = p.ftiles; b.gpart = (double *)p.gpart; b.gmax = p.gmax;
  b.zflag = (unsigned char *)p.zflag; b.landed = (float *)p.landed;
  return b;
}

inline unsigned blocks(size_t n, unsigned t) { return unsigned((n + t - 1) / t); }

// Named launch stages, for the per-stage timing.
struct Timer {
  hipStream_t s;
  std::vector<hipEvent_t> *events = nullptr;
  std::vector<std::string> *names = nullptr;
  void mark(const char *name) {
    if (!events) return;
    hipEvent_t e; hipEventCreate(&e); hipEventRecord(e, s); events->push_back(e); names->push_back(name);
  }
};

template <class R> void spectralSolve(const Pool &p, const Bufs<R> &b, int nr, hipStream_t s, int kind, int L, int invKind, Timer *t) {
  const Geo &g = p.g;
  size_t lds = specLdsBytes(g, sizeof(double));
  hipLaunchKernelGGL(k_spec_fwd<R>, dim3(L, nr), dim3(SPEC_THREADS), lds, s, g, b, p.plan, L, kind == 2 ? 1 : 0);
  if (t) t->mark("k_spec_fwd");
  hipLaunchKernelGGL(k_modes<R>, dim3(blocks(g.nm, MODE_THREADS), nr), dim3(MODE_THREADS), size_t(MODE_THREADS) * L * sizeof(double), s, g, b, kind, L);
  if (t) t->mark("k_modes");
  hipLaunchKernelGGL(k_spec_inv<R>, dim3(L, nr), dim3(SPEC_THREADS), lds, s, g, b, p.plan, L, invKind);
  if (t) t->mark("k_spec_inv");
}

// One RK3 stage of the transport of every carried field, in the batch's tuned layout.
// Fields fbase .. fbase + nc - 1.
template <class R> void launchStage(const Geo &g, const Bufs<R> &b, int nr, int nc, int fbase, int stage, double dt, hipStream_t s, const Tuning &tu, Timer *t = nullptr) {
  const unsigned zBlocks = unsigned((g.nz + tu.levels - 1) / tu.levels);
  if (tu.layout == 1) {
    hipLaunchKernelGGL(k_stage_level<R>, dim3(zBlocks * nc * nr), dim3(TPB), stageLevelLdsBytes(g, sizeof(R)), s, g, b, stage, dt, nr, nc, fbase, tu.xcd, tu.levels);
    if (t) t->mark("k_stage_level");
  } else {
    hipLaunchKernelGGL(k_stage_plane<R>, dim3(zBlocks * nc * nr), dim3(TPB), stagePlaneLdsBytes(g, sizeof(R)), s, g, b, stage, dt, nr, nc, fbase, tu.xcd, tu.levels);
    if (t) t->mark("k_stage_plane");
  }
}

// The parts of a model step, split where the reference's GPU step splits it (gpu/step.ts:12-31):
//   PART_DYNAMICS  beginStep's fall speeds, hail swath and sedimentation, then the transport (GpuDynamics.step);
//   PART_PHYSICS   physicsStep: the positivity fix from the last transport's sums, the column physics;
//   PART_MIX       mixStep and the projection with the wind clipping (GpuDynamics.mix, endGpuStep);
//   PART_AFTER     afterStep's lightning, then time += dt.
enum StepPart { PART_DYNAMICS, PART_PHYSICS, PART_MIX, PART_AFTER, N_PARTS };

// The lightning's discharge and conduction of a step (lightning.ts:261-276, 190-225). advanced: the run's clock has
// already moved on by dt (the discharge of a step run beside the next step's dynamics).
template <class R> void launchDischarge(const Pool &p, const Bufs<R> &b, int nr, double dt, hipStream_t s, Timer *t, int advanced) {
  const Geo &g = p.g;
  const unsigned ib = blocks(size_t(g.layer) * (g.nz - 2), 256);
  // flashLimit (lightning.ts:46): 10 flashes per 3 s of model time.
  const int limit = std::max(1, int(std::floor(10 * dt / 3 + .5)));
  hipLaunchKernelGGL(k_discharge<R>, dim3(nr), dim3(DISCHARGE_THREADS), solveLdsBytes(g, p.chunk), s, g, b, p.plan, dt, limit, p.chunk, advanced,
                     p.frontCap);
  if (t) t->mark("k_discharge");
  hipLaunchKernelGGL(k_conduct<R>, dim3(ib, nr), dim3(256), 0, s, g, b, dt);
  if (t) t->mark("k_conduct");
  hipLaunchKernelGGL(k_ground_total<R>, dim3(nr), dim3(256), 0, s, g, b, 1, int(ib));
  if (t) t->mark("k_ground_total");
}

// One part of a model step of the runs in `list` (device array of nr slots). With a timer, an event follows every
// kernel (named by the kernel).
// split (with lightning): the step is split around the lightning so that a step's discharge can run beside the next
// step's dynamics. The dynamics part then leaves out everything that reads the charges (the charge falls' Courant
// numbers, the charge falls and the charges' transport: PART_CHARGES, after the discharge), and the after-part leaves
// out the discharge and conduction (launchDischarge, with the step after or at the end of the call). Nothing else of a
// step reads or writes what the discharge does: the charges, the vertical field, the flash state and the electric
// solve's kept spectra; it reads the cloudiness the after-part snapshots (F_CLOUDY) and has spectra of its own.
enum : int { PART_CHARGES = 4 };
template <class R> void launchPart(const Pool &p, const int *list, int nr, double dt, hipStream_t s, Timer *t, const Tuning &tu, int part, bool split = false) {
  const Geo &g = p.g;
  Bufs<R> b = bufs<R>(p, list);
  const int el = p.electrified, nc = el ? N_CARRIED : F_COLD + 1;
  split = split && el;
  const unsigned nb = blocks(g.n, TPB), cb = blocks(size_t(g.layer) * (g.nz - 1), TPB);
  auto mark = [&](const char *n) { if (t) t->mark(n); };
  if (part == PART_DYNAMICS) {
    // beginStep: fall speeds, hail swath, sedimentation (model.ts:893-902).
    hipLaunchKernelGGL(k_begin<R>, dim3(nr), dim3(64), 0, s, g, b);
    mark("k_begin");
    hipLaunchKernelGGL(k_fall<R>, dim3(nb, nr), dim3(TPB), 0, s, g, b, dt, el, split ? 0 : el);
    mark("k_fall");
    hipLaunchKernelGGL(k_ground<R>, dim3(blocks(g.layer, TPB), nr), dim3(TPB), 0, s, g, b, dt, 0);
    mark("k_ground");
    hipLaunchKernelGGL(k_sediment<R>, dim3(blocks(g.layer, SED_COLS), nr), dim3(SED_COLS), size_t(2) * SED_COLS * g.nz * sizeof(float), s, g, b, dt, 0, el && !split ? N_FALLS : 4);
    mark("k_sediment");
    hipLaunchKernelGGL(k_ground<R>, dim3(blocks(g.layer, TPB), nr), dim3(TPB), 0, s, g, b, dt, 1);
    mark("k_ground");
    if (el && !split) {
      hipLaunchKernelGGL(k_ground_total<R>, dim3(nr), dim3(256), 0, s, g, b, 0, int(blocks(g.layer, SED_COLS)));
      mark("k_ground_total");
    }
    // Which levels of the fields transported next hold anything (a zero neighbourhood's transport is skipped).
    hipLaunchKernelGGL(k_zero_flags<R>, dim3(g.nz * (split ? int(F_CHARGE_ICE) : nc), nr), dim3(256), 0, s, g, b, 0, 0);
    mark("k_zero_flags");
    // dynamics: face projection (FluxTransport.setVelocity) and the three RK3 stages with the stage projections.
    hipLaunchKernelGGL(k_faces<R>, dim3(nb, nr), dim3(TPB), 0, s, g, b, 0);
    mark("k_faces");
    spectralSolve<R>(p, b, nr, s, 0, g.nz, 0, t);
    hipLaunchKernelGGL(k_faces<R>, dim3(nb, nr), dim3(TPB), 0, s, g, b, 1);
    mark("k_faces");
    for (int stage = 0; stage < 3; stage++) {
      launchStage<R>(g, b, nr, split ? int(F_CHARGE_ICE) : nc, 0, stage, dt, s, tu, t);
      if (stage < 2) {
        hipLaunchKernelGGL(k_pdiv<R>, dim3(cb, nr), dim3(TPB), 0, s, g, b, stage + 1, dt);
        mark("k_pdiv");
        spectralSolve<R>(p, b, nr, s, 1, g.nz - 1, 1, t);
        hipLaunchKernelGGL(k_pcorrect<R>, dim3(nb, nr), dim3(TPB), 0, s, g, b, stage + 1, dt, 0);
        mark("k_pcorrect");
      }
    }
  }
  if (part == PART_CHARGES) {
    // The charge falls with their carriers' speeds, then the charges' three RK3 stages with the step's mass fluxes.
    hipLaunchKernelGGL(k_charge_courant<R>, dim3(nb, nr), dim3(TPB), 0, s, g, b, dt);
    mark("k_charge_courant");
    hipLaunchKernelGGL(k_sediment<R>, dim3(blocks(g.layer, SED_COLS), nr), dim3(SED_COLS), size_t(2) * SED_COLS * g.nz * sizeof(float), s, g, b, dt, 4, N_FALLS);
    mark("k_sediment");
    hipLaunchKernelGGL(k_ground_total<R>, dim3(nr), dim3(256), 0, s, g, b, 0, int(blocks(g.layer, SED_COLS)));
    mark("k_ground_total");
    hipLaunchKernelGGL(k_zero_flags<R>, dim3(g.nz * (N_CARRIED - F_CHARGE_ICE), nr), dim3(256), 0, s, g, b, int(F_CHARGE_ICE), 0);
    mark("k_zero_flags");
    for (int stage = 0; stage < 3; stage++) {
      launchStage<R>(g, b, nr, N_CARRIED - F_CHARGE_ICE, F_CHARGE_ICE, stage, dt, s, tu, t);
    }
  }
  // endStep: positivity fix + physics (physicsStep), turbulence + drag (mixStep), projection (model.ts:1103-1113).
  if (part == PART_PHYSICS) {
    // The positivity-fix sums from the last transport stage's parts, in level order.
    hipLaunchKernelGGL(k_fill_total<R>, dim3(nr), dim3(64), 0, s, g, b, 1);
    mark("k_fill_total");
    hipLaunchKernelGGL((k_physics<R, double>), dim3(nb, nr), dim3(TPB), 0, s, g, b, dt, el);
    mark("k_physics");
  }
  if (part == PART_MIX) {
    hipLaunchKernelGGL(k_viscosity<R>, dim3(nb, nr), dim3(TPB), 0, s, g, b, dt);
    mark("k_viscosity");
    hipLaunchKernelGGL(k_mix<R>, dim3(nb, nr), dim3(TPB), 0, s, g, b, dt);
    mark("k_mix");
    hipLaunchKernelGGL(k_pdiv<R>, dim3(cb, nr), dim3(TPB), 0, s, g, b, 1, dt);
    mark("k_pdiv");
    spectralSolve<R>(p, b, nr, s, 1, g.nz - 1, 1, t);
    hipLaunchKernelGGL(k_pcorrect<R>, dim3(nb, nr), dim3(TPB), 0, s, g, b, 1, dt, 1);
    mark("k_pcorrect");
  }
  if (part != PART_AFTER) return;
  // afterStep: the rotation diagnostics (model.ts:1609-1648), then the lightning (model.ts:1537-1552).
  hipLaunchKernelGGL(k_rotation<R>, dim3(nr), dim3(ROTATION_THREADS), rotationLdsBytes(g), s, g, b, dt, p.rot, 0);
  mark("k_rotation");
  if (el) {
    unsigned ib = blocks(size_t(g.layer) * (g.nz - 2), 256);
    hipLaunchKernelGGL(k_charge_rhs<R>, dim3(nb, nr), dim3(TPB), 0, s, g, b);
    mark("k_charge_rhs");
    spectralSolve<R>(p, b, nr, s, 2, g.nz - 2, 2, t);
    hipLaunchKernelGGL(k_strongest<R>, dim3(ib, nr), dim3(256), 0, s, g, b);
    mark("k_strongest");
    if (!split) launchDischarge<R>(p, b, nr, dt, s, t, 0);
  }
  hipLaunchKernelGGL(k_end<R>, dim3(nr), dim3(64), 0, s, g, b, dt);
  mark("k_end");
}

// One step split around its lightning (with lightning only). prev: the discharge of the step before runs first on
// the side stream, beside this step's dynamics, and this step's charges wait for it. The step's own discharge is left
// to the next split step or to launchDischarge at the end of the call.
template <class R> void launchStepSplit(const Pool &p, const int *list, int nr, double dt, ml_ctx *c, const Tuning &tu, bool prev) {
  hipStream_t s = c->stream;
  if (prev) {
    hipEventRecord(c->fork, s);
    hipStreamWaitEvent(c->side, c->fork, 0);
    launchDischarge<R>(p, bufs<R>(p, list), nr, dt, c->side, nullptr, 1);
    hipEventRecord(c->join, c->side);
  }
  launchPart<R>(p, list, nr, dt, s, nullptr, tu, PART_DYNAMICS, true);
  if (prev) hipStreamWaitEvent(s, c->join, 0);
  launchPart<R>(p, list, nr, dt, s, nullptr, tu, PART_CHARGES, true);
  for (int part = PART_PHYSICS; part < N_PARTS; part++) launchPart<R>(p, list, nr, dt, s, nullptr, tu, part, true);
}

// One model step: AtmosphereModel.step (model.ts:873-877), the parts in order.
template <class R> void launchStep(const Pool &p, const int *list, int nr, double dt, hipStream_t s, Timer *t, const Tuning &tu) {
  if (t) t->mark("start");
  for (int part = 0; part < N_PARTS; part++) launchPart<R>(p, list, nr, dt, s, t, tu, part);
}

// One step with a timer after each kernel, in line: whole, or split around its lightning with the discharge after it.
template <class R> void launchStepTimed(const Pool &p, const int *list, int nr, double dt, hipStream_t s, Timer *t, const Tuning &tu, bool split) {
  if (!split) { launchStep<R>(p, list, nr, dt, s, t, tu); return; }
  t->mark("start");
  launchPart<R>(p, list, nr, dt, s, t, tu, PART_DYNAMICS, true);
  launchPart<R>(p, list, nr, dt, s, t, tu, PART_CHARGES, true);
  for (int part = PART_PHYSICS; part < N_PARTS; part++) launchPart<R>(p, list, nr, dt, s, t, tu, part, true);
  launchDischarge<R>(p, bufs<R>(p, list), nr, dt, s, t, 1);
}

void fillTables(Pool &p, std::vector<char> &sym, std::vector<char> &twx, std::vector<char> &twy) {
  typedef double R;
  const Geo &g = p.g;
  std::vector<R> s(4 * size_t(g.nm)), tx(2 * g.nx), ty(2 * g.ny);
  auto exact = [](double x) { return x < 1e-12 ? 0 : x > 1 - 1e-12 ? 1 : x; };
  for (int n = 0; n < g.ny; n++) for (int m = 0; m < g.mw; m++) {
    int mode = m + g.mw * n;
    double sxr = std::sin(M_PI * m / g.nx), syr = std::sin(M_PI * n / g.ny);
    double lam = 4 * std::pow(sxr, 2) / (g.dx * g.dx) + 4 * std::pow(syr, 2) / (g.dy * g.dy);
    double sx = exact(std::pow(std::sin(M_PI * m / g.nx), 2)), cx = exact(std::pow(std::cos(M_PI * m / g.nx), 2));
    double sy = exact(std::pow(std::sin(M_PI * n / g.ny), 2)), cy = exact(std::pow(std::cos(M_PI * n / g.ny), 2));
    s[mode] = R(m == 0 && n == 0 ? 0 : lam);
    s[g.nm + mode] = R(4 * (sx * cy / (g.dx * g.dx) + cx * sy / (g.dy * g.dy)));
    s[2 * g.nm + mode] = R(cx * cy);
    s[3 * g.nm + mode] = R(lam);
  }
  for (int k = 0; k < g.nx; k++) { tx[2 * k] = R(std::cos(2 * M_PI * k / g.nx)); tx[2 * k + 1] = R(-std::sin(2 * M_PI * k / g.nx)); }
  for (int k = 0; k < g.ny; k++) { ty[2 * k] = R(std::cos(2 * M_PI * k / g.ny)); ty[2 * k + 1] = R(-std::sin(2 * M_PI * k / g.ny)); }
  sym.assign((char *)s.data(), (char *)(s.data() + s.size())); twx.assign((char *)tx.data(), (char *)(tx.data() + tx.size()));
  twy.assign((char *)ty.data(), (char *)(ty.data() + ty.size()));
}

struct Sizes {
  size_t fields, stage, mass, node, node8, spec, mean0, cols, lev, par, scal, diag, dlist, lmark, lnode, chan, cdep, flog, fillp, gpart, zflag,
      landed;
  size_t slot() const {
    return fields + 2 * stage + mass + 3 * node + 4 * node8 + 3 * spec + 2 * mean0 + cols + lev + par + scal + diag + dlist + lmark + lnode + chan +
           cdep + flog + fillp + gpart + zflag + landed;
  }
};
// Blocks of a run in the launches that carry charge to the ground: the charge falls and the conduction.
int groundBlocks(const Geo &g) { return std::max(int(blocks(g.layer, SED_COLS)), int(blocks(size_t(g.layer) * (g.nz - 2), 256))); }
Sizes sizesOf(const Geo &g, size_t rb) {
  Sizes s;
  s.fields = size_t(N_FIELDS) * g.n * 4; s.stage = size_t(N_CARRIED) * g.n * 4; s.mass = 3 * size_t(g.n) * 8; s.node = size_t(g.n) * rb; s.node8 = size_t(g.n) * 8;
  s.spec = size_t(g.nz) * g.nm * 2 * 8; s.mean0 = size_t(g.nz) * 8; s.cols = size_t(N_COLUMNS) * g.layer * 4;
  s.lev = size_t(N_LEVELS) * g.nz * 8; s.par = N_PARAMS * 8; s.scal = N_SCALARS * 8; s.diag = size_t(DIAG_ROW) * 8; s.dlist = sizeof(int);
  s.lmark = size_t(N_LMARK) * g.n * 4; s.lnode = size_t(N_LNODE) * g.n * 8; s.chan = 2 * size_t(MAX_CHANNEL) * 4; s.cdep = size_t(MAX_CHANNEL) * 8;
  s.flog = size_t(FLASH_LOG) * FLASH_FIELDS * 8;
  s.fillp = size_t(N_CONSERVED) * g.nz * 3 * 8;
  s.gpart = 2 * size_t(groundBlocks(g)) * 8;
  s.zflag = (size_t(N_CARRIED) * g.nz + 15) / 16 * 16;
  s.landed = size_t(4) * LAND_PARTS * g.layer * 4;
  return s;
}
int geoOf(ml_ctx *c, const ml_grid *o, Geo &g, Plan &pl) {
  // The six-point stencils take one periodic wrap, so a level has at least four columns and rows.
  if (!o || o->nx < 4 || o->ny < 4 || o->nz < 4 || o->nz > 64 || o->nx % 2 || o->ny % 2 || !(o->width > 0) || !(o->depth > 0) || !(o->height > 0))
    return fail(c, ML_E_ARG, "grid: nx, ny even and >= 4, 4 <= nz <= 64, positive extents");
  if (!radices(o->nx, pl.rx, pl.nrx) || !radices(o->ny, pl.ry, pl.nry)) return fail(c, ML_E_UNSUPPORTED, "grid: nx and ny must factor into 2, 3 and 5");
  g.nx = o->nx; g.ny = o->ny; g.nz = o->nz; g.layer = g.nx * g.ny; g.n = g.layer * g.nz;
  g.mw = g.nx / 2 + 1; g.nm = g.mw * g.ny;
  g.dx = o->width / g.nx; g.dy = o->depth / g.ny; g.dz = o->height / (g.nz - 1); g.height = o->height;
  return 0;
}

Pool *poolOf(ml_ctx *c, int id) { return id >= 0 && id < int(c->pools.size()) ? c->pools[id] : nullptr; }
bool decode(ml_ctx *c, int run, Pool *&p, int &slot) {
  p = poolOf(c, run >> SLOT_BITS); slot = run & ((1 << SLOT_BITS) - 1);
  return p && slot < p->cap && p->used[slot];
}

void freeBatch(Batch *bt) {
  if (bt->exec) hipGraphExecDestroy(bt->exec);
  if (bt->graph) hipGraphDestroy(bt->graph);
  if (bt->first) hipGraphExecDestroy(bt->first);
  if (bt->firstGraph) hipGraphDestroy(bt->firstGraph);
  if (bt->next) hipGraphExecDestroy(bt->next);
  if (bt->nextGraph) hipGraphDestroy(bt->nextGraph);
  if (bt->dev) hipFree(bt->dev);
  delete bt;
}
void dropBatches(Pool *p) {
  for (auto &kv : p->batches) freeBatch(kv.second);
  p->batches.clear();
}

}  // namespace

extern "C" {

int ml_ctx_create(int device, int precision, ml_ctx **out) {
  if (!out || (precision != 32 && precision != 64)) return ML_E_ARG;
  ml_ctx *c = new ml_ctx();
  c->device = device; c->precision = precision;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipGetDeviceProperties(&c->prop, device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->join, hipEventDisableTiming);
  // The device address of this device's scratch array, with the device current.
  if (e == hipSuccess) e = hipGetSymbolAddress(&c->scratch, HIP_SYMBOL(scratch));
  if (e == hipSuccess && !c->scratch) e = hipErrorInvalidSymbol;
  if (e != hipSuccess) {
    if (c->join) hipEventDestroy(c->join);
    if (c->fork) hipEventDestroy(c->fork);
    if (c->side) hipStreamDestroy(c->side);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return ML_E_HIP;
  }
  // Room for any message, so that setting one later (ml_model_diagnostics_noalloc's errors) allocates nothing.
  c->err.reserve(1024);
  *out = c;
  return 0;
}

}  // extern "C"

namespace {
// Frees a pool's device memory and its captured batches (the stream must be idle).
void freePool(Pool *p) {
  dropBatches(p);
  void *all[] = {p->fields, p->stx, p->sty, p->mass, p->phi, p->cells, p->press, p->spec, p->mean0, p->km, p->kh, p->ephi, p->cols,
                 p->lev, p->par, p->scal, p->zs, p->dzs, p->hz, p->sym, p->twx, p->twy, p->single, p->diag, p->dlist, p->lrho,
                 p->lmark, p->lnode, p->chan, p->cdep, p->flog, p->efac, p->ehat, p->emean0, p->dspec, p->fillp,
                 p->gpart, p->pmap, p->zflag, p->landed};
  for (void *a : all) if (a) hipFree(a);
  delete p;
}
// The pool's arrays that hold a part per slot, each [cap][bytes]: what pool creation allocates and what a pool's growth
// reallocates with its slots' parts kept.
struct SlotArray { void **ptr; size_t bytes; };
std::vector<SlotArray> slotArrays(Pool *p, const Sizes &s) {
  return {{&p->fields, s.fields}, {&p->stx, s.stage}, {&p->sty, s.stage}, {&p->mass, s.mass}, {&p->phi, s.node8}, {&p->cells, s.node8},
          {&p->press, s.node8}, {&p->spec, s.spec}, {&p->ehat, s.spec}, {&p->dspec, s.spec}, {&p->mean0, s.mean0}, {&p->emean0, s.mean0},
          {&p->km, s.node}, {&p->kh, s.node}, {&p->ephi, s.node8}, {&p->cols, s.cols}, {&p->lev, s.lev}, {&p->par, s.par},
          {&p->scal, s.scal}, {(void **)&p->diag, s.diag}, {(void **)&p->dlist, s.dlist}, {&p->lrho, s.node},
          {&p->lmark, s.lmark}, {&p->lnode, s.lnode}, {&p->chan, s.chan}, {&p->cdep, s.cdep}, {&p->flog, s.flog}, {&p->fillp, s.fillp},
          {&p->gpart, s.gpart}, {&p->zflag, s.zflag}, {&p->landed, s.landed}};
}
}  // namespace

extern "C" {

void ml_ctx_destroy(ml_ctx *c) {
  if (!c) return;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  hipStreamSynchronize(c->side);
  for (Pool *p : c->pools) if (p) freePool(p);
  hipEventDestroy(c->join);
  hipEventDestroy(c->fork);
  hipStreamDestroy(c->side);
  hipStreamDestroy(c->stream);
  delete c;
}

int ml_pool_destroy(ml_ctx *c, int pool) {
  Pool *p = c ? poolOf(c, pool) : nullptr;
  if (!p) return fail(c, ML_E_ARG, "pool destroy: bad pool");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  freePool(p);
  c->pools[pool] = nullptr;
  return 0;
}

const char *ml_last_error(ml_ctx *c) { return c ? c->err.c_str() : "no context"; }

int ml_device_name(ml_ctx *c, char *out, int size) {
  if (!c || !out || size <= 0) return ML_E_ARG;
  std::snprintf(out, size, "%s (%s)", c->prop.name, c->prop.gcnArchName);
  return 0;
}

static_assert(int(N_COLUMNS) == int(ML_COLUMNS), "the ABI's ground columns");
static_assert(int(F_CLOUDY) == int(ML_FIELDS) && int(N_FIELDS) == int(ML_FIELDS) + 1, "the ABI's fields, then the internal cloudiness");

int64_t ml_slot_bytes(ml_ctx *c, const ml_grid *grid) {
  Geo g; Plan pl;
  if (!c || geoOf(c, grid, g, pl)) return ML_E_ARG;
  return int64_t(sizesOf(g, c->precision == 64 ? 8 : 4).slot());
}

// In pool creation: a HIP error frees what the pool holds so far.
#define POOLCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { freePool(p); return fail(c, ML_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); } } while (0)
int ml_pool_create(ml_ctx *c, const ml_grid *grid, int capacity, int electrified, int *pool) {
  if (!c || !pool || capacity < 1 || capacity >= (1 << SLOT_BITS)) return fail(c, ML_E_ARG, "pool: capacity 1 .. 65535");
  int id = int(std::find(c->pools.begin(), c->pools.end(), nullptr) - c->pools.begin());
  if (id >= MAX_POOLS) return fail(c, ML_E_FULL, "too many pools");
  Pool *p = new Pool();
  if (int r = geoOf(c, grid, p->g, p->plan)) { delete p; return r; }
  p->width = grid->width; p->depth = grid->depth;
  POOLCHK(hipSetDevice(c->device));
  const Geo &g = p->g;
  p->cap = capacity; p->electrified = electrified ? 1 : 0; p->rb = c->precision == 64 ? 8 : 4;
  // LDS the kernels need must fit the device's block limit.
  if (stageLevelLdsBytes(g, p->rb) > c->prop.sharedMemPerBlock ||
      specLdsBytes(g, 8) + 4096 > c->prop.sharedMemPerBlock) { delete p; return fail(c, ML_E_UNSUPPORTED, "grid too large for a level in LDS"); }
  Sizes s = sizesOf(g, p->rb);
  const size_t cap = capacity;
  p->ftiles = 1; p->gmax = groundBlocks(g);
  for (const SlotArray &a : slotArrays(p, s)) {
    hipError_t e = hipMalloc(a.ptr, a.bytes * cap);
    if (e == hipSuccess) e = hipMemset(*a.ptr, 0, a.bytes * cap);
    if (e != hipSuccess) { freePool(p); return fail(c, ML_E_HIP, std::string("pool allocation: ") + hipGetErrorString(e)); }
  }
  POOLCHK(hipMalloc(&p->pmap, size_t(g.layer) * 4));
  levels(*grid, p->zsH, p->dzsH, p->hzH);
  std::vector<char> sym, twx, twy;
  fillTables(*p, sym, twx, twy);
  std::vector<int> single(capacity); for (int k = 0; k < capacity; k++) single[k] = k;
  struct { void **ptr; const void *src; size_t bytes; } tabs[] = {
    {&p->zs, p->zsH.data(), p->zsH.size() * 8}, {&p->dzs, p->dzsH.data(), p->dzsH.size() * 8}, {&p->hz, p->hzH.data(), p->hzH.size() * 8},
    {&p->sym, sym.data(), sym.size()}, {&p->twx, twx.data(), twx.size()}, {&p->twy, twy.data(), twy.size()},
    {(void **)&p->single, single.data(), single.size() * sizeof(int)}};
  for (auto &t : tabs) {
    POOLCHK(hipMalloc(t.ptr, t.bytes));
    POOLCHK(hipMemcpy(*t.ptr, t.src, t.bytes, hipMemcpyHostToDevice));
  }
  // The electric modes' elimination, once for the grid; the discharge's solve takes as many levels a pass as the
  // block's LDS holds beside its own state.
  const int L = g.nz - 2;
  POOLCHK(hipMalloc(&p->efac, size_t(3) * L * g.nm * 8));
  hipLaunchKernelGGL(k_electric_factors, dim3(blocks(g.nm, 256)), dim3(256), 0, c->stream, g, (const double *)p->sym, (const double *)p->dzs,
                     (const double *)p->hz, (double *)p->efac);
  POOLCHK(hipGetLastError());
  POOLCHK(hipStreamSynchronize(c->stream));
  const size_t dischargeStatic = 8192, ldsMax = c->prop.sharedMemPerBlock;
  int chunk = 1;
  while (chunk < L && solveLdsBytes(g, chunk + 1) + dischargeStatic <= ldsMax) chunk++;
  if (solveLdsBytes(g, chunk) + dischargeStatic > ldsMax) { freePool(p); return fail(c, ML_E_UNSUPPORTED, "grid too large for the discharge's level in LDS"); }
  p->chunk = chunk;
  // METEOLAB_FRONT_LDS_ENTRIES lowers the front's LDS capacity (a check of the global overflow; results are the same).
  p->frontCap = frontCapacity(g, chunk);
  if (const char *env = std::getenv("METEOLAB_FRONT_LDS_ENTRIES")) p->frontCap = std::max(0, std::min(p->frontCap, std::atoi(env)));
  const int dyn = int(solveLdsBytes(g, chunk));
  if (p->rb == 8) POOLCHK(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_discharge<double>), hipFuncAttributeMaxDynamicSharedMemorySize, dyn));
  else POOLCHK(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_discharge<float>), hipFuncAttributeMaxDynamicSharedMemorySize, dyn));
  // The rotation diagnostics: the first level at or above 2 km to the last at or below 5 km (model.ts:389), the last
  // levels at or below 1 and 3 km (:1612), and the core-vorticity threshold of this spacing (mesoCoreZeta, :129).
  {
    RotLevels &r = p->rot;
    r.z0 = g.nz; r.z1 = -1; r.top1 = -1; r.top3 = -1;
    for (int z = 0; z < g.nz; z++) {
      const double h = p->zsH[z];
      if (h >= 2000 && r.z0 == g.nz) r.z0 = z;
      if (h <= 5000) r.z1 = z;
      if (h <= 1000) r.top1 = z;
      if (h <= 3000) r.top3 = z;
    }
    r.meso = 3e-3 * std::max(1.0, 1200.0 / g.dx);
    const int rotLds = int(rotationLdsBytes(g));
    if (size_t(rotLds) > ldsMax) { freePool(p); return fail(c, ML_E_UNSUPPORTED, "grid too large for the rotation diagnostics' columns in LDS"); }
    if (p->rb == 8) POOLCHK(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_rotation<double>), hipFuncAttributeMaxDynamicSharedMemorySize, rotLds));
    else POOLCHK(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_rotation<float>), hipFuncAttributeMaxDynamicSharedMemorySize, rotLds));
  }
  // The model diagnostics: cores between the first level at or above 1.9 km and the last at or below 6.6 km, at most
  // nz - 3 (model.ts:1653); the cold pool below 1.5 km (:1680).
  {
    DiagLevels &d = p->dlev;
    d.coreLo = g.nz; d.coreHi = -1; d.coldTop = 0;
    for (int z = 0; z < g.nz; z++) {
      const double h = p->zsH[z];
      if (h >= 1900 && d.coreLo == g.nz) d.coreLo = z;
      if (h <= 6600) d.coreHi = z;
      if (h < 1500) d.coldTop = z + 1;
    }
    d.coreHi = std::min(g.nz - 3, d.coreHi);
    const int diagLds = int(modelDiagLdsBytes(g));
    if (size_t(diagLds) > ldsMax) { freePool(p); return fail(c, ML_E_UNSUPPORTED, "grid too large for the diagnostics' columns in LDS"); }
    if (p->rb == 8) POOLCHK(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_model_diag<double>), hipFuncAttributeMaxDynamicSharedMemorySize, diagLds));
    else POOLCHK(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_model_diag<float>), hipFuncAttributeMaxDynamicSharedMemorySize, diagLds));
  }
  // The transport layouts' level planes in LDS (the plane layout only where it fits; the autotuner skips it otherwise).
  const int levelLds = int(stageLevelLdsBytes(g, p->rb)), planeLds = int(stagePlaneLdsBytes(g, p->rb));
  if (p->rb == 8) {
    POOLCHK(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_stage_level<double>), hipFuncAttributeMaxDynamicSharedMemorySize, levelLds));
    if (planeFits(g, p->rb, ldsMax)) POOLCHK(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_stage_plane<double>), hipFuncAttributeMaxDynamicSharedMemorySize, planeLds));
  } else {
    POOLCHK(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_stage_level<float>), hipFuncAttributeMaxDynamicSharedMemorySize, levelLds));
    if (planeFits(g, p->rb, ldsMax)) POOLCHK(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_stage_plane<float>), hipFuncAttributeMaxDynamicSharedMemorySize, planeLds));
  }
  p->used.assign(capacity, 0);
  if (id == int(c->pools.size())) c->pools.push_back(p); else c->pools[id] = p;
  *pool = id;
  return 0;
}
#undef POOLCHK

int ml_pool_free(ml_ctx *c, int pool) {
  Pool *p = c ? poolOf(c, pool) : nullptr;
  if (!p) return ML_E_ARG;
  return int(std::count(p->used.begin(), p->used.end(), 0));
}

}  // extern "C"

namespace {
// The parameters of a run's slot from its sounding and settings (the values of SimConfig the step reads).
void runParams(const Pool *p, const ml_profile *pr, const ml_run_config *rc, double *P) {
  for (int k = 0; k < N_PARAMS; k++) P[k] = 0;
  P[P_FRAME_U] = rc->frame_u; P[P_FRAME_V] = rc->frame_v;
  P[P_CORIOLIS] = 2 * 7.2921e-5 * std::sin(rc->latitude * M_PI / 180);
  P[P_CS] = .18 * rc->turbulence / .55;
  P[P_DRAG] = std::pow(.4 / std::log(2 * p->hzH[0] / 4 / rc->roughness), 2);
  P[P_SPONGE_START] = std::max(rc->tropopause_km * 1000 + 2500, 14000.0);
  P[P_SEED] = rc->seed; P[P_BUBBLE_AMP] = rc->bubble_amp; P[P_BUBBLE_X] = rc->bubble_x; P[P_BUBBLE_Y] = rc->bubble_y; P[P_BUBBLE_Z] = rc->bubble_z;
  P[P_BUBBLE_RH] = rc->bubble_rh; P[P_BUBBLE_RZ] = rc->bubble_rz; P[P_NOISE_SEED] = rc->noise_seed; P[P_NOISE_AMP] = rc->noise_amp; P[P_NOISE_TOP] = rc->noise_top;
  P[P_QVMAX] = pr->qv_max; P[P_SHEAR] = pr->shear; P[P_SHEAR_DEPTH] = pr->shear_depth; P[P_THETA0] = pr->theta0; P[P_DTHETA] = pr->dtheta;
  P[P_ZTR] = pr->ztr; P[P_TTR] = pr->ttr; P[P_RH_DROP] = pr->rh_drop; P[P_RH_TOP] = pr->rh_top; P[P_ELECTRIFIED] = p->electrified; P[P_TIME] = 0;
  for (int k = 0; k < 4; k++) { P[P_WIND_U + k] = pr->wind_u[k]; P[P_WIND_V + k] = pr->wind_v[k]; }
}

// `count` runs in free slots of pool `pool` (profiles[k], configs[k]): their parameters, cleared state, then one launch
// of the base-state integration (a block per run) and one of the initial state for all of them; one wait at the end.
int createRuns(ml_ctx *c, int pool, int count, const ml_profile *prs, const ml_run_config *rcs, int *runs) {
  Pool *p = c ? poolOf(c, pool) : nullptr;
  if (!p || !prs || !rcs || !runs || count < 1) return fail(c, ML_E_ARG, "run: bad pool or arguments");
  std::vector<int> slots;
  for (int k = 0; k < p->cap && int(slots.size()) < count; k++) if (!p->used[k]) slots.push_back(k);
  if (int(slots.size()) < count) return fail(c, ML_E_FULL, "pool full");
  HIPCHK(c, hipSetDevice(c->device));
  const Geo &g = p->g;
  Sizes s = sizesOf(g, p->rb);
  std::vector<double> P(size_t(count) * N_PARAMS);
  const double firstFlash = -1;
  for (int k = 0; k < count; k++) {
    const int slot = slots[k];
    runParams(p, prs + k, rcs + k, P.data() + size_t(k) * N_PARAMS);
    HIPCHK(c, hipMemcpyAsync((double *)p->par + size_t(slot) * N_PARAMS, P.data() + size_t(k) * N_PARAMS, N_PARAMS * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync((double *)p->scal + size_t(slot) * N_SCALARS, 0, N_SCALARS * 8, c->stream));
    HIPCHK(c, hipMemsetAsync((char *)p->stx + s.stage * slot, 0, s.stage, c->stream));
    HIPCHK(c, hipMemsetAsync((char *)p->sty + s.stage * slot, 0, s.stage, c->stream));
    HIPCHK(c, hipMemsetAsync((char *)p->ephi + s.node8 * slot, 0, s.node8, c->stream));
    HIPCHK(c, hipMemsetAsync((char *)p->press + s.node8 * slot, 0, s.node8, c->stream));
    HIPCHK(c, hipMemsetAsync((char *)p->lmark + s.lmark * slot, 0, s.lmark, c->stream));
    HIPCHK(c, hipMemsetAsync((char *)p->flog + s.flog * slot, 0, s.flog, c->stream));
    HIPCHK(c, hipMemcpyAsync((double *)p->scal + size_t(slot) * N_SCALARS + S_FIRST_FLASH, &firstFlash, 8, hipMemcpyHostToDevice, c->stream));
  }
  // The new slots' list: the pool's single-slot list where there is one run, else the diagnostics' run list (stream
  // ordered with every other use of it).
  const int *list = p->single + slots[0];
  if (count > 1) {
    HIPCHK(c, hipMemcpyAsync(p->dlist, slots.data(), count * sizeof(int), hipMemcpyHostToDevice, c->stream));
    list = p->dlist;
  }
  size_t lnp = (size_t(std::ceil((g.height + 2000) / 10)) + 2) * sizeof(double);
  if (p->rb == 8) {
    Bufs<double> b = bufs<double>(*p, list);
    hipLaunchKernelGGL(k_base_state<double>, dim3(count), dim3(64), lnp, c->stream, g, b);
    hipLaunchKernelGGL(k_init<double>, dim3(blocks(g.n, 256), count), dim3(256), 0, c->stream, g, b);
  } else {
    Bufs<float> b = bufs<float>(*p, list);
    hipLaunchKernelGGL(k_base_state<float>, dim3(count), dim3(64), lnp, c->stream, g, b);
    hipLaunchKernelGGL(k_init<float>, dim3(blocks(g.n, 256), count), dim3(256), 0, c->stream, g, b);
  }
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < count; k++) { p->used[slots[k]] = 1; runs[k] = (pool << SLOT_BITS) | slots[k]; }
  return 0;
}
}  // namespace

extern "C" {

int ml_run_create(ml_ctx *c, int pool, const ml_profile *pr, const ml_run_config *rc, int *run) {
  return createRuns(c, pool, 1, pr, rc, run);
}

int ml_runs_create(ml_ctx *c, int pool, int count, const ml_profile *profiles, const ml_run_config *configs, int *runs) {
  return createRuns(c, pool, count, profiles, configs, runs);
}

int ml_pool_grow(ml_ctx *c, int pool, int capacity) {
  Pool *p = c ? poolOf(c, pool) : nullptr;
  if (!p) return fail(c, ML_E_ARG, "pool grow: bad pool");
  if (capacity < p->cap || capacity >= (1 << SLOT_BITS)) return fail(c, ML_E_ARG, "pool grow: a capacity from the pool's own up to 65535");
  if (capacity == p->cap) return 0;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipStreamSynchronize(c->side));
  const Sizes s = sizesOf(p->g, p->rb);
  const std::vector<SlotArray> arrays = slotArrays(p, s);
  const size_t old = p->cap, cap = capacity;
  std::vector<void *> fresh(arrays.size(), nullptr);
  int *single = nullptr;
  std::vector<int> list(cap);
  for (size_t k = 0; k < cap; k++) list[k] = int(k);
  // The new arrays: the slots' parts copied, the new slots cleared. On an error the pool is left as it was.
  hipError_t e = hipSuccess;
  for (size_t k = 0; k < arrays.size() && e == hipSuccess; k++) {
    e = hipMalloc(&fresh[k], arrays[k].bytes * cap);
    if (e == hipSuccess) e = hipMemcpyAsync(fresh[k], *arrays[k].ptr, arrays[k].bytes * old, hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync((char *)fresh[k] + arrays[k].bytes * old, 0, arrays[k].bytes * (cap - old), c->stream);
  }
  if (e == hipSuccess) e = hipMalloc((void **)&single, cap * sizeof(int));
  if (e == hipSuccess) e = hipMemcpyAsync(single, list.data(), cap * sizeof(int), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    (void)hipStreamSynchronize(c->stream);
    for (void *f : fresh) if (f) (void)hipFree(f);
    if (single) (void)hipFree(single);
    return fail(c, ML_E_HIP, std::string("pool grow: ") + hipGetErrorString(e));
  }
  for (size_t k = 0; k < arrays.size(); k++) { (void)hipFree(*arrays[k].ptr); *arrays[k].ptr = fresh[k]; }
  (void)hipFree(p->single);
  p->single = single;
  p->cap = capacity;
  p->used.resize(cap, 0);
  // The captured steps hold the old arrays' addresses: each batch is captured again at its next step (its tuning kept).
  for (auto &kv : p->batches) kv.second->dt = 0;
  return 0;
}

int ml_mem_info(ml_ctx *c, int64_t *free_bytes, int64_t *total_bytes) {
  if (!c || !free_bytes || !total_bytes) return fail(c, ML_E_ARG, "mem info: null argument");
  HIPCHK(c, hipSetDevice(c->device));
  size_t f = 0, t = 0;
  HIPCHK(c, hipMemGetInfo(&f, &t));
  *free_bytes = int64_t(f); *total_bytes = int64_t(t);
  return 0;
}

int ml_ctx_query(ml_ctx *c) {
  if (!c) return ML_E_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  const hipError_t e = hipStreamQuery(c->stream);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  return fail(c, ML_E_HIP, std::string("query: ") + hipGetErrorString(e));
}

int ml_run_release(ml_ctx *c, int run) {
  Pool *p; int slot;
  if (!c || !decode(c, run, p, slot)) return fail(c, ML_E_ARG, "bad run");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  p->used[slot] = 0;
  return 0;
}

int ml_run_upload(ml_ctx *c, int run, int field, const float *host) {
  Pool *p; int slot;
  if (!c || !decode(c, run, p, slot) || field < 0 || field >= ML_FIELDS || !host) return fail(c, ML_E_ARG, "upload: bad run or field");
  size_t n = p->g.n;
  HIPCHK(c, hipMemcpyAsync((float *)p->fields + (size_t(slot) * N_FIELDS + field) * n, host, n * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}
int ml_run_download(ml_ctx *c, int run, int field, float *host) {
  Pool *p; int slot;
  if (!c || !decode(c, run, p, slot) || field < 0 || field >= ML_FIELDS || !host) return fail(c, ML_E_ARG, "download: bad run or field");
  size_t n = p->g.n;
  HIPCHK(c, hipMemcpyAsync(host, (float *)p->fields + (size_t(slot) * N_FIELDS + field) * n, n * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}
int ml_run_upload_column(ml_ctx *c, int run, int column, const float *host) {
  Pool *p; int slot;
  if (!c || !decode(c, run, p, slot) || column < 0 || column >= N_COLUMNS || !host) return fail(c, ML_E_ARG, "upload column: bad run or column");
  size_t l = p->g.layer;
  HIPCHK(c, hipMemcpyAsync((float *)p->cols + (size_t(slot) * N_COLUMNS + column) * l, host, l * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}
int ml_run_download_column(ml_ctx *c, int run, int column, float *host) {
  Pool *p; int slot;
  if (!c || !decode(c, run, p, slot) || column < 0 || column >= N_COLUMNS || !host) return fail(c, ML_E_ARG, "download column: bad run or column");
  size_t l = p->g.layer;
  HIPCHK(c, hipMemcpyAsync(host, (float *)p->cols + (size_t(slot) * N_COLUMNS + column) * l, l * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}
int ml_run_levels(ml_ctx *c, int run, double *out) {
  Pool *p; int slot;
  if (!c || !decode(c, run, p, slot) || !out) return fail(c, ML_E_ARG, "levels: bad run");
  HIPCHK(c, hipMemcpyAsync(out, (double *)p->lev + size_t(slot) * N_LEVELS * p->g.nz, size_t(ML_LEVELS) * p->g.nz * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}
int ml_run_pressure(ml_ctx *c, int run, double *out) {
  Pool *p; int slot;
  if (!c || !decode(c, run, p, slot) || !out) return fail(c, ML_E_ARG, "pressure: bad run");
  HIPCHK(c, hipSetDevice(c->device));
  // The cells (nz - 1 levels of nx ny) lead the slot's node-sized pressure array.
  HIPCHK(c, hipMemcpyAsync(out, (double *)p->press + size_t(slot) * p->g.n, size_t(p->g.nz - 1) * p->g.layer * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}
int ml_run_set_time(ml_ctx *c, int run, double time) {
  Pool *p; int slot;
  if (!c || !decode(c, run, p, slot)) return fail(c, ML_E_ARG, "set time: bad run");
  HIPCHK(c, hipMemcpyAsync((double *)p->par + size_t(slot) * N_PARAMS + P_TIME, &time, 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int ml_run_set_totals(ml_ctx *c, int run, double charge_ground, int64_t flashes_ic, int64_t flashes_cg, double persisted) {
  Pool *p; int slot;
  if (!c || !decode(c, run, p, slot) || flashes_ic < 0 || flashes_cg < 0 || !(persisted >= 0)) return fail(c, ML_E_ARG, "set totals: bad run or value");
  double v[3];
  v[0] = charge_ground; std::memcpy(v + 1, &flashes_ic, 8); std::memcpy(v + 2, &flashes_cg, 8);
  static_assert(S_FLASH_IC == S_CHARGE_GROUND + 1 && S_FLASH_CG == S_CHARGE_GROUND + 2, "the counts are adjacent scalars");
  double *sc = (double *)p->scal + size_t(slot) * N_SCALARS;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(sc + S_CHARGE_GROUND, v, sizeof(v), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(sc + S_ROT + ROT_PERSISTED, &persisted, 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int ml_run_intervene(ml_ctx *c, int run, int kind, double x, double y, double strength, double thermal_radius, double *out) {
  Pool *p; int slot;
  if (!c || !decode(c, run, p, slot)) return fail(c, ML_E_ARG, "intervene: bad run");
  if (kind < 0 || kind >= ML_INTERVENTIONS || !std::isfinite(x) || !std::isfinite(y) || !std::isfinite(strength) || !std::isfinite(thermal_radius))
    return fail(c, ML_E_ARG, "intervene: a kind 0-6 and finite position, strength and radius");
  // INTERVENTIONS (intervene.ts:22-30): radius (m, < 0: the thermal's own), depth (m), top (m, < 0: REACH = 3 depths),
  // theta (K), vapour (kg/kg), w (m/s), spin (m/s) per unit strength.
  static const double table[ML_INTERVENTIONS][7] = {
      {-1, 1800, 2700, 3.2, 0, 1.1, 0}, {5000, 1000, -1, 2, 0, 0, 0}, {5000, 1000, -1, -2, 0, 0, 0},
      {5000, 1200, -1, 0, .002, 0, 0}, {5000, 1200, -1, 0, -.002, 0, 0}, {4000, 700, -1, -6, 0, 0, 0},
      {3000, 3000, -1, 0, 0, 0, 10}};
  const double *t = table[kind];
  Patch k;
  k.radius = t[0] > 0 ? t[0] : (thermal_radius > 0 ? thermal_radius : 4.2) * 1000;
  k.depth = t[1]; k.theta = t[3]; k.vapour = t[4]; k.w = t[5]; k.spin = t[6];
  k.a0 = strength; k.x = x; k.y = y; k.width = p->width; k.depth_y = p->depth;
  const double top = t[2] > 0 ? t[2] : t[1] * 3;
  k.levels = 0;
  while (k.levels < p->g.nz && p->zsH[k.levels] < top) k.levels++;
  HIPCHK(c, hipSetDevice(c->device));
  if (p->rb == 8) hipLaunchKernelGGL(k_intervene<double>, dim3(1), dim3(INTERVENE_THREADS), 0, c->stream, p->g, bufs<double>(*p, p->single), slot, k);
  else hipLaunchKernelGGL(k_intervene<float>, dim3(1), dim3(INTERVENE_THREADS), 0, c->stream, p->g, bufs<float>(*p, p->single), slot, k);
  HIPCHK(c, hipGetLastError());
  double last[2];
  HIPCHK(c, hipMemcpyAsync(last, (double *)p->scal + size_t(slot) * N_SCALARS + S_LAST_ADDED, sizeof(last), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (out) { out[0] = last[0]; out[1] = last[1]; }
  return 0;
}

int ml_run_added(ml_ctx *c, int run, double *out) {
  Pool *p; int slot;
  if (!c || !decode(c, run, p, slot) || !out) return fail(c, ML_E_ARG, "added: bad run");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(out, (double *)p->scal + size_t(slot) * N_SCALARS + S_ADDED, 3 * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int ml_run_set_added(ml_ctx *c, int run, double heat, double vapour, double count) {
  Pool *p; int slot;
  if (!c || !decode(c, run, p, slot) || !std::isfinite(heat) || !std::isfinite(vapour) || !(count >= 0) || count != std::floor(count))
    return fail(c, ML_E_ARG, "set added: bad run or value");
  const double v[3] = {heat, vapour, count};
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync((double *)p->scal + size_t(slot) * N_SCALARS + S_ADDED, v, sizeof(v), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int ml_run_rotation(ml_ctx *c, int run, double *out) {
  Pool *p; int slot;
  if (!c || !decode(c, run, p, slot) || !out) return fail(c, ML_E_ARG, "rotation: bad run");
  static_assert(N_ROT == ML_ROT_COUNT, "the ABI's rotation record is the scalars'");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(out, (double *)p->scal + size_t(slot) * N_SCALARS + S_ROT, N_ROT * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int ml_run_set_swaths(ml_ctx *c, int run, int on) {
  Pool *p; int slot;
  if (!c || !decode(c, run, p, slot) || (on != 0 && on != 1)) return fail(c, ML_E_ARG, "set swaths: bad run or flag");
  const double v = on;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync((double *)p->par + size_t(slot) * N_PARAMS + P_SWATHS, &v, 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int ml_check_weno(ml_ctx *c, const float *stencils, int n, double *out) {
  if (!c || !stencils || !out || n < 1 || n > (1 << 24)) return fail(c, ML_E_ARG, "check weno: 1 .. 2^24 stencils and an output");
  HIPCHK(c, hipSetDevice(c->device));
  float *dst = nullptr; double *dout = nullptr;
  HIPCHK(c, hipMalloc(&dst, size_t(n) * 5 * sizeof(float)));
  hipError_t e = hipMalloc(&dout, size_t(n) * sizeof(double));
  if (e == hipSuccess) e = hipMemcpyAsync(dst, stencils, size_t(n) * 5 * sizeof(float), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    if (c->precision == 64) hipLaunchKernelGGL(k_check_weno<double>, dim3(blocks(n, 256)), dim3(256), 0, c->stream, dst, n, dout);
    else hipLaunchKernelGGL(k_check_weno<float>, dim3(blocks(n, 256)), dim3(256), 0, c->stream, dst, n, dout);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout, size_t(n) * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  hipFree(dst); if (dout) hipFree(dout);
  if (e != hipSuccess) return fail(c, ML_E_HIP, std::string("check weno: ") + hipGetErrorString(e));
  return 0;
}

int ml_probability_map(ml_ctx *c, const int *runs, int count, int column, double threshold, float *out) {
  if (!c || !runs || count < 1 || !out || column < 0 || column >= ML_COLUMNS) return fail(c, ML_E_ARG, "probability map: runs, a column and an output");
  Pool *p = nullptr;
  std::vector<int> slots(count);
  for (int k = 0; k < count; k++) {
    Pool *q; int slot;
    if (!decode(c, runs[k], q, slot)) return fail(c, ML_E_ARG, "probability map: bad run");
    if (p && q != p) return fail(c, ML_E_ARG, "probability map: runs of one pool per call");
    p = q; slots[k] = slot;
  }
  if (count > p->cap) return fail(c, ML_E_ARG, "probability map: more runs than the pool holds");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(p->dlist, slots.data(), count * sizeof(int), hipMemcpyHostToDevice, c->stream));
  float *dev = reinterpret_cast<float *>(p->pmap);
  const size_t need = size_t(p->g.layer) * sizeof(float);
  if (p->rb == 8) hipLaunchKernelGGL(k_probability<double>, dim3(blocks(p->g.layer, 256)), dim3(256), 0, c->stream, p->g, bufs<double>(*p, p->dlist), count, column, threshold, dev);
  else hipLaunchKernelGGL(k_probability<float>, dim3(blocks(p->g.layer, 256)), dim3(256), 0, c->stream, p->g, bufs<float>(*p, p->dlist), count, column, threshold, dev);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(out, dev, need, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int ml_run_refresh_rotation(ml_ctx *c, int run) {
  Pool *p; int slot;
  if (!c || !decode(c, run, p, slot)) return fail(c, ML_E_ARG, "refresh rotation: bad run");
  HIPCHK(c, hipSetDevice(c->device));
  const int *one = p->single + slot;
  if (p->rb == 8) hipLaunchKernelGGL(k_rotation<double>, dim3(1), dim3(ROTATION_THREADS), rotationLdsBytes(p->g), c->stream, p->g, bufs<double>(*p, one), 0.0, p->rot, 1);
  else hipLaunchKernelGGL(k_rotation<float>, dim3(1), dim3(ROTATION_THREADS), rotationLdsBytes(p->g), c->stream, p->g, bufs<float>(*p, one), 0.0, p->rot, 1);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int ml_run_diag(ml_ctx *c, int run, double *out) {
  Pool *p; int slot;
  if (!c || !decode(c, run, p, slot) || !out) return fail(c, ML_E_ARG, "diag: bad run");
  HIPCHK(c, hipSetDevice(c->device));
  const int *one = p->single + slot;
  if (p->rb == 8) hipLaunchKernelGGL(k_diag<double>, dim3(1), dim3(1024), 0, c->stream, p->g, bufs<double>(*p, one), p->diag);
  else hipLaunchKernelGGL(k_diag<float>, dim3(1), dim3(1024), 0, c->stream, p->g, bufs<float>(*p, one), p->diag);
  HIPCHK(c, hipGetLastError());
  double d[N_DIAG_DEV], sc[N_SCALARS], par[N_PARAMS];
  HIPCHK(c, hipMemcpyAsync(d, p->diag, sizeof(d), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(sc, (double *)p->scal + size_t(slot) * N_SCALARS, sizeof(sc), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(par, (double *)p->par + size_t(slot) * N_PARAMS, sizeof(par), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  auto i64 = [&](int k) { int64_t v; std::memcpy(&v, sc + k, 8); return double(v); };
  out[ML_D_TIME] = par[P_TIME]; out[ML_D_MAXW] = d[D_MAXW]; out[ML_D_MINW] = d[D_MINW]; out[ML_D_MAXTHETAP] = d[D_MAXTHETAP];
  out[ML_D_RAINTOTAL] = d[D_RAINTOTAL]; out[ML_D_WATER] = d[D_WATER]; out[ML_D_CHARGE] = d[D_CHARGE]; out[ML_D_CLIPPED] = i64(S_CLIPPED);
  out[ML_D_NEGFILLED] = sc[S_NEGATIVE]; out[ML_D_CHARGEGROUND] = sc[S_CHARGE_GROUND]; out[ML_D_FLASHES_IC] = i64(S_FLASH_IC);
  out[ML_D_FLASHES_CG] = i64(S_FLASH_CG); out[ML_D_FIRSTFLASH] = sc[S_FIRST_FLASH]; out[ML_D_ERROR] = i64(S_ERROR);
  return 0;
}

int ml_model_diagnostics(ml_ctx *c, const int *runs, int count, double *out) {
  if (!c || !runs || count < 1 || !out) return fail(c, ML_E_ARG, "model diagnostics: no runs");
  Pool *p = nullptr;
  std::vector<int> slots(count);
  for (int k = 0; k < count; k++) {
    Pool *q; int slot;
    if (!decode(c, runs[k], q, slot)) return fail(c, ML_E_ARG, "model diagnostics: bad run");
    if (p && q != p) return fail(c, ML_E_ARG, "model diagnostics: runs of one pool per call");
    p = q; slots[k] = slot;
  }
  if (count > p->cap) return fail(c, ML_E_ARG, "model diagnostics: more runs than the pool holds");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(p->dlist, slots.data(), count * sizeof(int), hipMemcpyHostToDevice, c->stream));
  const size_t lds = modelDiagLdsBytes(p->g);
  if (p->rb == 8) hipLaunchKernelGGL(k_model_diag<double>, dim3(count), dim3(MDIAG_THREADS), lds, c->stream, p->g, bufs<double>(*p, p->dlist), p->dlev, p->diag);
  else hipLaunchKernelGGL(k_model_diag<float>, dim3(count), dim3(MDIAG_THREADS), lds, c->stream, p->g, bufs<float>(*p, p->dlist), p->dlev, p->diag);
  HIPCHK(c, hipGetLastError());
  std::vector<double> rows(size_t(count) * DIAG_ROW);
  HIPCHK(c, hipMemcpyAsync(rows.data(), p->diag, rows.size() * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  static_assert(N_MD == ML_MD_COUNT, "the ABI's diagnostics record is the kernel's row");
  for (int k = 0; k < count; k++) std::memcpy(out + size_t(k) * ML_MD_COUNT, rows.data() + size_t(k) * DIAG_ROW, ML_MD_COUNT * 8);
  return 0;
}

// An error of ml_model_diagnostics_noalloc: the message goes into the context's reserved string without a temporary.
static int failFixed(ml_ctx *c, int code, const char *msg) { if (c) c->err.assign(msg); return code; }
#define HIPCHK_FIXED(c, x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return failFixed(c, ML_E_HIP, hipGetErrorString(e_)); } while (0)

int ml_model_diagnostics_noalloc(ml_ctx *c, const int *runs, int count, double *out) {
  if (!c || !runs || count < 1 || !out) return failFixed(c, ML_E_ARG, "model diagnostics (no allocation): no runs");
  if (count > ML_NOALLOC_RUNS) return failFixed(c, ML_E_ARG, "model diagnostics (no allocation): more than ML_NOALLOC_RUNS (1024) runs a call");
  int slots[ML_NOALLOC_RUNS];
  Pool *p = nullptr;
  for (int k = 0; k < count; k++) {
    Pool *q; int slot;
    if (!decode(c, runs[k], q, slot)) return failFixed(c, ML_E_ARG, "model diagnostics (no allocation): bad run");
    if (p && q != p) return failFixed(c, ML_E_ARG, "model diagnostics (no allocation): runs of one pool per call");
    p = q; slots[k] = slot;
  }
  if (count > p->cap) return failFixed(c, ML_E_ARG, "model diagnostics (no allocation): more runs than the pool holds");
  HIPCHK_FIXED(c, hipSetDevice(c->device));
  // This device's scratch: the rows first, then the run list (the kernel's Bufs::runs).
  double *rows = static_cast<double *>(c->scratch);
  int *list = reinterpret_cast<int *>(static_cast<unsigned char *>(c->scratch) + NOALLOC_ROWS_BYTES);
  HIPCHK_FIXED(c, hipMemcpyAsync(list, slots, count * sizeof(int), hipMemcpyHostToDevice, c->stream));
  const size_t lds = modelDiagLdsBytes(p->g);
  if (p->rb == 8) hipLaunchKernelGGL(k_model_diag<double>, dim3(count), dim3(MDIAG_THREADS), lds, c->stream, p->g, bufs<double>(*p, list), p->dlev, rows);
  else hipLaunchKernelGGL(k_model_diag<float>, dim3(count), dim3(MDIAG_THREADS), lds, c->stream, p->g, bufs<float>(*p, list), p->dlev, rows);
  HIPCHK_FIXED(c, hipGetLastError());
  // The rows' first ML_MD_COUNT doubles, straight into the caller's array.
  HIPCHK_FIXED(c, hipMemcpy2DAsync(out, ML_MD_COUNT * sizeof(double), rows, DIAG_ROW * sizeof(double), ML_MD_COUNT * sizeof(double), count,
                                   hipMemcpyDeviceToHost, c->stream));
  // The run list was copied from this function's stack: the stream is drained before it returns.
  HIPCHK_FIXED(c, hipStreamSynchronize(c->stream));
  return 0;
}

int ml_run_flashes(ml_ctx *c, int run, double *out, int max) {
  Pool *p; int slot;
  if (!c || !decode(c, run, p, slot) || (!out && max > 0)) return fail(c, ML_E_ARG, "flashes: bad run");
  double sc[N_SCALARS], log[FLASH_LOG * FLASH_FIELDS];
  HIPCHK(c, hipMemcpyAsync(sc, (double *)p->scal + size_t(slot) * N_SCALARS, sizeof(sc), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(log, (double *)p->flog + size_t(slot) * FLASH_LOG * FLASH_FIELDS, sizeof(log), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int64_t total; std::memcpy(&total, sc + S_FLASH_LOG, 8);
  int kept = int(std::min<int64_t>(total, FLASH_LOG)), n = std::min(kept, max);
  // Oldest first among the last `kept` flashes; the last n of them.
  for (int k = 0; k < n; k++) {
    int64_t idx = total - n + k;
    std::memcpy(out + size_t(k) * FLASH_FIELDS, log + size_t(idx % FLASH_LOG) * FLASH_FIELDS, FLASH_FIELDS * 8);
  }
  return n;
}

}  // extern "C"

namespace {
// The autotuner: the transport kernel's levels per block, with and without the XCD-per-run block order, by timing the
// stage-0 transport of this batch (it reads the state and writes only stage buffer X, which every step rewrites) for
// each candidate, fastest of 3 launches.
template <class R> float timeStage(ml_ctx *c, const Pool &p, const Batch &bt, const Tuning &tu) {
  const Geo &g = p.g;
  Bufs<R> b = bufs<R>(p, bt.dev);
  const int nr = int(bt.runs.size()), nc = p.electrified ? N_CARRIED : F_COLD + 1;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0); (void)hipEventCreate(&e1);
  hipLaunchKernelGGL(k_zero_flags<R>, dim3(g.nz * nc, nr), dim3(256), 0, c->stream, g, b, 0, 1);
  float best = 1e30f;
  for (int rep = 0; rep < 3; rep++) {
    (void)hipEventRecord(e0, c->stream);
    launchStage<R>(g, b, nr, nc, 0, 0, 3.0, c->stream, tu);
    (void)hipEventRecord(e1, c->stream);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    best = std::min(best, ms);
  }
  (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
  return best;
}
// The transport layout of a grid by a fixed rule: the padded-plane layout where its level planes fit a block's LDS, else
// whole levels. The layouts round differently in float32, so the layout is a function of the grid alone: a run's result
// does not depend on the batch it is stepped in, on the process or host that steps it, or on timing noise.
static int layoutOf(const ml_ctx *c, const Pool *p) { return planeFits(p->g, p->rb, c->prop.sharedMemPerBlock) ? 2 : 1; }
static int tune(ml_ctx *c, Pool *p, Batch *bt) {
  HIPCHK(c, hipSetDevice(c->device));
  const int nr = int(bt->runs.size()), layout = layoutOf(c, p);
  // Only what leaves the arithmetic unchanged is timed: levels per block and the XCD-per-run block order.
  std::vector<Tuning> cands;
  for (int xcd = 0; xcd < (nr % 8 == 0 ? 2 : 1); xcd++)
    for (int k : {1, 2, 5, 10}) {
      if (k > p->g.nz) continue;
      Tuning lv; lv.layout = layout; lv.xcd = xcd; lv.levels = k; cands.push_back(lv);
    }
  float fastest = 1e30f;
  for (Tuning &t : cands) {
    t.ms = p->rb == 8 ? timeStage<double>(c, *p, *bt, t) : timeStage<float>(c, *p, *bt, t);
    fastest = std::min(fastest, t.ms);
  }
  HIPCHK(c, hipGetLastError());
  // The first candidate within 3 % of the fastest: shapes that time alike within the launches' noise resolve the same
  // way on every run of the same batch.
  Tuning best = cands[0];
  for (const Tuning &t : cands) if (t.ms <= fastest * 1.03f) { best = t; break; }
  bt->tune = best;
  bt->tuned = true;
  return 0;
}
}  // namespace

extern "C" {

static int batchOf(ml_ctx *c, const int *runs, int count, Pool *&p, Batch *&bt) {
  if (!c || !runs || count < 1) return fail(c, ML_E_ARG, "step: no runs");
  std::vector<int> slots(count);
  p = nullptr;
  for (int k = 0; k < count; k++) {
    Pool *q; int slot;
    if (!decode(c, runs[k], q, slot)) return fail(c, ML_E_ARG, "step: bad run");
    if (p && q != p) return fail(c, ML_E_ARG, "step: runs of one pool per call");
    p = q; slots[k] = slot;
  }
  std::vector<int> sorted(slots);
  std::sort(sorted.begin(), sorted.end());
  if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end()) return fail(c, ML_E_ARG, "step: a run twice in one batch");
  auto it = p->batches.find(count);
  if (it != p->batches.end()) {
    bt = it->second;
    if (bt->runs != slots) {
      // The same launch shape for other runs: the new run list goes ahead of the next replay in the stream.
      HIPCHK(c, hipSetDevice(c->device));
      HIPCHK(c, hipMemcpyAsync(bt->dev, slots.data(), count * sizeof(int), hipMemcpyHostToDevice, c->stream));
      bt->runs = slots;
    }
    return 0;
  }
  bt = new Batch();
  bt->runs = slots;
  HIPCHK(c, hipMalloc((void **)&bt->dev, count * sizeof(int)));
  HIPCHK(c, hipMemcpy(bt->dev, slots.data(), count * sizeof(int), hipMemcpyHostToDevice));
  p->batches[count] = bt;
  return tune(c, p, bt);
}

int ml_ctx_set_overlap(ml_ctx *c, int on) {
  if (!c || (on != 0 && on != 1)) return fail(c, ML_E_ARG, "overlap: 0 or 1");
  c->overlap = on;
  return 0;
}

}  // extern "C"

namespace {
// Whether the batch steps split around the lightning (a pool with lightning, overlap on).
bool overlapped(const ml_ctx *c, const Pool *p) { return c->overlap && p->electrified; }

// The batch's step graphs for dt: one whole step, or with the overlap the split step alone (the first of a call) and
// beside the discharge of the step before (the others).
int captureSteps(ml_ctx *c, Pool *p, Batch *bt, int count, double dt) {
  const int ov = overlapped(c, p) ? 1 : 0;
  if (bt->dt == dt && bt->overlap == ov && (ov ? bt->first != nullptr : bt->exec != nullptr)) return 0;
  hipGraphExec_t *execs[3] = {&bt->exec, &bt->first, &bt->next};
  hipGraph_t *graphs[3] = {&bt->graph, &bt->firstGraph, &bt->nextGraph};
  for (int k = 0; k < 3; k++) {
    if (*execs[k]) { hipGraphExecDestroy(*execs[k]); *execs[k] = nullptr; }
    if (*graphs[k]) { hipGraphDestroy(*graphs[k]); *graphs[k] = nullptr; }
  }
  for (int k = ov ? 1 : 0; k < (ov ? 3 : 1); k++) {
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    if (k == 0) {
      if (p->rb == 8) launchStep<double>(*p, bt->dev, count, dt, c->stream, nullptr, bt->tune);
      else launchStep<float>(*p, bt->dev, count, dt, c->stream, nullptr, bt->tune);
    } else {
      if (p->rb == 8) launchStepSplit<double>(*p, bt->dev, count, dt, c, bt->tune, k == 2);
      else launchStepSplit<float>(*p, bt->dev, count, dt, c, bt->tune, k == 2);
    }
    HIPCHK(c, hipStreamEndCapture(c->stream, graphs[k]));
    HIPCHK(c, hipGraphInstantiate(execs[k], *graphs[k], nullptr, nullptr, 0));
  }
  bt->dt = dt; bt->overlap = ov;
  return 0;
}

// Step k (0-based) of a call of `steps`: the graph replay, and after the last the discharge of that step.
int launchStepOf(ml_ctx *c, Pool *p, Batch *bt, int count, double dt, int k, int steps) {
  if (!bt->overlap) { HIPCHK(c, hipGraphLaunch(bt->exec, c->stream)); return 0; }
  HIPCHK(c, hipGraphLaunch(k == 0 ? bt->first : bt->next, c->stream));
  if (k == steps - 1) {
    if (p->rb == 8) launchDischarge<double>(*p, bufs<double>(*p, bt->dev), count, dt, c->stream, nullptr, 1);
    else launchDischarge<float>(*p, bufs<float>(*p, bt->dev), count, dt, c->stream, nullptr, 1);
  }
  return 0;
}
}  // namespace

extern "C" {

int ml_step(ml_ctx *c, const int *runs, int count, int steps, double dt) {
  Pool *p; Batch *bt;
  if (int r = batchOf(c, runs, count, p, bt)) return r;
  if (steps < 0 || !(dt > 0)) return fail(c, ML_E_ARG, "step: steps >= 0, dt > 0");
  HIPCHK(c, hipSetDevice(c->device));
  if (int r = captureSteps(c, p, bt, count, dt)) return r;
  for (int k = 0; k < steps; k++) if (int r = launchStepOf(c, p, bt, count, dt, k, steps)) return r;
  HIPCHK(c, hipGetLastError());
  return 0;
}

static int stepPart(ml_ctx *c, const int *runs, int count, double dt, int part) {
  Pool *p; Batch *bt;
  if (int r = batchOf(c, runs, count, p, bt)) return r;
  if (!(dt > 0)) return fail(c, ML_E_ARG, "step part: dt > 0");
  HIPCHK(c, hipSetDevice(c->device));
  if (p->rb == 8) launchPart<double>(*p, bt->dev, count, dt, c->stream, nullptr, bt->tune, part);
  else launchPart<float>(*p, bt->dev, count, dt, c->stream, nullptr, bt->tune, part);
  HIPCHK(c, hipGetLastError());
  return 0;
}
int ml_dynamics(ml_ctx *c, const int *runs, int count, double dt) { return stepPart(c, runs, count, dt, PART_DYNAMICS); }
int ml_physics(ml_ctx *c, const int *runs, int count, double dt) { return stepPart(c, runs, count, dt, PART_PHYSICS); }
int ml_mix(ml_ctx *c, const int *runs, int count, double dt) { return stepPart(c, runs, count, dt, PART_MIX); }
int ml_after_step(ml_ctx *c, const int *runs, int count, double dt) { return stepPart(c, runs, count, dt, PART_AFTER); }

int ml_fall_inputs(ml_ctx *c, int run, int *parts, int max) {
  Pool *p; int slot;
  if (!c || !decode(c, run, p, slot) || (!parts && max > 0)) return fail(c, ML_E_ARG, "fall inputs: bad run");
  const int falls = p->electrified ? N_FALLS : 4;
  if (max < falls) return fail(c, ML_E_ARG, "fall inputs: room for " + std::to_string(falls) + " falls");
  double sc[N_SCALARS];
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(sc, (double *)p->scal + size_t(slot) * N_SCALARS, sizeof(sc), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  // fallParts (model.ts:575-580): sub-steps from the largest Courant number of the fall's non-zero nodes.
  for (int k = 0; k < falls; k++) parts[k] = std::max(1, int(std::ceil(sc[S_COURANT + k] / .8)));
  return falls;
}

int ml_batch_tuning(ml_ctx *c, const int *runs, int count, int *shape, float *ms) {
  Pool *p; Batch *bt;
  if (int r = batchOf(c, runs, count, p, bt)) return r;
  if (shape) { shape[0] = bt->tune.layout; shape[1] = bt->tune.groups; shape[2] = bt->tune.xcd; shape[3] = bt->tune.levels; }
  if (ms) *ms = bt->tune.ms;
  return 0;
}

int ml_batch_set_tuning(ml_ctx *c, const int *runs, int count, const int *shape) {
  Pool *p; Batch *bt;
  if (int r = batchOf(c, runs, count, p, bt)) return r;
  const int nc = p->electrified ? N_CARRIED : F_COLD + 1;
  if (!shape || shape[0] < 0 || shape[0] > 2 || shape[1] < 1 || shape[1] > nc || shape[2] < 0 || shape[2] > 1 || shape[3] < 1 || shape[3] > p->g.nz)
    return fail(c, ML_E_ARG, "set tuning: layout 1 or 2, 1 <= field groups <= carried fields, xcd 0 or 1, 1 <= levels <= nz");
  if (shape[0] == 0) return fail(c, ML_E_UNSUPPORTED, "set tuning: the strip layout (0) is not built; layouts 1 (levels) and 2 (planes)");
  if (shape[0] == 2 && !planeFits(p->g, p->rb, c->prop.sharedMemPerBlock)) return fail(c, ML_E_UNSUPPORTED, "set tuning: the plane layout needs nx, ny >= 6 and its LDS within a block's");
  bt->tune.layout = shape[0]; bt->tune.groups = 1; bt->tune.xcd = shape[2];
  bt->tune.levels = shape[3]; bt->tune.ms = 0;
  // The captured steps hold the old launch shape: captured again at the next step.
  bt->dt = 0;
  return 0;
}

int ml_step_timed(ml_ctx *c, const int *runs, int count, int steps, double dt, double *ms) {
  if (!ms || steps < 1) return fail(c, ML_E_ARG, "step timed: steps >= 1 and an output array");
  // Capture (or reuse) the batch's step graphs, then time each step between events on the stream (with the overlap, a
  // step's time holds the discharge of the step before it, and the last also its own).
  Pool *p; Batch *bt;
  if (int r = batchOf(c, runs, count, p, bt)) return r;
  HIPCHK(c, hipSetDevice(c->device));
  if (int r = captureSteps(c, p, bt, count, dt)) return r;
  std::vector<hipEvent_t> ev(steps + 1);
  for (auto &e : ev) HIPCHK(c, hipEventCreate(&e));
  HIPCHK(c, hipEventRecord(ev[0], c->stream));
  for (int k = 0; k < steps; k++) {
    if (int r = launchStepOf(c, p, bt, count, dt, k, steps)) return r;
    HIPCHK(c, hipEventRecord(ev[k + 1], c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < steps; k++) { float v = 0; HIPCHK(c, hipEventElapsedTime(&v, ev[k], ev[k + 1])); ms[k] = v; }
  for (auto &e : ev) (void)hipEventDestroy(e);
  return 0;
}

}  // extern "C"

namespace {
// A plain copy of float4s (the reference kernel of known traffic for the HBM counters and the achievable bandwidth).
__global__ __launch_bounds__(256) void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) dst[i] = src[i];
}
}  // namespace

extern "C" {

int ml_copy_probe(ml_ctx *c, int64_t bytes, int iters, double *gbps) {
  if (!c || bytes < 1024 || iters < 1 || !gbps) return fail(c, ML_E_ARG, "copy probe: bytes >= 1024, iters >= 1");
  HIPCHK(c, hipSetDevice(c->device));
  size_t n = size_t(bytes) / 16;
  float4 *a = nullptr, *b = nullptr;
  HIPCHK(c, hipMalloc(&a, n * 16));
  hipError_t e = hipMalloc(&b, n * 16);
  if (e != hipSuccess) { (void)hipFree(a); return fail(c, ML_E_HIP, "copy probe allocation"); }
  (void)hipMemsetAsync(a, 0, n * 16, c->stream);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0); (void)hipEventCreate(&e1);
  unsigned grid = unsigned(std::min<size_t>((n + 255) / 256, size_t(c->prop.multiProcessorCount) * 16));
  hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, c->stream, a, b, n);
  (void)hipEventRecord(e0, c->stream);
  for (int k = 0; k < iters; k++) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, c->stream, a, b, n);
  (void)hipEventRecord(e1, c->stream);
  e = hipStreamSynchronize(c->stream);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); (void)hipFree(a); (void)hipFree(b);
  if (e != hipSuccess) return fail(c, ML_E_HIP, hipGetErrorString(e));
  *gbps = 2.0 * double(n) * 16 * iters / (ms / 1e3) / 1e9;
  return 0;
}

int ml_sync(ml_ctx *c) {
  if (!c) return ML_E_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int ml_profile_step(ml_ctx *c, const int *runs, int count, int steps, double dt, double *ms, int max, char *names, int size) {
  Pool *p; Batch *bt;
  if (int r = batchOf(c, runs, count, p, bt)) return r;
  HIPCHK(c, hipSetDevice(c->device));
  std::vector<hipEvent_t> events; std::vector<std::string> labels;
  Timer t{c->stream, &events, &labels};
  std::vector<double> acc;
  std::vector<std::string> first;
  // The kernels ml_step runs, in line: with the overlap on, the step split around its lightning and the discharge
  // right after the step (what the overlap runs beside the next step).
  const bool split = overlapped(c, p);
  for (int k = 0; k < steps; k++) {
    events.clear(); labels.clear();
    if (p->rb == 8) launchStepTimed<double>(*p, bt->dev, count, dt, c->stream, &t, bt->tune, split);
    else launchStepTimed<float>(*p, bt->dev, count, dt, c->stream, &t, bt->tune, split);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (acc.empty()) { acc.assign(events.size(), 0.0); first = labels; }
    for (size_t e = 1; e < events.size(); e++) { float v = 0; hipEventElapsedTime(&v, events[e - 1], events[e]); acc[e] += v; }
    for (hipEvent_t e : events) hipEventDestroy(e);
  }
  int nstages = int(acc.size()) - 1;
  std::string list;
  for (int e = 1; e <= nstages && e - 1 < max; e++) { if (ms) ms[e - 1] = acc[e] / std::max(1, steps); list += first[e]; list += '\n'; }
  if (names && size > 0) { std::strncpy(names, list.c_str(), size - 1); names[size - 1] = 0; }
  return nstages;
}

}  // extern "C"
