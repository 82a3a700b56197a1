// The simulation context behind the C ABI: device allocation, the block
// loop, force evaluations, the Monte Carlo barostat, the minimiser and state
// readback. Mirrors GpuSimulation (webgpu/simulation.js:82-1341) method for
// method; the machinery is HIP streams, device-side rebuild decisions and a
// hipGraph per block of steps.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <initializer_list>
#include <map>
#include <unordered_map>
#include <numeric>
#include <string>
#include <vector>

#include "tuned.h"
#include "host.h"
#include "kernels.h"
#include "surrogate.h"

namespace rmd {
void setGlobalError(const std::string &m);
}

using namespace rmd;

#define HIP_CHECK(call)                                                                              \
  do {                                                                                               \
    hipError_t e_ = (call);                                                                          \
    if (e_ != hipSuccess)                                                                            \
      throw rmd::Error(RMD_E_DEVICE, std::string(#call) + " failed: " + hipGetErrorString(e_));     \
  } while (0)

namespace {

template <typename T>
T *devAlloc(size_t count, std::vector<void *> &held) {
  void *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc(&p, count * sizeof(T));
  if (e != hipSuccess)
    throw rmd::Error(e == hipErrorOutOfMemory ? RMD_E_BUDGET : RMD_E_DEVICE,
                     std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) + " bytes failed: " +
                         hipGetErrorString(e));
  // Zeroed and finished before returning: the context's stream is
  // non-blocking, so work queued on it later is not ordered after the null
  // stream, and a memset still in flight could clear an upload made after it.
  e = hipMemsetAsync(p, 0, count * sizeof(T), nullptr);
  if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
  if (e != hipSuccess) {
    (void)hipFree(p);
    throw rmd::Error(RMD_E_DEVICE, std::string("clearing a new buffer failed: ") + hipGetErrorString(e));
  }
  held.push_back(p);
  return static_cast<T *>(p);
}

template <typename T>
void upload(T *dst, const std::vector<T> &src, hipStream_t s) {
  if (src.empty()) return;
  HIP_CHECK(hipMemcpyAsync(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

std::vector<int> factorRadices(int K) {
  std::vector<int> r;
  int rem = K;
  while (rem % 4 == 0) {
    r.push_back(4);
    rem /= 4;
  }
  for (int p : {2, 3, 5, 7})
    while (rem % p == 0) {
      r.push_back(p);
      rem /= p;
    }
  if (rem != 1 || r.size() > 8) throw rmd::Error(RMD_E_INVALID, "PME grid size " + std::to_string(K) + " cannot be transformed");
  return r;
}

}  // namespace

// The streams and fork/join events of a context; a replica batch lends its
// own to every member context.
struct Streams {
  hipStream_t stream = nullptr;
  hipStream_t side[2] = {nullptr, nullptr};
  hipEvent_t fork = nullptr, join[2] = {nullptr, nullptr};
};

// The barostat's learned screen: the network, its training samples and
// normalisation, and the counts it reports. A context has one; a replica
// batch has one shared by its replicas.
struct ScreenState {
  rmd_screen_options options{};
  rmd_surrogate *net = nullptr;
  int phase = 0;
  double refVolume = 0;
  std::vector<float> x, t;  // raw features and residual targets
  float mean[RMD_SURROGATE_INPUTS] = {0}, scale[RMD_SURROGATE_INPUTS] = {0};
  double tMean = 0, tScale = 1;
  int64_t moves = 0, exact = 0, screened = 0, secondStage = 0, accepted = 0, residualCount = 0;
  double trainingLoss = 0, residual2 = 0, estimate2 = 0;
};

// What a block (and a trial appended to it) hands back to the host: the
// Scalars, the last step's energy terms, the trial's terms and the list's
// rebuild count when the trial started, gathered on the device at the end of
// the block's launch sequence so that one copy brings them over.
struct ReadBack {
  Scalars sc;
  double block[E_TERM_COUNT];
  double trial[E_TERM_COUNT];
  unsigned int rebuildsBefore;
};

struct rmd_ctx {
  System sys;
  // A member of a replica batch: the batch owns it (rmd_destroy leaves it
  // alone) and lends it its streams.
  bool batchMember = false;
  bool ownsStreams = true;
  rmd_options opt{};
  std::string lastError;
  // derived parameters
  double timestep = 0.002, temperature = 298, friction = 1, tolerance = 1e-5;
  uint32_t seed = 0;
  int maxIterations = 150;
  bool hasPressure = false;
  double pressure = 0;
  int barostatInterval = 25;
  bool dispersionOn = true;
  double skin = 0, slack = 2.0;
  int lanes = 64;
  size_t maxLds = 65536;
  CellPlan plan;
  rmd_pme_parameters pme{};
  std::vector<double> moduli[3];
  DispersionCoefficients disp;
  double box[9] = {0};
  double selfEnergy = 0, netEnergy = 0;
  double restraintK = 0;
  std::vector<int> restraintAtoms;
  // device
  int device = 0;
  hipStream_t stream = nullptr;
  // Force branches that need no neighbour list (PME, bonded) run on their own
  // streams, forked from and joined back to `stream` around every evaluation.
  bool concurrent = true;
  hipStream_t side[2] = {nullptr, nullptr};
  hipEvent_t fork = nullptr, join[2] = {nullptr, nullptr};
  std::vector<void *> held;
  Dev d{};
  PmeDev p{};
  BoxF *boxDev = nullptr;
  double *dModuli[3] = {nullptr, nullptr, nullptr};
  float4 *posSaved = nullptr, *velSaved = nullptr, *centroids = nullptr, *rebuildSaved = nullptr;
  long long *forceSaved = nullptr;
  Scalars *scSaved = nullptr;
  int *molStart = nullptr, *molAtoms = nullptr, *molOfAtom = nullptr;
  double *dTerms = nullptr;
  // Energy terms of the last step of the last block (E_TERM_COUNT doubles on
  // the device, copied to the host with the block's scalars), valid while
  // blockEndStep == step: the barostat's old potential without another
  // evaluation.
  double *dTermsBlock = nullptr;
  double blockEndTerms[E_TERM_COUNT] = {0};
  // A volume move's trial appended to the block before it (rmd_run): the
  // proposal on the device and its pinned staging copy, the state the trial
  // starts from, and the trial's energy terms (allocated at the first such
  // move).
  TrialParams *trialDev = nullptr, *trialHost = nullptr;
  float4 *trialPos = nullptr;
  long long *trialForce = nullptr;
  double *dTermsTrial = nullptr;
  // what a block hands back, gathered on the device for one copy
  ReadBack *readBack = nullptr;
  int64_t blockEndStep = -1;
  unsigned int *dMaxForce = nullptr;
  int nUnits = 0, waterUnits = 0, clusterUnits = 0, freeUnits = 0;
  int clusterLdsSlots = 0;  // largest per-block cluster slot count held in LDS
  int starMax = 0;          // the most constraints of any star unit
  // graphs per block length
  std::map<int, hipGraphExec_t> graphs;
  bool useGraphs = true;
  // host state
  int64_t step = 0;
  std::atomic<int> stop{0};
  int barostatAttempts = 0, barostatAccepted = 0, barostatRefusals = 0;
  int windowAttempts = 0, windowAccepted = 0;
  double volumeScale = -1;
  int64_t stepsSinceBarostat = 0;
  int pmeBinned = 0;   // whether the plane pipeline bins its atoms (set at creation)
  int pmeFolded = 0;   // whether force-only evaluations fold the x lines into the gathers (set at creation)
  int reachMax = 0;    // the cell plan's widest reach (0: CELL_REACH_MAX)
  int searchMode = 0;  // 0 by occupancy, 1 per atom, 2 per cell where the launch covers the cells
  int searchByCellCaptured = -1;  // the search form the captured steps were built for
  int neighbourGrowths = 0;
  unsigned int worstRequired = 0;
  unsigned int rebuildCount = 0;  // Scalars::rebuilds at the last read-back
  int64_t kineticSampledAtStep = -1;
  int64_t stepsRun = 0;
  bool initialised = false;
  int slotCounts[E_TERM_COUNT] = {0};
  IntegratorParams ip{};
  // The barostat's learned screen.
  ScreenState screen;
};

namespace {

double volumeOf(const double *b) { return std::fabs(b[0] * b[4] * b[8]); }

__global__ void gatherReadBackKernel(const Scalars *sc, const double *block, const double *trial,
                                     const unsigned int *rebuildsBefore, ReadBack *out) {
  const int t = threadIdx.x;
  if (t == 0) out->sc = *sc;
  if (t < E_TERM_COUNT) {
    out->block[t] = block[t];
    if (trial) out->trial[t] = trial[t];
  }
  if (t == 0 && rebuildsBefore) out->rebuildsBefore = *rebuildsBefore;
}

void invalidateGraphs(rmd_ctx *c) {
  for (auto &g : c->graphs) (void)hipGraphExecDestroy(g.second);
  c->graphs.clear();
}

// Write the device box: f32 rows and inverse widths for the pair kernels,
// FP64 rows for PME and restraints, the cell grid shape.
BoxF boxFor(const rmd_ctx *c, const double *box, const CellPlan &plan) {
  BoxF b{};
  double recip[9];
  reciprocalBoxVectors(box, recip);
  for (int k = 0; k < 3; ++k) {
    b.a[k] = (float)box[k];
    b.b[k] = (float)box[3 + k];
    b.c[k] = (float)box[6 + k];
  }
  b.inv[0] = (float)(1.0 / box[0]);
  b.inv[1] = (float)(1.0 / box[4]);
  b.inv[2] = (float)(1.0 / box[8]);
  for (int k = 0; k < 9; ++k) {
    b.recip[k] = (float)recip[k];
    b.recip64[k] = recip[k];
    b.box64[k] = box[k];
  }
  b.cutoff = (float)c->sys.cutoff;
  b.rc2 = (float)(c->sys.cutoff * c->sys.cutoff);
  b.rl2 = (float)(plan.listCutoff * plan.listCutoff);
  b.periodic = 1;
  for (int k = 0; k < 3; ++k) b.cells[k] = plan.counts[k];
  b.cellTotal = plan.cellCount;
  b.reach = plan.reach;
  b.cellSearch = searchByCell(c->searchMode, c->d.n, plan.cellCount) ? 1 : 0;
  return b;
}

BoxF boxFor(const rmd_ctx *c) { return boxFor(c, c->box, c->plan); }

bool rectangular(const double *box) {
  return box[1] == 0.0 && box[2] == 0.0 && box[3] == 0.0 && box[5] == 0.0 && box[6] == 0.0 && box[7] == 0.0;
}

void writeBox(rmd_ctx *c) {
  // the pair kernels have a rectangular-box form, chosen at launch: a box
  // that changes shape invalidates the captured launches
  const int rect = rectangular(c->box) ? 1 : 0;
  if (rect != c->d.rect) {
    c->d.rect = rect;
    invalidateGraphs(c);
  }
  c->p.xReal = rect;
  // the box travels as a kernel argument of the influence function's launch,
  // which also stores it: no staging buffer, no wait
  ewaldSelfEnergy(c->sys, c->pme.alpha, volumeOf(c->box), &c->selfEnergy, &c->netEnergy);
  const double *mods[3] = {c->dModuli[0], c->dModuli[1], c->dModuli[2]};
  pmeSetBoxAndInfluence(c->p, boxFor(c), c->boxDev, mods, c->pme.alpha, c->stream);
  HIP_CHECK(hipGetLastError());
}

void allocateCells(rmd_ctx *c, int cellCount) {
  if (cellCount <= c->d.cellCapacity) return;
  const int capacity = std::max(cellCount, c->d.cellCapacity * 2);
  c->d.cellCount = devAlloc<int>(capacity + 1, c->held);
  c->d.cellStart = devAlloc<int>(capacity + 1, c->held);
  c->d.cellCursor = devAlloc<int>(capacity + 1, c->held);
  c->d.cellCapacity = capacity;
  invalidateGraphs(c);
}

void allocateNeighbours(rmd_ctx *c, int cap) {
  c->d.cap = cap;
  c->d.nbrList = devAlloc<int>((size_t)c->d.nPad * cap, c->held);
  invalidateGraphs(c);
}

void buildUnits(rmd_ctx *c, std::vector<int4> &units, std::vector<double4> &geom, std::vector<int> &conStart,
                std::vector<int> &atomStart, std::vector<int> &atoms, std::vector<int4> &cons,
                std::vector<double> &conD) {
  const System &s = c->sys;
  const Topology &t = s.topology;
  const int n = t.atomCount;
  const int nc = (int)s.constraints.a.size();
  std::vector<int> parent(n);
  std::iota(parent.begin(), parent.end(), 0);
  std::function<int(int)> find = [&](int x) { return parent[x] == x ? x : parent[x] = find(parent[x]); };
  for (int k = 0; k < nc; ++k) {
    int a = find(s.constraints.a[k]), b = find(s.constraints.b[k]);
    if (a != b) parent[std::max(a, b)] = std::min(a, b);
  }
  std::map<int, std::vector<int>> consOf, atomsOf;
  for (int k = 0; k < nc; ++k) consOf[find(s.constraints.a[k])].push_back(k);
  std::vector<char> constrained(n, 0);
  for (int k = 0; k < nc; ++k) constrained[s.constraints.a[k]] = constrained[s.constraints.b[k]] = 1;
  for (int a = 0; a < n; ++a)
    if (constrained[a]) atomsOf[find(a)].push_back(a);
  conStart.assign(1, 0);
  atomStart.assign(1, 0);
  std::vector<char> emitted(n, 0);
  for (int a = 0; a < n; ++a) {
    if (!constrained[a]) {
      units.push_back(make_int4(0, a, 0, 0));
      geom.push_back(make_double4(0, 0, 0, 0));
      ++c->freeUnits;
      continue;
    }
    const int root = find(a);
    if (emitted[root]) continue;
    emitted[root] = 1;
    const std::vector<int> &ca = atomsOf[root];
    const std::vector<int> &ck = consOf[root];
    // A rigid water: three atoms, three constraints, heavy centre, equal
    // partner masses and equal centre-partner lengths (SETTLE's shape).
    bool water = false;
    if (ca.size() == 3 && ck.size() == 3) {
      int centre = ca[0];
      for (int x : ca)
        if (t.masses[x] > t.masses[centre]) centre = x;
      std::vector<int> partners;
      for (int x : ca)
        if (x != centre) partners.push_back(x);
      auto len = [&](int p, int q) {
        for (int k : ck)
          if ((s.constraints.a[k] == p && s.constraints.b[k] == q) || (s.constraints.a[k] == q && s.constraints.b[k] == p))
            return s.constraints.distance[k];
        return -1.0;
      };
      const double d1 = len(centre, partners[0]), d2 = len(centre, partners[1]), d3 = len(partners[0], partners[1]);
      if (d1 > 0 && d2 > 0 && d3 > 0 && d1 == d2 && t.masses[partners[0]] == t.masses[partners[1]] &&
          t.masses[partners[0]] > 0 && t.masses[centre] > 0) {
        water = true;
        units.push_back(make_int4(1, centre, partners[0], partners[1]));
        geom.push_back(make_double4(d1, d3, t.masses[centre], t.masses[partners[0]]));
        ++c->waterUnits;
      }
    }
    if (water) continue;
    const int cl = (int)conStart.size() - 1;
    const int base = (int)atoms.size();
    int centre = -1;
    if (ck.size() <= (size_t)STAR_MAX && ca.size() == ck.size() + 1)
      for (int x : ca) {
        bool inAll = true;
        for (int k : ck)
          if (s.constraints.a[k] != x && s.constraints.b[k] != x) inAll = false;
        if (inAll) {
          centre = x;
          break;
        }
      }
    if (centre >= 0) {
      atoms.push_back(centre);
      for (int k : ck) {
        atoms.push_back(s.constraints.a[k] == centre ? s.constraints.b[k] : s.constraints.a[k]);
        cons.push_back(make_int4((int)atoms.size() - 1, base, k, 0));
        conD.push_back(s.constraints.distance[k]);
      }
      conStart.push_back((int)cons.size());
      atomStart.push_back((int)atoms.size());
      units.push_back(make_int4(UNIT_STAR, cl, (int)ck.size(), 0));
      c->starMax = std::max(c->starMax, (int)ck.size());
      geom.push_back(make_double4(0, 0, 0, 0));
      ++c->clusterUnits;
      continue;
    }
    for (int x : ca) atoms.push_back(x);
    for (int k : ck) {
      int sa = base + (int)(std::find(ca.begin(), ca.end(), s.constraints.a[k]) - ca.begin());
      int sb = base + (int)(std::find(ca.begin(), ca.end(), s.constraints.b[k]) - ca.begin());
      cons.push_back(make_int4(sa, sb, k, 0));
      conD.push_back(s.constraints.distance[k]);
    }
    conStart.push_back((int)cons.size());
    atomStart.push_back((int)atoms.size());
    units.push_back(make_int4(2, cl, 0, 0));
    geom.push_back(make_double4(0, 0, 0, 0));
    ++c->clusterUnits;
  }
  // Wave-homogeneous layout: waters, free atoms, stars, then the clusters from
  // a block boundary, each with its slot offset in the block's LDS (clusters
  // too large for LDS keep global scratch). Padding units do nothing.
  std::vector<int4> waters, frees, stars, clusters;
  std::vector<double4> waterGeom;
  for (size_t u = 0; u < units.size(); ++u) {
    if (units[u].x == UNIT_WATER) {
      waters.push_back(units[u]);
      waterGeom.push_back(geom[u]);
    } else if (units[u].x == UNIT_FREE) {
      frees.push_back(units[u]);
    } else if (units[u].x == UNIT_STAR) {
      stars.push_back(units[u]);
    } else {
      clusters.push_back(units[u]);
    }
  }
  units.clear();
  geom.clear();
  const double4 none = make_double4(0, 0, 0, 0);
  auto pad = [&](size_t multiple) {
    while (units.size() % multiple) {
      units.push_back(make_int4(UNIT_NONE, 0, 0, 0));
      geom.push_back(none);
    }
  };
  for (size_t k = 0; k < waters.size(); ++k) {
    units.push_back(waters[k]);
    geom.push_back(waterGeom[k]);
  }
  pad(INTEGRATE_BLOCK);  // whole blocks of waters: their own kernel
  for (const int4 &u : frees) {
    units.push_back(u);
    geom.push_back(none);
  }
  pad(64);
  for (const int4 &u : stars) {
    units.push_back(u);
    geom.push_back(none);
  }
  c->clusterLdsSlots = 0;
  if (!clusters.empty()) pad(INTEGRATE_BLOCK);
  int used = 0;
  for (int4 u : clusters) {
    const int size = atomStart[u.y + 1] - atomStart[u.y];
    if (units.size() % INTEGRATE_BLOCK == 0) used = 0;
    if (size > CLUSTER_LDS_SLOTS) {
      u.z = 1;
      u.w = 0;
    } else {
      if (used + size > CLUSTER_LDS_SLOTS) {
        pad(INTEGRATE_BLOCK);
        used = 0;
      }
      u.z = 0;
      u.w = used;
      used += size;
      c->clusterLdsSlots = std::max(c->clusterLdsSlots, used);
    }
    units.push_back(u);
    geom.push_back(none);
  }
}

// Whether the step's integrator builds the cell list (IntegratorParams::
// cellsInIntegrator), from the current plan; the captured steps follow it.
void refreshCellBuilder(rmd_ctx *c) {
  const int builds = !c->batchMember && integratorBuildsCells(c->d, c->plan.cellCount) ? 1 : 0;
  // (the search's form, one wave per atom or per cell, decides whether the
  // joined force launch runs it: captured steps follow that too)
  const int byCell = searchByCell(c->searchMode, c->d.n, c->plan.cellCount) ? 1 : 0;
  if (builds != c->ip.cellsInIntegrator || byCell != c->searchByCellCaptured) {
    c->ip.cellsInIntegrator = builds;
    c->searchByCellCaptured = byCell;
    invalidateGraphs(c);
  }
}

void setupIntegrator(rmd_ctx *c, int suppliedNoise) {
  const double decay = std::exp(-c->friction * c->timestep);
  c->ip.halfStep = 0.5 * c->timestep;
  c->ip.decay = decay;
  c->ip.noiseScale = std::sqrt(std::max(0.0, 1.0 - decay * decay));
  c->ip.kT = BOLTZMANN * c->temperature;
  c->ip.tolerance = c->tolerance;
  c->ip.seed = c->seed;
  c->ip.maxIterations = c->maxIterations;
  c->ip.suppliedNoise = suppliedNoise;
  c->ip.clusterLdsBytes = (int)(c->clusterLdsSlots * CLUSTER_SLOT * sizeof(double));
  c->ip.starMax = c->starMax;
  c->ip.dispHalf = std::max(0.0, 0.5 * c->plan.skin - 1e-4);
  refreshCellBuilder(c);
}

// ---- force evaluation ---------------------------------------------------

// The start of an evaluation outside the step loop, in one launch: the
// rebuild decision (thread 0; this evaluation's list requirement starts
// here, unless `keepRequirement`: a trial appended to a block is checked with
// the block), the forces zeroed, the slots' positions refreshed from the
// atoms'.
__global__ void __launch_bounds__(256) beginEvaluationKernel(Dev d, double dispHalf, int force, int keepRequirement) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a == 0) {
    Scalars *sc = d.sc;
    if (!keepRequirement) sc->requiredMax = 0u;
    const float maxDisp2 = __uint_as_float(sc->maxDisp2Bits);
    const bool rebuild = force || sc->forceRebuild || displacementExceeded(sc, maxDisp2, dispHalf);
    sc->rebuild = rebuild ? 1u : 0u;
    if (rebuild) {
      sc->maxDisp2Bits = 0u;
      sc->slack = 0.f;
      retireRequirement(sc);
      sc->rebuilds += 1u;
      sc->forceRebuild = 0u;
    }
  }
  if (a < d.nPad) {
    d.force[a] = 0;
    d.force[d.nPad + a] = 0;
    d.force[2 * d.nPad + a] = 0;
  }
  if (a < d.n) d.slotPos[a] = d.pos[d.cellAtoms[a]];
}

__global__ void setForceRebuildKernel(Scalars *sc) { sc->forceRebuild = 1u; }
__global__ void setStepKernel(Scalars *sc, long long step) { sc->step = step; }

// One force evaluation. A small system runs its direct space, bonded terms
// and PME forward planes as one launch after the list rebuild, then the rest
// of PME, all on the context's stream: a captured step runs forked branches
// one after another, so separate launches cost the sum of their durations,
// where one launch overlaps them. Otherwise PME, the longest branch, stays on
// the context's stream, and the neighbour-list branch (rebuild + direct
// space) and the bonded terms, which read only the positions, run on side
// streams forked after the integrator and joined before anything that reads
// the forces. Forces are int64 atomics and every energy has its own slot row,
// so the results do not depend on the overlap.
// `molecular`: energies with the molecular scaling term; `cellsBuilt`: the
// step's integrator has built the cell list.
void enqueueForces(rmd_ctx *c, int terms, int energies, int slots[E_TERM_COUNT], bool molecular = true,
                   bool cellsBuilt = false) {
  Dev &d = c->d;
  for (int k = 0; k < E_TERM_COUNT; ++k) slots[k] = 0;
  const int bondedMask = RMD_TERM_BONDED | RMD_TERM_EXCEPTIONS | RMD_TERM_RESTRAINT;
  const int rational = ewaldRationalFits(c->pme.alpha, c->sys.cutoff) ? 1 : 0;
  auto setBondedSlots = [&](int b) {
    slots[E_BOND] = slots[E_ANGLE] = slots[E_TORSION] = slots[E_LJ_EXCEPTION] = slots[E_COUL_EXCEPTION] =
        slots[E_RESTRAINT] = slots[E_DSCALE_BONDED] = b;
  };
  const bool joined = c->concurrent && (terms & RMD_TERM_DIRECT) && (terms & RMD_TERM_RECIPROCAL) &&
                      (terms & bondedMask) && forcesJoin(d, c->p, c->lanes);
  if (joined) {
    // the one-wave-per-atom search runs in the joined launch
    const bool searchJoined = !searchByCell(c->searchMode, d.n, c->plan.cellCount);
    launchRebuild(d, c->stream, cellsBuilt, searchJoined);
    int pair = 0, bonded = 0;
    launchForcesJoined(d, c->p, (float)c->pme.alpha, rational, c->lanes, energies, c->pme.alpha, terms, c->stream,
                       &pair, &bonded, searchJoined ? 1 : 0);
    slots[E_LJ_DIRECT] = slots[E_COUL_DIRECT] = slots[E_DSCALE_DIRECT] = pair;
    setBondedSlots(bonded);
    slots[E_RECIPROCAL] = slots[E_DSCALE_RECIPROCAL] = launchPmeAfterForward(d, c->p, energies, c->stream);
  } else {
    const bool fork = c->concurrent;
    hipStream_t pmeStream = c->stream, listStream = fork ? c->side[0] : c->stream,
                bondedStream = fork ? c->side[1] : c->stream;
    if (fork) {
      HIP_CHECK(hipEventRecord(c->fork, c->stream));
      HIP_CHECK(hipStreamWaitEvent(c->side[0], c->fork, 0));
      HIP_CHECK(hipStreamWaitEvent(c->side[1], c->fork, 0));
    }
    if (terms & RMD_TERM_RECIPROCAL) {
      int b = 0;
      launchPme(d, c->p, energies, pmeStream, &b);
      slots[E_RECIPROCAL] = slots[E_DSCALE_RECIPROCAL] = b;
    }
    if (terms & bondedMask) {
      int b = 0;
      launchBonded(d, c->pme.alpha, terms, energies, bondedStream, &b);
      setBondedSlots(b);
    }
    launchRebuild(d, listStream, cellsBuilt);
    if (terms & RMD_TERM_DIRECT) {
      const int b = launchNonbonded(d, (float)c->pme.alpha, rational, c->lanes, energies, listStream);
      slots[E_LJ_DIRECT] = slots[E_COUL_DIRECT] = slots[E_DSCALE_DIRECT] = b;
    }
    if (fork) {
      for (int k = 0; k < 2; ++k) {
        HIP_CHECK(hipEventRecord(c->join[k], c->side[k]));
        HIP_CHECK(hipStreamWaitEvent(c->stream, c->join[k], 0));
      }
    }
  }
  if (energies && molecular)
    slots[E_DSCALE_MOLECULAR] =
        launchMolecularTerm(d, c->molStart, c->molAtoms, c->sys.topology.moleculeCount, c->stream);
}

[[noreturn]] void cellOverflowError() {
  throw Error(RMD_E_INVALID, "a cell held more atoms than the cell sort supports (4096); the box is too dense "
                             "or too small for this cutoff");
}

void checkScalars(rmd_ctx *c, Scalars &h) {
  HIP_CHECK(hipMemcpyAsync(&h, c->d.sc, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  if (h.cellOverflow) cellOverflowError();
}

// One evaluation at the current positions. force_rebuild mirrors the
// reference's #evaluatePotential, which always rebuilds.
void evaluate(rmd_ctx *c, int terms, bool forceRebuild, double terms_out[E_TERM_COUNT]) {
  for (int attempt = 0; attempt < 4; ++attempt) {
    beginEvaluationKernel<<<(c->d.nPad + 255) / 256, 256, 0, c->stream>>>(c->d, c->ip.dispHalf,
                                                                          forceRebuild ? 1 : 0, 0);
    int slots[E_TERM_COUNT];
    enqueueForces(c, terms, terms_out ? 1 : 0, slots);
    if (terms_out) {
      // reduced and copied before the one wait below (discarded if the list
      // overflowed and the evaluation runs again)
      SlotCounts counts;
      std::memcpy(counts.c, slots, sizeof slots);
      launchReduceEnergies(c->d, counts, c->dTerms, c->stream);
      HIP_CHECK(hipMemcpyAsync(terms_out, c->dTerms, sizeof(double) * E_TERM_COUNT, hipMemcpyDeviceToHost,
                               c->stream));
    }
    HIP_CHECK(hipGetLastError());
    Scalars h;
    checkScalars(c, h);
    c->rebuildCount = h.rebuilds;
    const unsigned int need = requirementSince(h);
    c->worstRequired = std::max(c->worstRequired, need);
    if (need > (unsigned int)c->d.cap) {
      const int grown = std::max(c->d.cap + 32, (int)std::ceil(need * 1.25));
      allocateNeighbours(c, grown);
      ++c->neighbourGrowths;
      forceRebuild = true;
      continue;
    }
    if (terms_out) std::memcpy(c->slotCounts, slots, sizeof slots);
    return;
  }
  throw Error(RMD_E_NEIGHBOUR_OVERFLOW, "the neighbour list kept overflowing after growth");
}

double dispersionNow(rmd_ctx *c) {
  return c->dispersionOn ? dispersionEnergy(c->disp, volumeOf(c->box), c->sys.cutoff) : 0.0;
}

double potentialFromTerms(rmd_ctx *c, const double *t) {
  return t[E_BOND] + t[E_ANGLE] + t[E_TORSION] + t[E_LJ_DIRECT] + t[E_LJ_EXCEPTION] + t[E_COUL_DIRECT] +
         t[E_COUL_EXCEPTION] + t[E_RECIPROCAL] + c->selfEnergy + c->netEnergy + dispersionNow(c) + t[E_RESTRAINT];
}

// ---- the barostat's learned screen ------------------------------------------

// d(energy)/d(ln lambda) when molecules are scaled about their centroids with
// their geometry kept: the atomic scaling derivatives of every term plus
// sum_i F_i . (r_i - c_mol(i)).
double molecularDerivative(const double *t) {
  return t[E_DSCALE_DIRECT] + t[E_DSCALE_RECIPROCAL] + t[E_DSCALE_BONDED] + t[E_DSCALE_MOLECULAR];
}

double dispersionAt(rmd_ctx *c, double volume) {
  return c->dispersionOn ? dispersionEnergy(c->disp, volume, c->sys.cutoff) : 0.0;
}

// The 16 raw features of a move from a state (its evaluated terms) at
// `volume` to `newVolume`, and the first-order estimate of the energy change
// from the scaling derivatives, dU/dln(lambda) * ln(lambda) with
// ln(lambda) = ln(V'/V) / 3.
void screenFeatures(rmd_ctx *c, const double *t, double volume, double newVolume, float *f, double *estimate) {
  const double kT = BOLTZMANN * c->temperature;
  const double nkT = std::max(1, c->sys.topology.moleculeCount) * kT;
  const double L = std::log(newVolume / volume);
  const double dD = t[E_DSCALE_DIRECT], dR = t[E_DSCALE_RECIPROCAL];
  // first order in ln(lambda) under the barostat's molecular scaling
  const double est = molecularDerivative(t) * L / 3.0;
  *estimate = est;
  double g[RMD_SURROGATE_INPUTS];
  g[0] = 100.0 * L;
  g[1] = 1e4 * L * L;
  g[2] = est / kT;
  g[3] = dD / nkT;
  g[4] = dR / nkT;
  g[5] = volume / c->screen.refVolume - 1.0;
  g[6] = t[E_LJ_DIRECT] / nkT;
  g[7] = t[E_COUL_DIRECT] / nkT;
  g[8] = t[E_RECIPROCAL] / nkT;
  g[9] = g[0] * g[3];
  g[10] = g[0] * g[4];
  g[11] = g[1] * g[3];
  g[12] = g[1] * g[4];
  g[13] = g[0] * g[6];
  g[14] = g[0] * g[7];
  g[15] = g[0] * g[8];
  for (int k = 0; k < RMD_SURROGATE_INPUTS; ++k) f[k] = (float)g[k];
}

// The network's residuals (in the targets' units) for n raw feature rows.
void predictScreen(ScreenState &sc, const float *raw, int n, double *out) {
  const int F = RMD_SURROGATE_INPUTS;
  std::vector<float> x((size_t)n * F), y(n);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < F; ++k) x[(size_t)i * F + k] = (raw[(size_t)i * F + k] - sc.mean[k]) * sc.scale[k];
  const int st = rmd_surrogate_forward(sc.net, x.data(), n, y.data());
  if (st != RMD_OK) throw Error(st, std::string("screen inference failed: ") + rmd_last_global_error());
  for (int i = 0; i < n; ++i) out[i] = sc.tMean + sc.tScale * (double)y[i];
}

double predictResidual(rmd_ctx *c, const float *raw) {
  double r = 0;
  predictScreen(c->screen, raw, 1, &r);
  return r;
}

// Normalise the kept samples, train the network on the GPU and start screening.
void trainScreen(ScreenState &sc, int device) {
  const size_t n = sc.t.size();
  const int F = RMD_SURROGATE_INPUTS;
  for (int k = 0; k < F; ++k) {
    double m = 0, v = 0;
    for (size_t i = 0; i < n; ++i) m += sc.x[i * F + k];
    m /= n;
    for (size_t i = 0; i < n; ++i) v += (sc.x[i * F + k] - m) * (sc.x[i * F + k] - m);
    const double sd = std::sqrt(v / n);
    sc.mean[k] = (float)m;
    sc.scale[k] = (float)(sd > 1e-12 ? 1.0 / sd : 0.0);
  }
  double tm = 0, tv = 0;
  for (size_t i = 0; i < n; ++i) tm += sc.t[i];
  tm /= n;
  for (size_t i = 0; i < n; ++i) tv += (sc.t[i] - tm) * (sc.t[i] - tm);
  sc.tMean = tm;
  sc.tScale = std::max(std::sqrt(tv / n), 1e-9);
  std::vector<float> x(n * F), t(n);
  for (size_t i = 0; i < n; ++i) {
    for (int k = 0; k < F; ++k) x[i * F + k] = (sc.x[i * F + k] - sc.mean[k]) * sc.scale[k];
    t[i] = (float)((sc.t[i] - sc.tMean) / sc.tScale);
  }
  float loss = 0.f;
  const int st = rmd_surrogate_train(sc.net, x.data(), t.data(), (int)n, sc.options.training_steps,
                                     sc.options.learning_rate, &loss);
  if (st != RMD_OK) throw Error(st, std::string("screen training failed: ") + rmd_last_global_error());
  HIP_CHECK(hipSetDevice(device));
  sc.trainingLoss = loss;
  sc.phase = 1;
}

void screenTrain(rmd_ctx *c) { trainScreen(c->screen, c->device); }

double evaluatePotential(rmd_ctx *c) {
  double t[E_TERM_COUNT];
  evaluate(c, RMD_TERM_ALL, true, t);
  return potentialFromTerms(c, t);
}

// ---- the block loop ---------------------------------------------------------

// Whether the energies a block ends with (and its appended trial's) include
// the molecular scaling term: only the barostat's screen reads it from them
// (its features; a batch's screen may read a member's), while rmd_energy and
// rmd_evaluate evaluate afresh.
bool blockEndMolecular(const rmd_ctx *c) { return c->screen.options.enabled || c->batchMember; }

// A volume move's trial after a block, as the barostat's own trial runs it
// (rmd_attempt_volume_change: the state saved, molecules scaled about their
// centroids, the box and the influence function, one evaluation with
// energies), from the proposal in trialDev; its energies go to dTermsTrial
// and its list requirement is checked with the block's.
void enqueueTrial(rmd_ctx *c) {
  const SnapshotBuffers from{c->trialPos, nullptr, c->rebuildSaved, c->trialForce, nullptr,
                             &c->trialDev->rebuildsBefore};
  launchSnapshot(c->d, from, SNAP_REBUILD, c->stream);
  launchScaleMoleculesFrom(c->d, c->molStart, c->molAtoms, c->sys.topology.moleculeCount, c->centroids,
                           c->molOfAtom, &c->trialDev->scale, c->stream);
  const double *mods[3] = {c->dModuli[0], c->dModuli[1], c->dModuli[2]};
  pmeSetBoxAndInfluenceFrom(c->p, &c->trialDev->box, c->boxDev, mods, c->pme.alpha, c->stream);
  beginEvaluationKernel<<<(c->d.nPad + 255) / 256, 256, 0, c->stream>>>(c->d, c->ip.dispHalf, 0, 1);
  int slots[E_TERM_COUNT];
  enqueueForces(c, RMD_TERM_ALL, 1, slots, blockEndMolecular(c));
  SlotCounts counts;
  std::memcpy(counts.c, slots, sizeof slots);
  launchReduceEnergies(c->d, counts, c->dTermsTrial, c->stream);
}

void enqueueBlock(rmd_ctx *c, int steps, bool trial) {
  const bool molecular = blockEndMolecular(c);
  SlotCounts counts{};
  for (int s = 0; s < steps; ++s) {
    const bool last = s == steps - 1;
    const int openBlocks = launchIntegrateOpen(c->d, c->ip, s > 0 ? 1 : 0, last ? 1 : 0, c->stream);
    int slots[E_TERM_COUNT];
    enqueueForces(c, RMD_TERM_ALL, last ? 1 : 0, slots, molecular, c->ip.cellsInIntegrator != 0);
    if (last) {
      std::memcpy(counts.c, slots, sizeof slots);
      counts.c[E_KINETIC] = 0;
      counts.c[E_KINETIC_EARLY] = openBlocks;
    }
  }
  // The last step's energies, reduced in the same launch sequence.
  launchReduceEnergies(c->d, counts, c->dTermsBlock, c->stream);
  launchIntegrateClose(c->d, c->ip, 0, c->stream);
  if (trial) enqueueTrial(c);
  static_assert(E_TERM_COUNT <= 64, "one thread per energy term");
  gatherReadBackKernel<<<1, 64, 0, c->stream>>>(c->d.sc, c->dTermsBlock, trial ? c->dTermsTrial : nullptr,
                                                trial ? &c->trialDev->rebuildsBefore : nullptr, c->readBack);
}

// Block lengths up to this many steps key the captured graphs; a block with
// a trial appended is another graph of the same length.
constexpr int GRAPH_TRIAL_KEY = 1 << 24;

void launchBlock(rmd_ctx *c, int steps, bool trial) {
  if (!c->useGraphs) {
    enqueueBlock(c, steps, trial);
    HIP_CHECK(hipGetLastError());
    return;
  }
  const int key = steps + (trial ? GRAPH_TRIAL_KEY : 0);
  auto it = c->graphs.find(key);
  if (it == c->graphs.end()) {
    hipGraph_t graph;
    HIP_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    enqueueBlock(c, steps, trial);
    HIP_CHECK(hipStreamEndCapture(c->stream, &graph));
    hipGraphExec_t exec;
    HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    HIP_CHECK(hipGraphDestroy(graph));
    it = c->graphs.emplace(key, exec).first;
  }
  HIP_CHECK(hipGraphLaunch(it->second, c->stream));
}

SnapshotBuffers savedBuffers(const rmd_ctx *c) {
  return SnapshotBuffers{c->posSaved, c->velSaved, c->rebuildSaved, c->forceSaved, c->scSaved};
}

// The state a block starts from (positions, velocities, forces, Scalars, with
// the block's list requirement cleared first), to run the block again if the
// list overflows in it; one launch each way.
void snapshot(rmd_ctx *c) {
  launchSnapshot(c->d, savedBuffers(c), SNAP_VELOCITIES | SNAP_SCALARS, c->stream);
}

void restore(rmd_ctx *c) {
  launchSnapshot(c->d, savedBuffers(c), SNAP_RESTORE | SNAP_VELOCITIES | SNAP_SCALARS, c->stream);
}

// With `trial`, the proposal staged in trialHost goes to the device first and
// the block runs with the trial appended: trialTerms and rebuildsBefore (the
// list's rebuild count when the trial started) come back with the block's
// scalars.
void runBlock(rmd_ctx *c, int steps, bool trial = false, double *trialTerms = nullptr,
              unsigned int *rebuildsBefore = nullptr) {
  if (trial)
    HIP_CHECK(hipMemcpyAsync(c->trialDev, c->trialHost, sizeof(TrialParams), hipMemcpyHostToDevice, c->stream));
  for (int attempt = 0; attempt < 4; ++attempt) {
    snapshot(c);
    launchBlock(c, steps, trial);
    ReadBack back;
    HIP_CHECK(hipMemcpyAsync(&back, c->readBack, sizeof back, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    const Scalars &h = back.sc;
    if (h.cellOverflow) cellOverflowError();
    std::memcpy(c->blockEndTerms, back.block, sizeof c->blockEndTerms);
    if (trial) {
      std::memcpy(trialTerms, back.trial, sizeof(double) * E_TERM_COUNT);
      *rebuildsBefore = back.rebuildsBefore;
    }
    if (h.nonFinite || std::isnan(__builtin_bit_cast(float, h.maxDisp2Bits)))
      throw Error(RMD_E_NONFINITE, "a coordinate became non-finite in the block of " + std::to_string(steps) +
                                       " steps that ended at step " + std::to_string(c->step + steps) +
                                       "; everything written during it should be discarded too.");
    if (h.shakeFailed) throw Error(RMD_E_CONSTRAINT, "constraints did not converge in the iteration limit");
    c->rebuildCount = h.rebuilds;
    const unsigned int need = requirementSince(h);
    c->worstRequired = std::max(c->worstRequired, need);
    if (need > (unsigned int)c->d.cap) {
      // The list overflowed somewhere in the block: grow it and run the
      // block again from its start, so no step ever used a truncated list.
      restore(c);
      if (trial) writeBox(c);  // the trial's box back to the block's
      const int grown = std::max(c->d.cap + 32, (int)std::ceil(need * 1.25));
      allocateNeighbours(c, grown);
      ++c->neighbourGrowths;
      setForceRebuildKernel<<<1, 1, 0, c->stream>>>(c->d.sc);
      continue;
    }
    c->step += steps;
    c->stepsRun += steps;
    c->kineticSampledAtStep = c->step;
    c->blockEndStep = c->step;
    return;
  }
  throw Error(RMD_E_NEIGHBOUR_OVERFLOW, "the neighbour list kept overflowing after growth");
}

// #updateBox (webgpu/simulation.js:681-731): adopt a new box, re-plan the
// cells and the capacity, update the self energy and the PME influence
// function; the PME grid and alpha stay fixed.
// The host half of a box change: the box, the cell plan and capacities, the
// self energy. The device box and the influence function follow separately.
bool updateBoxHost(rmd_ctx *c, const double *nb) {
  c->blockEndStep = -1;
  for (int k = 0; k < 9; ++k) c->box[k] = nb[k];
  CellPlan estimate = planCellList(c->sys, c->plan.skin, c->slack, 0, c->box, c->reachMax);
  const int capacity = std::max(c->d.cap, estimate.maxNeighbours);
  CellPlan plan = planCellList(c->sys, c->plan.skin, c->slack, capacity, c->box, c->reachMax);
  c->plan = plan;
  allocateCells(c, plan.cellCount);
  const bool replaced = capacity > c->d.cap;
  if (replaced) {
    allocateNeighbours(c, capacity);
    ++c->neighbourGrowths;
  }
  ewaldSelfEnergy(c->sys, c->pme.alpha, volumeOf(c->box), &c->selfEnergy, &c->netEnergy);
  refreshCellBuilder(c);
  return replaced;
}

void updateBox(rmd_ctx *c, const double *nb) {
  updateBoxHost(c, nb);
  writeBox(c);
  setForceRebuildKernel<<<1, 1, 0, c->stream>>>(c->d.sc);
}

double scaleVolume(rmd_ctx *c, double s, double *terms = nullptr) {
  launchScaleMolecules(c->d, c->molStart, c->molAtoms, c->sys.topology.moleculeCount, c->centroids, c->molOfAtom,
                       (float)s, c->stream);
  double nb[9];
  for (int k = 0; k < 9; ++k) nb[k] = c->box[k] * s;
  updateBox(c, nb);
  double t[E_TERM_COUNT];
  evaluate(c, RMD_TERM_ALL, true, t);
  if (terms) std::memcpy(terms, t, sizeof t);
  return potentialFromTerms(c, t);
}

// The barostat's trial state: molecules scaled about their centroids (the
// rebuild references with them, the move's share of the skin spent), the new
// box, and an evaluation that rebuilds only if the list is due or was replaced.
double moveVolume(rmd_ctx *c, double s, double *terms) {
  launchScaleMolecules(c->d, c->molStart, c->molAtoms, c->sys.topology.moleculeCount, c->centroids, c->molOfAtom,
                       (float)s, c->stream);
  double nb[9];
  for (int k = 0; k < 9; ++k) nb[k] = c->box[k] * s;
  const bool replaced = updateBoxHost(c, nb);
  writeBox(c);
  double t[E_TERM_COUNT];
  evaluate(c, RMD_TERM_ALL, replaced, t);
  if (terms) std::memcpy(terms, t, sizeof t);
  return potentialFromTerms(c, t);
}

void toDouble(rmd_ctx *c, double *forces) {
  const int n = c->sys.topology.atomCount;
  std::vector<long long> f(3 * (size_t)c->d.nPad);
  HIP_CHECK(hipMemcpyAsync(f.data(), c->d.force, f.size() * sizeof(long long), hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  for (int a = 0; a < n; ++a)
    for (int k = 0; k < 3; ++k) forces[3 * a + k] = (double)f[(size_t)k * c->d.nPad + a] * FORCE_UNSCALE;
}

void fillEnergy(rmd_ctx *c, const double *t, rmd_energy_terms *e, bool withKinetic, double kinetic, double early,
                bool earlyValid) {
  std::memset(e, 0, sizeof *e);
  e->step = c->step;
  e->time = c->step * c->timestep;
  e->harmonic_bond = t[E_BOND];
  e->harmonic_angle = t[E_ANGLE];
  e->periodic_torsion = t[E_TORSION];
  e->bonded = t[E_BOND] + t[E_ANGLE] + t[E_TORSION];
  e->lennard_jones = t[E_LJ_DIRECT] + t[E_LJ_EXCEPTION];
  e->dispersion = dispersionNow(c);
  e->coulomb_direct = t[E_COUL_DIRECT] + t[E_COUL_EXCEPTION];
  e->coulomb_reciprocal = t[E_RECIPROCAL];
  e->coulomb_self = c->selfEnergy + c->netEnergy;
  e->coulomb = e->coulomb_direct + e->coulomb_reciprocal + e->coulomb_self;
  e->position_restraint = t[E_RESTRAINT];
  e->restrained_atoms = (int32_t)c->restraintAtoms.size();
  e->potential = e->bonded + e->lennard_jones + e->dispersion + e->coulomb + e->position_restraint;
  e->dscale_direct = t[E_DSCALE_DIRECT];
  e->dscale_reciprocal = t[E_DSCALE_RECIPROCAL];
  e->dscale_molecular = molecularDerivative(t);
  if (withKinetic) {
    e->kinetic = kinetic;
    e->total = e->potential + kinetic;
    e->kinetic_for_temperature = earlyValid && (early > 0 || kinetic <= 0) ? early : kinetic;
    e->kinetic_sampled_in_step = earlyValid && (early > 0 || kinetic <= 0) ? 1 : 0;
    e->temperature = 2 * e->kinetic_for_temperature / (c->sys.degreesOfFreedom * BOLTZMANN);
  } else {
    e->total = e->potential;
  }
  e->volume = volumeOf(c->box);
  for (int k = 0; k < 9; ++k) e->box[k] = c->box[k];
}

void create(rmd_ctx *c, const System &sys, const rmd_options &o, const Streams *shared = nullptr) {
  c->sys = sys;
  c->opt = o;
  const Topology &t = sys.topology;
  if (!sys.isPeriodic) throw Error(RMD_E_INVALID, "this engine runs periodic PME systems; the system has no box");
  c->timestep = o.timestep > 0 ? o.timestep : 0.002;
  c->temperature = o.temperature > 0 ? o.temperature : 298;
  c->friction = o.friction >= 0 ? o.friction : 1.0;
  c->seed = o.seed;
  c->tolerance = o.tolerance > 0 ? o.tolerance : 1e-5;
  c->maxIterations = o.max_constraint_iterations > 0 ? o.max_constraint_iterations : 150;
  c->hasPressure = o.has_pressure != 0;
  c->pressure = o.pressure;
  c->barostatInterval = o.barostat_interval > 0 ? o.barostat_interval : 25;
  c->dispersionOn = o.dispersion_correction != 0;
  c->slack = o.neighbour_slack > 0 ? o.neighbour_slack : 2.0;
  // Shapes the caller leaves at 0 (and a negative skin) come from the
  // autotuner's table for this kind of workload.
  const TunedShape &tuned = tunedShape(shared ? 1 : 0, t.atomCount);
  c->lanes = o.lanes_per_atom > 0 ? o.lanes_per_atom : tuned.lanes;
  if (c->lanes & (c->lanes - 1) || c->lanes > 64) throw Error(RMD_E_INVALID, "lanes_per_atom must be a power of two <= 64");
  for (int k = 0; k < 9; ++k) c->box[k] = sys.box[k];
  if (o.skin > 0)
    c->skin = o.skin;
  else if (o.skin < 0 && tuned.skin > 0)
    c->skin = tuned.skin;
  else
    c->skin = skinFor(sys, c->timestep, c->temperature, nullptr);
  if (o.cell_reach < 0 || o.cell_reach > CELL_REACH_MAX)
    throw Error(RMD_E_INVALID, "cell_reach must be 0 (the default, 3) or 1..3");
  if (o.search_mode < 0 || o.search_mode > 3)
    throw Error(RMD_E_INVALID, "search_mode must be 0 (tuned), 1 (per atom), 2 (per cell) or 3 (by occupancy)");
  c->reachMax = o.cell_reach > 0 ? o.cell_reach : tuned.reach;
  // searchMode 0 is the rule by the plan's cell occupancy (option 3)
  c->searchMode = o.search_mode == 0 ? tuned.searchMode : o.search_mode == 3 ? 0 : o.search_mode;
  c->plan = planCellList(sys, c->skin, c->slack, o.max_neighbours, c->box, c->reachMax);
  {
    rmd_pme_parameters pp;
    const int32_t forced[3] = {o.grid[0], o.grid[1], o.grid[2]};
    if (rmd_pme_parameters_derive(c->box, sys.cutoff, o.ewald_tolerance > 0 ? o.ewald_tolerance : 0.0005,
                                  o.alpha > 0 ? o.alpha : 0, forced, &pp) != RMD_OK)
      throw Error(RMD_E_INVALID, rmd_last_global_error());
    c->pme = pp;
  }
  c->disp = dispersionCoefficients(sys);
  c->device = o.device;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
    throw Error(RMD_E_DEVICE, "no HIP device is available; this library runs on an MI355X (gfx950)");
  HIP_CHECK(hipSetDevice(c->device));
  if (shared) {
    c->ownsStreams = false;
    c->stream = shared->stream;
    c->fork = shared->fork;
    for (int k = 0; k < 2; ++k) {
      c->side[k] = shared->side[k];
      c->join[k] = shared->join[k];
    }
  } else {
    HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (int k = 0; k < 2; ++k) {
      HIP_CHECK(hipStreamCreateWithFlags(&c->side[k], hipStreamNonBlocking));
      HIP_CHECK(hipEventCreateWithFlags(&c->join[k], hipEventDisableTiming));
    }
    HIP_CHECK(hipEventCreateWithFlags(&c->fork, hipEventDisableTiming));
  }
  {
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) == hipSuccess && lds > 0)
      c->maxLds = (size_t)lds;
  }

  const int n = t.atomCount;
  Dev &d = c->d;
  d.n = n;
  d.nPad = (n + 63) / 64 * 64;
  d.pos = devAlloc<float4>(d.nPad, c->held);
  d.vel = devAlloc<float4>(d.nPad, c->held);
  d.force = devAlloc<long long>(3 * (size_t)d.nPad, c->held);
  d.posRebuild = devAlloc<float4>(d.nPad, c->held);
  c->posSaved = devAlloc<float4>(d.nPad, c->held);
  c->rebuildSaved = devAlloc<float4>(d.nPad, c->held);
  c->velSaved = devAlloc<float4>(d.nPad, c->held);
  c->forceSaved = devAlloc<long long>(3 * (size_t)d.nPad, c->held);
  c->boxDev = devAlloc<BoxF>(1, c->held);
  d.box = c->boxDev;
  d.type = devAlloc<int>(d.nPad, c->held);
  d.types = t.typeCount;
  {
    std::vector<float2> lj(t.ljA.size());
    for (size_t k = 0; k < lj.size(); ++k) lj[k] = make_float2((float)t.ljA[k], (float)t.ljB[k]);
    d.lj = devAlloc<float2>(lj.size(), c->held);
    upload(d.lj, lj, c->stream);
    std::vector<double2> lj64(t.ljA.size());
    for (size_t k = 0; k < lj64.size(); ++k) lj64[k] = make_double2(t.ljA[k], t.ljB[k]);
    d.lj64 = devAlloc<double2>(lj64.size(), c->held);
    upload(d.lj64, lj64, c->stream);
    HIP_CHECK(hipStreamSynchronize(c->stream));
    upload(d.type, t.atomTypes, c->stream);
  }
  d.cellOfAtom = devAlloc<int>(d.nPad, c->held);
  d.cellAtoms = devAlloc<int>(d.nPad, c->held);
  d.cellPos = devAlloc<float4>(d.nPad, c->held);
  d.slotPos = devAlloc<float4>(d.nPad, c->held);
  d.slotType = devAlloc<int>(d.nPad, c->held);
  d.slotOf = devAlloc<int>(d.nPad, c->held);
  d.cellCapacity = 0;
  allocateCells(c, c->plan.cellCount);
  d.nbrCount = devAlloc<int>(d.nPad, c->held);
  allocateNeighbours(c, c->plan.maxNeighbours);
  d.exclStart = devAlloc<int>(sys.exclusionStart.size(), c->held);
  d.exclPartners = devAlloc<int>(sys.exclusionPartners.size(), c->held);
  upload(d.exclStart, sys.exclusionStart, c->stream);
  upload(d.exclPartners, sys.exclusionPartners, c->stream);

  // Bonded terms.
  d.nBond = (int)sys.harmonicBonds.a.size();
  d.nAngle = (int)t.angles.a.size();
  d.nTorsion = (int)t.dihedrals.a.size();
  d.nException = (int)sys.exceptions.a.size();
  {
    std::vector<int2> ba(d.nBond);
    std::vector<double2> bp(d.nBond);
    for (int k = 0; k < d.nBond; ++k) {
      ba[k] = make_int2(sys.harmonicBonds.a[k], sys.harmonicBonds.b[k]);
      bp[k] = make_double2(sys.harmonicBonds.k[k], sys.harmonicBonds.r0[k]);
    }
    d.bondAtoms = devAlloc<int2>(ba.size(), c->held);
    d.bondParams = devAlloc<double2>(bp.size(), c->held);
    upload(d.bondAtoms, ba, c->stream);
    upload(d.bondParams, bp, c->stream);
    std::vector<int4> aa(d.nAngle);
    std::vector<double2> ap(d.nAngle);
    for (int k = 0; k < d.nAngle; ++k) {
      aa[k] = make_int4(t.angles.a[k], t.angles.b[k], t.angles.c[k], 0);
      ap[k] = make_double2(t.angles.k[k], t.angles.theta0[k]);
    }
    d.angleAtoms = devAlloc<int4>(aa.size(), c->held);
    d.angleParams = devAlloc<double2>(ap.size(), c->held);
    upload(d.angleAtoms, aa, c->stream);
    upload(d.angleParams, ap, c->stream);
    std::vector<int4> ta(d.nTorsion);
    std::vector<double4> tp(d.nTorsion);
    for (int k = 0; k < d.nTorsion; ++k) {
      ta[k] = make_int4(t.dihedrals.a[k], t.dihedrals.b[k], t.dihedrals.c[k], t.dihedrals.d[k]);
      tp[k] = make_double4(t.dihedrals.k[k], t.dihedrals.periodicity[k], t.dihedrals.phase[k], 0);
    }
    d.torsionAtoms = devAlloc<int4>(ta.size(), c->held);
    d.torsionParams = devAlloc<double4>(tp.size(), c->held);
    upload(d.torsionAtoms, ta, c->stream);
    upload(d.torsionParams, tp, c->stream);
    // The exceptions of each three-atom molecule whose three pairs are all
    // exceptions (a rigid water) come first, three to a molecule: the
    // bonded kernel takes them together. The rest keep their order.
    std::vector<int> order;
    order.reserve(d.nException);
    {
      std::unordered_map<long long, int> pairIndex;
      const long long n = t.atomCount;
      auto key = [n](int i, int j) { return (long long)std::min(i, j) * n + std::max(i, j); };
      for (int k = 0; k < d.nException; ++k) pairIndex.emplace(key(sys.exceptions.a[k], sys.exceptions.b[k]), k);
      std::vector<char> taken(d.nException, 0);
      d.nExcTriplets = 0;
      for (size_t m = 0; m + 1 < sys.moleculeStart.size(); ++m) {
        if (sys.moleculeStart[m + 1] - sys.moleculeStart[m] != 3) continue;
        const int *at = &sys.moleculeAtoms[sys.moleculeStart[m]];
        const long long keys[3] = {key(at[0], at[1]), key(at[0], at[2]), key(at[1], at[2])};
        int found[3], count = 0;
        for (int q = 0; q < 3; ++q) {
          const auto it = pairIndex.find(keys[q]);
          if (it != pairIndex.end() && !taken[it->second]) found[count++] = it->second;
        }
        if (count != 3) continue;
        for (int q = 0; q < 3; ++q) {
          taken[found[q]] = 1;
          order.push_back(found[q]);
        }
        ++d.nExcTriplets;
      }
      for (int k = 0; k < d.nException; ++k)
        if (!taken[k]) order.push_back(k);
    }
    std::vector<int2> ea(d.nException);
    std::vector<double4> ep(d.nException);
    for (int q = 0; q < d.nException; ++q) {
      const int k = order[q];
      const int i = sys.exceptions.a[k], j = sys.exceptions.b[k];
      ea[q] = make_int2(i, j);
      ep[q] = make_double4(ONE_4PI_EPS0 * sys.exceptions.chargeProduct[k], sys.exceptions.ljA[k],
                           sys.exceptions.ljB[k], ONE_4PI_EPS0 * t.charges[i] * t.charges[j]);
    }
    d.excAtoms = devAlloc<int2>(ea.size(), c->held);
    d.excParams = devAlloc<double4>(ep.size(), c->held);
    upload(d.excAtoms, ea, c->stream);
    upload(d.excParams, ep, c->stream);
  }
  // Restraints (k*d^2, reference positions in FP64).
  c->restraintK = o.restraint_force > 0 ? o.restraint_force : 0;
  if (c->restraintK > 0 && o.restraint_count > 0) {
    if (!o.restraint_atoms || !o.restraint_reference)
      throw Error(RMD_E_INVALID, "restraint atoms and reference are required when restraint_force > 0");
    std::vector<double4> ref;
    for (int k = 0; k < o.restraint_count; ++k) {
      const int a = o.restraint_atoms[k];
      if (a < 0 || a >= n) throw Error(RMD_E_INVALID, "restraint atom out of range");
      c->restraintAtoms.push_back(a);
      ref.push_back(make_double4(o.restraint_reference[3 * a], o.restraint_reference[3 * a + 1],
                                 o.restraint_reference[3 * a + 2], c->restraintK));
    }
    d.nRestraint = (int)c->restraintAtoms.size();
    d.restraintAtoms = devAlloc<int>(d.nRestraint, c->held);
    d.restraintRef = devAlloc<double4>(d.nRestraint, c->held);
    upload(d.restraintAtoms, c->restraintAtoms, c->stream);
    upload(d.restraintRef, ref, c->stream);
  }
  // Integrator units.
  {
    std::vector<int4> units, cons;
    std::vector<double4> geom;
    std::vector<int> conStart, atomStart, atoms;
    std::vector<double> conD;
    buildUnits(c, units, geom, conStart, atomStart, atoms, cons, conD);
    d.nUnits = (int)units.size();
    c->nUnits = d.nUnits;
    int waterUnits = 0;
    for (const int4 &u : units) waterUnits += u.x == UNIT_WATER ? 1 : 0;
    d.waterBlocks = (waterUnits + INTEGRATE_BLOCK - 1) / INTEGRATE_BLOCK;
    d.units = devAlloc<int4>(units.size(), c->held);
    d.waterGeom = devAlloc<double4>(geom.size(), c->held);
    d.clusterConStart = devAlloc<int>(conStart.size(), c->held);
    d.clusterAtomStart = devAlloc<int>(atomStart.size(), c->held);
    d.clusterAtoms = devAlloc<int>(atoms.size(), c->held);
    d.clusterCons = devAlloc<int4>(cons.size(), c->held);
    d.clusterConD = devAlloc<double>(conD.size(), c->held);
    d.clusterScratch = devAlloc<double>(CLUSTER_SLOT * std::max<size_t>(1, atoms.size()), c->held);
    upload(d.units, units, c->stream);
    upload(d.waterGeom, geom, c->stream);
    upload(d.clusterConStart, conStart, c->stream);
    upload(d.clusterAtomStart, atomStart, c->stream);
    upload(d.clusterAtoms, atoms, c->stream);
    upload(d.clusterCons, cons, c->stream);
    upload(d.clusterConD, conD, c->stream);
  }
  d.noise = devAlloc<float>(3 * (size_t)d.nPad, c->held);
  d.slots = devAlloc<double>((size_t)E_TERM_COUNT * MAX_SLOTS, c->held);
  d.sc = devAlloc<Scalars>(1, c->held);
  c->scSaved = devAlloc<Scalars>(1, c->held);
  c->dTerms = devAlloc<double>(E_TERM_COUNT, c->held);
  c->dTermsBlock = devAlloc<double>(E_TERM_COUNT, c->held);
  c->readBack = devAlloc<ReadBack>(1, c->held);
  c->dMaxForce = devAlloc<unsigned int>(1, c->held);
  // Molecules for the barostat.
  c->molStart = devAlloc<int>(sys.moleculeStart.size(), c->held);
  c->molAtoms = devAlloc<int>(sys.moleculeAtoms.size(), c->held);
  c->molOfAtom = devAlloc<int>(t.moleculeOfAtom.size(), c->held);
  c->centroids = devAlloc<float4>(std::max(1, t.moleculeCount), c->held);
  upload(c->molStart, sys.moleculeStart, c->stream);
  upload(c->molAtoms, sys.moleculeAtoms, c->stream);
  upload(c->molOfAtom, t.moleculeOfAtom, c->stream);
  // PME.
  PmeDev &p = c->p;
  for (int k = 0; k < 3; ++k) p.K[k] = c->pme.grid[k];
  p.K2h = p.K[2] / 2 + 1;
  for (int k = 0; k < 3; ++k) {
    if (p.K[k] > 1024) throw Error(RMD_E_INVALID, "PME grid sizes above 1024 are not supported");
    std::vector<int> r = factorRadices(p.K[k]);
    for (int q = 0; q < 8; ++q) p.radices[k][q] = q < (int)r.size() ? r[q] : 0;
    std::vector<float2> tw(p.K[k]);
    for (int q = 0; q < p.K[k]; ++q) {
      const double ang = 2.0 * M_PI * q / p.K[k];
      tw[q] = make_float2((float)std::cos(ang), (float)std::sin(ang));
    }
    p.twiddle[k] = devAlloc<float2>(tw.size(), c->held);
    upload(p.twiddle[k], tw, c->stream);
    c->moduli[k] = bsplineModuli(p.K[k], PME_ORDER);
    c->dModuli[k] = devAlloc<double>(p.K[k], c->held);
    upload(c->dModuli[k], c->moduli[k], c->stream);
  }
  const size_t grid = (size_t)p.K[0] * p.K[1] * p.K[2];
  const size_t half = (size_t)p.K[0] * p.K[1] * p.K2h;
  p.qgrid = devAlloc<long long>(grid, c->held);
  p.rgrid = devAlloc<float>(grid, c->held);
  p.cgrid = devAlloc<float2>(half, c->held);
  p.influence = devAlloc<float>(half, c->held);
  p.piOverAlpha2 = M_PI * M_PI / (c->pme.alpha * c->pme.alpha);
  pmeConfigure(p, PME_AUTO, c->maxLds);
  // The plane pipeline's atom bins. A small single system scans every atom
  // per plane (one launch fewer on a short critical path); replica batches
  // bin their members in one workgroup each, and a large single system bins
  // its atoms in groups over the whole chip (the same results: see pme.hip).
  p.binned = 0;
  p.binGroups = 0;
  if (o.pme_bins < 0 || o.pme_bins > 3)
    throw Error(RMD_E_INVALID, "pme_bins must be 0 (tuned), 1 (bins), 2 (scan) or 3 (by atom count)");
  // bins 0 is the rule by atom count (option 3)
  const int bins = o.pme_bins == 0 ? tuned.pmeBins : o.pme_bins == 3 ? 0 : o.pme_bins;
  const bool wide = !shared && pmeBinsFitWide(p) && (bins == 1 || (bins == 0 && d.n >= BIN_SINGLE_MIN_ATOMS));
  if ((shared && pmeBinsFit(p, d.n)) || wide) {
    p.binned = 1;
    p.binStart = devAlloc<int>(p.K[0] + 1, c->held);
    p.atomBin = devAlloc<int2>(d.nPad, c->held);
    p.binAtoms = devAlloc<int>(d.nPad, c->held);
    if (wide) {
      p.binGroups = (d.n + BIN_GROUP_ATOMS - 1) / BIN_GROUP_ATOMS;
      p.binTable = devAlloc<int>((size_t)p.binGroups * p.K[0], c->held);
    }
  }
  c->pmeBinned = p.binned;
  // A small single system folds its x lines into the plane gathers on
  // force-only evaluations (PmeDev::xFolded); the kernel X is kept whenever
  // the grid allows it, so the pipeline can be switched either way.
  p.xFolded = 0;
  p.xReal = 1;
  p.xKernel = nullptr;
  if (!shared && p.K[0] <= X_FOLD_MAX_K) {
    p.xKernel = devAlloc<float>(2 * half, c->held);
    p.xFolded = p.planes && !p.binned ? 1 : 0;
  }
  c->pmeFolded = p.xFolded;
  writeBox(c);
  setupIntegrator(c, 0);
  // Initial state from the system's own coordinates, if it has them.
  HIP_CHECK(hipStreamSynchronize(c->stream));
}

void setState(rmd_ctx *c, const double *pos, const double *vel, int64_t step, int supplied) {
  c->blockEndStep = -1;
  const System &s = c->sys;
  const int n = s.topology.atomCount;
  for (int k = 0; k < 3 * n; ++k)
    if (!std::isfinite(pos[k]) || (vel && !std::isfinite(vel[k])))
      throw Error(RMD_E_NONFINITE, "setState was given a non-finite coordinate or velocity for atom " +
                                       std::to_string(k / 3));
  std::vector<float4> p(c->d.nPad, make_float4(0, 0, 0, 0)), v(c->d.nPad, make_float4(0, 0, 0, 0));
  for (int a = 0; a < n; ++a) {
    p[a] = make_float4((float)pos[3 * a], (float)pos[3 * a + 1], (float)pos[3 * a + 2], (float)s.topology.charges[a]);
    const double m = s.topology.masses[a];
    const float im = m > 0 ? (float)(1.0 / m) : 0.f;
    if (vel)
      v[a] = make_float4((float)vel[3 * a], (float)vel[3 * a + 1], (float)vel[3 * a + 2], im);
    else
      v[a] = make_float4(0, 0, 0, im);
  }
  upload(c->d.pos, p, c->stream);
  upload(c->d.vel, v, c->stream);
  HIP_CHECK(hipMemsetAsync(c->d.sc, 0, sizeof(Scalars), c->stream));
  setStepKernel<<<1, 1, 0, c->stream>>>(c->d.sc, step);
  setForceRebuildKernel<<<1, 1, 0, c->stream>>>(c->d.sc);
  c->step = step;
  c->kineticSampledAtStep = -1;
  c->initialised = false;
  if (supplied != c->ip.suppliedNoise) {
    setupIntegrator(c, supplied);
    invalidateGraphs(c);
  }
  HIP_CHECK(hipStreamSynchronize(c->stream));
}

}  // namespace

// ---- C ABI ----------------------------------------------------------------

#define CTX_GUARD(ctx, ...)                                   \
  do {                                                         \
    if (!(ctx)) {                                              \
      rmd::setGlobalError("null context");                     \
      return RMD_E_INVALID;                                    \
    }                                                          \
    try {                                                      \
      HIP_CHECK(hipSetDevice((ctx)->device));                  \
      __VA_ARGS__;                                             \
      return RMD_OK;                                           \
    } catch (const rmd::Error &e) {                            \
      (ctx)->lastError = e.what();                             \
      rmd::setGlobalError(e.what());                           \
      return e.code;                                           \
    } catch (const std::exception &e) {                        \
      (ctx)->lastError = e.what();                             \
      rmd::setGlobalError(e.what());                           \
      return RMD_E_INVALID;                                    \
    }                                                          \
  } while (0)

extern "C" {

void rmd_options_default(rmd_options *o) {
  std::memset(o, 0, sizeof *o);
  o->timestep = 0.002;
  o->temperature = 298;
  o->friction = 1.0;
  o->seed = 0;
  o->tolerance = 1e-5;
  o->max_constraint_iterations = 150;
  o->rebuild_interval = 20;
  o->adaptive_rebuild = 1;
  o->barostat_interval = 25;
  o->dispersion_correction = 1;
  o->neighbour_slack = 2.0;
  o->ewald_tolerance = 0.0005;
}

void rmd_min_options_default(rmd_min_options *o) {
  o->max_iterations = 200;
  o->force_tolerance = 500;
  o->max_step = 0.002;
  o->max_displacement = 0.05;
}

int rmd_create(const rmd_system *system, const rmd_options *options, rmd_ctx **out) {
  if (!system || !out) {
    setGlobalError("null argument");
    return RMD_E_INVALID;
  }
  rmd_ctx *c = new rmd_ctx();
  try {
    rmd_options d;
    rmd_options_default(&d);
    create(c, system->s, options ? *options : d);
    if (!system->s.positions.empty())
      setState(c, system->s.positions.data(), system->s.velocities.empty() ? nullptr : system->s.velocities.data(), 0, 0);
    *out = c;
    return RMD_OK;
  } catch (const rmd::Error &e) {
    setGlobalError(e.what());
    rmd_destroy(c);
    return e.code;
  } catch (const std::exception &e) {
    setGlobalError(e.what());
    rmd_destroy(c);
    return RMD_E_INVALID;
  }
}

namespace {
void destroyContext(rmd_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  invalidateGraphs(c);
  if (c->screen.net) rmd_surrogate_free(c->screen.net);
  (void)hipSetDevice(c->device);
  for (void *p : c->held) (void)hipFree(p);
  if (c->trialHost) (void)hipHostFree(c->trialHost);
  if (c->ownsStreams) {
    for (int k = 0; k < 2; ++k) {
      if (c->side[k]) (void)hipStreamDestroy(c->side[k]);
      if (c->join[k]) (void)hipEventDestroy(c->join[k]);
    }
    if (c->fork) (void)hipEventDestroy(c->fork);
    if (c->stream) (void)hipStreamDestroy(c->stream);
  }
  delete c;
}
}  // namespace

void rmd_destroy(rmd_ctx *c) {
  // a batch member belongs to its batch (rmd_batch_destroy)
  if (!c || c->batchMember) return;
  destroyContext(c);
}

const char *rmd_last_error(const rmd_ctx *c) { return c ? c->lastError.c_str() : rmd_last_global_error(); }

int rmd_set_state(rmd_ctx *c, const double *pos, const double *vel, int64_t step, int32_t supplied) {
  CTX_GUARD(c, {
    if (!pos) throw Error(RMD_E_INVALID, "positions are required");
    setState(c, pos, vel, step, supplied);
  });
}

int rmd_set_noise(rmd_ctx *c, const float *noise) {
  CTX_GUARD(c, {
    if (!noise) throw Error(RMD_E_INVALID, "noise is required");
    HIP_CHECK(hipMemcpyAsync(c->d.noise, noise, 3 * (size_t)c->d.n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

int rmd_initialise(rmd_ctx *c) {
  CTX_GUARD(c, {
    evaluate(c, RMD_TERM_ALL, true, nullptr);
    c->initialised = true;
  });
}

int rmd_request_stop(rmd_ctx *c) {
  if (!c) return RMD_E_INVALID;
  c->stop.store(1);
  return RMD_OK;
}

int rmd_attempt_volume_change(rmd_ctx *c, rmd_move *m);

// ---- volume moves appended to the block before them -------------------------
//
// rmd_run's volume move comes right after a block, and everything about its
// proposal is known before that block runs: the draw (a counter hash of the
// attempt number), the window, and so the scale factor and the trial box. So
// the trial (enqueueTrial) runs in the block's own launch sequence and its
// energies come back with the block's, one wait instead of two and none of
// the trial's launches from the host. The host then takes the plain
// barostat's decision with rmd_attempt_volume_change's arithmetic, and
// restores the state if the move is turned down; positions, box, energies and
// counters are those of that function bit for bit. Moves the screen decides,
// and proposals that are refused or would grow the cell or list buffers, take
// the host path.
struct FusedMove {
  double volume = 0, deltaVolume = 0, newVolume = 0;
  double box[9] = {0};
};

SnapshotBuffers trialBuffers(const rmd_ctx *c) {
  return SnapshotBuffers{c->trialPos, nullptr, c->rebuildSaved, c->trialForce, nullptr, nullptr};
}

bool proposeFusedMove(rmd_ctx *c, FusedMove &mv) {
  if (!c->hasPressure || c->screen.options.enabled) return false;
  const double volume = volumeOf(c->box);
  const double window = c->volumeScale < 0 ? 0.01 * volume : c->volumeScale;
  const uint32_t bseed = c->seed ^ 0x5a17u;
  const double draw = uniformAt(bseed, (uint32_t)(c->barostatAttempts + 1), 0, 0);
  const double deltaVolume = window * 2 * (draw - 0.5);
  const double newVolume = volume + deltaVolume;
  if (!(newVolume > 0)) return false;
  const double s = std::cbrt(newVolume / volume);
  const double smallest = s * std::min({std::fabs(c->box[0]), std::fabs(c->box[4]), std::fabs(c->box[8])});
  if (c->sys.cutoff > 0.5 * smallest) return false;
  double nb[9];
  for (int k = 0; k < 9; ++k) nb[k] = c->box[k] * s;
  // the trial box's plan as updateBoxHost will make it, without buffer growth
  const CellPlan estimate = planCellList(c->sys, c->plan.skin, c->slack, 0, nb, c->reachMax);
  const int capacity = std::max(c->d.cap, estimate.maxNeighbours);
  const CellPlan plan = planCellList(c->sys, c->plan.skin, c->slack, capacity, nb, c->reachMax);
  if (capacity > c->d.cap || plan.cellCount > c->d.cellCapacity) return false;
  if (!c->trialDev) {
    c->trialDev = devAlloc<TrialParams>(1, c->held);
    c->trialPos = devAlloc<float4>(c->d.nPad, c->held);
    c->trialForce = devAlloc<long long>(3 * (size_t)c->d.nPad, c->held);
    c->dTermsTrial = devAlloc<double>(E_TERM_COUNT, c->held);
    void *host = nullptr;
    HIP_CHECK(hipHostMalloc(&host, sizeof(TrialParams), hipHostMallocDefault));
    c->trialHost = static_cast<TrialParams *>(host);
  }
  c->trialHost->box = boxFor(c, nb, plan);
  c->trialHost->scale = (float)s;
  c->trialHost->rebuildsBefore = 0u;
  mv.volume = volume;
  mv.deltaVolume = deltaVolume;
  mv.newVolume = newVolume;
  std::memcpy(mv.box, nb, sizeof nb);
  return true;
}

// Back to the block's end state (a move that is not taken after all).
void undoFusedTrial(rmd_ctx *c, unsigned int rebuildsBefore) {
  const bool relist = c->rebuildCount != rebuildsBefore;
  launchSnapshot(c->d, trialBuffers(c), SNAP_RESTORE | (relist ? 0 : SNAP_REBUILD), c->stream);
  writeBox(c);
  if (relist) setForceRebuildKernel<<<1, 1, 0, c->stream>>>(c->d.sc);
}

void finishFusedMove(rmd_ctx *c, const FusedMove &mv, const double *trialTerms, unsigned int rebuildsBefore) {
  const double volume = mv.volume;
  if (c->volumeScale < 0) c->volumeScale = 0.01 * volume;
  const double kT = BOLTZMANN * c->temperature;
  const double pressure = c->pressure * BAR_TO_KJ_PER_MOL_NM3;
  double oldTerms[E_TERM_COUNT];
  std::memcpy(oldTerms, c->blockEndTerms, sizeof oldTerms);
  const double oldPotential = potentialFromTerms(c, oldTerms);
  double oldBox[9];
  std::memcpy(oldBox, c->box, sizeof oldBox);
  c->barostatAttempts += 1;
  c->windowAttempts += 1;
  const uint32_t bseed = c->seed ^ 0x5a17u;
  const double lnRatio = std::log(mv.newVolume / volume);
  const double work = pressure * mv.deltaVolume - c->sys.topology.moleculeCount * kT * lnRatio;
  (void)updateBoxHost(c, mv.box);  // the trial box's plan and self energy, as the trial itself had them
  const double newPotential = potentialFromTerms(c, trialTerms);
  const double weight = (newPotential - oldPotential) + work;
  const bool accept =
      weight <= 0 || uniformAt(bseed, (uint32_t)c->barostatAttempts, 1, 0) < std::exp(-weight / kT);
  if (accept) {
    std::memcpy(c->blockEndTerms, trialTerms, sizeof c->blockEndTerms);
    c->barostatAccepted += 1;
    c->windowAccepted += 1;
  } else {
    (void)updateBoxHost(c, oldBox);
    undoFusedTrial(c, rebuildsBefore);
    std::memcpy(c->blockEndTerms, oldTerms, sizeof oldTerms);
  }
  c->blockEndStep = c->step;
  if (c->windowAttempts >= 10) {
    const double rate = (double)c->windowAccepted / c->windowAttempts;
    if (rate < 0.25)
      c->volumeScale /= 1.1;
    else if (rate > 0.75)
      c->volumeScale = std::min(c->volumeScale * 1.1, 0.3 * volume);
    c->windowAttempts = c->windowAccepted = 0;
  }
}

int rmd_run(rmd_ctx *c, int64_t steps, int32_t block, int64_t *stepOut) {
  CTX_GUARD(c, {
    if (!c->initialised) {
      evaluate(c, RMD_TERM_ALL, true, nullptr);
      c->initialised = true;
    }
    const int blockSize = block > 0 ? block : (c->hasPressure ? c->barostatInterval : 25);
    int64_t remaining = steps;
    while (remaining > 0) {
      if (c->stop.exchange(0)) throw Error(RMD_E_STOPPED, "the simulation was stopped between blocks");
      if (c->hasPressure && c->stepsSinceBarostat >= c->barostatInterval) {
        rmd_move mv;
        const int rc = rmd_attempt_volume_change(c, &mv);
        if (rc != RMD_OK) throw Error(rc, c->lastError);
        c->stepsSinceBarostat = 0;
      }
      const int n = (int)std::min<int64_t>(blockSize, remaining);
      // the move due after this block (there are steps left to run) rides
      // in the block's launch sequence where it can
      FusedMove mv;
      const bool fused = remaining > n && c->hasPressure && c->stepsSinceBarostat + n >= c->barostatInterval &&
                         proposeFusedMove(c, mv);
      double trialTerms[E_TERM_COUNT];
      unsigned int rebuildsBefore = 0;
      runBlock(c, n, fused, trialTerms, &rebuildsBefore);
      remaining -= n;
      c->stepsSinceBarostat += n;
      if (fused) {
        if (c->stop.load()) {
          undoFusedTrial(c, rebuildsBefore);  // stopped before the move, as without the trial
        } else {
          finishFusedMove(c, mv, trialTerms, rebuildsBefore);
          c->stepsSinceBarostat = 0;
        }
      }
    }
    if (stepOut) *stepOut = c->step;
  });
}

int rmd_step_once(rmd_ctx *c) {
  CTX_GUARD(c, {
    if (!c->initialised) {
      evaluate(c, RMD_TERM_ALL, true, nullptr);
      c->initialised = true;
    }
    const bool graphs = c->useGraphs;
    c->useGraphs = false;
    try {
      runBlock(c, 1);
    } catch (...) {
      c->useGraphs = graphs;
      throw;
    }
    c->useGraphs = graphs;
  });
}

int rmd_energy(rmd_ctx *c, rmd_energy_terms *out) {
  CTX_GUARD(c, {
    if (!out) throw Error(RMD_E_INVALID, "null output");
    double t[E_TERM_COUNT];
    evaluate(c, RMD_TERM_ALL, false, t);
    int slots[E_TERM_COUNT];
    std::memcpy(slots, c->slotCounts, sizeof slots);
    slots[E_KINETIC] = launchKinetic(c->d, E_KINETIC, c->stream);
    const bool earlyValid = c->kineticSampledAtStep == c->step;
    slots[E_KINETIC_EARLY] = earlyValid ? singleOpenSlotCount(c->d) : 0;
    SlotCounts counts;
    std::memcpy(counts.c, slots, sizeof slots);
    launchReduceEnergies(c->d, counts, c->dTerms, c->stream);
    double all[E_TERM_COUNT];
    HIP_CHECK(hipMemcpyAsync(all, c->dTerms, sizeof all, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    fillEnergy(c, t, out, true, all[E_KINETIC], all[E_KINETIC_EARLY], earlyValid);
    if (!std::isfinite(out->potential) || !std::isfinite(out->kinetic))
      throw Error(RMD_E_NONFINITE, c->step == 0 ? "the first force evaluation is non-finite: the input itself "
                                                  "cannot be evaluated"
                                                : "the energy is non-finite at a reporting interval");
  });
}

int rmd_evaluate(rmd_ctx *c, int32_t terms, rmd_energy_terms *energies, double *forces) {
  CTX_GUARD(c, {
    double t[E_TERM_COUNT];
    // without energies, the force-only launches of a step
    evaluate(c, terms ? terms : RMD_TERM_ALL, true, energies ? t : nullptr);
    if (energies) fillEnergy(c, t, energies, false, 0, 0, false);
    if (forces) toDouble(c, forces);
  });
}

int32_t rmd_atom_count(const rmd_ctx *c) { return c ? c->d.n : RMD_E_INVALID; }

int rmd_read_positions(rmd_ctx *c, float *out) {
  CTX_GUARD(c, {
    std::vector<float4> p(c->d.nPad);
    HIP_CHECK(hipMemcpyAsync(p.data(), c->d.pos, p.size() * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    for (int a = 0; a < c->d.n; ++a) {
      out[3 * a] = p[a].x;
      out[3 * a + 1] = p[a].y;
      out[3 * a + 2] = p[a].z;
    }
  });
}

int rmd_read_state(rmd_ctx *c, float *pos, float *vel, float *forces, double box[9], double *volume) {
  CTX_GUARD(c, {
    std::vector<float4> p(c->d.nPad), v(c->d.nPad);
    HIP_CHECK(hipMemcpyAsync(p.data(), c->d.pos, p.size() * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipMemcpyAsync(v.data(), c->d.vel, v.size() * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    for (int a = 0; a < c->d.n; ++a) {
      if (pos) {
        pos[3 * a] = p[a].x;
        pos[3 * a + 1] = p[a].y;
        pos[3 * a + 2] = p[a].z;
      }
      if (vel) {
        vel[3 * a] = v[a].x;
        vel[3 * a + 1] = v[a].y;
        vel[3 * a + 2] = v[a].z;
      }
    }
    if (forces) {
      std::vector<double> f(3 * (size_t)c->d.n);
      toDouble(c, f.data());
      for (size_t k = 0; k < f.size(); ++k) forces[k] = (float)f[k];
    }
    if (box)
      for (int k = 0; k < 9; ++k) box[k] = c->box[k];
    if (volume) *volume = volumeOf(c->box);
  });
}

int rmd_scale_volume(rmd_ctx *c, double s, double *potential) {
  CTX_GUARD(c, {
    const double u = scaleVolume(c, s);
    if (potential) *potential = u;
  });
}

int rmd_attempt_volume_change(rmd_ctx *c, rmd_move *m) {
  CTX_GUARD(c, {
    if (!m) throw Error(RMD_E_INVALID, "null output");
    std::memset(m, 0, sizeof *m);
    if (!c->hasPressure) return RMD_OK;
    const double volume = volumeOf(c->box);
    if (c->volumeScale < 0) c->volumeScale = 0.01 * volume;
    const double kT = BOLTZMANN * c->temperature;
    const double pressure = c->pressure * BAR_TO_KJ_PER_MOL_NM3;
    // The old potential: the last block's own energies when the state has not
    // changed since (same positions and box), else a fresh evaluation.
    double oldTerms[E_TERM_COUNT];
    if (c->blockEndStep == c->step) {
      std::memcpy(oldTerms, c->blockEndTerms, sizeof oldTerms);
    } else {
      evaluate(c, RMD_TERM_ALL, true, oldTerms);
    }
    const double oldPotential = potentialFromTerms(c, oldTerms);
    double oldBox[9];
    std::memcpy(oldBox, c->box, sizeof oldBox);
    launchSnapshot(c->d, savedBuffers(c), SNAP_REBUILD, c->stream);  // positions, rebuild positions, forces
    const unsigned int rebuildsBefore = c->rebuildCount;
    c->barostatAttempts += 1;
    c->windowAttempts += 1;
    const uint32_t bseed = c->seed ^ 0x5a17u;
    const double draw = uniformAt(bseed, (uint32_t)c->barostatAttempts, 0, 0);
    const double deltaVolume = c->volumeScale * 2 * (draw - 0.5);
    const double newVolume = volume + deltaVolume;
    m->attempted = 1;
    if (!(newVolume > 0)) return RMD_OK;
    const double s = std::cbrt(newVolume / volume);
    const double smallest = s * std::min({std::fabs(c->box[0]), std::fabs(c->box[4]), std::fabs(c->box[8])});
    m->delta_volume = deltaVolume;
    m->length_scale = s;
    m->volume = volume;
    if (c->sys.cutoff > 0.5 * smallest) {
      c->barostatRefusals += 1;
      m->refused = 1;
      m->volume_scale = c->volumeScale;
      m->acceptance_rate = (double)c->barostatAccepted / c->barostatAttempts;
      return RMD_OK;
    }
    auto &sc = c->screen;
    const bool screening = sc.options.enabled && sc.phase == 1 && sc.net;
    const double lnRatio = std::log(newVolume / volume);
    const double work = pressure * deltaVolume - c->sys.topology.moleculeCount * kT * lnRatio;
    const double dispersionChange = dispersionAt(c, newVolume) - dispersionAt(c, volume);
    float fx[RMD_SURROGATE_INPUTS];
    double estimateX = 0.0, predictedX = 0.0, firstStage = 1.0;
    if (sc.options.enabled) screenFeatures(c, oldTerms, volume, newVolume, fx, &estimateX);
    bool evaluated = true, accept = false;
    double newPotential = oldPotential, weight = 0.0;
    if (screening) {
      // Stage one: the predicted weight alone; a move it turns down is never
      // evaluated and the state is untouched.
      sc.moves += 1;
      predictedX = dispersionChange + estimateX + predictResidual(c, fx);
      const double predictedWeight = predictedX + work;
      firstStage = predictedWeight <= 0 ? 1.0 : std::exp(-predictedWeight / kT);
      if (uniformAt(bseed, (uint32_t)c->barostatAttempts, 2, 0) >= firstStage) {
        sc.screened += 1;
        evaluated = false;
        weight = predictedWeight;
      }
    }
    if (evaluated) {
      double newTerms[E_TERM_COUNT];
      newPotential = moveVolume(c, s, newTerms);
      weight = (newPotential - oldPotential) + work;
      const double change = newPotential - oldPotential;
      if (sc.options.enabled) sc.exact += 1;
      if (screening) {
        // Stage two (delayed acceptance): exp(-w/kT) times the reverse move's
        // stage-one probability over the forward one's.
        float fy[RMD_SURROGATE_INPUTS];
        double estimateY = 0.0;
        screenFeatures(c, newTerms, newVolume, volume, fy, &estimateY);
        const double predictedReverse = -dispersionChange + estimateY + predictResidual(c, fy) - work;
        const double reverseStage = predictedReverse <= 0 ? 1.0 : std::exp(-predictedReverse / kT);
        const double ratio = std::exp(-weight / kT) * reverseStage / firstStage;
        accept = ratio >= 1.0 || uniformAt(bseed, (uint32_t)c->barostatAttempts, 1, 0) < ratio;
        if (!accept) sc.secondStage += 1;
        sc.residual2 += (change - predictedX) * (change - predictedX);
        sc.estimate2 += (change - dispersionChange - estimateX) * (change - dispersionChange - estimateX);
        sc.residualCount += 1;
      } else {
        accept = weight <= 0 || uniformAt(bseed, (uint32_t)c->barostatAttempts, 1, 0) < std::exp(-weight / kT);
        if (sc.options.enabled && sc.phase == 0) {
          // both directions of the move: the network sees the pair's antisymmetry
          float fy[RMD_SURROGATE_INPUTS];
          double estimateY = 0.0;
          screenFeatures(c, newTerms, newVolume, volume, fy, &estimateY);
          sc.x.insert(sc.x.end(), fx, fx + RMD_SURROGATE_INPUTS);
          sc.t.push_back((float)(change - dispersionChange - estimateX));
          sc.x.insert(sc.x.end(), fy, fy + RMD_SURROGATE_INPUTS);
          sc.t.push_back((float)(-change + dispersionChange - estimateY));
        }
      }
      if (accept) {
        std::memcpy(c->blockEndTerms, newTerms, sizeof newTerms);
      } else {
        // Back to the old state exactly: positions, box (copied, not divided
        // back), forces; the old energies describe it again. The list stays
        // with its rebuild references restored, unless the trial rebuilt it
        // (then it describes the trial state) or replaced it.
        const bool replaced = updateBoxHost(c, oldBox);
        const bool relist = replaced || c->rebuildCount != rebuildsBefore;
        launchSnapshot(c->d, savedBuffers(c), SNAP_RESTORE | (relist ? 0 : SNAP_REBUILD), c->stream);
        writeBox(c);
        if (relist) setForceRebuildKernel<<<1, 1, 0, c->stream>>>(c->d.sc);
        std::memcpy(c->blockEndTerms, oldTerms, sizeof oldTerms);
      }
    }
    if (accept) {
      c->barostatAccepted += 1;
      c->windowAccepted += 1;
      if (sc.options.enabled) sc.accepted += 1;
    }
    c->blockEndStep = c->step;
    if (sc.options.enabled && sc.phase == 0 && sc.net &&
        (int64_t)sc.t.size() >= 2 * (int64_t)sc.options.training_moves)
      screenTrain(c);
    if (c->windowAttempts >= 10) {
      const double rate = (double)c->windowAccepted / c->windowAttempts;
      if (rate < 0.25)
        c->volumeScale /= 1.1;
      else if (rate > 0.75)
        c->volumeScale = std::min(c->volumeScale * 1.1, 0.3 * volume);
      c->windowAttempts = c->windowAccepted = 0;
    }
    m->accepted = accept;
    m->weight = weight;
    m->old_potential = oldPotential;
    m->new_potential = evaluated ? newPotential : NAN;
    m->volume = volumeOf(c->box);
    m->volume_scale = c->volumeScale;
    m->acceptance_rate = (double)c->barostatAccepted / c->barostatAttempts;
  });
}

int rmd_minimise(rmd_ctx *c, const rmd_min_options *options, rmd_min_result *out) {
  CTX_GUARD(c, {
    c->blockEndStep = -1;
    rmd_min_options o;
    rmd_min_options_default(&o);
    if (options) o = *options;
    double maxStep = o.max_step;
    const size_t np = c->d.nPad;
    auto maxForce = [&]() {
      HIP_CHECK(hipMemsetAsync(c->dMaxForce, 0, sizeof(unsigned int), c->stream));
      launchMaxForce(c->d, c->dMaxForce, c->stream);
      unsigned int bits = 0;
      HIP_CHECK(hipMemcpyAsync(&bits, c->dMaxForce, sizeof bits, hipMemcpyDeviceToHost, c->stream));
      HIP_CHECK(hipStreamSynchronize(c->stream));
      return (double)__builtin_bit_cast(float, bits);
    };
    double potential = evaluatePotential(c);
    int iterations = 0, rejected = 0;
    bool stalled = false;
    double worst = 0;
    for (int it = 0; it < o.max_iterations; ++it) {
      iterations = it + 1;
      worst = maxForce();
      if (worst <= o.force_tolerance) break;
      HIP_CHECK(hipMemcpyAsync(c->posSaved, c->d.pos, np * sizeof(float4), hipMemcpyDeviceToDevice, c->stream));
      HIP_CHECK(hipMemcpyAsync(c->forceSaved, c->d.force, 3 * np * sizeof(long long), hipMemcpyDeviceToDevice, c->stream));
      bool accepted = false;
      for (int attempt = 0; attempt < 14; ++attempt) {
        launchDisplace(c->d, (float)(maxStep / std::max(worst, 1e-6)), (float)o.max_displacement, c->stream);
        launchConstrainPositions(c->d, c->posSaved, c->ip, c->stream);
        const double trial = evaluatePotential(c);
        if (trial < potential) {
          potential = trial;
          maxStep *= 1.3;
          accepted = true;
          break;
        }
        rejected += 1;
        // Back to the pre-move positions AND forces: the next, smaller step
        // goes along the gradient at the base point.
        HIP_CHECK(hipMemcpyAsync(c->d.pos, c->posSaved, np * sizeof(float4), hipMemcpyDeviceToDevice, c->stream));
        HIP_CHECK(hipMemcpyAsync(c->d.force, c->forceSaved, 3 * np * sizeof(long long), hipMemcpyDeviceToDevice, c->stream));
        maxStep *= 0.5;
        if (maxStep < 1e-9) break;
      }
      if (!accepted) {
        stalled = true;
        break;
      }
    }
    worst = maxForce();
    // Velocities are zeroed, as the reference does after minimising.
    std::vector<float4> v(np);
    HIP_CHECK(hipMemcpyAsync(v.data(), c->d.vel, np * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    for (auto &x : v) x.x = x.y = x.z = 0.f;
    HIP_CHECK(hipMemcpyAsync(c->d.vel, v.data(), np * sizeof(float4), hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    c->initialised = true;
    if (out) {
      out->iterations = iterations;
      out->rejected = rejected;
      out->stalled = stalled;
      out->converged = worst <= o.force_tolerance;
      out->potential = potential;
      out->max_force = worst;
      out->max_step = maxStep;
      out->energy_resolution = std::fabs(potential) * std::pow(2.0, -24) * std::sqrt((double)c->d.n);
    }
  });
}

int rmd_diagnostics_get(rmd_ctx *c, rmd_diagnostics *o) {
  CTX_GUARD(c, {
    std::memset(o, 0, sizeof *o);
    Scalars h;
    HIP_CHECK(hipMemcpyAsync(&h, c->d.sc, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    o->max_shake_iterations = (int32_t)h.shakeMaxIterations;
    o->shake_hit_iteration_cap = h.shakeFailed ? 1 : 0;
    o->rattle_hit_iteration_cap = h.rattleFailed ? 1 : 0;
    o->neighbour_required = (int32_t)h.required;
    o->neighbour_capacity = c->d.cap;
    o->worst_neighbour_required = (int32_t)c->worstRequired;
    o->neighbour_growths = c->neighbourGrowths;
    o->skin = c->plan.skin;
    o->worst_displacement = std::sqrt(std::max(0.0f, __builtin_bit_cast(float, h.maxDisp2Bits)));
    o->rebuilds = h.rebuilds;
    o->steps = c->stepsRun;
    for (int k = 0; k < 3; ++k) {
      o->cell_counts[k] = c->plan.counts[k];
      o->pme_grid[k] = c->pme.grid[k];
    }
    o->volume = volumeOf(c->box);
    o->barostat_attempts = c->barostatAttempts;
    o->barostat_accepted = c->barostatAccepted;
    o->barostat_refused = c->barostatRefusals;
    o->volume_scale = c->volumeScale;
    o->constraint_tolerance = c->tolerance;
    o->pme_alpha = c->pme.alpha;
    o->lanes_per_atom = c->lanes;
    o->pme_planes = c->p.planes;
    o->pme_x_folded = c->p.planes && c->p.xFolded && !c->p.binned ? 1 : 0;
    o->water_units = c->waterUnits;
    o->cluster_units = c->clusterUnits;
    o->free_units = c->freeUnits;
    o->cell_reach = c->plan.reach;
    o->search_by_cell = searchByCell(c->searchMode, c->d.n, c->plan.cellCount) ? 1 : 0;
    o->steps_since_barostat = c->stepsSinceBarostat;
  });
}

int rmd_read_cells(rmd_ctx *c, int32_t *cellOfAtom, int32_t *cellStart, int32_t *cellAtoms) {
  CTX_GUARD(c, {
    HIP_CHECK(hipStreamSynchronize(c->stream));
    if (cellOfAtom) HIP_CHECK(hipMemcpy(cellOfAtom, c->d.cellOfAtom, c->d.n * sizeof(int), hipMemcpyDeviceToHost));
    if (cellStart)
      HIP_CHECK(hipMemcpy(cellStart, c->d.cellStart, (c->plan.cellCount + 1) * sizeof(int), hipMemcpyDeviceToHost));
    if (cellAtoms) HIP_CHECK(hipMemcpy(cellAtoms, c->d.cellAtoms, c->d.n * sizeof(int), hipMemcpyDeviceToHost));
  });
}

int rmd_read_neighbours(rmd_ctx *c, int32_t *counts, int32_t *lists) {
  CTX_GUARD(c, {
    // rows and entries are cell slots on the device; the caller gets them
    // by atom
    const int n = c->d.n, cap = c->d.cap;
    HIP_CHECK(hipStreamSynchronize(c->stream));
    std::vector<int> atomOf(n), rowCount(n);
    HIP_CHECK(hipMemcpy(atomOf.data(), c->d.cellAtoms, n * sizeof(int), hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(rowCount.data(), c->d.nbrCount, n * sizeof(int), hipMemcpyDeviceToHost));
    if (counts)
      for (int r = 0; r < n; ++r) counts[atomOf[r]] = rowCount[r];
    if (lists) {
      std::vector<int> rows((size_t)n * cap);
      HIP_CHECK(hipMemcpy(rows.data(), c->d.nbrList, rows.size() * sizeof(int), hipMemcpyDeviceToHost));
      for (int r = 0; r < n; ++r) {
        int32_t *dst = lists + (size_t)atomOf[r] * cap;
        for (int k = 0; k < cap; ++k) dst[k] = k < rowCount[r] ? atomOf[rows[(size_t)r * cap + k]] : 0;
      }
    }
  });
}

int rmd_device_info_get(int32_t device, rmd_device_info *out) {
  if (!out) return RMD_E_INVALID;
  std::memset(out, 0, sizeof *out);
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess) count = 0;
  out->device_count = count;
  if (count <= 0 || device >= count) {
    setGlobalError("no HIP device is available");
    return RMD_E_DEVICE;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return RMD_E_DEVICE;
  std::strncpy(out->name, prop.name, sizeof out->name - 1);
  std::strncpy(out->arch, prop.gcnArchName, sizeof out->arch - 1);
  out->compute_units = prop.multiProcessorCount;
  out->global_memory = (int64_t)prop.totalGlobalMem;
  return RMD_OK;
}

// Verification entry: the forward real-to-complex 3D transform of a real grid
// with the context's PME plan. out holds K0*K1*(K2/2+1) complex values.
int rmd_debug_fft_forward(rmd_ctx *c, const float *real, float *out) {
  CTX_GUARD(c, {
    const size_t n = (size_t)c->p.K[0] * c->p.K[1] * c->p.K[2];
    const size_t nc = (size_t)c->p.K[0] * c->p.K[1] * c->p.K2h;
    std::vector<void *> tmp;
    float *dr = devAlloc<float>(n, tmp);
    float2 *dc = devAlloc<float2>(nc, tmp);
    HIP_CHECK(hipMemcpyAsync(dr, real, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    launchFftForward(c->d, c->p, dr, dc, c->stream);
    HIP_CHECK(hipMemcpyAsync(out, dc, nc * sizeof(float2), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    for (void *p : tmp) (void)hipFree(p);
  });
}

// Verification entry: the thermostat's counter-hash draws for every atom at
// one step with the context's seed: 4 uniforms per atom (may be null) and the
// 3 Box-Muller normals per atom, from the integrator's own device functions.
int rmd_debug_noise(rmd_ctx *c, int64_t step, float *uniforms, float *normals) {
  CTX_GUARD(c, {
    if (!normals) throw Error(RMD_E_INVALID, "normals are required");
    const int n = c->d.n;
    std::vector<void *> tmp;
    try {
      float *du = uniforms ? devAlloc<float>(4 * (size_t)n, tmp) : nullptr;
      float *dg = devAlloc<float>(3 * (size_t)n, tmp);
      launchNoise(n, c->ip.seed, (unsigned int)step, du, dg, c->stream);
      HIP_CHECK(hipGetLastError());
      if (uniforms)
        HIP_CHECK(hipMemcpyAsync(uniforms, du, 4 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
      HIP_CHECK(hipMemcpyAsync(normals, dg, 3 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
      HIP_CHECK(hipStreamSynchronize(c->stream));
    } catch (...) {
      for (void *p : tmp) (void)hipFree(p);
      throw;
    }
    for (void *p : tmp) (void)hipFree(p);
  });
}

int rmd_set_lanes(rmd_ctx *c, int32_t lanes) {
  CTX_GUARD(c, {
    if (lanes <= 0 || lanes > 64 || (lanes & (lanes - 1))) throw Error(RMD_E_INVALID, "lanes must be a power of two <= 64");
    c->lanes = lanes;
    invalidateGraphs(c);
  });
}

int rmd_set_graphs(rmd_ctx *c, int32_t enabled) {
  CTX_GUARD(c, {
    c->useGraphs = enabled != 0;
    invalidateGraphs(c);
  });
}

// Average duration of one launch of a kernel family, timed with HIP events
// on the context's stream at the current state (the list is rebuilt first so
// the nonbonded kernel sees a fresh list). kernel: 0 nonbonded, 1 PME chain,
// 2 integrator (open), 3 bonded/exceptions, 4 neighbour rebuild chain.
int rmd_profile_kernel(rmd_ctx *c, int32_t kernel, int32_t reps, double *avg_ms) {
  CTX_GUARD(c, {
    if (!avg_ms || reps <= 0) throw Error(RMD_E_INVALID, "bad arguments");
    evaluate(c, RMD_TERM_ALL, true, nullptr);
    hipEvent_t a, b;
    HIP_CHECK(hipEventCreate(&a));
    HIP_CHECK(hipEventCreate(&b));
    int slots[E_TERM_COUNT];
    auto once = [&]() {
      switch (kernel) {
        case 0:
          launchNonbonded(c->d, (float)c->pme.alpha, ewaldRationalFits(c->pme.alpha, c->sys.cutoff) ? 1 : 0, c->lanes,
                          0, c->stream);
          break;
        case 1: launchPme(c->d, c->p, 0, c->stream, nullptr); break;
        case 2: {
          // the integrator alone: repeated launches drift into rebuilds,
          // whose cell list a step's integrator may build (cellsInIntegrator)
          IntegratorParams ip = c->ip;
          ip.cellsInIntegrator = 0;
          launchIntegrateOpen(c->d, ip, 1, 0, c->stream);
          break;
        }
        case 3: launchBonded(c->d, c->pme.alpha, RMD_TERM_ALL, 0, c->stream, nullptr); break;
        default:
          launchForceRebuildFlag(c->d, c->stream);
          launchRebuild(c->d, c->stream);
          break;
      }
    };
    // The integrator moves atoms; time it on a copy of the state and put the
    // state back afterwards.
    if (kernel == 2) snapshot(c);
    once();
    HIP_CHECK(hipEventRecord(a, c->stream));
    for (int r = 0; r < reps; ++r) once();
    HIP_CHECK(hipEventRecord(b, c->stream));
    HIP_CHECK(hipEventSynchronize(b));
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, a, b));
    *avg_ms = ms / reps;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (kernel == 2) restore(c);
    (void)slots;
    evaluate(c, RMD_TERM_ALL, true, nullptr);
  });
}

// Average durations of the PME chain's three stages (spread with the forward
// transforms it carries; the remaining transforms and the convolution; the
// inverse transforms and the gather), each over reps back-to-back launches of
// the stage on the context's stream (the middle stage as the first stage with
// it less the first stage alone: the x lines transform in place), at the
// current state; the middle stage is 0 where the gathers fold it in.
int rmd_profile_pme_stages(rmd_ctx *c, int32_t reps, double *stage_ms) {
  CTX_GUARD(c, {
    if (!stage_ms || reps <= 0) throw Error(RMD_E_INVALID, "bad arguments");
    evaluate(c, RMD_TERM_ALL, true, nullptr);
    hipEvent_t a, b;
    HIP_CHECK(hipEventCreate(&a));
    HIP_CHECK(hipEventCreate(&b));
    auto timed = [&](std::initializer_list<int> stages) {
      for (int st : stages) launchPmeStage(c->d, c->p, c->stream, st);  // warm
      HIP_CHECK(hipEventRecord(a, c->stream));
      for (int r = 0; r < reps; ++r)
        for (int st : stages) launchPmeStage(c->d, c->p, c->stream, st);
      HIP_CHECK(hipEventRecord(b, c->stream));
      HIP_CHECK(hipEventSynchronize(b));
      float ms = 0;
      HIP_CHECK(hipEventElapsedTime(&ms, a, b));
      return (double)ms / reps;
    };
    const double forward = timed({0});
    stage_ms[0] = forward;
    stage_ms[1] = pmeTransformsFolded(c->p) ? 0.0 : std::max(0.0, timed({0, 1}) - forward);
    stage_ms[2] = timed({2});
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    evaluate(c, RMD_TERM_ALL, true, nullptr);
  });
}

void rmd_screen_options_default(rmd_screen_options *o) {
  if (!o) return;
  o->enabled = 1;
  o->training_moves = 2000;
  o->training_steps = 4000;
  o->learning_rate = 1e-3f;
  o->seed = 0x5eedu;
}

int rmd_screen_configure(rmd_ctx *c, const rmd_screen_options *options) {
  CTX_GUARD(c, {
    rmd_screen_options o;
    rmd_screen_options_default(&o);
    if (options) o = *options;
    if (o.enabled && (o.training_moves <= 0 || o.training_steps <= 0 || !(o.learning_rate > 0)))
      throw Error(RMD_E_INVALID, "the screen needs training_moves, training_steps and learning_rate > 0");
    auto &sc = c->screen;
    if (sc.net) rmd_surrogate_free(sc.net);
    sc = ScreenState();
    sc.options = o;
    invalidateGraphs(c);  // the blocks' end energies change with the screen (blockEndMolecular)
    if (o.enabled) {
      rmd_surrogate *net = nullptr;
      const int st = rmd_surrogate_create(c->device, o.seed, &net);
      if (st != RMD_OK) throw Error(st, std::string("screen network: ") + rmd_last_global_error());
      HIP_CHECK(hipSetDevice(c->device));
      sc.net = net;
      sc.refVolume = volumeOf(c->box);
    }
  });
}

int rmd_screen_stats_get(rmd_ctx *c, rmd_screen_stats *out) {
  CTX_GUARD(c, {
    if (!out) throw Error(RMD_E_INVALID, "null output");
    const auto &sc = c->screen;
    std::memset(out, 0, sizeof *out);
    out->phase = sc.phase;
    out->samples = (int64_t)sc.t.size();
    out->moves = sc.moves;
    out->exact_evaluations = sc.exact;
    out->screened_out = sc.screened;
    out->second_stage_rejections = sc.secondStage;
    out->accepted = sc.accepted;
    out->training_loss = sc.trainingLoss;
    out->residual_rms = sc.residualCount ? std::sqrt(sc.residual2 / sc.residualCount) : 0.0;
    out->estimate_rms = sc.residualCount ? std::sqrt(sc.estimate2 / sc.residualCount) : 0.0;
  });
}

int rmd_screen_samples(rmd_ctx *c, float *features, float *targets, int32_t capacity, int32_t *count) {
  CTX_GUARD(c, {
    const auto &sc = c->screen;
    const int32_t n = (int32_t)std::min<size_t>(sc.t.size(), (size_t)std::max(0, capacity));
    if (features) std::memcpy(features, sc.x.data(), (size_t)n * RMD_SURROGATE_INPUTS * sizeof(float));
    if (targets) std::memcpy(targets, sc.t.data(), (size_t)n * sizeof(float));
    if (count) *count = (int32_t)sc.t.size();
  });
}

int rmd_set_concurrent_forces(rmd_ctx *c, int32_t enabled) {
  CTX_GUARD(c, {
    HIP_CHECK(hipStreamSynchronize(c->stream));
    c->concurrent = enabled != 0;
    invalidateGraphs(c);
  });
}

int rmd_set_pme_pipeline(rmd_ctx *c, int32_t mode) {
  CTX_GUARD(c, {
    if (mode < PME_AUTO || mode > PME_PLANES_FOLDED)
      throw Error(RMD_E_INVALID, "PME pipeline must be 0, 1, 2, 3 or 4");
    const bool scan = mode == PME_PLANES_SCAN || mode == PME_PLANES_FOLDED;
    if (!pmeConfigure(c->p, scan ? PME_PLANES : mode, c->maxLds))
      throw Error(RMD_E_INVALID, "one x plane of this PME grid does not fit in LDS");
    if (mode == PME_PLANES_FOLDED && !c->p.xKernel)
      throw Error(RMD_E_INVALID, "folded x lines need a single system of at most 32 x planes");
    // planes without atom bins (every plane scans every atom), or the bins
    // the context was created with; folded x lines as asked, or as created
    c->p.binned = scan ? 0 : c->pmeBinned;
    c->p.xFolded = mode == PME_PLANES_FOLDED ? 1 : mode == PME_AUTO ? c->pmeFolded : 0;
    invalidateGraphs(c);
  });
}

__global__ void copyProbeKernel(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

int rmd_copy_probe(rmd_ctx *c, int64_t bytes, int32_t reps, double *avg_ms) {
  CTX_GUARD(c, {
    if (bytes <= 0 || bytes % 16 || reps <= 0 || !avg_ms) throw Error(RMD_E_INVALID, "bytes must be a positive multiple of 16");
    const size_t n = (size_t)bytes / 16;
    std::vector<void *> tmp;
    float4 *src = devAlloc<float4>(n, tmp), *dst = devAlloc<float4>(n, tmp);
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 8192);
    copyProbeKernel<<<blocks, 256, 0, c->stream>>>(src, dst, n);
    hipEvent_t a, b;
    HIP_CHECK(hipEventCreate(&a));
    HIP_CHECK(hipEventCreate(&b));
    HIP_CHECK(hipEventRecord(a, c->stream));
    for (int r = 0; r < reps; ++r) copyProbeKernel<<<blocks, 256, 0, c->stream>>>(src, dst, n);
    HIP_CHECK(hipEventRecord(b, c->stream));
    HIP_CHECK(hipEventSynchronize(b));
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, a, b));
    *avg_ms = ms / reps;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    for (void *p : tmp) (void)hipFree(p);
  });
}

int rmd_get_stream(rmd_ctx *c, void **stream) {
  CTX_GUARD(c, { *stream = (void *)c->stream; });
}

}  // extern "C"

// ============================================================================
// Replica batches
// ============================================================================
//
// A batch is R member contexts, each a complete single-system context (its
// own buffers, Scalars, energy slots, PME descriptor and seeds), stepped by
// batched launches: one launch per kernel covers every replica's workgroups
// through a task table, so a step of the whole batch is the same dozen
// launches as a step of one system, however many replicas there are. Each
// replica computes exactly what its own context would compute alone (the
// kernel bodies and per-replica grids are the single-system ones), so a
// batch run reproduces the single runs bit for bit. Member contexts stay
// usable through the single-system calls (set state, minimise, energy,
// read state, diagnostics, and a run of that replica alone).

struct rmd_batch {
  std::string lastError;
  int device = 0;
  int replicas = 0;
  Streams streams;
  std::vector<rmd_ctx *> reps;
  std::vector<void *> held;
  std::vector<BatchView> views;  // host copy of the device views
  BatchView *dViews = nullptr;
  TaskTable tables[BK_COUNT];
  int *searchQueue = nullptr;  // the due replicas' list search blocks (searchQueueInts ints)
  int searchGrid = 0;          // workgroups of the batched list search
  Scalars *scAll = nullptr, *scSaved = nullptr;  // every replica's scalars, contiguous
  BoxF *boxAll = nullptr;
  double *termsAll = nullptr;  // replicas * E_TERM_COUNT
  std::vector<Scalars> hostScalars;
  std::vector<double> hostTerms;
  size_t ldsForward = 0, ldsXlines = 0, ldsGather = 0;
  int ldsOpen = 0;
  int starMax = 0;  // the widest star unit of any member
  double alpha = 0;
  int lanes = 16;  // direct-space lanes per atom of every member
  int maxTypes = 0;  // the largest LJ type count of the members
  int allRect = 1;   // every member's box is rectangular
  int rational = 1;  // the direct-space Coulomb force from the rational fit for every member
  bool concurrent = true, useGraphs = true;
  std::map<int, hipGraphExec_t> graphs;
  std::atomic<int> stop{0};
  int64_t blocksRun = 0;
  ScreenState screen;  // the learned screen shared by the replicas
};

namespace {

int replicaBlocks(const rmd_ctx *c, int kind) {
  const Dev &d = c->d;
  switch (kind) {
    case BK_OPEN: return generalBlockCount(d);
    case BK_WATER: return d.waterBlocks;
    case BK_UNITS: return unitBlockCount(d);
    case BK_BONDED: return bondedBlocks(d);
    case BK_PLANES: return c->p.K[0];
    case BK_XLINES: return pmeXLineBlocks(c->p);
    case BK_REPLICA: return 1;
    case BK_ATOMS256: return (d.nPad + 255) / 256;
    case BK_MOLECULES: return std::max(1, (c->sys.topology.moleculeCount + 255) / 256);
    case BK_INFLUENCE: return (int)(((size_t)c->p.K[0] * c->p.K[1] * c->p.K2h + 255) / 256);
    case BK_NONBONDED: return (int)(((long long)d.n * c->lanes + 255) / 256);
  }
  return 0;
}

// Workgroups reach the eight XCDs round robin (workgroup b runs on XCD
// b % 8), and each XCD has its own L2. The table gives all of a replica's
// workgroups to one XCD, so a replica's positions, lists and grids stay in
// one L2; replicas go to the least loaded XCD, largest first, and the
// shorter XCD queues are padded with empty tasks.
TaskTable buildTasks(rmd_batch *b, int kind) {
  const int R = b->replicas;
  std::vector<int> per(R), order(R);
  for (int r = 0; r < R; ++r) per[r] = replicaBlocks(b->reps[r], kind);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return per[x] > per[y]; });
  std::vector<std::vector<int2>> bins(XCD_COUNT);
  std::vector<long long> load(XCD_COUNT, 0);
  for (int r : order) {
    const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    for (int k = 0; k < per[r]; ++k) bins[x].push_back(make_int2(r, k));
    load[x] += per[r];
  }
  size_t longest = 0;
  for (auto &bin : bins) longest = std::max(longest, bin.size());
  std::vector<int2> task(longest * XCD_COUNT, make_int2(-1, 0));
  for (int x = 0; x < XCD_COUNT; ++x)
    for (size_t k = 0; k < bins[x].size(); ++k) task[k * XCD_COUNT + x] = bins[x][k];
  TaskTable t;
  t.blocks = (int)task.size();
  t.task = devAlloc<int2>(std::max<size_t>(1, task.size()), b->held);
  upload(t.task, task, b->streams.stream);
  HIP_CHECK(hipStreamSynchronize(b->streams.stream));
  return t;
}

// Refresh the device views from the member contexts (their buffers move when
// a list grows or the cell grid changes) with the flags the host has set.
// The batch launches the rectangular-box form of the pair kernels when every
// member's box is rectangular; a change drops the captured launches.
void refreshRect(rmd_batch *b) {
  int all = 1;
  for (rmd_ctx *c : b->reps) {
    c->d.rect = rectangular(c->box) ? 1 : 0;
    all = all && c->d.rect;
  }
  if (all != b->allRect) {
    b->allRect = all;
    for (auto &g : b->graphs) (void)hipGraphExecDestroy(g.second);
    b->graphs.clear();
  }
}

void uploadViews(rmd_batch *b) {
  refreshRect(b);
  for (int r = 0; r < b->replicas; ++r) {
    const rmd_ctx *c = b->reps[r];
    BatchView &v = b->views[r];
    v.d = c->d;
    v.p = c->p;
    v.ip = c->ip;
    v.openBlocks = replicaBlocks(c, BK_OPEN);
    v.openSlots = openSlotCount(c->d);
    std::memset(&v.counts, 0, sizeof v.counts);
    const int nb = replicaBlocks(c, BK_NONBONDED), bb = replicaBlocks(c, BK_BONDED), xb = replicaBlocks(c, BK_XLINES);
    v.counts.c[E_LJ_DIRECT] = v.counts.c[E_COUL_DIRECT] = v.counts.c[E_DSCALE_DIRECT] = nb;
    v.counts.c[E_BOND] = v.counts.c[E_ANGLE] = v.counts.c[E_TORSION] = v.counts.c[E_LJ_EXCEPTION] =
        v.counts.c[E_COUL_EXCEPTION] = v.counts.c[E_RESTRAINT] = v.counts.c[E_DSCALE_BONDED] = bb;
    v.counts.c[E_DSCALE_MOLECULAR] = replicaBlocks(c, BK_MOLECULES);
    v.counts.c[E_RECIPROCAL] = v.counts.c[E_DSCALE_RECIPROCAL] = xb;
    v.terms = b->termsAll + (size_t)r * E_TERM_COUNT;
    v.molStart = c->molStart;
    v.molAtoms = c->molAtoms;
    v.molOfAtom = c->molOfAtom;
    v.nMol = c->sys.topology.moleculeCount;
    v.centroids = c->centroids;
    v.posSaved = c->posSaved;
    v.velSaved = c->velSaved;
    v.rebuildSaved = c->rebuildSaved;
    v.forceSaved = c->forceSaved;
    for (int k = 0; k < 3; ++k) v.moduli[k] = c->dModuli[k];
    v.alpha = c->pme.alpha;
  }
  HIP_CHECK(hipMemcpyAsync(b->dViews, b->views.data(), b->views.size() * sizeof(BatchView), hipMemcpyHostToDevice,
                           b->streams.stream));
  HIP_CHECK(hipStreamSynchronize(b->streams.stream));
}

void clearFlags(rmd_batch *b) {
  for (auto &v : b->views) {
    v.rebuild = 0;
    v.restore = 0;
    v.boxChanged = 0;
    v.scale = 1.0f;
  }
}

// Every replica's device box from its host box, in one copy.
void uploadBoxes(rmd_batch *b) {
  refreshRect(b);
  std::vector<BoxF> boxes(b->replicas);
  for (int r = 0; r < b->replicas; ++r) boxes[r] = boxFor(b->reps[r]);
  HIP_CHECK(hipMemcpyAsync(b->boxAll, boxes.data(), boxes.size() * sizeof(BoxF), hipMemcpyHostToDevice,
                           b->streams.stream));
  HIP_CHECK(hipStreamSynchronize(b->streams.stream));
}

void readBack(rmd_batch *b, bool terms) {
  hipStream_t s = b->streams.stream;
  HIP_CHECK(hipMemcpyAsync(b->hostScalars.data(), b->scAll, b->hostScalars.size() * sizeof(Scalars),
                           hipMemcpyDeviceToHost, s));
  if (terms)
    HIP_CHECK(hipMemcpyAsync(b->hostTerms.data(), b->termsAll, b->hostTerms.size() * sizeof(double),
                             hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  for (int r = 0; r < b->replicas; ++r)
    if (b->hostScalars[r].cellOverflow)
      throw Error(RMD_E_INVALID, "replica " + std::to_string(r) +
                                     ": a cell held more atoms than the cell sort supports; the box is too dense "
                                     "or too small for this cutoff");
}

std::string replicaLabel(int r) { return "replica " + std::to_string(r) + ": "; }

// Grow every overflowed replica's list; true when any grew.
bool growOverflowed(rmd_batch *b) {
  bool grew = false;
  for (int r = 0; r < b->replicas; ++r) {
    rmd_ctx *c = b->reps[r];
    const Scalars &h = b->hostScalars[r];
    const unsigned int need = requirementSince(h);
    c->worstRequired = std::max(c->worstRequired, need);
    if (need > (unsigned int)c->d.cap) {
      allocateNeighbours(c, std::max(c->d.cap + 32, (int)std::ceil(need * 1.25)));
      ++c->neighbourGrowths;
      b->views[r].rebuild = 1;
      grew = true;
    }
  }
  return grew;
}

// One force evaluation of every replica (PME, bonded, then the lists and the
// direct space), forked and joined like the single-system evaluation.
void enqueueBatchForces(rmd_batch *b, int energies, int terms = RMD_TERM_ALL) {
  const Streams &st = b->streams;
  const bool fork = b->concurrent;
  hipStream_t pmeStream = fork ? st.side[0] : st.stream, bondedStream = fork ? st.side[1] : st.stream;
  if (fork) {
    HIP_CHECK(hipEventRecord(st.fork, st.stream));
    HIP_CHECK(hipStreamWaitEvent(st.side[0], st.fork, 0));
    HIP_CHECK(hipStreamWaitEvent(st.side[1], st.fork, 0));
  }
  if (terms & RMD_TERM_RECIPROCAL)
    launchPmeBatch(b->dViews, b->tables[BK_PLANES], b->tables[BK_XLINES], b->tables[BK_REPLICA],
                   b->tables[BK_ATOMS256], energies, b->ldsForward, b->ldsXlines, b->ldsGather, pmeStream);
  if (terms & (RMD_TERM_BONDED | RMD_TERM_EXCEPTIONS | RMD_TERM_RESTRAINT))
    launchBondedBatch(b->dViews, b->tables[BK_BONDED], b->alpha, terms, energies, bondedStream);
  launchRebuildBatch(b->dViews, b->tables[BK_REPLICA], b->replicas, b->searchQueue, b->searchGrid, b->allRect,
                     st.stream);
  if (terms & RMD_TERM_DIRECT)
    launchNonbondedBatch(b->dViews, b->tables[BK_NONBONDED], b->lanes, (float)b->alpha, b->rational, energies,
                         b->maxTypes, b->allRect, st.stream);
  if (fork) {
    for (int k = 0; k < 2; ++k) {
      HIP_CHECK(hipEventRecord(st.join[k], st.side[k]));
      HIP_CHECK(hipStreamWaitEvent(st.stream, st.join[k], 0));
    }
  }
  if (energies) launchMolecularTermBatch(b->dViews, b->tables[BK_MOLECULES], st.stream);
}

void enqueueBatchBlock(rmd_batch *b, int steps) {
  hipStream_t s = b->streams.stream;
  for (int k = 0; k < steps; ++k) {
    const bool last = k == steps - 1;
    launchIntegrateOpenBatch(b->dViews, b->tables[BK_WATER], b->tables[BK_OPEN], k > 0 ? 1 : 0, last ? 1 : 0,
                             b->ldsOpen, b->starMax, s);
    enqueueBatchForces(b, last ? 1 : 0);
  }
  launchReduceEnergiesBatch(b->dViews, b->replicas, 1, s);
  launchIntegrateCloseBatch(b->dViews, b->tables[BK_UNITS], b->ldsOpen, b->starMax, s);
}

void launchBatchBlock(rmd_batch *b, int steps) {
  hipStream_t s = b->streams.stream;
  if (!b->useGraphs) {
    enqueueBatchBlock(b, steps);
    HIP_CHECK(hipGetLastError());
    return;
  }
  auto it = b->graphs.find(steps);
  if (it == b->graphs.end()) {
    hipGraph_t graph;
    HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    enqueueBatchBlock(b, steps);
    HIP_CHECK(hipStreamEndCapture(s, &graph));
    hipGraphExec_t exec;
    HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    HIP_CHECK(hipGraphDestroy(graph));
    it = b->graphs.emplace(steps, exec).first;
  }
  HIP_CHECK(hipGraphLaunch(it->second, s));
}

// One evaluation of every replica at its current positions; the replicas
// whose view has `rebuild` set (or all, with forceAll) rebuild their lists.
// With energies the reduced terms land in hostTerms.
void batchEvaluate(rmd_batch *b, bool energies, bool forceAll, int terms = RMD_TERM_ALL) {
  hipStream_t s = b->streams.stream;
  for (int attempt = 0; attempt < 4; ++attempt) {
    uploadViews(b);
    launchClearRequiredBatch(b->scAll, b->replicas, s);
    launchDecideRebuildBatch(b->dViews, b->replicas, forceAll ? 1 : 0, s);
    launchZeroForcesBatch(b->dViews, b->tables[BK_ATOMS256], s);
    launchRefreshSlotsBatch(b->dViews, b->tables[BK_ATOMS256], s);
    enqueueBatchForces(b, energies ? 1 : 0, terms);
    if (energies) launchReduceEnergiesBatch(b->dViews, b->replicas, 0, s);
    HIP_CHECK(hipGetLastError());
    readBack(b, energies);
    if (growOverflowed(b)) continue;
    for (auto &v : b->views) v.rebuild = 0;
    for (rmd_ctx *c : b->reps) c->initialised = true;
    return;
  }
  throw Error(RMD_E_NEIGHBOUR_OVERFLOW, "the neighbour list kept overflowing after growth");
}

// A block of `steps` steps of every replica, with the single-system block's
// overflow contract: a list that overflowed anywhere re-runs the whole block
// from its saved start after growing.
void batchRunBlock(rmd_batch *b, int steps) {
  hipStream_t s = b->streams.stream;
  const int R = b->replicas;
  for (int attempt = 0; attempt < 4; ++attempt) {
    uploadViews(b);
    launchClearRequiredBatch(b->scAll, R, s);
    HIP_CHECK(hipMemcpyAsync(b->scSaved, b->scAll, R * sizeof(Scalars), hipMemcpyDeviceToDevice, s));
    launchSnapshotBatch(b->dViews, b->tables[BK_ATOMS256], SNAP_VELOCITIES, s);
    launchBatchBlock(b, steps);
    readBack(b, true);
    for (int r = 0; r < R; ++r) {
      const Scalars &h = b->hostScalars[r];
      rmd_ctx *c = b->reps[r];
      if (h.nonFinite || std::isnan(__builtin_bit_cast(float, h.maxDisp2Bits)))
        throw Error(RMD_E_NONFINITE, replicaLabel(r) + "a coordinate became non-finite in the block of " +
                                         std::to_string(steps) + " steps that ended at step " +
                                         std::to_string(c->step + steps));
      if (h.shakeFailed)
        throw Error(RMD_E_CONSTRAINT, replicaLabel(r) + "constraints did not converge in the iteration limit");
    }
    if (growOverflowed(b)) {
      launchSnapshotBatch(b->dViews, b->tables[BK_ATOMS256], SNAP_RESTORE | SNAP_VELOCITIES, s);
      HIP_CHECK(hipMemcpyAsync(b->scAll, b->scSaved, R * sizeof(Scalars), hipMemcpyDeviceToDevice, s));
      for (int r = 0; r < R; ++r)
        if (b->views[r].rebuild) setForceRebuildKernel<<<1, 1, 0, s>>>(b->reps[r]->d.sc);
      HIP_CHECK(hipStreamSynchronize(s));
      for (auto &v : b->views) v.rebuild = 0;
      continue;
    }
    for (int r = 0; r < R; ++r) {
      rmd_ctx *c = b->reps[r];
      c->step += steps;
      c->stepsRun += steps;
      c->kineticSampledAtStep = c->step;
      c->blockEndStep = c->step;
      c->stepsSinceBarostat += steps;
      std::memcpy(c->blockEndTerms, &b->hostTerms[(size_t)r * E_TERM_COUNT], sizeof c->blockEndTerms);
    }
    ++b->blocksRun;
    return;
  }
  throw Error(RMD_E_NEIGHBOUR_OVERFLOW, "the neighbour list kept overflowing after growth");
}

// The screen's features for a replica of a batch: the single system's
// (already per molecule and kT) with the first-order estimate per molecule
// too, so one network serves replicas of every size; its targets are the
// residuals per molecule and kT.
void batchFeatures(rmd_ctx *c, const double *t, double volume, double newVolume, float *f, double *estimate,
                   double *nkT) {
  screenFeatures(c, t, volume, newVolume, f, estimate);
  *nkT = std::max(1, c->sys.topology.moleculeCount) * BOLTZMANN * c->temperature;
  f[2] = (float)(*estimate / *nkT);
}

// The Monte Carlo volume move of every replica that is due one, with the
// single-system move's arithmetic, random numbers and bookkeeping
// (webgpu/simulation.js:995-1070): the trial boxes of all of them are
// evaluated together, then each replica accepts or goes back on its own.
// With the batch's learned screen, a move goes through the same two stages
// as the single system's (section "learned screen" of the single move): the
// network's prediction first, the exact evaluation only for the moves it
// lets through, and the delayed-acceptance ratio with the reverse move's
// prediction. All predictions of a round are one batched forward pass.
void batchBarostat(rmd_batch *b) {
  const int R = b->replicas;
  const int F = RMD_SURROGATE_INPUTS;
  hipStream_t s = b->streams.stream;
  ScreenState &sc = b->screen;
  const bool screenOn = sc.options.enabled && sc.net;
  const bool screening = screenOn && sc.phase == 1;
  struct Trial {
    bool active = false, candidate = false, moved = false;
    double volume = 0, newVolume = 0, kT = 0, work = 0, oldPotential = 0, dispersionChange = 0;
    double estimateX = 0, predictedX = 0, firstStage = 1.0, nkT = 1.0;
    double oldBox[9];
    double oldTerms[E_TERM_COUNT];
    float fx[RMD_SURROGATE_INPUTS];
  };
  std::vector<Trial> tr(R);
  bool any = false, stale = false;
  for (int r = 0; r < R; ++r) {
    const rmd_ctx *c = b->reps[r];
    if (!c->hasPressure || c->stepsSinceBarostat < c->barostatInterval) continue;
    tr[r].active = any = true;
    stale = stale || c->blockEndStep != c->step;
  }
  if (!any) return;
  if (stale) batchEvaluate(b, true, true);
  std::vector<int> candidates;
  for (int r = 0; r < R; ++r) {
    if (!tr[r].active) continue;
    rmd_ctx *c = b->reps[r];
    Trial &t = tr[r];
    c->stepsSinceBarostat = 0;
    const double volume = volumeOf(c->box);
    if (c->volumeScale < 0) c->volumeScale = 0.01 * volume;
    t.volume = volume;
    t.kT = BOLTZMANN * c->temperature;
    const double pressure = c->pressure * BAR_TO_KJ_PER_MOL_NM3;
    std::memcpy(t.oldTerms, c->blockEndStep == c->step ? c->blockEndTerms : &b->hostTerms[(size_t)r * E_TERM_COUNT],
                sizeof t.oldTerms);
    t.oldPotential = potentialFromTerms(c, t.oldTerms);
    std::memcpy(t.oldBox, c->box, sizeof t.oldBox);
    c->barostatAttempts += 1;
    c->windowAttempts += 1;
    const uint32_t bseed = c->seed ^ 0x5a17u;
    const double draw = uniformAt(bseed, (uint32_t)c->barostatAttempts, 0, 0);
    const double deltaVolume = c->volumeScale * 2 * (draw - 0.5);
    const double newVolume = volume + deltaVolume;
    if (!(newVolume > 0)) continue;
    const double sc3 = std::cbrt(newVolume / volume);
    const double smallest = sc3 * std::min({std::fabs(c->box[0]), std::fabs(c->box[4]), std::fabs(c->box[8])});
    if (c->sys.cutoff > 0.5 * smallest) {
      c->barostatRefusals += 1;
      continue;
    }
    t.newVolume = newVolume;
    t.work = pressure * deltaVolume - c->sys.topology.moleculeCount * t.kT * std::log(newVolume / volume);
    t.dispersionChange = dispersionAt(c, newVolume) - dispersionAt(c, volume);
    if (screenOn) batchFeatures(c, t.oldTerms, volume, newVolume, t.fx, &t.estimateX, &t.nkT);
    t.candidate = true;
    candidates.push_back(r);
  }
  // Stage one: the predicted weight alone; a move it turns down is never
  // evaluated and its replica is untouched.
  if (screening && !candidates.empty()) {
    std::vector<float> rows(candidates.size() * F);
    for (size_t i = 0; i < candidates.size(); ++i) std::memcpy(&rows[i * F], tr[candidates[i]].fx, F * sizeof(float));
    std::vector<double> pred(candidates.size());
    predictScreen(sc, rows.data(), (int)candidates.size(), pred.data());
    for (size_t i = 0; i < candidates.size(); ++i) {
      const int r = candidates[i];
      Trial &t = tr[r];
      const rmd_ctx *c = b->reps[r];
      sc.moves += 1;
      t.predictedX = t.dispersionChange + t.estimateX + t.nkT * pred[i];
      const double predictedWeight = t.predictedX + t.work;
      t.firstStage = predictedWeight <= 0 ? 1.0 : std::exp(-predictedWeight / t.kT);
      if (uniformAt(c->seed ^ 0x5a17u, (uint32_t)c->barostatAttempts, 2, 0) >= t.firstStage) {
        sc.screened += 1;
        t.candidate = false;
      }
    }
  }
  // Rebuild counts before the trial: a replica whose trial evaluation rebuilt
  // its list cannot keep it when the move is turned down.
  std::vector<unsigned int> rebuildsBefore(R);
  for (int r = 0; r < R; ++r) rebuildsBefore[r] = b->hostScalars[r].rebuilds;
  bool moved = false;
  for (int r = 0; r < R; ++r) {
    if (!tr[r].candidate) continue;
    rmd_ctx *c = b->reps[r];
    const double sc3 = std::cbrt(tr[r].newVolume / tr[r].volume);
    double nb[9];
    for (int k = 0; k < 9; ++k) nb[k] = c->box[k] * sc3;
    const bool replaced = updateBoxHost(c, nb);
    if (!fitsSmallCellBuild(c->d.n, c->plan.cellCount))
      throw Error(RMD_E_INVALID, replicaLabel(r) + "the trial box needs more cells than a batch replica supports");
    b->views[r].scale = (float)sc3;
    b->views[r].boxChanged = 1;
    b->views[r].rebuild = replaced ? 1 : 0;  // else the list is kept (rebuild references scaled, slack spent)
    tr[r].moved = moved = true;
  }
  if (moved) {
    uploadViews(b);
    launchSnapshotBatch(b->dViews, b->tables[BK_ATOMS256], SNAP_REBUILD, s);
    launchScaleBatch(b->dViews, b->tables[BK_MOLECULES], b->tables[BK_ATOMS256], s);
    uploadBoxes(b);
    launchInfluenceBatch(b->dViews, b->tables[BK_INFLUENCE], s);
    batchEvaluate(b, true, false);
  }
  // Stage two (or the plain test): the reverse move's features from the
  // evaluated trial state, predicted together.
  std::vector<int> evaluated;
  for (int r = 0; r < R; ++r)
    if (tr[r].moved) evaluated.push_back(r);
  std::vector<float> reverseRows(evaluated.size() * F);
  std::vector<double> reverseEstimate(evaluated.size()), reversePred(evaluated.size(), 0.0);
  if (screenOn) {
    for (size_t i = 0; i < evaluated.size(); ++i) {
      const int r = evaluated[i];
      double nkT = 1.0;
      batchFeatures(b->reps[r], &b->hostTerms[(size_t)r * E_TERM_COUNT], tr[r].newVolume, tr[r].volume,
                    &reverseRows[i * F], &reverseEstimate[i], &nkT);
    }
    if (screening && !evaluated.empty()) predictScreen(sc, reverseRows.data(), (int)evaluated.size(), reversePred.data());
  }
  bool restore = false;
  for (size_t i = 0; i < evaluated.size(); ++i) {
    const int r = evaluated[i];
    rmd_ctx *c = b->reps[r];
    const Trial &t = tr[r];
    const double *newTerms = &b->hostTerms[(size_t)r * E_TERM_COUNT];
    const double change = potentialFromTerms(c, newTerms) - t.oldPotential;
    const double weight = change + t.work;
    const uint32_t bseed = c->seed ^ 0x5a17u;
    const double u = uniformAt(bseed, (uint32_t)c->barostatAttempts, 1, 0);
    bool accept;
    if (screening) sc.exact += 1;
    if (screening) {
      const double predictedReverse = -t.dispersionChange + reverseEstimate[i] + t.nkT * reversePred[i] - t.work;
      const double reverseStage = predictedReverse <= 0 ? 1.0 : std::exp(-predictedReverse / t.kT);
      const double ratio = std::exp(-weight / t.kT) * reverseStage / t.firstStage;
      accept = ratio >= 1.0 || u < ratio;
      if (!accept) sc.secondStage += 1;
      sc.residual2 += (change - t.predictedX) * (change - t.predictedX);
      const double first = change - t.dispersionChange - t.estimateX;
      sc.estimate2 += first * first;
      sc.residualCount += 1;
    } else {
      accept = weight <= 0 || u < std::exp(-weight / t.kT);
      if (screenOn && sc.phase == 0) {
        // both directions of the move, per molecule and kT
        sc.x.insert(sc.x.end(), t.fx, t.fx + F);
        sc.t.push_back((float)((change - t.dispersionChange - t.estimateX) / t.nkT));
        sc.x.insert(sc.x.end(), &reverseRows[i * F], &reverseRows[i * F] + F);
        sc.t.push_back((float)((-change + t.dispersionChange - reverseEstimate[i]) / t.nkT));
      }
    }
    if (accept) {
      std::memcpy(c->blockEndTerms, newTerms, sizeof c->blockEndTerms);
      c->barostatAccepted += 1;
      c->windowAccepted += 1;
      if (screening) sc.accepted += 1;
    } else {
      // back to the old state exactly: positions, forces, rebuild references,
      // the box (copied); a list the trial rebuilt or replaced is rebuilt
      const bool replaced = updateBoxHost(c, t.oldBox);
      std::memcpy(c->blockEndTerms, t.oldTerms, sizeof c->blockEndTerms);
      b->views[r].restore = 1;
      b->views[r].boxChanged = 1;
      b->views[r].rebuild = (replaced || b->hostScalars[r].rebuilds != rebuildsBefore[r]) ? 1 : 0;
      restore = true;
    }
  }
  // The adaptive window of every move that got past the refusals (screened
  // out or evaluated); a refused move or a non-positive trial volume leaves
  // it alone, as in the single move.
  for (int r = 0; r < R; ++r) {
    if (!tr[r].active || !(tr[r].newVolume > 0)) continue;
    rmd_ctx *c = b->reps[r];
    c->blockEndStep = c->step;
    if (c->windowAttempts >= 10) {
      const double rate = (double)c->windowAccepted / c->windowAttempts;
      if (rate < 0.25)
        c->volumeScale /= 1.1;
      else if (rate > 0.75)
        c->volumeScale = std::min(c->volumeScale * 1.1, 0.3 * tr[r].volume);
      c->windowAttempts = c->windowAccepted = 0;
    }
  }
  if (restore) {
    uploadViews(b);
    launchSnapshotBatch(b->dViews, b->tables[BK_ATOMS256], SNAP_RESTORE | SNAP_FLAGGED | SNAP_REBUILD, s);
    uploadBoxes(b);
    launchInfluenceBatch(b->dViews, b->tables[BK_INFLUENCE], s);
    launchForceRebuildFlaggedBatch(b->dViews, R, s);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(s));
  }
  clearFlags(b);
  if (screenOn && sc.phase == 0 && (int64_t)sc.t.size() >= 2 * (int64_t)sc.options.training_moves)
    trainScreen(sc, b->device);
}

void destroyBatch(rmd_batch *b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->streams.stream) (void)hipStreamSynchronize(b->streams.stream);
  for (auto &g : b->graphs) (void)hipGraphExecDestroy(g.second);
  if (b->screen.net) rmd_surrogate_free(b->screen.net);
  for (rmd_ctx *c : b->reps) destroyContext(c);
  for (void *p : b->held) (void)hipFree(p);
  for (int k = 0; k < 2; ++k) {
    if (b->streams.side[k]) (void)hipStreamDestroy(b->streams.side[k]);
    if (b->streams.join[k]) (void)hipEventDestroy(b->streams.join[k]);
  }
  if (b->streams.fork) (void)hipEventDestroy(b->streams.fork);
  if (b->streams.stream) (void)hipStreamDestroy(b->streams.stream);
  delete b;
}

void createBatch(rmd_batch *b, const rmd_system *const *systems, int count, const rmd_options &o,
                 const uint32_t *seeds) {
  int devices = 0;
  if (hipGetDeviceCount(&devices) != hipSuccess || devices <= 0)
    throw Error(RMD_E_DEVICE, "no HIP device is available; this library runs on an MI355X (gfx950)");
  b->device = o.device;
  HIP_CHECK(hipSetDevice(b->device));
  Streams &st = b->streams;
  HIP_CHECK(hipStreamCreateWithFlags(&st.stream, hipStreamNonBlocking));
  for (int k = 0; k < 2; ++k) {
    HIP_CHECK(hipStreamCreateWithFlags(&st.side[k], hipStreamNonBlocking));
    HIP_CHECK(hipEventCreateWithFlags(&st.join[k], hipEventDisableTiming));
  }
  HIP_CHECK(hipEventCreateWithFlags(&st.fork, hipEventDisableTiming));
  b->replicas = count;
  // Lanes per atom: with hundreds of thousands of atoms in flight, fewer
  // lanes per atom and longer per-lane loops hide the gather latency better
  // than the single system's 64.
  b->lanes = o.lanes_per_atom > 0 ? o.lanes_per_atom : tunedShape(1, 0).lanes;
  if (b->lanes != 8 && b->lanes != 16 && b->lanes != 32 && b->lanes != 64)
    throw Error(RMD_E_INVALID, "a batch runs 8, 16, 32 or 64 lanes per atom");
  b->scAll = devAlloc<Scalars>(count, b->held);
  b->scSaved = devAlloc<Scalars>(count, b->held);
  b->boxAll = devAlloc<BoxF>(count, b->held);
  b->termsAll = devAlloc<double>((size_t)count * E_TERM_COUNT, b->held);
  b->dViews = devAlloc<BatchView>(count, b->held);
  b->hostScalars.resize(count);
  b->hostTerms.assign((size_t)count * E_TERM_COUNT, 0.0);
  b->views.assign(count, BatchView{});
  clearFlags(b);
  for (int r = 0; r < count; ++r) {
    if (!systems[r]) throw Error(RMD_E_INVALID, replicaLabel(r) + "null system");
    rmd_options ro = o;
    ro.seed = seeds ? seeds[r] : o.seed;
    ro.lanes_per_atom = b->lanes;
    rmd_ctx *c = new rmd_ctx();
    c->batchMember = true;
    b->reps.push_back(c);
    try {
      create(c, systems[r]->s, ro, &st);
    } catch (const Error &e) {
      throw Error(e.code, replicaLabel(r) + e.what());
    }
    // the member's scalars and box live in the batch's contiguous arrays
    c->d.sc = b->scAll + r;
    c->boxDev = b->boxAll + r;
    c->d.box = c->boxDev;
    writeBox(c);
    HIP_CHECK(hipMemsetAsync(c->d.sc, 0, sizeof(Scalars), st.stream));
    setForceRebuildKernel<<<1, 1, 0, st.stream>>>(c->d.sc);
    HIP_CHECK(hipStreamSynchronize(st.stream));
    const System &sys = systems[r]->s;
    if (!sys.positions.empty())
      setState(c, sys.positions.data(), sys.velocities.empty() ? nullptr : sys.velocities.data(), 0, 0);
    if (!c->p.planes)
      throw Error(RMD_E_INVALID, replicaLabel(r) + "its PME grid needs the line pipeline, which a batch does not run");
    if (!fitsSmallCellBuild(c->d.n, c->plan.cellCount))
      throw Error(RMD_E_INVALID, replicaLabel(r) + "a batch replica takes at most 16384 atoms and 4096 cells");
    if (r == 0)
      b->alpha = c->pme.alpha;
    else if (c->pme.alpha != b->alpha)
      throw Error(RMD_E_INVALID, replicaLabel(r) + "every replica of a batch needs the same Ewald alpha (the same "
                                                   "cutoff and tolerance)");
    b->maxTypes = std::max(b->maxTypes, c->d.types);
    b->allRect = b->allRect && c->d.rect;
    b->rational = b->rational && ewaldRationalFits(c->pme.alpha, c->sys.cutoff);
    b->ldsForward = std::max(b->ldsForward, pmePlaneForwardLds(c->p));
    b->ldsXlines = std::max(b->ldsXlines, pmeXLinesLds(c->p));
    b->ldsGather = std::max(b->ldsGather, pmePlaneGatherLds(c->p));
    b->ldsOpen = std::max(b->ldsOpen, c->ip.clusterLdsBytes);
    b->starMax = std::max(b->starMax, c->ip.starMax);
  }
  pmeBatchLdsLimits(b->ldsForward, b->ldsXlines, b->ldsGather);
  for (int k = 0; k < BK_COUNT; ++k) b->tables[k] = buildTasks(b, k);
  // The list search: at most 1024 workgroups per XCD (enough to fill the chip
  // when every replica rebuilds; a search block is four cells or four atoms),
  // fewer when the replicas hold fewer blocks.
  long long searchBlocks = 0;
  for (rmd_ctx *c : b->reps) searchBlocks += (c->d.n + 3) / 4;
  b->searchGrid = XCD_COUNT * (int)std::min<long long>(1024, (searchBlocks + XCD_COUNT - 1) / XCD_COUNT);
  b->searchQueue = devAlloc<int>(searchQueueInts(b->replicas), b->held);
  uploadViews(b);
}

}  // namespace

#define BATCH_GUARD(batch, ...)                               \
  do {                                                         \
    if (!(batch)) {                                            \
      rmd::setGlobalError("null batch");                       \
      return RMD_E_INVALID;                                    \
    }                                                          \
    try {                                                      \
      HIP_CHECK(hipSetDevice((batch)->device));                \
      __VA_ARGS__;                                             \
      return RMD_OK;                                           \
    } catch (const rmd::Error &e) {                            \
      (batch)->lastError = e.what();                           \
      rmd::setGlobalError(e.what());                           \
      return e.code;                                           \
    } catch (const std::exception &e) {                        \
      (batch)->lastError = e.what();                           \
      rmd::setGlobalError(e.what());                           \
      return RMD_E_INVALID;                                    \
    }                                                          \
  } while (0)

extern "C" {

int rmd_batch_create(const rmd_system *const *systems, int32_t count, const rmd_options *options,
                     const uint32_t *seeds, rmd_batch **out) {
  if (!systems || !out || count <= 0) {
    setGlobalError(count <= 0 ? "a batch needs at least one replica" : "null argument");
    return RMD_E_INVALID;
  }
  if (count > RMD_BATCH_MAX_REPLICAS) {
    setGlobalError("a batch holds at most " + std::to_string(RMD_BATCH_MAX_REPLICAS) +
                   " replicas (the list search stages two ints per replica in LDS)");
    return RMD_E_INVALID;
  }
  rmd_batch *b = new rmd_batch();
  try {
    rmd_options d;
    rmd_options_default(&d);
    createBatch(b, systems, count, options ? *options : d, seeds);
    *out = b;
    return RMD_OK;
  } catch (const rmd::Error &e) {
    setGlobalError(e.what());
    destroyBatch(b);
    return e.code;
  } catch (const std::exception &e) {
    setGlobalError(e.what());
    destroyBatch(b);
    return RMD_E_INVALID;
  }
}

void rmd_batch_destroy(rmd_batch *b) { destroyBatch(b); }

const char *rmd_batch_last_error(const rmd_batch *b) { return b ? b->lastError.c_str() : rmd_last_global_error(); }

rmd_ctx *rmd_batch_replica(rmd_batch *b, int32_t replica) {
  if (!b || replica < 0 || replica >= b->replicas) return nullptr;
  return b->reps[replica];
}

int rmd_batch_info_get(rmd_batch *b, rmd_batch_info *out) {
  BATCH_GUARD(b, {
    if (!out) throw Error(RMD_E_INVALID, "null output");
    std::memset(out, 0, sizeof *out);
    out->replicas = b->replicas;
    for (rmd_ctx *c : b->reps) out->atoms += c->d.n;
    out->blocks_run = b->blocksRun;
    out->step_workgroups = b->tables[BK_WATER].blocks + b->tables[BK_OPEN].blocks + b->searchGrid + b->tables[BK_NONBONDED].blocks +
                           b->tables[BK_BONDED].blocks + 2 * b->tables[BK_PLANES].blocks +
                           b->tables[BK_XLINES].blocks + b->tables[BK_REPLICA].blocks;
    out->nonbonded_workgroups = b->tables[BK_NONBONDED].blocks;
  });
}

int rmd_batch_initialise(rmd_batch *b) {
  BATCH_GUARD(b, { batchEvaluate(b, false, true); });
}

int rmd_batch_evaluate(rmd_batch *b, int32_t terms) {
  BATCH_GUARD(b, {
    batchEvaluate(b, true, true, terms ? terms : RMD_TERM_ALL);
    // the replicas' forces now hold only these terms: the next run starts
    // from a full evaluation
    for (rmd_ctx *c : b->reps) {
      c->initialised = false;
      c->blockEndStep = -1;
    }
  });
}

int rmd_batch_run(rmd_batch *b, int64_t steps, int32_t block, int64_t *stepOut) {
  BATCH_GUARD(b, {
    bool ready = true, pressure = false;
    for (rmd_ctx *c : b->reps) {
      ready = ready && c->initialised;
      pressure = pressure || c->hasPressure;
    }
    if (!ready) batchEvaluate(b, false, true);
    const int blockSize = block > 0 ? block : (pressure ? b->reps[0]->barostatInterval : 25);
    int64_t remaining = steps;
    while (remaining > 0) {
      if (b->stop.exchange(0)) throw Error(RMD_E_STOPPED, "the batch was stopped between blocks");
      batchBarostat(b);
      const int n = (int)std::min<int64_t>(blockSize, remaining);
      batchRunBlock(b, n);
      remaining -= n;
    }
    if (stepOut) *stepOut = b->reps[0]->step;
  });
}

int rmd_batch_screen_configure(rmd_batch *b, const rmd_screen_options *options) {
  BATCH_GUARD(b, {
    rmd_screen_options o;
    rmd_screen_options_default(&o);
    if (options) o = *options;
    if (o.enabled && (o.training_moves <= 0 || o.training_steps <= 0 || !(o.learning_rate > 0)))
      throw Error(RMD_E_INVALID, "the screen needs training_moves, training_steps and learning_rate > 0");
    auto &sc = b->screen;
    if (sc.net) rmd_surrogate_free(sc.net);
    sc = ScreenState();
    sc.options = o;
    if (o.enabled) {
      rmd_surrogate *net = nullptr;
      const int st = rmd_surrogate_create(b->device, o.seed, &net);
      if (st != RMD_OK) throw Error(st, std::string("screen network: ") + rmd_last_global_error());
      HIP_CHECK(hipSetDevice(b->device));
      sc.net = net;
      // each replica's features measure its volume against its own
      for (rmd_ctx *c : b->reps) c->screen.refVolume = volumeOf(c->box);
    }
  });
}

int rmd_batch_screen_stats_get(rmd_batch *b, rmd_screen_stats *out) {
  BATCH_GUARD(b, {
    if (!out) throw Error(RMD_E_INVALID, "null output");
    const auto &sc = b->screen;
    std::memset(out, 0, sizeof *out);
    out->phase = sc.phase;
    out->samples = (int64_t)sc.t.size();
    out->moves = sc.moves;
    out->exact_evaluations = sc.exact;
    out->screened_out = sc.screened;
    out->second_stage_rejections = sc.secondStage;
    out->accepted = sc.accepted;
    out->training_loss = sc.trainingLoss;
    out->residual_rms = sc.residualCount ? std::sqrt(sc.residual2 / sc.residualCount) : 0.0;
    out->estimate_rms = sc.residualCount ? std::sqrt(sc.estimate2 / sc.residualCount) : 0.0;
  });
}

int rmd_batch_request_stop(rmd_batch *b) {
  if (!b) return RMD_E_INVALID;
  b->stop.store(1);
  return RMD_OK;
}

int rmd_batch_set_graphs(rmd_batch *b, int32_t enabled) {
  BATCH_GUARD(b, { b->useGraphs = enabled != 0; });
}

int rmd_batch_set_concurrent_forces(rmd_batch *b, int32_t enabled) {
  BATCH_GUARD(b, {
    b->concurrent = enabled != 0;
    for (auto &g : b->graphs) (void)hipGraphExecDestroy(g.second);
    b->graphs.clear();
  });
}

// Average duration of the batched nonbonded launch (every replica's direct
// space in one launch), timed with events on the batch's stream.
int rmd_batch_profile_nonbonded(rmd_batch *b, int32_t reps, double *avg_ms) {
  BATCH_GUARD(b, {
    if (reps <= 0 || !avg_ms) throw Error(RMD_E_INVALID, "reps must be positive");
    hipStream_t s = b->streams.stream;
    uploadViews(b);
    launchNonbondedBatch(b->dViews, b->tables[BK_NONBONDED], b->lanes, (float)b->alpha, b->rational, 0, b->maxTypes,
                         b->allRect, s);
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    HIP_CHECK(hipEventRecord(e0, s));
    for (int k = 0; k < reps; ++k)
      launchNonbondedBatch(b->dViews, b->tables[BK_NONBONDED], b->lanes, (float)b->alpha, b->rational, 0, b->maxTypes,
                         b->allRect, s);
    HIP_CHECK(hipEventRecord(e1, s));
    HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    *avg_ms = ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // the timed launches added forces on top of the state's: evaluate again
    batchEvaluate(b, false, false);
  });
}

}  // extern "C"
