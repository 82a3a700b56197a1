// Host-side launchers for every kernel. Each kernel file defines its own
// launchers; the context calls only these.
#pragma once

#include <hip/hip_runtime.h>

#include "device_state.h"

namespace rmd {

// ---- cell list and neighbour list (neighbours.hip) ----
// The five rebuild stages are launched every step and return at once unless
// the integrator's last block (or the host) asked for a rebuild; with
// `cellsBuilt` a small system's cell build is left out (the step's integrator
// builds its cell list: IntegratorParams::cellsInIntegrator), with `searched`
// the list search (the joined force launch runs it).
void launchRebuild(const Dev &d, hipStream_t s, bool cellsBuilt = false, bool searched = false);
void launchForceRebuildFlag(const Dev &d, hipStream_t s);

// ---- nonbonded direct space (nonbonded.hip) ----
// lanes: lanes per atom (1, 2, 4, ..., 64). energies != 0 fills the
// E_LJ_DIRECT / E_COUL_DIRECT slot rows and returns the slot count used.
// rational: the screened Coulomb force from the rational fit (alpha * cutoff
// within its range, ewaldRationalFits), else from erfc.
bool ewaldRationalFits(double alpha, double cutoff);
int launchNonbonded(const Dev &d, float alpha, int rational, int lanes, int energies, hipStream_t s);

// ---- bonded, exceptions, restraints (bonded.hip) ----
int launchBonded(const Dev &d, double alpha, int terms, int energies, hipStream_t s, int *slotsUsed);

// ---- PME (pme.hip) ----
struct PmeDev {
  int K[3];
  int K2h;                  // K[2]/2 + 1
  long long *qgrid;         // K0*K1*K2 fixed-point charges
  float2 *cgrid;            // K0*K1*K2h complex half spectrum
  float *rgrid;             // K0*K1*K2 real potential
  float *influence;         // K0*K1*K2h: pref * kernel(m) (0 at m = 0)
  float2 *twiddle[3];       // per axis, exp(+2 pi i k / K)
  int radices[3][8];        // factorisation per axis (0-terminated)
  double piOverAlpha2;      // pi^2 / alpha^2 (the scaling derivative of the reciprocal energy)
  int planeSpan;            // max(ceil(K1/2) * K2, K2h * K1): one plane's complex work buffer
  int planeRegion;          // float2 slots of the second buffer: max(planeSpan, the int64 plane)
  int planes;               // 1: plane pipeline, 0: line pipeline
  // The plane pipeline's atoms by the base index of their x support (K0
  // bins): a plane reads the ORDER bins that reach it instead of every atom.
  int binned;               // 1: the bins are built each evaluation (replica batches, large systems)
  int *binStart;            // K0 + 1 offsets into binAtoms
  int2 *atomBin;            // per atom: (bin or -1 for no charge, rank in the bin)
  int *binAtoms;            // atoms grouped by bin
  // A large single system builds its bins in three launches over groups of
  // BIN_GROUP_ATOMS atoms (binGroups > 0): per-group counts, their offsets,
  // the scatter. 0: one workgroup builds them (batch members).
  int binGroups;
  int *binTable;            // K0 x binGroups counts, then offsets within each bin
  // A small single system's force-only evaluations fold the x lines into the
  // plane gathers: the forward x transform, the influence function and the
  // inverse x transform of every (y, z) column are one circular convolution
  // along x with the kernel
  //   X(l, c) = sum_k 2 g(k, c) exp(-2 pi i k l / K0),  l in [0, K0),
  // g the influence, so gather workgroup i forms column c of its plane as
  // sum_x X(i - x, c) F(x, c) from the forward planes F - one launch and
  // its dependency fewer on the PME chain. Evaluations with energies run the
  // x lines (the energy and its scaling derivative are sums over the full
  // spectrum).
  int xFolded;              // 1: force-only evaluations fold the x lines into the gathers
  int xReal;                // 1: X is real (a rectangular box, where g is even in k)
  float *xKernel;           // X as K0 x (K1 * K2h) floats (xReal) or float2s; null when not allocated
};
// Grids up to this many x planes may fold their x lines into the gathers
// (each gather workgroup reads all K0 forward planes and K0 kernel rows).
constexpr int X_FOLD_MAX_K = 32;
enum { PME_AUTO = 0, PME_PLANES = 1, PME_LINES = 2, PME_PLANES_SCAN = 3, PME_PLANES_FOLDED = 4 };
// Chooses the pipeline (PME_AUTO: planes when a plane fits in maxLds bytes of
// LDS). Returns false when PME_PLANES was asked for and does not fit.
bool pmeConfigure(PmeDev &p, int mode, size_t maxLds);
// Whether a plane-pipeline system can bin its atoms in one workgroup.
constexpr int BIN_MAX_ATOMS = 16384;
bool pmeBinsFit(const PmeDev &p, int atoms);
// Single systems of at least this many atoms bin their plane-pipeline atoms
// (smaller ones scan every atom per plane: one launch fewer on a short
// critical path), in groups of BIN_GROUP_ATOMS.
constexpr int BIN_SINGLE_MIN_ATOMS = 8192;
constexpr int BIN_GROUP_ATOMS = 1024;
bool pmeBinsFitWide(const PmeDev &p);
// The device box set from its value and the influence function computed from
// it, in one launch (then the folded x lines' kernel where it is kept).
void pmeSetBoxAndInfluence(const PmeDev &p, const BoxF &box, BoxF *dst, const double *moduli[3], double alpha,
                           hipStream_t s);
// The same from a box in device memory.
void pmeSetBoxAndInfluenceFrom(const PmeDev &p, const BoxF *box, BoxF *dst, const double *moduli[3], double alpha,
                               hipStream_t s);
int launchPme(const Dev &d, const PmeDev &p, int energies, hipStream_t s, int *slotsUsed);
// A small system's direct space, bonded terms and PME forward planes in one
// launch (where forcesJoin allows), with the energy slots each part used;
// launchPmeAfterForward then runs the rest of the plane pipeline and returns
// its slots. With `search` (the one-wave-per-atom list search) a rebuild's
// list search runs there too: each direct-space workgroup searches its own
// rows before it computes them, and launchRebuild leaves its search out.
bool forcesJoin(const Dev &d, const PmeDev &p, int lanes);
void launchForcesJoined(const Dev &d, const PmeDev &p, float alpha, int rational, int lanes, int energies,
                        double alphaD, int terms, hipStream_t s, int *pairSlots, int *bondedSlots, int search = 0);
int launchPmeAfterForward(const Dev &d, const PmeDev &p, int energies, hipStream_t s);
void launchPmeStage(const Dev &d, const PmeDev &p, hipStream_t s, int stage);
bool pmeTransformsFolded(const PmeDev &p);
// Debug entry: forward R2C 3D FFT of a real grid into the half spectrum.
void launchFftForward(const Dev &d, const PmeDev &p, const float *real, float2 *out, hipStream_t s);

// ---- integrator and reductions (integrate.hip) ----
// Unit kinds: a free atom, a rigid water (SETTLE), a general constraint
// cluster (iterative SHAKE/RATTLE), padding, and a star cluster (one centre
// atom in every constraint, at most STAR_MAX constraints: the shape of every
// H-bond cluster) solved in registers. A cluster unit is (UNIT_CLUSTER,
// cluster, 1 = global scratch / 0 = LDS, LDS slot offset); a star unit is
// (UNIT_STAR, cluster, constraint count, 0) with the centre first in the
// cluster's atoms and constraint k joining atom k + 1 to the centre.
enum { UNIT_FREE = 0, UNIT_WATER = 1, UNIT_CLUSTER = 2, UNIT_NONE = 3, UNIT_STAR = 4 };
constexpr int STAR_MAX = 4;
constexpr int INTEGRATE_BLOCK = 256;
constexpr int CLUSTER_SLOT = 13;        // doubles per cluster atom: x, ref, v, corr, inverse mass
constexpr int CLUSTER_LDS_SLOTS = 512;  // cluster atom slots per block held in LDS
struct IntegratorParams {
  double halfStep, decay, noiseScale, kT, tolerance;
  double dispHalf;  // the displacement since the last rebuild that triggers one: half the skin (less a margin)
  unsigned int seed;
  int maxIterations;
  int suppliedNoise;
  int clusterLdsBytes;  // dynamic LDS per integrator block
  int starMax;          // the most constraints of any star unit (the star solver's width: 3 or 4)
  // 1: the last workgroup of the step's integrator, which decides the
  // rebuild, also builds the cell list when one is due (a small single system
  // in one integrator launch whose plan has at most FEW_MAX_CELLS cells), and
  // the step launches no cell build of its own (integratorBuildsCells)
  int cellsInIntegrator;
};
// closing: apply the previous step's closing kick + RATTLE first.
// early: record the in-step kinetic energy (E_KINETIC_EARLY).
// The unit table's workgroups: the leading water blocks (their own kernel),
// then the general blocks (at least one: its last workgroup decides the
// rebuild); the in-step kinetic slots of a step cover both.
inline int unitBlockCount(const Dev &d) { return (d.nUnits + INTEGRATE_BLOCK - 1) / INTEGRATE_BLOCK; }
inline int generalBlockCount(const Dev &d) {
  const int rest = unitBlockCount(d) - d.waterBlocks;
  return rest > 1 ? rest : 1;
}
inline int openSlotCount(const Dev &d) { return d.waterBlocks + generalBlockCount(d); }
// A single system runs its water blocks in their own kernel when there are
// many (the water path's occupancy), else one kernel for every unit (one
// launch, and the general units' serial solves overlap the waters). A batch
// always splits.
constexpr int SPLIT_WATER_BLOCKS = 32;
inline bool splitWater(const Dev &d) { return d.waterBlocks >= SPLIT_WATER_BLOCKS; }
inline int singleOpenSlotCount(const Dev &d) { return splitWater(d) ? openSlotCount(d) : unitBlockCount(d); }
int launchIntegrateOpen(const Dev &d, const IntegratorParams &ip, int closing, int early, hipStream_t s);
// Whether a single system's integrator can build its cell list (one
// integrator launch, the one-workgroup build, a plan of few cells).
bool integratorBuildsCells(const Dev &d, int cellCount);
int launchIntegrateClose(const Dev &d, const IntegratorParams &ip, int kinetic, hipStream_t s);
int launchKinetic(const Dev &d, int term, hipStream_t s);
// Sums each term's slot row in a fixed order; the counts travel by value so a
// captured graph keeps them.
struct SlotCounts {
  int c[E_TERM_COUNT];
};
void launchReduceEnergies(const Dev &d, const SlotCounts &counts, double *out, hipStream_t s);
// Barostat: mass-weighted molecule centroids and the scaling about them. The
// rebuild references are scaled with the positions, and the move's share of
// the skin is added to Scalars::slack, so the neighbour list stays valid
// without a rebuild.
void launchScaleMolecules(const Dev &d, const int *molStart, const int *molAtoms, int nMol, float4 *centroids,
                          const int *molOfAtom, float scale, hipStream_t s);
// The molecular-scaling correction, sum_i F_i . (r_i - c_mol(i)), into the
// E_DSCALE_MOLECULAR slot row (after every force kernel); returns the rows used.
int launchMolecularTerm(const Dev &d, const int *molStart, const int *molAtoms, int nMol, hipStream_t s);
// Minimiser helpers.
void launchMaxForce(const Dev &d, unsigned int *out, hipStream_t s);
void launchNoise(int n, unsigned int seed, unsigned int step, float *uniforms, float *normals, hipStream_t s);
void launchDisplace(const Dev &d, float scale, float cap, hipStream_t s);
void launchConstrainPositions(const Dev &d, const float4 *reference, const IntegratorParams &ip, hipStream_t s);

// ---- replica batches ----------------------------------------------------------
// A batch runs R independent systems, each a complete single-system context
// with buffers of its own (atom indices stay local to the replica). A batched
// launch covers every replica's workgroups at once: a task table maps each
// workgroup to (replica, the workgroup's index within that replica's own
// grid), so the kernel bodies are the single-system ones. A task with
// replica -1 is padding and returns at once.
struct BatchView {
  Dev d;
  PmeDev p;
  IntegratorParams ip;
  int openBlocks;        // the integrator's general blocks for this replica (its last-block ticket)
  int openSlots;         // in-step kinetic slots (water and general blocks)
  int rebuild;           // 1: this replica's lists are rebuilt at the coming evaluation
  float scale;           // the barostat's length scale for this replica (1: no move)
  int restore;           // 1: the barostat puts this replica back
  int boxChanged;        // 1: recompute this replica's PME influence function
  SlotCounts counts;     // energy slot rows a block's last step fills
  double *terms;         // E_TERM_COUNT reduced energies
  const int *molStart, *molAtoms, *molOfAtom;
  int nMol;
  float4 *centroids, *posSaved, *velSaved, *rebuildSaved;
  long long *forceSaved;
  const double *moduli[3];
  double alpha;
};

struct TaskTable {
  int2 *task = nullptr;  // (replica, block within the replica), device memory
  int blocks = 0;        // grid size (padding included)
};

// MI355X's XCDs: workgroup b of a launch runs on XCD b % XCD_COUNT.
constexpr int XCD_COUNT = 8;

enum BatchKernel {
  BK_OPEN = 0,      // integrator general blocks (after the water blocks)
  BK_BONDED,        // bonded, exception and restraint terms
  BK_PLANES,        // PME plane kernels (one workgroup per x plane)
  BK_XLINES,        // PME x lines
  BK_REPLICA,       // one workgroup per replica (cell lists)
  BK_ATOMS256,      // 256 atoms per workgroup (zeroing, scaling, copies)
  BK_MOLECULES,     // 256 molecules per workgroup (centroids)
  BK_INFLUENCE,     // 256 half-spectrum points per workgroup
  BK_NONBONDED,     // the direct space at the batch's lanes per atom
  BK_WATER,         // integrator water blocks
  BK_UNITS,         // every integrator unit block (the closing kick)
  BK_COUNT
};

// Batched launchers. `lds` is the largest dynamic LDS any replica needs.
void launchIntegrateOpenBatch(const BatchView *v, const TaskTable &water, const TaskTable &general, int closing,
                              int early, int lds, int starMax, hipStream_t s);
void launchIntegrateCloseBatch(const BatchView *v, const TaskTable &t, int lds, int starMax, hipStream_t s);
// withEarly: a block's last step (the in-step kinetic energy is summed too).
void launchReduceEnergiesBatch(const BatchView *v, int replicas, int withEarly, hipStream_t s);
void launchZeroForcesBatch(const BatchView *v, const TaskTable &t, hipStream_t s);
void launchDecideRebuildBatch(const BatchView *v, int replicas, int force, hipStream_t s);
void launchScaleBatch(const BatchView *v, const TaskTable &molecules, const TaskTable &atoms, hipStream_t s);
void launchMolecularTermBatch(const BatchView *v, const TaskTable &molecules, hipStream_t s);
// Saves (or restores) positions and forces, with SNAP_VELOCITIES the
// velocities too, of every replica (with SNAP_FLAGGED, of those whose
// `restore` is set).
enum { SNAP_RESTORE = 1, SNAP_VELOCITIES = 2, SNAP_FLAGGED = 4, SNAP_REBUILD = 8, SNAP_SCALARS = 16 };
void launchSnapshotBatch(const BatchView *v, const TaskTable &atoms, int mode, hipStream_t s);
// A single system's save (or restore, SNAP_RESTORE) of positions and forces,
// with SNAP_VELOCITIES the velocities, SNAP_REBUILD the rebuild positions and
// SNAP_SCALARS the Scalars too (a save with SNAP_SCALARS clears
// Scalars::requiredMax first), in one launch.
struct SnapshotBuffers {
  float4 *pos, *vel, *rebuild;
  long long *force;
  Scalars *sc;
  unsigned int *rebuilds;  // saving, if set: Scalars::rebuilds stored here
};
void launchSnapshot(const Dev &d, const SnapshotBuffers &saved, int mode, hipStream_t s);
// A volume move's trial appended to a block: what the host proposes before
// the block runs (the trial box and the scale factor), and the list's rebuild
// count when the trial starts.
struct TrialParams {
  BoxF box;
  float scale;
  unsigned int rebuildsBefore;
};
// The barostat's scaling with the factor read from device memory.
void launchScaleMoleculesFrom(const Dev &d, const int *molStart, const int *molAtoms, int nMol, float4 *centroids,
                              const int *molOfAtom, const float *scale, hipStream_t s);
// Sets Scalars::forceRebuild of the replicas whose `restore` and `rebuild` are set.
void launchForceRebuildFlaggedBatch(const BatchView *v, int replicas, hipStream_t s);
// Clears Scalars::requiredMax of `replicas` consecutive scalar blocks.
void launchClearRequiredBatch(Scalars *sc, int replicas, hipStream_t s);
// Whether a system's cell list fits the single-workgroup build (a batch needs it).
bool fitsSmallCellBuild(int atoms, int cells);
// Whether the list search runs one wave per cell: small cells (a few dozen
// atoms at most on average), and a cell count the per-atom launch's blocks
// cover at four cells each.
bool cellSearchFits(int atoms, int cells);
// The search path for a search mode (0 by occupancy, 1 per atom, 2 per cell
// where the launch covers the cells).
bool searchByCell(int mode, int atoms, int cells);
// Refresh every replica's cell-ordered position copy from its positions
// (evaluations outside the step loop, where positions may have been set or
// restored).
void launchRefreshSlotsBatch(const BatchView *v, const TaskTable &atoms256, hipStream_t s);
// The cell lists of the replicas due a rebuild, then their neighbour lists;
// `queue` holds searchQueueInts(replicas) ints of device memory, `grid` is a
// multiple of XCD_COUNT.
void launchRebuildBatch(const BatchView *v, const TaskTable &replicaTable, int replicas, int *queue, int grid,
                        int rect, hipStream_t s);
int searchQueueInts(int replicas);
// lanes: lanes per atom (8, 16, 32 or 64), the table's grid at that count.
void launchNonbondedBatch(const BatchView *v, const TaskTable &t, int lanes, float alpha, int rational, int energies,
                          int types, int rect, hipStream_t s);
void launchBondedBatch(const BatchView *v, const TaskTable &t, double alpha, int terms, int energies, hipStream_t s);
void launchPmeBatch(const BatchView *v, const TaskTable &planes, const TaskTable &xlines, const TaskTable &replicas,
                    const TaskTable &atoms256, int energies, size_t forwardLds, size_t xlinesLds, size_t gatherLds,
                    hipStream_t s);
void launchInfluenceBatch(const BatchView *v, const TaskTable &t, hipStream_t s);
void pmeBatchLdsLimits(size_t forwardLds, size_t xlinesLds, size_t gatherLds);
// Per-replica grid sizes and LDS of the PME plane pipeline (for the task tables).
int pmeXLineBlocks(const PmeDev &p);
size_t pmePlaneForwardLds(const PmeDev &p);
size_t pmeXLinesLds(const PmeDev &p);
size_t pmePlaneGatherLds(const PmeDev &p);
int bondedBlocks(const Dev &d);

}  // namespace rmd
