// Host-side data structures behind the C ABI: the parsed topology, the built
// system and the derived plans. Pure C++ (no HIP); compiled into the same
// shared library as the kernels.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/rainmd.h"

namespace rmd {

// Unit constants (shared/units.js:22-53).
constexpr double KCAL_TO_KJ = 4.184;
constexpr double NM_PER_ANGSTROM = 0.1;
constexpr double AMBER_CHARGE_SCALE = 18.2223;
constexpr double AMBER_VELOCITY_SCALE = 20.455;
constexpr double ONE_4PI_EPS0 = 138.935456;
constexpr double BOLTZMANN = 0.008314462618;
constexpr double BAR_TO_KJ_PER_MOL_NM3 = 0.06022140857;
constexpr double BOND_K_SCALE = 2.0 * KCAL_TO_KJ / (NM_PER_ANGSTROM * NM_PER_ANGSTROM);
constexpr double ANGLE_K_SCALE = 2.0 * KCAL_TO_KJ;
constexpr int PME_ORDER = 5;

// An error carrying one of the RMD_E_* codes.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// The message rmd_last_global_error() returns (per thread).
void setGlobalError(const std::string &m);

struct Topology {
  int atomCount = 0, typeCount = 0, residueCount = 0;
  std::vector<std::string> atomNames, residueNames, atomTypeNames;
  std::vector<int> residueFirstAtom, residueOfAtom;
  std::vector<double> charges, masses;
  std::vector<int> atomicNumbers, atomTypes;
  std::vector<double> ljA, ljB;  // T*T internal units
  struct {
    std::vector<int> a, b;
    std::vector<double> k, r0;
    std::vector<uint8_t> containsHydrogen;
  } bonds;
  struct {
    std::vector<int> a, b, c;
    std::vector<double> k, theta0;
  } angles;
  struct {
    std::vector<int> a, b, c, d;
    std::vector<double> k, periodicity, phase;
    std::vector<uint8_t> isImproper;
  } dihedrals;
  struct {
    std::vector<int> a, b;
    std::vector<double> scee, scnb;
  } pairs14;
  std::vector<int> exclusionStart, exclusionPartners;  // one-sided, as in the file
  bool isPeriodic = false;
  bool hasBox = false;
  double boxLengths[3] = {0, 0, 0}, boxAngles[3] = {0, 0, 0};
  int moleculeCount = 0;
  std::vector<int> moleculeOfAtom;
  std::vector<std::string> warnings;
};

struct Coordinates {
  int atomCount = 0;
  std::vector<double> positions, velocities;
  bool hasVelocities = false, hasBox = false;
  double boxLengths[3] = {0, 0, 0}, boxAngles[3] = {0, 0, 0};
  double time = 0.0;
  bool hasTime = false;
};

struct System {
  Topology topology;
  int constraintMode = RMD_CONSTRAINTS_HBONDS;
  bool rigidWater = true;
  double cutoff = 1.0;
  bool isPeriodic = false;
  double box[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  struct {
    std::vector<int> a, b;
    std::vector<double> k, r0;
  } harmonicBonds;
  struct {
    std::vector<int> a, b;
    std::vector<double> distance;
  } constraints;
  struct {
    std::vector<int> a, b;
    std::vector<double> chargeProduct, ljA, ljB;
    std::vector<uint8_t> isScaled;
  } exceptions;
  std::vector<int> moleculeStart, moleculeAtoms;
  int degreesOfFreedom = 0;
  // Symmetric sorted exclusion CSR (shared/topology/cells.js:201-237).
  std::vector<int> exclusionStart, exclusionPartners;
  std::vector<double> positions, velocities;
  int waterCount = 0;
};

// Parsers (shared/amber/*).
Topology parsePrmtop(const std::string &text);
Coordinates parseInpcrd(const std::string &text, int expectedAtomCount);
void boxVectorsFromLengthsAndAngles(const double lengths[3], const double anglesDeg[3],
                                    double out[9]);

// buildSystem (shared/topology/system.js:40-236).
System buildSystem(const Topology &topology, const Coordinates *coordinates,
                   const rmd_build_options &options);

// Seeded solvated builder (see rainmd.h).
System buildSolvated(const Topology &templ, const Coordinates &solute,
                     const rmd_solvated_options &options);

// PME setup (shared/pme/parameters.js).
double alphaFromTolerance(double cutoff, double tolerance);
bool isTransformableSize(int value);
int nextTransformableSize(int minimum);
void bsplineWeights(double offset, int order, double *weights, double *derivatives);
std::vector<double> bsplineModuli(int gridSize, int order);

// Cell plan (shared/topology/cells.js).
double skinFor(const System &system, double timestep, double temperature, const double *box);
struct CellPlan {
  double skin = 0, cutoff = 0, listCutoff = 0;
  int counts[3] = {1, 1, 1};
  int cellCount = 1;
  // Cells are at least listCutoff / reach wide and the search visits reach
  // cells either side (reach 1 is the reference's plan).
  int reach = 1;
  int maxNeighbours = 32;
  double estimatedNeighbours = 0;
};
// The neighbour search's limits: the widest stencil reach, and the cell and
// atom counts of the single-workgroup cell build.
constexpr int CELL_REACH_MAX = 3;
constexpr int CELL_LIMIT_SMALL = 4096;
constexpr int CELL_ATOMS_SMALL = 16384;
// reachMax limits the plan's reach (1 gives the reference's plan; values
// outside 1..CELL_REACH_MAX mean CELL_REACH_MAX).
CellPlan planCellList(const System &system, double skin, double slack, int maxNeighbours,
                      const double *box, int reachMax = 0);

// The autotuner's shapes (tuned.h) for a single system (kind 0) or a batch
// member (kind 1) of `atoms` atoms: the row of that kind with the largest
// minAtoms not above it.
struct TunedShape;
const TunedShape &tunedShape(int kind, int atoms);

// Dispersion (shared/dispersion.js:37-82) and Ewald self energy
// (cpu/forces/pme.js:270-286).
struct DispersionCoefficients {
  double a = 0, b = 0;
};
DispersionCoefficients dispersionCoefficients(const System &system);
double dispersionEnergy(const DispersionCoefficients &c, double volume, double cutoff);
void ewaldSelfEnergy(const System &system, double alpha, double volume, double *self,
                     double *net);

// The uniform/Gaussian counter hash (shared/random.js:24-87).
uint32_t hashMix(uint32_t value);
double uniformAt(uint32_t seed, uint32_t a, uint32_t b, uint32_t c);
void gaussianTriple(uint32_t seed, uint32_t atom, uint32_t step, double out[3]);

// Geometry helpers (cpu/forces/geometry.js).
void reciprocalBoxVectors(const double box[9], double out[9]);
void faceWidths(const double box[9], double out[3]);
inline double boxVolume(const double box[9]) { return box[0] * box[4] * box[8]; }

}  // namespace rmd

// The opaque C handles (see rainmd.h).
struct rmd_topology {
  rmd::Topology t;
};
struct rmd_coordinates {
  rmd::Coordinates c;
};
struct rmd_system {
  rmd::System s;
};
