// The device-resident state every kernel reads, passed by value as a kernel
// argument (a few hundred bytes of pointers and sizes).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rainmd.h"
#include "device_common.h"

namespace rmd {

// Energy accumulation: each kernel block that computes energies writes one
// FP64 partial into its own slot of a term's row; a reduction kernel sums
// each row in a fixed order, so energies are bit-reproducible.
enum EnergyTerm {
  E_LJ_DIRECT = 0,
  E_COUL_DIRECT,
  E_LJ_EXCEPTION,
  E_COUL_EXCEPTION,
  E_BOND,
  E_ANGLE,
  E_TORSION,
  E_RESTRAINT,
  E_RECIPROCAL,
  E_KINETIC,
  E_KINETIC_EARLY,
  // d(energy)/d(ln lambda) under a uniform scaling of every coordinate by
  // lambda: direct-space pairs (sum r dU/dr) and the reciprocal sum
  // (sum_k E_k (2 pi^2 m^2 / alpha^2 - 1)). Filled with the energies; the
  // barostat's learned screen reads them as features.
  E_DSCALE_DIRECT,
  E_DSCALE_RECIPROCAL,
  // The same for the bonded and exception terms (all within molecules), and
  // sum_i F_i . (r_i - c_mol(i)) with the total forces: together with the
  // two above, the derivative under the barostat's molecular scaling, where
  // intramolecular geometry is kept.
  E_DSCALE_BONDED,
  E_DSCALE_MOLECULAR,
  E_TERM_COUNT
};
constexpr int MAX_SLOTS = 65536;

// Scalars that live on the device so a captured block never needs the host.
struct Scalars {
  unsigned int maxDisp2Bits;  // max |x - x_rebuild|^2 since the last rebuild (float bits)
  unsigned int rebuild;       // this step rebuilds the cell and neighbour lists
  unsigned int forceRebuild;  // set by the host: rebuild at the next decision
  unsigned int required;      // neighbour requirement of the current list (the longest list found)
  unsigned int ticket;        // last-block ticket of the integrator
  unsigned int rebuilds;      // rebuild count
  unsigned int shakeMaxIterations;
  unsigned int shakeFailed;
  unsigned int rattleFailed;
  unsigned int cellOverflow;  // a cell held more atoms than the sort can take
  unsigned int nonFinite;     // a non-finite coordinate was seen
  unsigned int requiredMax;   // largest requirement of the lists retired since the host last cleared
                              // it (with `required`, every list's since then: requirementSince)
  float slack;                // nm of the skin spent by volume moves since the last rebuild
  unsigned int spanBits;      // largest atom-to-centroid distance of any molecule seen (float bits)
  long long step;             // the noise counter (completed steps)
};

// A rebuild retires the current list's requirement into requiredMax before
// the search starts counting the next one.
__device__ __forceinline__ void retireRequirement(Scalars *sc) {
  sc->requiredMax = sc->required > sc->requiredMax ? sc->required : sc->requiredMax;
  sc->required = 0u;
}

// The largest requirement of every list since the host last cleared
// requiredMax: the retired ones and the current one.
inline unsigned int requirementSince(const Scalars &h) {
  return h.required > h.requiredMax ? h.required : h.requiredMax;
}

// The rebuild test: the largest displacement since the last rebuild against
// half the skin less half of what volume moves have spent of it (NaN rebuilds).
__device__ __forceinline__ bool displacementExceeded(const Scalars *sc, float maxDisp2, double dispHalf) {
  const float half = (float)dispHalf - 0.5f * sc->slack;
  return !(half > 0.f) || !(maxDisp2 <= half * half);
}

struct Dev {
  int n;          // atoms
  int nPad;       // stride of the three force planes
  float4 *pos;    // x, y, z, charge
  float4 *vel;    // vx, vy, vz, inverse mass
  long long *force;  // 3 planes of nPad fixed-point components
  float4 *posRebuild;
  int *type;
  float2 *lj;     // (A, B) per type pair
  double2 *lj64;  // the same in FP64, for the energies
  int types;
  const BoxF *box;  // device-resident geometry (see BoxF)
  // cells (arrays sized for cellCapacity cells)
  int *cellOfAtom, *cellCount, *cellStart, *cellCursor, *cellAtoms;
  float4 *cellPos;  // positions in cell order at the last rebuild, the atom index in .w (int bits)
  int cellCapacity;
  // The current positions (.w the charge) and types in cell order ("slots":
  // slot k holds atom cellAtoms[k]); the integrator keeps slotPos current
  // through slotOf. The neighbour list's rows and entries are slots, so the
  // direct space reads a cell's atoms from consecutive addresses.
  float4 *slotPos;
  int *slotType, *slotOf;
  // neighbours: full list, `cap` slots per atom
  int *nbrCount, *nbrList;
  int cap;
  int *exclStart, *exclPartners;
  // bonded terms (f32 parameters)
  int nBond, nAngle, nTorsion, nException, nRestraint;
  int nExcTriplets;         // leading exception triplets: the three pairs of a three-atom molecule
  int2 *bondAtoms;
  double2 *bondParams;      // k, r0 (FP64: r - r0 cancels)
  int4 *angleAtoms;         // a, b, c
  double2 *angleParams;     // k, theta0
  int4 *torsionAtoms;
  double4 *torsionParams;   // k, n, phase
  int2 *excAtoms;
  double4 *excParams;       // qq scaled (with 1/4pi eps0), A, B, full qq (with 1/4pi eps0)
  int *restraintAtoms;
  double4 *restraintRef;    // reference xyz and k, FP64
  // integrator units
  int rect;                 // the box is rectangular (and stays so: volume moves are isotropic)
  int nUnits;
  int waterBlocks;          // leading unit blocks holding water units only
  int4 *units;              // kind, a, b, c (see integrate.hip)
  double4 *waterGeom;       // dOH, dHH, mO, mH per unit (used by water units)
  int *clusterConStart, *clusterAtomStart, *clusterAtoms;
  int4 *clusterCons;        // local slot a, local slot b, constraint index, unused
  double *clusterConD;      // constrained length per cluster constraint
  double *clusterScratch;   // CLUSTER_SLOT doubles per cluster atom slot (clusters too large for LDS)
  // noise
  float *noise;             // 3N supplied normals (optional)
  // energies and scalars
  double *slots;            // E_TERM_COUNT * MAX_SLOTS
  Scalars *sc;
};

}  // namespace rmd
