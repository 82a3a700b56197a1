// buildSystem, the derived host-side setup, and the seeded solvated builder.
//
// buildSystem follows shared/topology/system.js:40-236 (constrained bonds
// leave the harmonic term, every exclusion becomes an exception, 1-4 scaling
// comes from the torsion that produced the pair). The setup helpers follow
// shared/pme/parameters.js, shared/topology/cells.js, shared/dispersion.js,
// shared/random.js and cpu/forces/pme.js:270-286.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <numeric>
#include <set>

#include "host.h"
#include "tuned.h"

namespace rmd {

namespace {
const std::set<std::string> &waterResidues() {
  static const std::set<std::string> s = {"WAT", "HOH", "H2O", "SOL", "TIP", "TP3", "SPC"};
  return s;
}
}  // namespace

// ---- random (shared/random.js:24-87) -------------------------------------

uint32_t hashMix(uint32_t state) {
  state ^= state >> 16;
  state *= 0x7feb352du;
  state ^= state >> 15;
  state *= 0x846ca68bu;
  state ^= state >> 16;
  return state;
}

double uniformAt(uint32_t seed, uint32_t a, uint32_t b, uint32_t c) {
  uint32_t state = hashMix(seed ^ (a * 0x9e3779b1u));
  state = hashMix(state ^ (b * 0x85ebca6bu));
  state = hashMix(state ^ (c * 0xc2b2ae35u));
  return (static_cast<double>(state >> 9) + 0.5) * (1.0 / 8388608.0);
}

void gaussianTriple(uint32_t seed, uint32_t atom, uint32_t step, double out[3]) {
  double u1 = uniformAt(seed, atom, step, 0), u2 = uniformAt(seed, atom, step, 1);
  double u3 = uniformAt(seed, atom, step, 2), u4 = uniformAt(seed, atom, step, 3);
  double r1 = std::sqrt(-2.0 * std::log(u1)), r2 = std::sqrt(-2.0 * std::log(u3));
  out[0] = r1 * std::cos(2.0 * M_PI * u2);
  out[1] = r1 * std::sin(2.0 * M_PI * u2);
  out[2] = r2 * std::cos(2.0 * M_PI * u4);
}

// ---- geometry (cpu/forces/geometry.js:47-69) ------------------------------

// Distances between the box's opposite faces, 1/|column k of the reciprocal
// box|: the widths that bound a cell's thickness along each fractional axis.
// For a rectangular box they are the diagonal; a sheared box's faces are
// closer than its diagonal entries.
void faceWidths(const double box[9], double out[3]) {
  double r[9];
  reciprocalBoxVectors(box, r);
  for (int k = 0; k < 3; ++k) out[k] = 1.0 / std::sqrt(r[k] * r[k] + r[3 + k] * r[3 + k] + r[6 + k] * r[6 + k]);
}

void reciprocalBoxVectors(const double box[9], double out[9]) {
  const double det = box[0] * box[4] * box[8];
  if (!(std::fabs(det) > 0)) throw Error(RMD_E_INVALID, "box vectors are singular; volume is zero");
  const double ax = box[0], bx = box[3], by = box[4], cx = box[6], cy = box[7], cz = box[8];
  const double r[9] = {1 / ax, 0, 0, -bx / (ax * by), 1 / by, 0,
                       (bx * cy - by * cx) / (ax * by * cz), -cy / (by * cz), 1 / cz};
  for (int k = 0; k < 9; ++k) out[k] = r[k];
}

// ---- buildSystem ----------------------------------------------------------

static void buildSymmetricExclusions(System &s) {
  // buildExclusionGather (shared/topology/cells.js:201-237).
  const Topology &t = s.topology;
  const int n = t.atomCount;
  std::vector<int> counts(n, 0);
  for (int atom = 0; atom < n; ++atom)
    for (int c = t.exclusionStart[atom]; c < t.exclusionStart[atom + 1]; ++c) {
      counts[atom] += 1;
      counts[t.exclusionPartners[c]] += 1;
    }
  s.exclusionStart.assign(n + 1, 0);
  for (int a = 0; a < n; ++a) s.exclusionStart[a + 1] = s.exclusionStart[a] + counts[a];
  s.exclusionPartners.assign(s.exclusionStart[n], 0);
  std::vector<int> cursor(s.exclusionStart.begin(), s.exclusionStart.end() - 1);
  for (int atom = 0; atom < n; ++atom)
    for (int c = t.exclusionStart[atom]; c < t.exclusionStart[atom + 1]; ++c) {
      int partner = t.exclusionPartners[c];
      s.exclusionPartners[cursor[atom]++] = partner;
      s.exclusionPartners[cursor[partner]++] = atom;
    }
  for (int a = 0; a < n; ++a)
    std::sort(s.exclusionPartners.begin() + s.exclusionStart[a], s.exclusionPartners.begin() + s.exclusionStart[a + 1]);
}

static int countWaters(const System &s) {
  // A rigid water: a three-constraint triangle (cpu/integrate/constraints.js:36-76).
  std::vector<int> perMolecule(s.topology.moleculeCount, 0);
  for (size_t k = 0; k < s.constraints.a.size(); ++k) perMolecule[s.topology.moleculeOfAtom[s.constraints.a[k]]] += 1;
  int waters = 0;
  for (int m = 0; m < s.topology.moleculeCount; ++m)
    if (perMolecule[m] == 3 && s.moleculeStart[m + 1] - s.moleculeStart[m] == 3) ++waters;
  return waters;
}

System buildSystem(const Topology &t, const Coordinates *coordinates, const rmd_build_options &options) {
  System s;
  s.topology = t;
  if (options.constraints < RMD_CONSTRAINTS_NONE || options.constraints > RMD_CONSTRAINTS_ALLBONDS)
    throw Error(RMD_E_SYSTEM, "unsupported constraints mode; this port implements none, HBonds and "
                              "AllBonds (HAngles is refused rather than approximated).");
  s.constraintMode = options.constraints;
  s.rigidWater = options.rigid_water != 0;
  s.cutoff = options.cutoff;

  bool haveBox = false;
  if (options.has_box_vectors) {
    for (int k = 0; k < 9; ++k) s.box[k] = options.box_vectors[k];
    haveBox = true;
  } else if (options.has_box) {
    boxVectorsFromLengthsAndAngles(options.box_lengths, options.box_angles, s.box);
    haveBox = true;
  } else if (coordinates && coordinates->hasBox) {
    boxVectorsFromLengthsAndAngles(coordinates->boxLengths, coordinates->boxAngles, s.box);
    haveBox = true;
  } else if (t.hasBox) {
    boxVectorsFromLengthsAndAngles(t.boxLengths, t.boxAngles, s.box);
    haveBox = true;
  }
  if (t.isPeriodic && !haveBox)
    throw Error(RMD_E_SYSTEM, "topology declares a periodic box (IFBOX > 0) but no box vectors were "
                              "supplied by either the topology or the coordinate file");
  s.isPeriodic = haveBox;
  if (haveBox) {
    double smallest = std::min({std::fabs(s.box[0]), std::fabs(s.box[4]), std::fabs(s.box[8])});
    if (s.cutoff > 0.5 * smallest) {
      char buf[256];
      std::snprintf(buf, sizeof buf,
                    "a %g nm cutoff does not fit in a box whose smallest width is %.4f nm: the cutoff "
                    "must be at most half of it (%.4f nm).",
                    s.cutoff, smallest, 0.5 * smallest);
      throw Error(RMD_E_SYSTEM, buf);
    }
  }

  // Which bonds are constraints (system.js:90-121).
  std::vector<uint8_t> isWater(t.atomCount, 0);
  if (s.rigidWater)
    for (int a = 0; a < t.atomCount; ++a)
      if (waterResidues().count(t.residueNames[t.residueOfAtom[a]])) isWater[a] = 1;
  const size_t nb = t.bonds.a.size();
  for (size_t i = 0; i < nb; ++i) {
    int a = t.bonds.a[i], b = t.bonds.b[i];
    bool touchesH = t.atomicNumbers[a] == 1 || t.atomicNumbers[b] == 1;
    bool constrained = false;
    if (s.constraintMode == RMD_CONSTRAINTS_ALLBONDS) constrained = true;
    else if (s.constraintMode == RMD_CONSTRAINTS_HBONDS && touchesH) constrained = true;
    if (s.rigidWater && isWater[a] && isWater[b]) constrained = true;
    if (constrained) {
      s.constraints.a.push_back(a);
      s.constraints.b.push_back(b);
      s.constraints.distance.push_back(t.bonds.r0[i]);
    } else {
      s.harmonicBonds.a.push_back(a);
      s.harmonicBonds.b.push_back(b);
      s.harmonicBonds.k.push_back(t.bonds.k[i]);
      s.harmonicBonds.r0.push_back(t.bonds.r0[i]);
    }
  }

  // Exceptions (system.js:141-178): every excluded pair, scaled where it is a 1-4.
  std::map<long long, size_t> pair14;
  for (size_t i = 0; i < t.pairs14.a.size(); ++i)
    pair14[static_cast<long long>(t.pairs14.a[i]) * 0x4000000LL + t.pairs14.b[i]] = i;
  for (int atom = 0; atom < t.atomCount; ++atom)
    for (int c = t.exclusionStart[atom]; c < t.exclusionStart[atom + 1]; ++c) {
      int partner = t.exclusionPartners[c];
      int low = std::min(atom, partner), high = std::max(atom, partner);
      auto found = pair14.find(static_cast<long long>(low) * 0x4000000LL + high);
      int typePair = t.typeCount * t.atomTypes[atom] + t.atomTypes[partner];
      s.exceptions.a.push_back(low);
      s.exceptions.b.push_back(high);
      if (found == pair14.end()) {
        s.exceptions.chargeProduct.push_back(0);
        s.exceptions.ljA.push_back(0);
        s.exceptions.ljB.push_back(0);
        s.exceptions.isScaled.push_back(0);
      } else {
        double scee = t.pairs14.scee[found->second], scnb = t.pairs14.scnb[found->second];
        s.exceptions.chargeProduct.push_back(t.charges[atom] * t.charges[partner] / scee);
        s.exceptions.ljA.push_back(t.ljA[typePair] / scnb);
        s.exceptions.ljB.push_back(t.ljB[typePair] / scnb);
        s.exceptions.isScaled.push_back(1);
      }
    }

  // Molecules as index ranges (system.js:181-197).
  s.moleculeStart.assign(t.moleculeCount + 1, 0);
  std::vector<int> counts(t.moleculeCount, 0);
  for (int a = 0; a < t.atomCount; ++a) counts[t.moleculeOfAtom[a]] += 1;
  for (int m = 0; m < t.moleculeCount; ++m) s.moleculeStart[m + 1] = s.moleculeStart[m] + counts[m];
  s.moleculeAtoms.assign(t.atomCount, 0);
  std::vector<int> cursor(s.moleculeStart.begin(), s.moleculeStart.end() - 1);
  for (int a = 0; a < t.atomCount; ++a) s.moleculeAtoms[cursor[t.moleculeOfAtom[a]]++] = a;

  int massive = 0;
  for (int a = 0; a < t.atomCount; ++a)
    if (t.masses[a] > 0) ++massive;
  s.degreesOfFreedom = 3 * massive - static_cast<int>(s.constraints.a.size()) - 3;

  buildSymmetricExclusions(s);
  if (coordinates) {
    if (coordinates->atomCount != t.atomCount)
      throw Error(RMD_E_PARSE, "coordinate file has " + std::to_string(coordinates->atomCount) +
                                   " atoms but the topology has " + std::to_string(t.atomCount) +
                                   ". These two files are not from the same tleap run - re-save the pair together.");
    s.positions = coordinates->positions;
    if (coordinates->hasVelocities) s.velocities = coordinates->velocities;
  }
  s.waterCount = countWaters(s);
  return s;
}

// ---- seeded solvated builder ---------------------------------------------

System buildSolvated(const Topology &templ, const Coordinates &solute, const rmd_solvated_options &o) {
  if (o.water_count < 0) throw Error(RMD_E_INVALID, "water_count must be >= 0");
  if (!(o.box_edge > 0)) throw Error(RMD_E_INVALID, "box_edge must be positive");
  // Residue classes in the template: waters, single-atom ions, and the solute.
  int waterResidue = -1;
  std::vector<int> soluteResidues;
  for (int r = 0; r < templ.residueCount; ++r) {
    int start = templ.residueFirstAtom[r];
    int end = r + 1 < templ.residueCount ? templ.residueFirstAtom[r + 1] : templ.atomCount;
    bool water = waterResidues().count(templ.residueNames[r]) > 0;
    if (water) {
      if (waterResidue < 0) waterResidue = r;
      continue;
    }
    if (end - start == 1) continue;  // counter-ion
    soluteResidues.push_back(r);
  }
  if (waterResidue < 0) throw Error(RMD_E_INVALID, "the template topology holds no water residue");
  std::vector<int> map(templ.atomCount, -1);  // template atom -> new atom (solute only)
  std::vector<int> soluteAtoms;
  if (o.include_solute)
    for (int r : soluteResidues) {
      int start = templ.residueFirstAtom[r];
      int end = r + 1 < templ.residueCount ? templ.residueFirstAtom[r + 1] : templ.atomCount;
      for (int a = start; a < end; ++a) soluteAtoms.push_back(a);
    }
  if (o.include_solute && solute.atomCount != static_cast<int>(soluteAtoms.size()))
    throw Error(RMD_E_INVALID, "solute coordinates hold " + std::to_string(solute.atomCount) +
                                   " atoms but the template's solute has " + std::to_string(soluteAtoms.size()));
  const int wStart = templ.residueFirstAtom[waterResidue];
  const int wEnd = waterResidue + 1 < templ.residueCount ? templ.residueFirstAtom[waterResidue + 1] : templ.atomCount;
  const int wAtoms = wEnd - wStart;
  if (wAtoms != 3) throw Error(RMD_E_INVALID, "the template water is not a 3-site model");

  Topology t;
  t.typeCount = templ.typeCount;
  t.ljA = templ.ljA;
  t.ljB = templ.ljB;
  t.isPeriodic = true;
  auto addAtom = [&](int src) {
    t.atomNames.push_back(templ.atomNames[src]);
    if (!templ.atomTypeNames.empty()) t.atomTypeNames.push_back(templ.atomTypeNames[src]);
    t.charges.push_back(templ.charges[src]);
    t.masses.push_back(templ.masses[src]);
    t.atomicNumbers.push_back(templ.atomicNumbers[src]);
    t.atomTypes.push_back(templ.atomTypes[src]);
    return t.atomCount++;
  };
  // Solute atoms and residues.
  int lastResidue = -1;
  for (int a : soluteAtoms) {
    int r = templ.residueOfAtom[a];
    if (r != lastResidue) {
      t.residueNames.push_back(templ.residueNames[r]);
      t.residueFirstAtom.push_back(t.atomCount);
      lastResidue = r;
    }
    map[a] = addAtom(a);
    t.residueOfAtom.push_back(static_cast<int>(t.residueNames.size()) - 1);
  }
  // Copy every bonded term whose atoms all map through `m`.
  auto copyTerms = [&](const std::vector<int> &m) {
    for (size_t i = 0; i < templ.bonds.a.size(); ++i) {
      int a = m[templ.bonds.a[i]], b = m[templ.bonds.b[i]];
      if (a < 0 || b < 0) continue;
      t.bonds.a.push_back(a);
      t.bonds.b.push_back(b);
      t.bonds.k.push_back(templ.bonds.k[i]);
      t.bonds.r0.push_back(templ.bonds.r0[i]);
      t.bonds.containsHydrogen.push_back(templ.bonds.containsHydrogen[i]);
    }
    for (size_t i = 0; i < templ.angles.a.size(); ++i) {
      int a = m[templ.angles.a[i]], b = m[templ.angles.b[i]], c = m[templ.angles.c[i]];
      if (a < 0 || b < 0 || c < 0) continue;
      t.angles.a.push_back(a);
      t.angles.b.push_back(b);
      t.angles.c.push_back(c);
      t.angles.k.push_back(templ.angles.k[i]);
      t.angles.theta0.push_back(templ.angles.theta0[i]);
    }
    for (size_t i = 0; i < templ.dihedrals.a.size(); ++i) {
      int a = m[templ.dihedrals.a[i]], b = m[templ.dihedrals.b[i]], c = m[templ.dihedrals.c[i]],
          d = m[templ.dihedrals.d[i]];
      if (a < 0 || b < 0 || c < 0 || d < 0) continue;
      t.dihedrals.a.push_back(a);
      t.dihedrals.b.push_back(b);
      t.dihedrals.c.push_back(c);
      t.dihedrals.d.push_back(d);
      t.dihedrals.k.push_back(templ.dihedrals.k[i]);
      t.dihedrals.periodicity.push_back(templ.dihedrals.periodicity[i]);
      t.dihedrals.phase.push_back(templ.dihedrals.phase[i]);
      t.dihedrals.isImproper.push_back(templ.dihedrals.isImproper[i]);
    }
    for (size_t i = 0; i < templ.pairs14.a.size(); ++i) {
      int a = m[templ.pairs14.a[i]], b = m[templ.pairs14.b[i]];
      if (a < 0 || b < 0) continue;
      t.pairs14.a.push_back(std::min(a, b));
      t.pairs14.b.push_back(std::max(a, b));
      t.pairs14.scee.push_back(templ.pairs14.scee[i]);
      t.pairs14.scnb.push_back(templ.pairs14.scnb[i]);
    }
  };
  copyTerms(map);

  // Water copies: positions first (a seeded jittered lattice), then topology.
  const double L = o.box_edge;
  std::vector<double> solutePos;
  if (o.include_solute) {
    double centroid[3] = {0, 0, 0};
    for (int i = 0; i < solute.atomCount; ++i)
      for (int k = 0; k < 3; ++k) centroid[k] += solute.positions[3 * i + k] / solute.atomCount;
    for (int i = 0; i < solute.atomCount; ++i)
      for (int k = 0; k < 3; ++k) solutePos.push_back(solute.positions[3 * i + k] - centroid[k] + 0.5 * L);
  }
  auto minImage = [L](double d) { return d - L * std::floor(d / L + 0.5); };
  int perAxis = std::max(1, static_cast<int>(std::ceil(std::cbrt(static_cast<double>(o.water_count)))));
  std::vector<int> sites;
  for (;; ++perAxis) {
    sites.clear();
    const double spacing = L / perAxis;
    for (int i = 0; i < perAxis; ++i)
      for (int j = 0; j < perAxis; ++j)
        for (int k = 0; k < perAxis; ++k) {
          double p[3] = {(i + 0.5) * spacing, (j + 0.5) * spacing, (k + 0.5) * spacing};
          bool clear = true;
          for (size_t s = 0; s + 2 < solutePos.size() && clear; s += 3) {
            double dx = minImage(p[0] - solutePos[s]), dy = minImage(p[1] - solutePos[s + 1]),
                   dz = minImage(p[2] - solutePos[s + 2]);
            if (dx * dx + dy * dy + dz * dz < o.solute_clearance * o.solute_clearance) clear = false;
          }
          if (clear) sites.push_back((i * perAxis + j) * perAxis + k);
        }
    if (static_cast<int>(sites.size()) >= o.water_count) break;
  }
  // Seeded choice of which free sites hold a water (partial Fisher-Yates),
  // then lattice order so neighbouring waters have neighbouring indices.
  for (int i = 0; i < o.water_count; ++i) {
    int remaining = static_cast<int>(sites.size()) - i;
    int pick = i + std::min(remaining - 1, static_cast<int>(uniformAt(o.seed, i, 0x51e5, 0) * remaining));
    std::swap(sites[i], sites[pick]);
  }
  sites.resize(o.water_count);
  std::sort(sites.begin(), sites.end());

  // Template water geometry from its constraint lengths (O-H, O-H, H-H).
  int oxygen = wStart;
  for (int a = wStart; a < wEnd; ++a)
    if (templ.masses[a] > templ.masses[oxygen]) oxygen = a;
  int h1 = -1, h2 = -1;
  for (int a = wStart; a < wEnd; ++a)
    if (a != oxygen) (h1 < 0 ? h1 : h2) = a;
  double rOH = 0, rHH = 0;
  for (size_t i = 0; i < templ.bonds.a.size(); ++i) {
    int a = templ.bonds.a[i], b = templ.bonds.b[i];
    if ((a == oxygen && b == h1) || (a == h1 && b == oxygen)) rOH = templ.bonds.r0[i];
    if ((a == h1 && b == h2) || (a == h2 && b == h1)) rHH = templ.bonds.r0[i];
  }
  if (!(rOH > 0 && rHH > 0)) throw Error(RMD_E_INVALID, "the template water has no O-H / H-H bonds");
  const double cosTheta = (2 * rOH * rOH - rHH * rHH) / (2 * rOH * rOH);
  const double sinTheta = std::sqrt(std::max(0.0, 1 - cosTheta * cosTheta));
  const double local[3][3] = {{0, 0, 0}, {rOH, 0, 0}, {rOH * cosTheta, rOH * sinTheta, 0}};
  // Centre the template on its centre of mass before rotating.
  double mo = templ.masses[oxygen], mh1 = templ.masses[h1], mh2 = templ.masses[h2];
  double com[3];
  for (int k = 0; k < 3; ++k) com[k] = (mo * local[0][k] + mh1 * local[1][k] + mh2 * local[2][k]) / (mo + mh1 + mh2);

  std::vector<double> positions = solutePos;
  const double spacing = L / perAxis;
  for (int w = 0; w < o.water_count; ++w) {
    int site = sites[w];
    int i = site / (perAxis * perAxis), j = (site / perAxis) % perAxis, k = site % perAxis;
    double centre[3] = {(i + 0.5) * spacing, (j + 0.5) * spacing, (k + 0.5) * spacing};
    for (int c = 0; c < 3; ++c) centre[c] += o.jitter * (2.0 * uniformAt(o.seed, w, 0x717e, c) - 1.0);
    // Uniform random rotation (Shoemake's quaternion).
    double u1 = uniformAt(o.seed, w, 0x907a, 0), u2 = uniformAt(o.seed, w, 0x907a, 1),
           u3 = uniformAt(o.seed, w, 0x907a, 2);
    double qx = std::sqrt(1 - u1) * std::sin(2 * M_PI * u2), qy = std::sqrt(1 - u1) * std::cos(2 * M_PI * u2);
    double qz = std::sqrt(u1) * std::sin(2 * M_PI * u3), qw = std::sqrt(u1) * std::cos(2 * M_PI * u3);
    double R[3][3] = {{1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)},
                      {2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)},
                      {2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)}};
    std::vector<int> wmap(templ.atomCount, -1);
    t.residueNames.push_back(templ.residueNames[waterResidue]);
    t.residueFirstAtom.push_back(t.atomCount);
    for (int a = wStart; a < wEnd; ++a) {
      int which = a == oxygen ? 0 : a == h1 ? 1 : 2;
      double rel[3] = {local[which][0] - com[0], local[which][1] - com[1], local[which][2] - com[2]};
      wmap[a] = addAtom(a);
      t.residueOfAtom.push_back(static_cast<int>(t.residueNames.size()) - 1);
      for (int c = 0; c < 3; ++c)
        positions.push_back(centre[c] + R[c][0] * rel[0] + R[c][1] * rel[1] + R[c][2] * rel[2]);
    }
    copyTerms(wmap);
    // Exclusions of this copy, one-sided as in a prmtop, filled below.
  }
  t.residueCount = static_cast<int>(t.residueNames.size());

  // One-sided exclusions: remap the template lists of the atoms kept.
  std::vector<std::vector<int>> excl(t.atomCount);
  auto addExclusions = [&](const std::vector<int> &m) {
    for (int a = 0; a < templ.atomCount; ++a) {
      if (m[a] < 0) continue;
      for (int c = templ.exclusionStart[a]; c < templ.exclusionStart[a + 1]; ++c) {
        int p = m[templ.exclusionPartners[c]];
        if (p >= 0) excl[m[a]].push_back(p);
      }
    }
  };
  addExclusions(map);
  for (int w = 0; w < o.water_count; ++w) {
    std::vector<int> wmap(templ.atomCount, -1);
    int base = static_cast<int>(soluteAtoms.size()) + 3 * w;
    for (int a = wStart; a < wEnd; ++a) wmap[a] = base + (a - wStart);
    addExclusions(wmap);
  }
  t.exclusionStart.assign(t.atomCount + 1, 0);
  for (int a = 0; a < t.atomCount; ++a) {
    t.exclusionStart[a] = static_cast<int>(t.exclusionPartners.size());
    t.exclusionPartners.insert(t.exclusionPartners.end(), excl[a].begin(), excl[a].end());
  }
  t.exclusionStart[t.atomCount] = static_cast<int>(t.exclusionPartners.size());

  // Molecules from the bond graph (each water its own).
  {
    std::vector<int> parent(t.atomCount);
    std::iota(parent.begin(), parent.end(), 0);
    std::function<int(int)> find = [&](int x) { return parent[x] == x ? x : parent[x] = find(parent[x]); };
    for (size_t i = 0; i < t.bonds.a.size(); ++i) {
      int a = find(t.bonds.a[i]), b = find(t.bonds.b[i]);
      if (a != b) parent[b] = a;
    }
    std::map<int, int> rootToMolecule;
    t.moleculeOfAtom.assign(t.atomCount, 0);
    for (int a = 0; a < t.atomCount; ++a) {
      int root = find(a);
      auto it = rootToMolecule.find(root);
      if (it == rootToMolecule.end()) it = rootToMolecule.emplace(root, static_cast<int>(rootToMolecule.size())).first;
      t.moleculeOfAtom[a] = it->second;
    }
    t.moleculeCount = static_cast<int>(rootToMolecule.size());
  }

  rmd_build_options b;
  rmd_build_options_default(&b);
  b.constraints = o.constraints;
  b.rigid_water = o.rigid_water;
  b.cutoff = o.cutoff;
  b.has_box_vectors = 1;
  for (int k = 0; k < 9; ++k) b.box_vectors[k] = 0;
  b.box_vectors[0] = b.box_vectors[4] = b.box_vectors[8] = L;
  Coordinates c;
  c.atomCount = t.atomCount;
  c.positions = positions;
  return buildSystem(t, &c, b);
}

// ---- PME setup (shared/pme/parameters.js) ---------------------------------

double alphaFromTolerance(double cutoff, double tolerance) {
  if (!(cutoff > 0)) throw Error(RMD_E_INVALID, "cutoff must be positive");
  if (!(tolerance > 0 && tolerance < 0.5)) throw Error(RMD_E_INVALID, "ewald tolerance must be in (0, 0.5)");
  return std::sqrt(-std::log(2 * tolerance)) / cutoff;
}

// The reference rule for a DERIVED grid: 2^a 3^b 5^c (parameters.js:63-69).
bool isTransformableSize(int value) {
  int r = value;
  for (int radix : {2, 3, 5})
    while (r % radix == 0) r /= radix;
  return r == 1;
}

// What this build's FFT can transform: 2^a 3^b 5^c 7^d.
static bool isSevenSmooth(int value) {
  int r = value;
  for (int radix : {2, 3, 5, 7})
    while (r % radix == 0) r /= radix;
  return r == 1;
}

int nextTransformableSize(int minimum) {
  int c = std::max(2, minimum);
  while (!isTransformableSize(c)) ++c;
  return c;
}

// De Boor recursion with the derivative taken at order-1
// (shared/pme/parameters.js:121-149).
void bsplineWeights(double offset, int order, double *w, double *d) {
  for (int k = 0; k < order; ++k) w[k] = 0;
  w[0] = 1 - offset;
  w[1] = offset;
  for (int k = 3; k < order; ++k) {
    double div = 1.0 / (k - 1);
    w[k - 1] = div * offset * w[k - 2];
    for (int j = 1; j < k - 1; ++j) w[k - j - 1] = div * ((offset + j) * w[k - j - 2] + (k - j - offset) * w[k - j - 1]);
    w[0] = div * (1 - offset) * w[0];
  }
  if (d) {
    d[0] = -w[0];
    for (int j = 1; j < order; ++j) d[j] = w[j - 1] - w[j];
  }
  double div = 1.0 / (order - 1);
  w[order - 1] = div * offset * w[order - 2];
  for (int j = 1; j < order - 1; ++j)
    w[order - j - 1] = div * ((offset + j) * w[order - j - 2] + (order - j - offset) * w[order - j - 1]);
  w[0] = div * (1 - offset) * w[0];
}

// Explicit DFT of the integer-point spline values, Nyquist patched
// (parameters.js:164-215).
std::vector<double> bsplineModuli(int gridSize, int order) {
  std::vector<double> atZero(order), dummy(order);
  bsplineWeights(0.0, order, atZero.data(), dummy.data());
  std::vector<double> m(gridSize);
  for (int i = 0; i < gridSize; ++i) {
    double re = 0, im = 0;
    for (int k = 0; k < order; ++k) {
      double angle = 2 * M_PI * i * k / gridSize;
      re += atZero[k] * std::cos(angle);
      im += atZero[k] * std::sin(angle);
    }
    m[i] = re * re + im * im;
  }
  std::vector<double> patched = m;
  for (int i = 0; i < gridSize; ++i)
    if (m[i] < 1e-7) patched[i] = 0.5 * (m[(i - 1 + gridSize) % gridSize] + m[(i + 1) % gridSize]);
  for (int i = 0; i < gridSize; ++i)
    if (!(patched[i] > 0))
      throw Error(RMD_E_INVALID, "B-spline modulus is still zero after the neighbour substitution; the grid "
                                 "is too small for this interpolation order.");
  return patched;
}

// ---- cells (shared/topology/cells.js) -------------------------------------

double skinFor(const System &s, double timestep, double temperature, const double *box) {
  double lightest = INFINITY;
  for (double m : s.topology.masses)
    if (m > 0 && m < lightest) lightest = m;
  if (!std::isfinite(lightest)) lightest = 1.008;
  const double anchor = 0.40;
  const double speedRatio = std::sqrt((temperature / 298.0) * (1.008 / lightest));
  const double scaled = anchor * (timestep / 0.002) * speedRatio;
  const double *b = box ? box : s.box;
  double room = INFINITY;
  if (s.isPeriodic) {
    // half the smallest face-to-face width (the reference takes the diagonal,
    // which is the same for a rectangular box and too wide for a sheared one)
    double w[3];
    faceWidths(b, w);
    room = std::min({w[0], w[1], w[2]}) / 2 - s.cutoff;
  }
  return std::max(0.1, std::min(scaled, room > 0 ? room * 0.9 : 0.1));
}

CellPlan planCellList(const System &s, double skin, double slack, int maxNeighbours, const double *box,
                      int reachMax) {
  if (!s.isPeriodic) throw Error(RMD_E_INVALID, "a cell list needs a periodic box");
  const double *b = box ? box : s.box;
  CellPlan p;
  p.skin = skin;
  p.cutoff = s.cutoff;
  p.listCutoff = s.cutoff + skin;
  // Face-to-face widths: a cell of width w along fractional axis k is
  // w/counts[k] thick only when w is the distance between the box's faces.
  // For a rectangular box this is the diagonal the reference uses; for a
  // sheared box the diagonal overstates it and a stencil over cells that thin
  // would miss listed pairs.
  double widths[3];
  faceWidths(b, widths);
  // The reference's cells are at least one list cutoff wide (reach 1), so in
  // a box under four cutoffs wide the 3-cell stencil covers every cell and
  // the search is all-pairs. Cells of listCutoff / reach with a stencil of
  // reach cells either side find the same pairs; the plan takes the reach
  // (up to 3) that leaves the fewest candidates, keeping the reference's plan
  // when a finer one would not cull, and at most CELL_LIMIT_SMALL cells for
  // systems the single-workgroup build takes.
  const int n = s.topology.atomCount;
  double best = 2.0;
  const int widest = reachMax >= 1 && reachMax <= CELL_REACH_MAX ? reachMax : CELL_REACH_MAX;
  for (int reach = 1; reach <= widest; ++reach) {
    int counts[3];
    double fraction = 1.0;
    long long cells = 1;
    for (int k = 0; k < 3; ++k) {
      counts[k] = std::max(1, static_cast<int>(std::floor(std::fabs(widths[k]) * reach / p.listCutoff)));
      fraction *= (double)std::min(2 * reach + 1, counts[k]) / counts[k];
      cells *= counts[k];
    }
    if (reach > 1 && n <= CELL_ATOMS_SMALL && cells > CELL_LIMIT_SMALL) break;
    if (reach == 1 || fraction < best - 1e-12) {
      best = fraction;
      p.reach = reach;
      for (int k = 0; k < 3; ++k) p.counts[k] = counts[k];
    }
  }
  p.cellCount = p.counts[0] * p.counts[1] * p.counts[2];
  const double volume = std::fabs(b[0] * b[4] * b[8]);
  const double density = s.topology.atomCount / volume;
  p.estimatedNeighbours = density * (4.0 / 3.0) * M_PI * std::pow(p.listCutoff, 3);
  p.maxNeighbours = maxNeighbours > 0 ? maxNeighbours
                                      : std::max(32, static_cast<int>(std::ceil(p.estimatedNeighbours * slack)));
  return p;
}

const TunedShape &tunedShape(int kind, int atoms) {
  const TunedShape *best = nullptr;
  for (const TunedShape &t : TUNED_SHAPES)
    if (t.kind == kind && t.minAtoms <= atoms && (!best || t.minAtoms > best->minAtoms)) best = &t;
  if (!best) throw Error(RMD_E_INVALID, "the tuning table has no row for this workload");
  return *best;
}

// ---- dispersion / self ----------------------------------------------------

DispersionCoefficients dispersionCoefficients(const System &s) {
  const Topology &t = s.topology;
  std::vector<double> counts(t.typeCount, 0.0);
  for (int a = 0; a < t.atomCount; ++a) counts[t.atomTypes[a]] += 1;
  double oA = 0, oB = 0, dA = 0, dB = 0;
  for (int i = 0; i < t.typeCount; ++i) {
    for (int j = 0; j < t.typeCount; ++j) {
      oA += counts[i] * counts[j] * t.ljA[t.typeCount * i + j];
      oB += counts[i] * counts[j] * t.ljB[t.typeCount * i + j];
    }
    dA += counts[i] * t.ljA[t.typeCount * i + i];
    dB += counts[i] * t.ljB[t.typeCount * i + i];
  }
  const double scale = static_cast<double>(t.atomCount) / (t.atomCount + 1);
  return {(oA + dA) * scale, (oB + dB) * scale};
}

double dispersionEnergy(const DispersionCoefficients &c, double volume, double cutoff) {
  if (!(volume > 0) || !(cutoff > 0)) return 0;
  const double c3 = cutoff * cutoff * cutoff, c9 = c3 * c3 * c3;
  return (2 * M_PI / volume) * (c.a / (9 * c9) - c.b / (3 * c3));
}

void ewaldSelfEnergy(const System &s, double alpha, double volume, double *self, double *net) {
  double q2 = 0, q = 0;
  for (double c : s.topology.charges) {
    q2 += c * c;
    q += c;
  }
  *self = -ONE_4PI_EPS0 * alpha * q2 / std::sqrt(M_PI);
  *net = -ONE_4PI_EPS0 * M_PI * q * q / (2 * alpha * alpha * volume);
}

}  // namespace rmd

// ---- C ABI: host-side entry points ----------------------------------------

using namespace rmd;

namespace rmd {
thread_local std::string g_lastError;
void setGlobalError(const std::string &m) { g_lastError = m; }
}  // namespace rmd

#define RMD_GUARD(body)                              \
  try {                                              \
    body;                                            \
    return RMD_OK;                                   \
  } catch (const rmd::Error &e) {                    \
    rmd::setGlobalError(e.what());                   \
    return e.code;                                   \
  } catch (const std::exception &e) {                \
    rmd::setGlobalError(e.what());                   \
    return RMD_E_INVALID;                            \
  }

extern "C" {

const char *rmd_last_global_error(void) { return rmd::g_lastError.c_str(); }
int32_t rmd_abi_version(void) { return RMD_ABI_VERSION; }

int rmd_topology_parse_prmtop(const char *text, size_t length, rmd_topology **out) {
  RMD_GUARD({
    if (!text || !out) throw Error(RMD_E_INVALID, "null argument");
    auto *t = new rmd_topology{parsePrmtop(std::string(text, length))};
    *out = t;
  });
}
void rmd_topology_free(rmd_topology *t) { delete t; }

int rmd_coordinates_parse_inpcrd(const char *text, size_t length, int32_t expected, rmd_coordinates **out) {
  RMD_GUARD({
    if (!text || !out) throw Error(RMD_E_INVALID, "null argument");
    *out = new rmd_coordinates{parseInpcrd(std::string(text, length), expected)};
  });
}
void rmd_coordinates_free(rmd_coordinates *c) { delete c; }

int rmd_coordinates_view_get(const rmd_coordinates *c, rmd_coordinates_view *v) {
  RMD_GUARD({
    if (!c || !v) throw Error(RMD_E_INVALID, "null argument");
    v->atom_count = c->c.atomCount;
    v->positions = c->c.positions.data();
    v->velocities = c->c.hasVelocities ? c->c.velocities.data() : nullptr;
    v->has_box = c->c.hasBox;
    for (int k = 0; k < 3; ++k) {
      v->box_lengths[k] = c->c.boxLengths[k];
      v->box_angles[k] = c->c.boxAngles[k];
    }
    v->time = c->c.hasTime ? c->c.time : NAN;
  });
}

int rmd_box_vectors_from_lengths_and_angles(const double lengths[3], const double angles[3], double out[9]) {
  RMD_GUARD({ boxVectorsFromLengthsAndAngles(lengths, angles, out); });
}

void rmd_build_options_default(rmd_build_options *o) {
  std::memset(o, 0, sizeof *o);
  o->constraints = RMD_CONSTRAINTS_HBONDS;
  o->rigid_water = 1;
  o->cutoff = 1.0;
}

int rmd_system_build(const rmd_topology *t, const rmd_coordinates *c, const rmd_build_options *o,
                     rmd_system **out) {
  RMD_GUARD({
    if (!t || !out) throw Error(RMD_E_INVALID, "null argument");
    rmd_build_options d;
    rmd_build_options_default(&d);
    *out = new rmd_system{buildSystem(t->t, c ? &c->c : nullptr, o ? *o : d)};
  });
}
void rmd_system_free(rmd_system *s) { delete s; }

void rmd_solvated_options_default(rmd_solvated_options *o) {
  std::memset(o, 0, sizeof *o);
  o->water_count = 389;
  o->box_edge = 2.2816;
  o->seed = 101;
  o->include_solute = 1;
  o->solute_clearance = 0.25;
  o->jitter = 0.02;
  o->cutoff = 0.9;
  o->constraints = RMD_CONSTRAINTS_HBONDS;
  o->rigid_water = 1;
}

int rmd_system_build_solvated(const char *tp, size_t tl, const char *si, size_t sl, const rmd_solvated_options *o,
                              rmd_system **out) {
  RMD_GUARD({
    if (!tp || !out || !o) throw Error(RMD_E_INVALID, "null argument");
    Topology templ = parsePrmtop(std::string(tp, tl));
    Coordinates solute;
    if (o->include_solute) {
      if (!si) throw Error(RMD_E_INVALID, "solute coordinates are required when include_solute is set");
      solute = parseInpcrd(std::string(si, sl), 0);
    }
    *out = new rmd_system{buildSolvated(templ, solute, *o)};
  });
}

int rmd_system_view_get(const rmd_system *sys, rmd_system_view *v) {
  RMD_GUARD({
    if (!sys || !v) throw Error(RMD_E_INVALID, "null argument");
    const System &s = sys->s;
    const Topology &t = s.topology;
    std::memset(v, 0, sizeof *v);
    v->atom_count = t.atomCount;
    v->type_count = t.typeCount;
    v->molecule_count = t.moleculeCount;
    v->residue_count = t.residueCount;
    v->charges = t.charges.data();
    v->masses = t.masses.data();
    v->atom_types = t.atomTypes.data();
    v->atomic_numbers = t.atomicNumbers.data();
    v->residue_of_atom = t.residueOfAtom.data();
    v->lj_a = t.ljA.data();
    v->lj_b = t.ljB.data();
    v->bond_count = static_cast<int32_t>(s.harmonicBonds.a.size());
    v->bond_a = s.harmonicBonds.a.data();
    v->bond_b = s.harmonicBonds.b.data();
    v->bond_k = s.harmonicBonds.k.data();
    v->bond_r0 = s.harmonicBonds.r0.data();
    v->angle_count = static_cast<int32_t>(t.angles.a.size());
    v->angle_a = t.angles.a.data();
    v->angle_b = t.angles.b.data();
    v->angle_c = t.angles.c.data();
    v->angle_k = t.angles.k.data();
    v->angle_theta0 = t.angles.theta0.data();
    v->torsion_count = static_cast<int32_t>(t.dihedrals.a.size());
    v->torsion_a = t.dihedrals.a.data();
    v->torsion_b = t.dihedrals.b.data();
    v->torsion_c = t.dihedrals.c.data();
    v->torsion_d = t.dihedrals.d.data();
    v->torsion_k = t.dihedrals.k.data();
    v->torsion_periodicity = t.dihedrals.periodicity.data();
    v->torsion_phase = t.dihedrals.phase.data();
    v->constraint_count = static_cast<int32_t>(s.constraints.a.size());
    v->constraint_a = s.constraints.a.data();
    v->constraint_b = s.constraints.b.data();
    v->constraint_distance = s.constraints.distance.data();
    v->exception_count = static_cast<int32_t>(s.exceptions.a.size());
    v->exception_a = s.exceptions.a.data();
    v->exception_b = s.exceptions.b.data();
    v->exception_charge_product = s.exceptions.chargeProduct.data();
    v->exception_lj_a = s.exceptions.ljA.data();
    v->exception_lj_b = s.exceptions.ljB.data();
    v->exclusion_total = static_cast<int32_t>(s.exclusionPartners.size());
    v->exclusion_start = s.exclusionStart.data();
    v->exclusion_partners = s.exclusionPartners.data();
    v->molecule_start = s.moleculeStart.data();
    v->molecule_atoms = s.moleculeAtoms.data();
    v->molecule_of_atom = t.moleculeOfAtom.data();
    v->is_periodic = s.isPeriodic;
    for (int k = 0; k < 9; ++k) v->box[k] = s.box[k];
    v->cutoff = s.cutoff;
    v->degrees_of_freedom = s.degreesOfFreedom;
    v->constraint_mode = s.constraintMode;
    v->rigid_water = s.rigidWater;
    v->positions = s.positions.empty() ? nullptr : s.positions.data();
    v->velocities = s.velocities.empty() ? nullptr : s.velocities.data();
    v->water_count = s.waterCount;
  });
}

int rmd_pme_parameters_derive(const double box[9], double cutoff, double tolerance, double forcedAlpha,
                              const int32_t forcedGrid[3], rmd_pme_parameters *out) {
  RMD_GUARD({
    if (!box || !out) throw Error(RMD_E_INVALID, "null argument");
    double tol = tolerance > 0 ? tolerance : 0.0005;
    double alpha = forcedAlpha > 0 ? forcedAlpha : alphaFromTolerance(cutoff, tol);
    for (int k = 0; k < 3; ++k) {
      int g;
      if (forcedGrid && forcedGrid[k] > 0) {
        g = forcedGrid[k];
        if (!isSevenSmooth(g))
          throw Error(RMD_E_INVALID, "PME grid size " + std::to_string(g) +
                                         " is not a product of 2, 3, 5 and 7, which is what this build's FFT can transform.");
      } else {
        double length = std::sqrt(box[3 * k] * box[3 * k] + box[3 * k + 1] * box[3 * k + 1] + box[3 * k + 2] * box[3 * k + 2]);
        g = nextTransformableSize(static_cast<int>(std::ceil(2 * alpha * length / (3 * std::pow(tol, 0.2)))));
      }
      if (g < PME_ORDER)
        throw Error(RMD_E_INVALID, "PME grid size " + std::to_string(g) + " is smaller than the interpolation order 5.");
      out->grid[k] = g;
    }
    out->alpha = alpha;
    out->order = PME_ORDER;
    out->tolerance = tol;
  });
}

int rmd_pme_moduli(int32_t gridSize, int32_t order, double *moduli) {
  RMD_GUARD({
    if (!moduli || gridSize < order) throw Error(RMD_E_INVALID, "bad moduli arguments");
    auto m = bsplineModuli(gridSize, order);
    std::copy(m.begin(), m.end(), moduli);
  });
}

namespace {
// The selection's residue classes (shared/analysis/selection.js:9-24,70-77):
// a residue is solute unless its trimmed, upper-cased name is a water or an
// ion. The selection's water set is wider than buildSystem's rigid-water set
// by T3P.
bool isSoluteResidue(std::string name) {
  static const std::set<std::string> solvent = {
      "WAT", "HOH", "H2O", "SOL", "TIP", "TP3", "SPC", "T3P",                                   // water
      "NA",  "NA+", "CL",  "CL-", "K",   "K+",  "MG",  "MG2", "CA", "CA2", "ZN", "ZN2", "LI",   // ions
      "LI+", "RB",  "RB+", "CS",  "CS+", "F",   "F-",  "BR",  "BR-", "I",  "I-", "IB",  "CIO",
      "MN",  "FE",  "FE2", "CU",  "CU1"};
  const size_t first = name.find_first_not_of(" \t");
  if (first == std::string::npos) return true;
  name = name.substr(first, name.find_last_not_of(" \t") + 1 - first);
  for (char &ch : name) ch = (char)std::toupper((unsigned char)ch);
  return solvent.count(name) == 0;
}
}  // namespace

int rmd_tuned_shape_get(int32_t kind, int32_t atoms, rmd_tuned_shape *out) {
  if (!out || (kind != 0 && kind != 1) || atoms < 0) {
    rmd::setGlobalError("bad arguments");
    return RMD_E_INVALID;
  }
  try {
    const rmd::TunedShape &t = rmd::tunedShape(kind, atoms);
    out->lanes_per_atom = t.lanes;
    out->cell_reach = t.reach;
    out->search_mode = t.searchMode;
    out->skin = t.skin;
    out->pme_bins = t.pmeBins;
    return RMD_OK;
  } catch (const rmd::Error &e) {
    rmd::setGlobalError(e.what());
    return e.code;
  }
}

int rmd_system_rebox(const rmd_system *sys, const double box[9], rmd_system **out) {
  RMD_GUARD({
    if (!sys || !box || !out) throw Error(RMD_E_INVALID, "null argument");
    if (!sys->s.isPeriodic) throw Error(RMD_E_SYSTEM, "a non-periodic system has no box to replace");
    for (int k = 0; k < 9; ++k)
      if (!std::isfinite(box[k])) throw Error(RMD_E_INVALID, "box vectors must be finite");
    if (box[1] != 0 || box[2] != 0 || box[5] != 0 || !(box[0] > 0) || !(box[4] > 0) || !(box[8] > 0))
      throw Error(RMD_E_INVALID, "box vectors must be reduced: a along x, b in the xy plane, positive diagonal");
    const double smallest = std::min({box[0], box[4], box[8]});
    if (sys->s.cutoff > 0.5 * smallest) {
      char buf[256];
      std::snprintf(buf, sizeof buf,
                    "a %g nm cutoff does not fit in a box whose smallest width is %.4f nm: the cutoff "
                    "must be at most half of it (%.4f nm).",
                    sys->s.cutoff, smallest, 0.5 * smallest);
      throw Error(RMD_E_SYSTEM, buf);
    }
    rmd_system *copy = new rmd_system{sys->s};
    std::memcpy(copy->s.box, box, sizeof copy->s.box);
    *out = copy;
  });
}

int rmd_system_solute_heavy_atoms(const rmd_system *sys, int32_t *out, int32_t capacity, int32_t *count) {
  RMD_GUARD({
    if (!sys || !count || (capacity > 0 && !out)) throw Error(RMD_E_INVALID, "null argument");
    const Topology &t = sys->s.topology;
    int32_t n = 0;
    for (int a = 0; a < t.atomCount; ++a) {
      if (!isSoluteResidue(t.residueNames[t.residueOfAtom[a]]) || t.atomicNumbers[a] == 1) continue;
      if (n < capacity) out[n] = a;
      ++n;
    }
    *count = n;
  });
}

double rmd_skin_for(const rmd_system *s, double timestep, double temperature) {
  try {
    return skinFor(s->s, timestep > 0 ? timestep : 0.002, temperature > 0 ? temperature : 298.0, nullptr);
  } catch (...) {
    return NAN;
  }
}

double rmd_dispersion_energy(const rmd_system *s, double volume) {
  return dispersionEnergy(dispersionCoefficients(s->s), volume, s->s.cutoff);
}

}  // extern "C"
