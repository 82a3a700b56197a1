// AMBER parm7 topology and ASCII coordinate readers.
//
// Follows shared/amber/sections.js (fixed-width %FLAG/%FORMAT reader),
// shared/amber/flags.js (every flag audited; energy-bearing unknowns are
// refused), shared/amber/prmtop.js (units converted once, here) and
// shared/amber/inpcrd.js (content decided by counting).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <numeric>
#include <set>
#include <sstream>

#include "host.h"

namespace rmd {
namespace {

struct Format {
  int perLine = 1;
  char kind = 'i';  // 'i' integer, 'f' float, 's' string
  int width = 8;
};

struct Section {
  Format format;
  std::vector<double> numbers;
  std::vector<std::string> strings;
};

std::string trim(const std::string &s) {
  size_t b = 0, e = s.size();
  while (b < e && std::isspace(static_cast<unsigned char>(s[b]))) ++b;
  while (e > b && std::isspace(static_cast<unsigned char>(s[e - 1]))) --e;
  return s.substr(b, e - b);
}

// parseFortranFormat (shared/amber/sections.js:47-59).
Format parseFormat(const std::string &descriptor) {
  std::string d = trim(descriptor);
  size_t i = 0;
  std::string repeat;
  while (i < d.size() && std::isdigit(static_cast<unsigned char>(d[i]))) repeat += d[i++];
  if (i >= d.size()) throw Error(RMD_E_PARSE, "unsupported AMBER %FORMAT descriptor: " + descriptor);
  char letter = static_cast<char>(std::tolower(static_cast<unsigned char>(d[i++])));
  if (std::string("aiefg").find(letter) == std::string::npos)
    throw Error(RMD_E_PARSE, "unsupported AMBER %FORMAT descriptor: " + descriptor);
  std::string width;
  while (i < d.size() && std::isdigit(static_cast<unsigned char>(d[i]))) width += d[i++];
  if (width.empty()) throw Error(RMD_E_PARSE, "unsupported AMBER %FORMAT descriptor: " + descriptor);
  if (i < d.size() && d[i] == '.') {
    ++i;
    while (i < d.size() && std::isdigit(static_cast<unsigned char>(d[i]))) ++i;
  }
  if (i != d.size()) throw Error(RMD_E_PARSE, "unsupported AMBER %FORMAT descriptor: " + descriptor);
  Format f;
  f.perLine = repeat.empty() ? 1 : std::atoi(repeat.c_str());
  f.kind = letter == 'a' ? 's' : letter == 'i' ? 'i' : 'f';
  f.width = std::atoi(width.c_str());
  return f;
}

std::vector<std::string> splitLines(const std::string &text) {
  std::vector<std::string> lines;
  size_t start = 0;
  while (start <= text.size()) {
    size_t end = text.find('\n', start);
    if (end == std::string::npos) end = text.size();
    std::string line = text.substr(start, end - start);
    if (!line.empty() && line.back() == '\r') line.pop_back();
    lines.push_back(line);
    if (end == text.size()) break;
    start = end + 1;
  }
  return lines;
}

bool startsWith(const std::string &s, const char *prefix) {
  return s.compare(0, std::strlen(prefix), prefix) == 0;
}

double parseNumber(const std::string &field, const std::string &flag) {
  std::string t = trim(field);
  char *end = nullptr;
  double v = std::strtod(t.c_str(), &end);
  if (end == t.c_str() || *end != '\0' || !std::isfinite(v))
    throw Error(RMD_E_PARSE, "non-numeric field \"" + field + "\" in %FLAG " + flag);
  return v;
}

// readAmberSections (shared/amber/sections.js:87-147): fields are sliced by
// their declared width; blank fields are padding; %COMMENT may appear anywhere.
std::map<std::string, Section> readSections(const std::string &text, std::vector<std::string> &order) {
  std::map<std::string, Section> sections;
  std::string flag;
  bool haveFormat = false;
  Section current;
  auto commit = [&]() {
    if (!flag.empty() && haveFormat) {
      sections[flag] = current;
      order.push_back(flag);
    }
    flag.clear();
    haveFormat = false;
    current = Section();
  };
  for (const std::string &line : splitLines(text)) {
    if (startsWith(line, "%VERSION")) continue;
    if (startsWith(line, "%COMMENT")) continue;
    if (startsWith(line, "%FLAG")) {
      commit();
      flag = trim(line.substr(5));
      continue;
    }
    if (startsWith(line, "%FORMAT")) {
      size_t open = line.find('('), close = line.rfind(')');
      if (open == std::string::npos || close == std::string::npos || close < open)
        throw Error(RMD_E_PARSE, "malformed %FORMAT line: " + line);
      current.format = parseFormat(line.substr(open + 1, close - open - 1));
      haveFormat = true;
      continue;
    }
    if (flag.empty() || !haveFormat) continue;
    const Format &f = current.format;
    for (int index = 0; index < f.perLine; ++index) {
      size_t start = static_cast<size_t>(index) * f.width;
      if (start >= line.size()) break;
      std::string field = line.substr(start, f.width);
      if (f.kind == 's') {
        current.strings.push_back(field);
        continue;
      }
      if (trim(field).empty()) continue;
      current.numbers.push_back(parseNumber(field, flag));
    }
  }
  commit();
  return sections;
}

const std::set<std::string> &handledFlags() {
  static const std::set<std::string> s = {
      "VERSION", "TITLE", "POINTERS", "ATOM_NAME", "CHARGE", "MASS", "ATOM_TYPE_INDEX",
      "AMBER_ATOM_TYPE", "ATOMIC_NUMBER", "RESIDUE_LABEL", "RESIDUE_POINTER",
      "NUMBER_EXCLUDED_ATOMS", "EXCLUDED_ATOMS_LIST", "NONBONDED_PARM_INDEX",
      "LENNARD_JONES_ACOEF", "LENNARD_JONES_BCOEF", "BOND_FORCE_CONSTANT", "BOND_EQUIL_VALUE",
      "BONDS_INC_HYDROGEN", "BONDS_WITHOUT_HYDROGEN", "ANGLE_FORCE_CONSTANT",
      "ANGLE_EQUIL_VALUE", "ANGLES_INC_HYDROGEN", "ANGLES_WITHOUT_HYDROGEN",
      "DIHEDRAL_FORCE_CONSTANT", "DIHEDRAL_PERIODICITY", "DIHEDRAL_PHASE",
      "DIHEDRALS_INC_HYDROGEN", "DIHEDRALS_WITHOUT_HYDROGEN", "SCEE_SCALE_FACTOR",
      "SCNB_SCALE_FACTOR", "BOX_DIMENSIONS", "ATOMS_PER_MOLECULE", "SOLVENT_POINTERS", "IPOL",
      "HBOND_ACOEF", "HBOND_BCOEF", "HBCUT"};
  return s;
}

const std::set<std::string> &ignoredFlags() {
  static const std::set<std::string> s = {
      "SOLTY", "TREE_CHAIN_CLASSIFICATION", "JOIN_ARRAY", "IROTAT", "RADIUS_SET", "RADII",
      "SCREEN", "RESIDUE_CHAINID", "RESIDUE_NUMBER", "RESIDUE_ICODE", "ATOM_NUMBER",
      "ATOM_BFACTOR", "ATOM_OCCUPANCY", "ATOM_ELEMENT", "FORCE_FIELD_TYPE_COMMENT",
      "CHARMM_FORCE_FIELD_NAME"};
  return s;
}

// REFUSED_FLAGS (shared/amber/flags.js:92-130): the flag families that change
// the energy and are not implemented, each with what to do instead.
std::string refusalFor(const std::string &name) {
  struct Family {
    const char *prefix;
    const char *message;
  };
  static const char *cmap =
      "this topology carries CMAP backbone-correction terms (%FLAG CMAP_*), which this engine "
      "does not implement (ff19SB or a CHARMM conversion). Rebuild with leaprc.protein.ff14SB.";
  static const char *chamber =
      "this topology was converted from CHARMM by chamber (%FLAG CTITLE / CHARMM_*); its "
      "Urey-Bradley terms, impropers, CMAP and 1-4 LJ table are not implemented.";
  static const char *cap =
      "this topology uses a solvent cap (%FLAG CAP_INFO), which this engine does not implement. "
      "Re-solvate with solvateBox or solvateOct.";
  static const Family families[] = {
      {"CMAP_COUNT", cmap}, {"CMAP_RESOLUTION", cmap}, {"CMAP_PARAMETER", cmap},
      {"CMAP_INDEX", cmap},
      {"LENNARD_JONES_CCOEF",
       "this topology uses the 12-6-4 ion model (%FLAG LENNARD_JONES_CCOEF), whose r^-4 term "
       "is not implemented. Rebuild the ions with ordinary 12-6 parameters."},
      {"CTITLE", chamber}, {"CHARMM_UREY_BRADLEY_COUNT", chamber},
      {"CHARMM_UREY_BRADLEY", chamber}, {"CHARMM_UREY_BRADLEY_FORCE_CONSTANT", chamber},
      {"CHARMM_UREY_BRADLEY_EQUIL_VALUE", chamber}, {"CHARMM_NUM_IMPROPERS", chamber},
      {"CHARMM_IMPROPERS", chamber}, {"CHARMM_NUM_IMPR_TYPES", chamber},
      {"CHARMM_IMPROPER_FORCE_CONSTANT", chamber}, {"CHARMM_IMPROPER_PHASE", chamber},
      {"CHARMM_CMAP_COUNT", chamber}, {"CHARMM_CMAP_RESOLUTION", chamber},
      {"CHARMM_CMAP_PARAMETER", chamber}, {"CHARMM_CMAP_INDEX", chamber},
      {"LENNARD_JONES_14_ACOEF", chamber}, {"LENNARD_JONES_14_BCOEF", chamber},
      {"POLARIZABILITY",
       "this topology carries atomic polarizabilities (%FLAG POLARIZABILITY); this engine is "
       "fixed-charge. Rebuild it as fixed-charge."},
      {"CAP_INFO", cap}, {"CAP_INFO2", cap},
      {"AMOEBA_FORCEFIELD",
       "this is an AMOEBA topology (%FLAG AMOEBA_FORCEFIELD); nothing in this engine applies."},
  };
  for (const Family &f : families) {
    std::string p = f.prefix;
    if (name == p || name.compare(0, p.size() + 1, p + "_") == 0) return f.message;
  }
  return "";
}

const std::vector<double> &require(const std::map<std::string, Section> &s, const char *flag) {
  auto it = s.find(flag);
  if (it == s.end()) throw Error(RMD_E_PARSE, std::string("AMBER topology is missing required %FLAG ") + flag);
  return it->second.numbers;
}

const std::vector<std::string> &requireStrings(const std::map<std::string, Section> &s, const char *flag) {
  auto it = s.find(flag);
  if (it == s.end()) throw Error(RMD_E_PARSE, std::string("AMBER topology is missing required %FLAG ") + flag);
  return it->second.strings;
}

const Section *optional(const std::map<std::string, Section> &s, const char *flag) {
  auto it = s.find(flag);
  return it == s.end() ? nullptr : &it->second;
}

// An index in a bonded array is 3*atom and may carry a sign
// (shared/amber/prmtop.js:49-51).
int atomFromStride(double value) { return static_cast<int>(std::fabs(value)) / 3; }

void buildMolecules(Topology &t) {
  std::vector<int> parent(t.atomCount);
  std::iota(parent.begin(), parent.end(), 0);
  auto find = [&](int node) {
    int root = node;
    while (parent[root] != root) root = parent[root];
    while (parent[node] != root) {
      int next = parent[node];
      parent[node] = root;
      node = next;
    }
    return root;
  };
  for (size_t i = 0; i < t.bonds.a.size(); ++i) {
    int a = find(t.bonds.a[i]), b = find(t.bonds.b[i]);
    if (a != b) parent[b] = a;
  }
  std::map<int, int> rootToMolecule;
  t.moleculeOfAtom.assign(t.atomCount, 0);
  for (int i = 0; i < t.atomCount; ++i) {
    int root = find(i);
    auto it = rootToMolecule.find(root);
    int m;
    if (it == rootToMolecule.end()) {
      m = static_cast<int>(rootToMolecule.size());
      rootToMolecule[root] = m;
    } else {
      m = it->second;
    }
    t.moleculeOfAtom[i] = m;
  }
  t.moleculeCount = static_cast<int>(rootToMolecule.size());
}

}  // namespace

// parsePrmtop (shared/amber/prmtop.js:111-385).
Topology parsePrmtop(const std::string &text) {
  std::vector<std::string> order;
  auto sections = readSections(text, order);
  if (!sections.count("POINTERS"))
    throw Error(RMD_E_PARSE, "not an AMBER parm7 topology: no %FLAG POINTERS");
  const auto &p = require(sections, "POINTERS");
  auto ptr = [&](int index) { return index < static_cast<int>(p.size()) ? static_cast<int>(p[index]) : 0; };
  const int NATOM = ptr(0), NTYPES = ptr(1), NRES = ptr(11), IFPERT = ptr(20), IFBOX = ptr(27),
            IFCAP = ptr(29), NUMEXTRA = ptr(30), NCOPY = ptr(31);
  if (NATOM <= 0) throw Error(RMD_E_PARSE, "AMBER topology declares NATOM = " + std::to_string(NATOM));
  if (IFPERT != 0)
    throw Error(RMD_E_PARSE, "this topology carries perturbation (IFPERT != 0) parameters, which "
                             "this engine does not read.");

  std::vector<std::string> refusals, unrecognised;
  std::set<std::string> seen;
  for (const std::string &name : order) {
    if (handledFlags().count(name) || ignoredFlags().count(name)) continue;
    std::string r = refusalFor(name);
    if (!r.empty()) {
      if (!seen.count(r)) {
        seen.insert(r);
        refusals.push_back(r);
      }
      continue;
    }
    unrecognised.push_back(name);
  }
  if (!refusals.empty()) {
    std::string m = refusals[0];
    for (size_t i = 1; i < refusals.size(); ++i) m += " ALSO: " + refusals[i];
    throw Error(RMD_E_PARSE, m);
  }
  if (!unrecognised.empty()) {
    std::string m = "this topology carries " + std::to_string(unrecognised.size()) +
                    " %FLAG section(s) this engine does not know:";
    for (auto &u : unrecognised) m += " " + u;
    throw Error(RMD_E_PARSE, m + ". It is refused rather than ignored.");
  }
  if (NUMEXTRA > 0)
    throw Error(RMD_E_PARSE, "this topology has " + std::to_string(NUMEXTRA) +
                                 " extra points (NUMEXTRA): a 4- or 5-point water. This engine has no "
                                 "virtual-site support; rebuild with a 3-point water.");
  if (const Section *ipol = optional(sections, "IPOL")) {
    if (!ipol->numbers.empty() && ipol->numbers[0] != 0)
      throw Error(RMD_E_PARSE, "this topology declares IPOL != 0 (polarizable force field).");
  }
  if (IFCAP != 0) throw Error(RMD_E_PARSE, "this topology declares IFCAP != 0, a solvent cap.");
  if (NCOPY > 1) throw Error(RMD_E_PARSE, "this topology declares NCOPY > 1 (LES).");

  Topology t;
  t.atomCount = NATOM;
  t.typeCount = NTYPES;
  t.residueCount = NRES;
  const auto &names = requireStrings(sections, "ATOM_NAME");
  for (int i = 0; i < NATOM; ++i) t.atomNames.push_back(i < static_cast<int>(names.size()) ? trim(names[i]) : "");
  const auto &charge = require(sections, "CHARGE");
  const auto &mass = require(sections, "MASS");
  if (static_cast<int>(charge.size()) < NATOM || static_cast<int>(mass.size()) < NATOM)
    throw Error(RMD_E_PARSE, "CHARGE or MASS is shorter than NATOM");
  t.charges.resize(NATOM);
  t.masses.resize(NATOM);
  for (int i = 0; i < NATOM; ++i) {
    t.charges[i] = charge[i] / AMBER_CHARGE_SCALE;
    t.masses[i] = mass[i];
  }
  t.atomicNumbers.assign(NATOM, 0);
  if (const Section *z = optional(sections, "ATOMIC_NUMBER")) {
    for (int i = 0; i < NATOM; ++i) t.atomicNumbers[i] = static_cast<int>(z->numbers.at(i));
  } else {
    // Pre-2010 files: hydrogens by mass (shared/amber/prmtop.js:228-236).
    for (int i = 0; i < NATOM; ++i) t.atomicNumbers[i] = (t.masses[i] > 0 && t.masses[i] < 3.5) ? 1 : 0;
  }
  const auto &typeIndex = require(sections, "ATOM_TYPE_INDEX");
  t.atomTypes.resize(NATOM);
  for (int i = 0; i < NATOM; ++i) t.atomTypes[i] = static_cast<int>(typeIndex.at(i)) - 1;
  if (const Section *tn = optional(sections, "AMBER_ATOM_TYPE"))
    for (int i = 0; i < NATOM && i < static_cast<int>(tn->strings.size()); ++i)
      t.atomTypeNames.push_back(trim(tn->strings[i]));

  const auto &rlabel = requireStrings(sections, "RESIDUE_LABEL");
  const auto &rptr = require(sections, "RESIDUE_POINTER");
  for (int r = 0; r < NRES; ++r) {
    t.residueNames.push_back(trim(rlabel.at(r)));
    t.residueFirstAtom.push_back(static_cast<int>(rptr.at(r)) - 1);
  }
  t.residueOfAtom.assign(NATOM, 0);
  for (int r = 0; r < NRES; ++r) {
    int start = t.residueFirstAtom[r], end = r + 1 < NRES ? t.residueFirstAtom[r + 1] : NATOM;
    for (int a = start; a < end; ++a) t.residueOfAtom[a] = r;
  }

  // Lennard-Jones: a dense T*T table of (A, B) in internal units
  // (shared/amber/prmtop.js:258-294; shared/units.js:99-104).
  const auto &ico = require(sections, "NONBONDED_PARM_INDEX");
  const auto &rawA = require(sections, "LENNARD_JONES_ACOEF");
  const auto &rawB = require(sections, "LENNARD_JONES_BCOEF");
  t.ljA.assign(static_cast<size_t>(NTYPES) * NTYPES, 0.0);
  t.ljB.assign(static_cast<size_t>(NTYPES) * NTYPES, 0.0);
  int hbondPairs = 0;
  const double scaleA = KCAL_TO_KJ * std::pow(NM_PER_ANGSTROM, 12);
  const double scaleB = KCAL_TO_KJ * std::pow(NM_PER_ANGSTROM, 6);
  for (int i = 0; i < NTYPES; ++i)
    for (int j = 0; j < NTYPES; ++j) {
      int parm = static_cast<int>(ico.at(static_cast<size_t>(NTYPES) * i + j));
      if (parm > 0) {
        t.ljA[NTYPES * i + j] = rawA.at(parm - 1) * scaleA;
        t.ljB[NTYPES * i + j] = rawB.at(parm - 1) * scaleB;
      } else if (parm < 0) {
        ++hbondPairs;
      }
    }
  if (hbondPairs > 0)
    throw Error(RMD_E_PARSE, "topology uses the legacy 10-12 hydrogen-bond term on " +
                                 std::to_string(hbondPairs) + " type pairs, which is not implemented.");

  // Bonds: INC_HYDROGEN first, then WITHOUT (prmtop.js:299-305, 398-418).
  const auto &bondK = require(sections, "BOND_FORCE_CONSTANT");
  const auto &bondR0 = require(sections, "BOND_EQUIL_VALUE");
  auto readBonds = [&](const char *flag, bool withH) {
    const Section *s = optional(sections, flag);
    if (!s) return;
    if (s->numbers.size() % 3 != 0) throw Error(RMD_E_PARSE, std::string(flag) + " length is not a multiple of 3");
    for (size_t i = 0; i < s->numbers.size(); i += 3) {
      int parm = static_cast<int>(s->numbers[i + 2]) - 1;
      t.bonds.a.push_back(atomFromStride(s->numbers[i]));
      t.bonds.b.push_back(atomFromStride(s->numbers[i + 1]));
      t.bonds.k.push_back(bondK.at(parm) * BOND_K_SCALE);
      t.bonds.r0.push_back(bondR0.at(parm) * NM_PER_ANGSTROM);
      t.bonds.containsHydrogen.push_back(withH ? 1 : 0);
    }
  };
  readBonds("BONDS_INC_HYDROGEN", true);
  readBonds("BONDS_WITHOUT_HYDROGEN", false);

  const auto &angleK = require(sections, "ANGLE_FORCE_CONSTANT");
  const auto &angleT = require(sections, "ANGLE_EQUIL_VALUE");
  auto readAngles = [&](const char *flag) {
    const Section *s = optional(sections, flag);
    if (!s) return;
    if (s->numbers.size() % 4 != 0) throw Error(RMD_E_PARSE, std::string(flag) + " length is not a multiple of 4");
    for (size_t i = 0; i < s->numbers.size(); i += 4) {
      int parm = static_cast<int>(s->numbers[i + 3]) - 1;
      t.angles.a.push_back(atomFromStride(s->numbers[i]));
      t.angles.b.push_back(atomFromStride(s->numbers[i + 1]));
      t.angles.c.push_back(atomFromStride(s->numbers[i + 2]));
      t.angles.k.push_back(angleK.at(parm) * ANGLE_K_SCALE);
      t.angles.theta0.push_back(angleT.at(parm));
    }
  };
  readAngles("ANGLES_INC_HYDROGEN");
  readAngles("ANGLES_WITHOUT_HYDROGEN");

  // Dihedrals with the 1-4 bookkeeping their signs carry
  // (prmtop.js:451-511): a negative third atom skips the 1-4, a negative
  // fourth marks an improper (which also skips it).
  const auto &dk = require(sections, "DIHEDRAL_FORCE_CONSTANT");
  const auto &dn = require(sections, "DIHEDRAL_PERIODICITY");
  const auto &dp = require(sections, "DIHEDRAL_PHASE");
  const Section *scee = optional(sections, "SCEE_SCALE_FACTOR");
  const Section *scnb = optional(sections, "SCNB_SCALE_FACTOR");
  std::set<long long> seen14;
  auto readDihedrals = [&](const char *flag) {
    const Section *s = optional(sections, flag);
    if (!s) return;
    if (s->numbers.size() % 5 != 0) throw Error(RMD_E_PARSE, std::string(flag) + " length is not a multiple of 5");
    for (size_t i = 0; i < s->numbers.size(); i += 5) {
      double rawC = s->numbers[i + 2], rawD = s->numbers[i + 3];
      int a = atomFromStride(s->numbers[i]), b = atomFromStride(s->numbers[i + 1]);
      int c = atomFromStride(rawC), d = atomFromStride(rawD);
      int parm = static_cast<int>(s->numbers[i + 4]) - 1;
      t.dihedrals.a.push_back(a);
      t.dihedrals.b.push_back(b);
      t.dihedrals.c.push_back(c);
      t.dihedrals.d.push_back(d);
      t.dihedrals.k.push_back(dk.at(parm) * KCAL_TO_KJ);
      t.dihedrals.periodicity.push_back(std::fabs(dn.at(parm)));
      t.dihedrals.phase.push_back(dp.at(parm));
      t.dihedrals.isImproper.push_back(rawD < 0 ? 1 : 0);
      if (rawC < 0 || rawD < 0) continue;
      int low = std::min(a, d), high = std::max(a, d);
      long long key = static_cast<long long>(low) * 0x4000000LL + high;
      if (seen14.count(key)) continue;
      seen14.insert(key);
      t.pairs14.a.push_back(low);
      t.pairs14.b.push_back(high);
      t.pairs14.scee.push_back(scee ? scee->numbers.at(parm) : 1.2);
      t.pairs14.scnb.push_back(scnb ? scnb->numbers.at(parm) : 2.0);
    }
  };
  readDihedrals("DIHEDRALS_INC_HYDROGEN");
  readDihedrals("DIHEDRALS_WITHOUT_HYDROGEN");

  // Exclusions: a count of one holding zero means none (prmtop.js:524-543).
  const auto &nex = require(sections, "NUMBER_EXCLUDED_ATOMS");
  const auto &exl = require(sections, "EXCLUDED_ATOMS_LIST");
  t.exclusionStart.assign(NATOM + 1, 0);
  size_t cursor = 0;
  for (int atom = 0; atom < NATOM; ++atom) {
    t.exclusionStart[atom] = static_cast<int>(t.exclusionPartners.size());
    int declared = static_cast<int>(nex.at(atom));
    for (int o = 0; o < declared; ++o) {
      int value = static_cast<int>(exl.at(cursor + o));
      if (value == 0) continue;
      t.exclusionPartners.push_back(value - 1);
    }
    cursor += declared;
  }
  t.exclusionStart[NATOM] = static_cast<int>(t.exclusionPartners.size());

  // BOX_DIMENSIONS is [beta, a, b, c] (prmtop.js:331-347).
  t.isPeriodic = IFBOX > 0;
  if (const Section *bd = optional(sections, "BOX_DIMENSIONS")) {
    if (IFBOX > 0 && bd->numbers.size() >= 4) {
      t.hasBox = true;
      for (int k = 0; k < 3; ++k) {
        t.boxLengths[k] = bd->numbers[1 + k] * NM_PER_ANGSTROM;
        t.boxAngles[k] = bd->numbers[0];
      }
    }
  }
  buildMolecules(t);
  if (const Section *apm = optional(sections, "ATOMS_PER_MOLECULE")) {
    if (static_cast<int>(apm->numbers.size()) != t.moleculeCount)
      t.warnings.push_back("ATOMS_PER_MOLECULE disagrees with the bond graph; using the bond graph.");
  }
  return t;
}

// parseInpcrd (shared/amber/inpcrd.js:34-132).
Coordinates parseInpcrd(const std::string &text, int expectedAtomCount) {
  if (text.size() >= 3 && text.compare(0, 3, "CDF") == 0)
    throw Error(RMD_E_PARSE, "this is a NetCDF restart (binary), not an ASCII inpcrd. Re-save it as "
                             "ASCII (ntxo = 1).");
  std::vector<std::string> lines = splitLines(text);
  if (lines.size() < 2) throw Error(RMD_E_PARSE, "inpcrd is too short to hold an atom count");
  std::istringstream header(lines[1]);
  std::string countField, timeField;
  header >> countField >> timeField;
  char *end = nullptr;
  long n = std::strtol(countField.c_str(), &end, 10);
  if (countField.empty() || n <= 0)
    throw Error(RMD_E_PARSE, "inpcrd declares an unusable atom count: \"" + lines[1] + "\"");
  Coordinates c;
  c.atomCount = static_cast<int>(n);
  if (!timeField.empty()) {
    c.time = std::strtod(timeField.c_str(), nullptr);
    c.hasTime = true;
  }
  if (expectedAtomCount > 0 && expectedAtomCount != c.atomCount)
    throw Error(RMD_E_PARSE, "coordinate file has " + std::to_string(c.atomCount) +
                                 " atoms but the topology has " + std::to_string(expectedAtomCount) +
                                 ". These two files are not from the same tleap run - re-save the pair together.");
  std::vector<double> numbers;
  for (size_t li = 2; li < lines.size(); ++li) {
    const std::string &line = lines[li];
    if (trim(line).empty()) continue;
    for (int field = 0; field < 6; ++field) {
      size_t start = static_cast<size_t>(field) * 12;
      if (start >= line.size()) break;
      std::string f = trim(line.substr(start, 12));
      if (f.empty()) continue;
      char *e = nullptr;
      double v = std::strtod(f.c_str(), &e);
      if (e == f.c_str() || *e != '\0' || !std::isfinite(v))
        throw Error(RMD_E_PARSE, "non-numeric coordinate field \"" + f + "\" on line " + std::to_string(li + 1));
      numbers.push_back(v);
    }
  }
  const size_t coordinateCount = static_cast<size_t>(c.atomCount) * 3;
  const size_t total = numbers.size();
  if (total == coordinateCount) {
  } else if (total == coordinateCount + 6) {
    c.hasBox = true;
  } else if (total == coordinateCount * 2) {
    c.hasVelocities = true;
  } else if (total == coordinateCount * 2 + 6) {
    c.hasVelocities = true;
    c.hasBox = true;
  } else {
    throw Error(RMD_E_PARSE, "inpcrd holds " + std::to_string(total) + " numbers, which is none of " +
                                 std::to_string(coordinateCount) + ", " + std::to_string(coordinateCount + 6) +
                                 ", " + std::to_string(coordinateCount * 2) + " or " +
                                 std::to_string(coordinateCount * 2 + 6) + " for " +
                                 std::to_string(c.atomCount) + " atoms.");
  }
  c.positions.resize(coordinateCount);
  for (size_t i = 0; i < coordinateCount; ++i) c.positions[i] = numbers[i] * NM_PER_ANGSTROM;
  if (c.hasVelocities) {
    c.velocities.resize(coordinateCount);
    for (size_t i = 0; i < coordinateCount; ++i)
      c.velocities[i] = numbers[coordinateCount + i] * AMBER_VELOCITY_SCALE * NM_PER_ANGSTROM;
  }
  if (c.hasBox) {
    size_t o = total - 6;
    for (int k = 0; k < 3; ++k) {
      c.boxLengths[k] = numbers[o + k] * NM_PER_ANGSTROM;
      c.boxAngles[k] = numbers[o + 3 + k];
    }
  }
  return c;
}

// JavaScript's Math.round: halves go towards +infinity.
static double jsRound(double x) {
  double f = std::floor(x);
  return (x - f >= 0.5) ? f + 1.0 : f;
}

// boxVectorsFromLengthsAndAngles (shared/amber/inpcrd.js:154-196), with
// OpenMM's reduction.
void boxVectorsFromLengthsAndAngles(const double lengths[3], const double anglesDeg[3], double out[9]) {
  const double a = lengths[0], b = lengths[1], c = lengths[2];
  const double toRad = M_PI / 180.0;
  const double alpha = anglesDeg[0] * toRad, beta = anglesDeg[1] * toRad, gamma = anglesDeg[2] * toRad;
  double va[3] = {a, 0, 0};
  double vb[3] = {b * std::cos(gamma), b * std::sin(gamma), 0};
  double cx = c * std::cos(beta);
  double cy = c * (std::cos(alpha) - std::cos(beta) * std::cos(gamma)) / std::sin(gamma);
  double cz2 = c * c - cx * cx - cy * cy;
  if (cz2 <= 0) throw Error(RMD_E_PARSE, "box lengths and angles do not describe a realisable cell");
  double vc[3] = {cx, cy, std::sqrt(cz2)};
  double *vectors[3] = {va, vb, vc};
  for (double *v : vectors)
    for (int k = 0; k < 3; ++k)
      if (std::fabs(v[k]) < 1e-12 * a) v[k] = 0;
  auto reduce = [](double *target, const double *source, int component) {
    double shift = jsRound(target[component] / source[component]);
    if (shift == 0) return;
    for (int k = 0; k < 3; ++k) target[k] -= shift * source[k];
  };
  reduce(vc, vb, 1);
  reduce(vc, va, 0);
  reduce(vb, va, 0);
  for (int k = 0; k < 3; ++k) {
    out[k] = va[k];
    out[3 + k] = vb[k];
    out[6 + k] = vc[k];
  }
}

}  // namespace rmd
