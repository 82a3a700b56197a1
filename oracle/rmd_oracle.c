/*
 * rmd_oracle.c - TEST INFRASTRUCTURE ONLY: the FP64 checker.
 *
 * A plain-C restatement of RainMD's host reference (the cpu/ JavaScript and the shared/
 * math it uses). Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it, and only to check or to time the reference
 * algorithm; the product never does.
 *
 * What pins each function (tests/test_oracle_pins.py, test_oracle_openmm.py,
 * test_oracle_pins_host.py; no function here is unpinned):
 *   oracle_erfc / oracle_erf      SciPy special.erfc/erf values, incl. the
 *                                 reference's erfc(2.628261) known answer.
 *   oracle_alpha_from_tolerance   alpha(1 nm, 5e-4) = 2.628261.
 *   oracle_potential (bonded)     OpenMM force-group energies in the fixtures
 *                                 (1e-10); central differences of itself;
 *                                 bonded forces sum to zero.
 *   oracle_potential (LJ, Coulomb, OpenMM ljOnly / coulombOnly (1e-9 / 1e-7)
 *     self and net-charge terms)  on ala-vac (NoCutoff) and ala-wat (PME);
 *                                 forces vs OpenMM (1e-6 of rms).
 *   reciprocal (naive DFT PME)    the plain Ewald k-sum (oracle_ewald_reciprocal)
 *                                 at a fine grid, on a rectangular and a sheared
 *                                 box; OpenMM Ewald fixture for the k-sum.
 *   oracle_bspline / moduli       partition of unity, zero derivative sum,
 *                                 closed-form moduli of [1,11,11,1]/24.
 *   oracle_dispersion             the reference's -41.537158 kJ/mol and 1/V.
 *   oracle_mix / uniform /        the reference module's own output vectors
 *     gaussian                    (tests/golden/random), bit for bit; mix is
 *                                 lowbias32 (its published inverse undoes it).
 *   oracle_langevin_*             one step against the closed-form B/A/O
 *                                 update with supplied noise; the thermostat's
 *                                 decay closed form; constraint lengths.
 *   oracle_shake / oracle_rattle  constraint lengths and velocity projections
 *                                 to the tolerance; SETTLE as their limit.
 *   oracle_settle                 converged SHAKE (tolerance 1e-13) on the
 *                                 same input; exact constraint lengths.
 *   oracle_kinetic                sum of m v^2 / 2 for unit velocities by
 *                                 hand; T = 2 KE / (dof k_B).
 *   oracle_velocities             per-type Maxwell-Boltzmann moments
 *                                 (mean 0, variance kT/m) over many atoms.
 *   oracle_skin_for /             the reference's rules restated by hand at
 *     oracle_plan_cells           the workload sizes (cell counts floor(w/rl),
 *                                 stencil, capacity), the skin's cap and floor.
 *   oracle_minimise               reaches scipy's L-BFGS minimum of the same
 *                                 potential; monotone descent, displacement
 *                                 cap, first trial moves the largest force by
 *                                 max_step, constraints projected.
 *   oracle_cells_f32 / pairs_f32  brute force FP64 distances away from the
 *                                 boundary; partition and sortedness.
 *   oracle_pairs_f32_binned       the all-pairs finder's sets (rectangular
 *                                 and triclinic boxes).
 *   oracle_scale_molecules        molecule geometry preserved, centroids scaled.
 *   oracle_metropolis_weight      the reference's sign known-answers.
 */
#include "rmd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ONE_4PI_EPS0 138.935456
#define BOLTZMANN 0.008314462618
#define PME_ORDER 5

/* ---- erfc (shared/erf.js:17-107, W. J. Cody's CALERF) -------------------- */

static const double CA[5] = {3.16112374387056560e+00, 1.13864154151050156e+02, 3.77485237685302021e+02,
                             3.20937758913846947e+03, 1.85777706184603153e-01};
static const double CB[4] = {2.36012909523441209e+01, 2.44024637934444173e+02, 1.28261652607737228e+03,
                             2.84423683343917062e+03};
static const double CC[9] = {5.64188496988670089e-01, 8.88314979438837594e+00, 6.61191906371416295e+01,
                             2.98635138197400131e+02, 8.81952221241769090e+02, 1.71204761263407058e+03,
                             2.05107837782607147e+03, 1.23033935479799725e+03, 2.15311535474403846e-08};
static const double CD[8] = {1.57449261107098347e+01, 1.17693950891312499e+02, 5.37181101862009858e+02,
                             1.62138957456669019e+03, 3.29079923573345963e+03, 4.36261909014324716e+03,
                             3.43936767414372164e+03, 1.23033935480374942e+03};
static const double CP[6] = {3.05326634961232344e-01, 3.60344899949804439e-01, 1.25781726111229246e-01,
                             1.60837851487422766e-02, 6.58749161529837803e-04, 1.63153871373020978e-02};
static const double CQ[5] = {2.56852019228982242e+00, 1.87295284992346047e+00, 5.27905102951428412e-01,
                             6.05183413124413191e-02, 2.33520497626869185e-03};

static double scaled_exp(double y) {
  double z = trunc(y * 16.0) / 16.0;
  double del = (y - z) * (y + z);
  return exp(-z * z) * exp(-del);
}

double oracle_erfc(double x) {
  double y = fabs(x), result;
  if (y <= 0.46875) {
    double z = y > 1e-20 ? y * y : 0;
    double num = CA[4] * z, den = z;
    for (int i = 0; i < 3; ++i) {
      num = (num + CA[i]) * z;
      den = (den + CB[i]) * z;
    }
    return 1.0 - (x * (num + CA[3])) / (den + CB[3]);
  }
  if (y <= 4.0) {
    double num = CC[8] * y, den = y;
    for (int i = 0; i < 7; ++i) {
      num = (num + CC[i]) * y;
      den = (den + CD[i]) * y;
    }
    result = ((num + CC[7]) / (den + CD[7])) * scaled_exp(y);
  } else if (y >= 26.543) {
    result = 0;
  } else {
    double z = 1.0 / (y * y);
    double num = CP[5] * z, den = z;
    for (int i = 0; i < 4; ++i) {
      num = (num + CP[i]) * z;
      den = (den + CQ[i]) * z;
    }
    result = (z * (num + CP[4])) / (den + CQ[4]);
    result = ((5.6418958354775628695e-1 - result) / y) * scaled_exp(y);
  }
  return x < 0 ? 2.0 - result : result;
}

double oracle_erf(double x) { return 1.0 - oracle_erfc(x); }

/* ---- geometry (cpu/forces/geometry.js:23-69) ----------------------------- */

/* JavaScript's Math.round: halves towards +infinity. */
static double js_round(double x) {
  double f = floor(x);
  return (x - f >= 0.5) ? f + 1.0 : f;
}

static void minimum_image(double d[3], const double *box) {
  if (!box) return;
  double s2 = js_round(d[2] / box[8]);
  d[0] -= s2 * box[6];
  d[1] -= s2 * box[7];
  d[2] -= s2 * box[8];
  double s1 = js_round(d[1] / box[4]);
  d[0] -= s1 * box[3];
  d[1] -= s1 * box[4];
  double s0 = js_round(d[0] / box[0]);
  d[0] -= s0 * box[0];
}

static void reciprocal_box(const double *box, double *r) {
  double ax = box[0], bx = box[3], by = box[4], cx = box[6], cy = box[7], cz = box[8];
  r[0] = 1 / ax;
  r[1] = 0;
  r[2] = 0;
  r[3] = -bx / (ax * by);
  r[4] = 1 / by;
  r[5] = 0;
  r[6] = (bx * cy - by * cx) / (ax * by * cz);
  r[7] = -cy / (by * cz);
  r[8] = 1 / cz;
}

/* ---- FP32 distance test, this build's fixed operation order ---------------
 * d = p_i - p_j; image along c, then b, then a with s = rint(d*inv), the
 * nearest integer (the reference's floor(d/L + 0.5) differs only at
 * half-box separations, beyond the list cutoff);
 * r2 = (dx*dx + dy*dy) + dz*dz. No fused multiply-add anywhere (the file is
 * compiled with -ffp-contract=off). Emulates the device's list test and the
 * reference's minimumImage order (webgpu/kernels/wgsl-common.js:81-88).     */
float oracle_r2_f32(const float *pi, const float *pj, const float box9[9], const float inv3[3]) {
  float dx = pi[0] - pj[0], dy = pi[1] - pj[1], dz = pi[2] - pj[2];
  float s2 = rintf(dz * inv3[2]);
  dx = dx - s2 * box9[6];
  dy = dy - s2 * box9[7];
  dz = dz - s2 * box9[8];
  float s1 = rintf(dy * inv3[1]);
  dx = dx - s1 * box9[3];
  dy = dy - s1 * box9[4];
  float s0 = rintf(dx * inv3[0]);
  dx = dx - s0 * box9[0];
  float xx = dx * dx, yy = dy * dy, zz = dz * dz;
  float sum = xx + yy;
  return sum + zz;
}

/* cellOfPosition (webgpu/kernels/wgsl-common.js:112-123) in f32. */
void oracle_cells_f32(int32_t n, const float *pos, const float recip[9], const int32_t counts[3],
                      int32_t *cell_of_atom) {
  for (int32_t i = 0; i < n; ++i) {
    const float x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
    int c[3];
    for (int d = 0; d < 3; ++d) {
      float a = x * recip[d], b = y * recip[3 + d], cz = z * recip[6 + d];
      float f = (a + b) + cz;
      float w = f - floorf(f);
      float scaled = w * (float)counts[d];
      int ci = (int)scaled;
      if (ci > counts[d] - 1) ci = counts[d] - 1;
      if (ci < 0) ci = 0;
      c[d] = ci;
    }
    cell_of_atom[i] = (c[0] * counts[1] + c[1]) * counts[2] + c[2];
  }
}

static int is_excluded(const int32_t *start, const int32_t *partners, int32_t i, int32_t j) {
  int32_t lo = start[i], hi = start[i + 1];
  while (lo < hi) {
    int32_t mid = lo + (hi - lo) / 2;
    if (partners[mid] == j) return 1;
    if (partners[mid] < j) lo = mid + 1;
    else hi = mid;
  }
  return 0;
}

int32_t oracle_pairs_f32(int32_t n, const float *pos, const float box9[9], const float inv3[3], float rl2,
                         const int32_t *excl_start, const int32_t *excl_partners, int32_t cap, int32_t *counts,
                         int32_t *lists) {
  int32_t widest = 0;
  for (int32_t i = 0; i < n; ++i) {
    int32_t c = 0;
    for (int32_t j = 0; j < n; ++j) {
      if (j == i) continue;
      if (oracle_r2_f32(pos + 3 * i, pos + 3 * j, box9, inv3) > rl2) continue;
      if (is_excluded(excl_start, excl_partners, i, j)) continue;
      if (lists && c < cap) lists[(int64_t)i * cap + c] = j;
      ++c;
    }
    counts[i] = c;
    if (c > widest) widest = c;
  }
  return widest;
}

/* ---- bonded (cpu/forces/bonded.js:22-207) --------------------------------- */

static double bonds(const oracle_system *s, const double *p, double *f) {
  double e = 0;
  for (int i = 0; i < s->nbond; ++i) {
    int a = 3 * s->bond_a[i], b = 3 * s->bond_b[i];
    double dx = p[a] - p[b], dy = p[a + 1] - p[b + 1], dz = p[a + 2] - p[b + 2];
    double r = sqrt(dx * dx + dy * dy + dz * dz);
    double st = r - s->bond_r0[i];
    e += 0.5 * s->bond_k[i] * st * st;
    double sc = r > 0 ? s->bond_k[i] * st / r : 0;
    f[a] -= sc * dx;
    f[a + 1] -= sc * dy;
    f[a + 2] -= sc * dz;
    f[b] += sc * dx;
    f[b + 1] += sc * dy;
    f[b + 2] += sc * dz;
  }
  return e;
}

static double angles(const oracle_system *s, const double *p, double *f) {
  double e = 0;
  for (int i = 0; i < s->nangle; ++i) {
    int a = 3 * s->angle_a[i], b = 3 * s->angle_b[i], c = 3 * s->angle_c[i];
    double v1x = p[a] - p[b], v1y = p[a + 1] - p[b + 1], v1z = p[a + 2] - p[b + 2];
    double v2x = p[c] - p[b], v2y = p[c + 1] - p[b + 1], v2z = p[c + 2] - p[b + 2];
    double n1 = v1x * v1x + v1y * v1y + v1z * v1z, n2 = v2x * v2x + v2y * v2y + v2z * v2z;
    double nx = v1y * v2z - v1z * v2y, ny = v1z * v2x - v1x * v2z, nz = v1x * v2y - v1y * v2x;
    double cm = sqrt(nx * nx + ny * ny + nz * nz);
    double dot = v1x * v2x + v1y * v2y + v1z * v2z;
    double theta = atan2(cm, dot);
    double dev = theta - s->angle_theta0[i];
    e += 0.5 * s->angle_k[i] * dev * dev;
    double dE = s->angle_k[i] * dev;
    double safe = cm > 1e-12 ? cm : 1e-12;
    double s1 = -dE / (n1 * safe), s2 = dE / (n2 * safe);
    double fax = s1 * (v1y * nz - v1z * ny), fay = s1 * (v1z * nx - v1x * nz), faz = s1 * (v1x * ny - v1y * nx);
    double fcx = s2 * (v2y * nz - v2z * ny), fcy = s2 * (v2z * nx - v2x * nz), fcz = s2 * (v2x * ny - v2y * nx);
    f[a] += fax;
    f[a + 1] += fay;
    f[a + 2] += faz;
    f[c] += fcx;
    f[c + 1] += fcy;
    f[c + 2] += fcz;
    f[b] -= fax + fcx;
    f[b + 1] -= fay + fcy;
    f[b + 2] -= faz + fcz;
  }
  return e;
}

static double torsions(const oracle_system *s, const double *p, double *f) {
  double e = 0;
  for (int i = 0; i < s->ntorsion; ++i) {
    int a = 3 * s->tor_a[i], b = 3 * s->tor_b[i], c = 3 * s->tor_c[i], d = 3 * s->tor_d[i];
    double b1x = p[b] - p[a], b1y = p[b + 1] - p[a + 1], b1z = p[b + 2] - p[a + 2];
    double b2x = p[c] - p[b], b2y = p[c + 1] - p[b + 1], b2z = p[c + 2] - p[b + 2];
    double b3x = p[d] - p[c], b3y = p[d + 1] - p[c + 1], b3z = p[d + 2] - p[c + 2];
    double n1x = b1y * b2z - b1z * b2y, n1y = b1z * b2x - b1x * b2z, n1z = b1x * b2y - b1y * b2x;
    double n2x = b2y * b3z - b2z * b3y, n2y = b2z * b3x - b2x * b3z, n2z = b2x * b3y - b2y * b3x;
    double n1s = n1x * n1x + n1y * n1y + n1z * n1z, n2s = n2x * n2x + n2y * n2y + n2z * n2z;
    double b2l = sqrt(b2x * b2x + b2y * b2y + b2z * b2z);
    if (n1s < 1e-24 || n2s < 1e-24 || b2l < 1e-12) continue;
    double dn = n1x * n2x + n1y * n2y + n1z * n2z;
    double cr = (n1y * n2z - n1z * n2y) * b2x + (n1z * n2x - n1x * n2z) * b2y + (n1x * n2y - n1y * n2x) * b2z;
    double phi = atan2(cr / b2l, dn);
    double n = s->tor_n[i];
    double arg = n * phi - s->tor_phase[i];
    e += s->tor_k[i] * (1 + cos(arg));
    double dE = -s->tor_k[i] * n * sin(arg);
    double fA = dE * b2l / n1s, fD = -dE * b2l / n2s;
    double fax = fA * n1x, fay = fA * n1y, faz = fA * n1z;
    double fdx = fD * n2x, fdy = fD * n2y, fdz = fD * n2z;
    double inv = 1 / (b2l * b2l);
    double t = (b1x * b2x + b1y * b2y + b1z * b2z) * inv;
    double u = (b3x * b2x + b3y * b2y + b3z * b2z) * inv;
    f[a] += fax;
    f[a + 1] += fay;
    f[a + 2] += faz;
    f[d] += fdx;
    f[d + 1] += fdy;
    f[d + 2] += fdz;
    f[b] += -(1 + t) * fax + u * fdx;
    f[b + 1] += -(1 + t) * fay + u * fdy;
    f[b + 2] += -(1 + t) * faz + u * fdz;
    f[c] += t * fax - (1 + u) * fdx;
    f[c + 1] += t * fay - (1 + u) * fdy;
    f[c + 2] += t * faz - (1 + u) * fdz;
  }
  return e;
}

/* Positional restraint k*|minImage(r - r0)|^2 (shared/topology/restraints.js:98-126). */
static double restraint(const oracle_system *s, const oracle_options *o, const double *p, double *f) {
  if (!(o->restraint_k > 0) || o->restraint_count <= 0) return 0;
  double e = 0;
  for (int k = 0; k < o->restraint_count; ++k) {
    int a = o->restraint_atoms[k];
    double d[3] = {p[3 * a] - o->restraint_reference[3 * k], p[3 * a + 1] - o->restraint_reference[3 * k + 1],
                   p[3 * a + 2] - o->restraint_reference[3 * k + 2]};
    if (s->periodic) minimum_image(d, s->box);
    e += o->restraint_k * (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int c = 0; c < 3; ++c) f[3 * a + c] -= 2 * o->restraint_k * d[c];
  }
  return e;
}

/* ---- direct space and exceptions (cpu/forces/nonbonded.js:24-162) -------- */

typedef struct {
  const int32_t *start, *partners;
} excl_set;

static int pair_excluded(const oracle_system *s, const excl_set *x, int i, int j) {
  (void)s;
  return is_excluded(x->start, x->partners, i, j);
}

static void build_exclusions(const oracle_system *s, excl_set *x, int32_t **start_out, int32_t **part_out) {
  int n = s->n;
  int32_t *count = (int32_t *)calloc(n + 1, sizeof(int32_t));
  for (int e = 0; e < s->nexception; ++e) {
    count[s->exc_a[e]]++;
    count[s->exc_b[e]]++;
  }
  int32_t *start = (int32_t *)calloc(n + 1, sizeof(int32_t));
  for (int i = 0; i < n; ++i) start[i + 1] = start[i] + count[i];
  int32_t *part = (int32_t *)malloc(sizeof(int32_t) * (start[n] + 1));
  int32_t *cur = (int32_t *)malloc(sizeof(int32_t) * (n + 1));
  memcpy(cur, start, sizeof(int32_t) * n);
  for (int e = 0; e < s->nexception; ++e) {
    part[cur[s->exc_a[e]]++] = s->exc_b[e];
    part[cur[s->exc_b[e]]++] = s->exc_a[e];
  }
  for (int i = 0; i < n; ++i) /* insertion sort per atom */
    for (int a = start[i] + 1; a < start[i + 1]; ++a) {
      int32_t v = part[a];
      int b = a - 1;
      while (b >= start[i] && part[b] > v) {
        part[b + 1] = part[b];
        --b;
      }
      part[b + 1] = v;
    }
  free(count);
  free(cur);
  x->start = start;
  x->partners = part;
  *start_out = start;
  *part_out = part;
}

static void pair_energy(const oracle_system *s, int i, int j, const double *d, double r2, double alpha,
                        double *lj, double *coul, double *f) {
  double r = sqrt(r2), ir = 1 / r, ir2 = ir * ir, ir6 = ir2 * ir2 * ir2;
  int tp = s->types * s->atom_types[i] + s->atom_types[j];
  double A = s->lj_a[tp], B = s->lj_b[tp];
  *lj += A * ir6 * ir6 - B * ir6;
  double sc = (12 * A * ir6 * ir6 - 6 * B * ir6) * ir2;
  double qq = ONE_4PI_EPS0 * s->charges[i] * s->charges[j];
  if (alpha <= 0) {
    *coul += qq * ir;
    sc += qq * ir * ir2;
  } else {
    double ar = alpha * r, scr = oracle_erfc(ar);
    *coul += qq * scr * ir;
    double g = 2 * alpha * exp(-ar * ar) / sqrt(M_PI);
    sc += qq * (scr * ir + g) * ir2;
  }
  if (f) {
    for (int c = 0; c < 3; ++c) {
      f[3 * i + c] += sc * d[c];
      f[3 * j + c] -= sc * d[c];
    }
  }
}

typedef struct {
  float box9[9], inv3[3], rc2;
} f32_geometry;

static void make_f32_geometry(const oracle_system *s, f32_geometry *g) {
  for (int k = 0; k < 9; ++k) g->box9[k] = (float)s->box[k];
  g->inv3[0] = (float)(1.0 / s->box[0]);
  g->inv3[1] = (float)(1.0 / s->box[4]);
  g->inv3[2] = (float)(1.0 / s->box[8]);
  g->rc2 = (float)(s->cutoff * s->cutoff);
}

static int within_cutoff(const oracle_system *s, const oracle_options *o, const f32_geometry *g, const double *p,
                         int i, int j, double r2, double rc2) {
  if (!o->f32_membership) return r2 <= rc2;
  (void)s;
  float pi[3] = {(float)p[3 * i], (float)p[3 * i + 1], (float)p[3 * i + 2]};
  float pj[3] = {(float)p[3 * j], (float)p[3 * j + 1], (float)p[3 * j + 2]};
  return oracle_r2_f32(pi, pj, g->box9, g->inv3) <= g->rc2;
}

static void direct_all_pairs(const oracle_system *s, const oracle_options *o, const double *p, double alpha,
                             double cutoff, const excl_set *x, double *lj, double *coul, double *f) {
  const double *box = s->periodic ? s->box : NULL;
  double rc2 = cutoff * cutoff;
  f32_geometry g;
  if (s->periodic) make_f32_geometry(s, &g);
  for (int i = 0; i < s->n; ++i)
    for (int j = i + 1; j < s->n; ++j) {
      if (pair_excluded(s, x, i, j)) continue;
      double d[3] = {p[3 * i] - p[3 * j], p[3 * i + 1] - p[3 * j + 1], p[3 * i + 2] - p[3 * j + 2]};
      minimum_image(d, box);
      double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (isfinite(cutoff) && !within_cutoff(s, o, &g, p, i, j, r2, rc2)) continue;
      pair_energy(s, i, j, d, r2, alpha, lj, coul, f);
    }
}

/* An FP64 cell-list direct space for large rectangular boxes: the same pair
 * test and pair arithmetic as direct_all_pairs, visiting only the 27
 * neighbouring cells (requires >= 3 cells per axis so no cell repeats). */
static int direct_cells(const oracle_system *s, const oracle_options *o, const double *p, double alpha,
                        const excl_set *x, double *lj, double *coul, double *f) {
  if (!s->periodic || s->box[3] != 0 || s->box[6] != 0 || s->box[7] != 0) return -1;
  int nc[3];
  for (int k = 0; k < 3; ++k) {
    nc[k] = (int)floor(s->box[4 * k] / s->cutoff);
    if (nc[k] < 3) return -1;
  }
  int ncell = nc[0] * nc[1] * nc[2];
  int *head = (int *)malloc(sizeof(int) * ncell), *next = (int *)malloc(sizeof(int) * s->n);
  for (int c = 0; c < ncell; ++c) head[c] = -1;
  int *cellOf = (int *)malloc(sizeof(int) * s->n);
  for (int i = s->n - 1; i >= 0; --i) {
    int c3[3];
    for (int k = 0; k < 3; ++k) {
      double L = s->box[4 * k];
      double w = p[3 * i + k] / L;
      w -= floor(w);
      int ci = (int)(w * nc[k]);
      if (ci >= nc[k]) ci = nc[k] - 1;
      c3[k] = ci;
    }
    int c = (c3[0] * nc[1] + c3[1]) * nc[2] + c3[2];
    cellOf[i] = c;
    next[i] = head[c];
    head[c] = i;
  }
  double rc2 = s->cutoff * s->cutoff;
  f32_geometry g;
  make_f32_geometry(s, &g);
  for (int i = 0; i < s->n; ++i) {
    int c = cellOf[i];
    int cx = c / (nc[1] * nc[2]), cy = (c / nc[2]) % nc[1], cz = c % nc[2];
    for (int dx = -1; dx <= 1; ++dx)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dz = -1; dz <= 1; ++dz) {
          int ox = (cx + dx + nc[0]) % nc[0], oy = (cy + dy + nc[1]) % nc[1], oz = (cz + dz + nc[2]) % nc[2];
          for (int j = head[(ox * nc[1] + oy) * nc[2] + oz]; j >= 0; j = next[j]) {
            if (j <= i) continue;
            if (pair_excluded(s, x, i, j)) continue;
            double d[3] = {p[3 * i] - p[3 * j], p[3 * i + 1] - p[3 * j + 1], p[3 * i + 2] - p[3 * j + 2]};
            minimum_image(d, s->box);
            double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (!within_cutoff(s, o, &g, p, i, j, r2, rc2)) continue;
            pair_energy(s, i, j, d, r2, alpha, lj, coul, f);
          }
        }
  }
  free(head);
  free(next);
  free(cellOf);
  return 0;
}

static void exceptions(const oracle_system *s, const double *p, double alpha, double *lj, double *coul, double *f) {
  const double *box = s->periodic ? s->box : NULL;
  for (int e = 0; e < s->nexception; ++e) {
    int i = s->exc_a[e], j = s->exc_b[e];
    double d[3] = {p[3 * i] - p[3 * j], p[3 * i + 1] - p[3 * j + 1], p[3 * i + 2] - p[3 * j + 2]};
    minimum_image(d, box);
    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    double r = sqrt(r2), ir = 1 / r, ir2 = ir * ir, ir6 = ir2 * ir2 * ir2, sc = 0;
    double A = s->exc_lja[e], B = s->exc_ljb[e];
    if (A != 0 || B != 0) {
      *lj += A * ir6 * ir6 - B * ir6;
      sc += (12 * A * ir6 * ir6 - 6 * B * ir6) * ir2;
    }
    if (s->exc_qq[e] != 0) {
      double qq = ONE_4PI_EPS0 * s->exc_qq[e];
      *coul += qq * ir;
      sc += qq * ir * ir2;
    }
    if (alpha > 0) {
      double qf = ONE_4PI_EPS0 * s->charges[i] * s->charges[j];
      if (qf != 0) {
        double ar = alpha * r, un = oracle_erf(ar);
        *coul -= qf * un * ir;
        double g = 2 * alpha * exp(-ar * ar) / sqrt(M_PI);
        sc -= qf * (un * ir - g) * ir2;
      }
    }
    if (f)
      for (int c = 0; c < 3; ++c) {
        f[3 * i + c] += sc * d[c];
        f[3 * j + c] -= sc * d[c];
      }
  }
}

/* ---- PME (shared/pme/parameters.js, cpu/forces/pme.js) ------------------- */

double oracle_alpha_from_tolerance(double cutoff, double tolerance) {
  return sqrt(-log(2 * tolerance)) / cutoff;
}

void oracle_bspline(double offset, int32_t order, double *w, double *d) {
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

int oracle_moduli(int32_t size, int32_t order, double *out) {
  double z[16], dd[16];
  oracle_bspline(0.0, order, z, dd);
  double *m = (double *)malloc(sizeof(double) * size);
  for (int i = 0; i < size; ++i) {
    double re = 0, im = 0;
    for (int k = 0; k < order; ++k) {
      double ang = 2 * M_PI * i * k / size;
      re += z[k] * cos(ang);
      im += z[k] * sin(ang);
    }
    m[i] = re * re + im * im;
  }
  int ok = 0;
  for (int i = 0; i < size; ++i) {
    out[i] = m[i] < 1e-7 ? 0.5 * (m[(i - 1 + size) % size] + m[(i + 1) % size]) : m[i];
    if (!(out[i] > 0)) ok = -1;
  }
  free(m);
  return ok;
}

/* dftAlongAxis (cpu/forces/pme.js:28-70): naive O(n^2) per line. */
static void dft_axis(double *data, const int *grid, int axis, int sign) {
  int nx = grid[0], ny = grid[1], nz = grid[2];
  int n = grid[axis];
  int strides[3] = {ny * nz, nz, 1};
  int stride = strides[axis];
  double *line = (double *)malloc(sizeof(double) * 2 * n);
  double *ct = (double *)malloc(sizeof(double) * n), *st = (double *)malloc(sizeof(double) * n);
  for (int k = 0; k < n; ++k) {
    double ang = sign * 2 * M_PI * k / n;
    ct[k] = cos(ang);
    st[k] = sin(ang);
  }
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z) {
        int coord[3] = {x, y, z};
        if (coord[axis] != 0) continue;
        int base = x * strides[0] + y * strides[1] + z * strides[2];
        for (int k = 0; k < n; ++k) {
          line[2 * k] = data[2 * (base + k * stride)];
          line[2 * k + 1] = data[2 * (base + k * stride) + 1];
        }
        for (int m = 0; m < n; ++m) {
          double re = 0, im = 0;
          for (int k = 0; k < n; ++k) {
            int tw = (int)(((long long)m * k) % n);
            re += line[2 * k] * ct[tw] - line[2 * k + 1] * st[tw];
            im += line[2 * k] * st[tw] + line[2 * k + 1] * ct[tw];
          }
          data[2 * (base + m * stride)] = re;
          data[2 * (base + m * stride) + 1] = im;
        }
      }
  free(line);
  free(ct);
  free(st);
}

/* reciprocalEnergy (cpu/forces/pme.js:137-259). */
static double reciprocal(const oracle_system *s, const double *p, double alpha, const int *grid, double *f) {
  const int nx = grid[0], ny = grid[1], nz = grid[2], order = PME_ORDER;
  const int n = s->n;
  double recip[9];
  reciprocal_box(s->box, recip);
  const double volume = s->box[0] * s->box[4] * s->box[8];
  double *q = (double *)calloc((size_t)2 * nx * ny * nz, sizeof(double));
  int *base = (int *)malloc(sizeof(int) * 3 * n);
  double *w = (double *)malloc(sizeof(double) * 3 * n * order), *dw = (double *)malloc(sizeof(double) * 3 * n * order);
  for (int a = 0; a < n; ++a) {
    for (int d = 0; d < 3; ++d) {
      double fr = p[3 * a] * recip[d] + p[3 * a + 1] * recip[3 + d] + p[3 * a + 2] * recip[6 + d];
      double wr = fr - floor(fr);
      double sc = wr * grid[d];
      double in = floor(sc);
      base[3 * a + d] = (int)in;
      oracle_bspline(sc - in, order, w + (3 * a + d) * order, dw + (3 * a + d) * order);
    }
    double qa = s->charges[a];
    for (int kx = 0; kx < order; ++kx) {
      int ix = ((base[3 * a] + kx - order + 1) % nx + nx) % nx;
      double wx = w[(3 * a) * order + kx];
      for (int ky = 0; ky < order; ++ky) {
        int iy = ((base[3 * a + 1] + ky - order + 1) % ny + ny) % ny;
        double wxy = wx * w[(3 * a + 1) * order + ky];
        for (int kz = 0; kz < order; ++kz) {
          int iz = ((base[3 * a + 2] + kz - order + 1) % nz + nz) % nz;
          q[2 * ((ix * ny + iy) * nz + iz)] += qa * wxy * w[(3 * a + 2) * order + kz];
        }
      }
    }
  }
  dft_axis(q, grid, 0, 1);
  dft_axis(q, grid, 1, 1);
  dft_axis(q, grid, 2, 1);
  double *mod[3];
  for (int d = 0; d < 3; ++d) {
    mod[d] = (double *)malloc(sizeof(double) * grid[d]);
    oracle_moduli(grid[d], order, mod[d]);
  }
  const double factor = M_PI * M_PI / (alpha * alpha);
  const double pref = ONE_4PI_EPS0 / (2 * M_PI * volume);
  double energy = 0;
  for (int mx = 0; mx < nx; ++mx) {
    int kx = mx <= nx / 2 ? mx : mx - nx;
    for (int my = 0; my < ny; ++my) {
      int ky = my <= ny / 2 ? my : my - ny;
      for (int mz = 0; mz < nz; ++mz) {
        int kz = mz <= nz / 2 ? mz : mz - nz;
        size_t idx = (size_t)(mx * ny + my) * nz + mz;
        if (kx == 0 && ky == 0 && kz == 0) {
          q[2 * idx] = q[2 * idx + 1] = 0;
          continue;
        }
        double mxc = kx * recip[0] + ky * recip[1] + kz * recip[2];
        double myc = kx * recip[3] + ky * recip[4] + kz * recip[5];
        double mzc = kx * recip[6] + ky * recip[7] + kz * recip[8];
        double m2 = mxc * mxc + myc * myc + mzc * mzc;
        double modulus = mod[0][mx] * mod[1][my] * mod[2][mz];
        double kern = exp(-factor * m2) / (m2 * modulus);
        double re = q[2 * idx], im = q[2 * idx + 1];
        energy += pref * kern * (re * re + im * im);
        q[2 * idx] = re * kern * 2 * pref;
        q[2 * idx + 1] = im * kern * 2 * pref;
      }
    }
  }
  if (f) {
    dft_axis(q, grid, 0, -1);
    dft_axis(q, grid, 1, -1);
    dft_axis(q, grid, 2, -1);
    for (int a = 0; a < n; ++a) {
      double qa = s->charges[a];
      if (qa == 0) continue;
      double gu = 0, gv = 0, gw = 0;
      for (int kx = 0; kx < order; ++kx) {
        int ix = ((base[3 * a] + kx - order + 1) % nx + nx) % nx;
        double wx = w[(3 * a) * order + kx], dx = dw[(3 * a) * order + kx];
        for (int ky = 0; ky < order; ++ky) {
          int iy = ((base[3 * a + 1] + ky - order + 1) % ny + ny) % ny;
          double wy = w[(3 * a + 1) * order + ky], dy = dw[(3 * a + 1) * order + ky];
          for (int kz = 0; kz < order; ++kz) {
            int iz = ((base[3 * a + 2] + kz - order + 1) % nz + nz) % nz;
            double wz = w[(3 * a + 2) * order + kz], dz = dw[(3 * a + 2) * order + kz];
            double v = q[2 * ((ix * ny + iy) * nz + iz)];
            gu += dx * wy * wz * v;
            gv += wx * dy * wz * v;
            gw += wx * wy * dz * v;
          }
        }
      }
      gu *= grid[0] * qa;
      gv *= grid[1] * qa;
      gw *= grid[2] * qa;
      f[3 * a] -= gu * recip[0] + gv * recip[1] + gw * recip[2];
      f[3 * a + 1] -= gu * recip[3] + gv * recip[4] + gw * recip[5];
      f[3 * a + 2] -= gu * recip[6] + gv * recip[7] + gw * recip[8];
    }
  }
  for (int d = 0; d < 3; ++d) free(mod[d]);
  free(q);
  free(base);
  free(w);
  free(dw);
  return energy;
}

/* Plain Ewald reciprocal sum (textbook; independent of the grid). */
double oracle_ewald_reciprocal(const oracle_system *s, const double *p, double alpha, int32_t kmax, double *f) {
  double recip[9];
  reciprocal_box(s->box, recip);
  const double volume = s->box[0] * s->box[4] * s->box[8];
  const double factor = M_PI * M_PI / (alpha * alpha);
  const double pref = ONE_4PI_EPS0 / (2 * M_PI * volume);
  double energy = 0;
  for (int kx = -kmax; kx <= kmax; ++kx)
    for (int ky = -kmax; ky <= kmax; ++ky)
      for (int kz = -kmax; kz <= kmax; ++kz) {
        if (kx == 0 && ky == 0 && kz == 0) continue;
        double m[3] = {kx * recip[0] + ky * recip[1] + kz * recip[2], kx * recip[3] + ky * recip[4] + kz * recip[5],
                       kx * recip[6] + ky * recip[7] + kz * recip[8]};
        double m2 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
        double kern = exp(-factor * m2) / m2;
        if (kern < 1e-30) continue;
        double sr = 0, si = 0;
        for (int a = 0; a < s->n; ++a) {
          double ph = 2 * M_PI * (m[0] * p[3 * a] + m[1] * p[3 * a + 1] + m[2] * p[3 * a + 2]);
          sr += s->charges[a] * cos(ph);
          si += s->charges[a] * sin(ph);
        }
        energy += pref * kern * (sr * sr + si * si);
        if (f)
          for (int a = 0; a < s->n; ++a) {
            double ph = 2 * M_PI * (m[0] * p[3 * a] + m[1] * p[3 * a + 1] + m[2] * p[3 * a + 2]);
            /* dE/dr_a = pref*kern*2*(sr*dsr + si*dsi), dsr = -q sin(ph) 2pi m, dsi = q cos(ph) 2pi m */
            double g = pref * kern * 2 * s->charges[a] * 2 * M_PI * (-sr * sin(ph) + si * cos(ph));
            for (int c = 0; c < 3; ++c) f[3 * a + c] -= g * m[c];
          }
      }
  return energy;
}

void oracle_self_energy(const oracle_system *s, double alpha, double volume, double *self, double *net) {
  double q2 = 0, q = 0;
  for (int a = 0; a < s->n; ++a) {
    q2 += s->charges[a] * s->charges[a];
    q += s->charges[a];
  }
  *self = -ONE_4PI_EPS0 * alpha * q2 / sqrt(M_PI);
  *net = -ONE_4PI_EPS0 * M_PI * q * q / (2 * alpha * alpha * volume);
}

double oracle_dispersion(const oracle_system *s, double volume, double cutoff) {
  double *counts = (double *)calloc(s->types, sizeof(double));
  for (int a = 0; a < s->n; ++a) counts[s->atom_types[a]] += 1;
  double oA = 0, oB = 0, dA = 0, dB = 0;
  for (int i = 0; i < s->types; ++i) {
    for (int j = 0; j < s->types; ++j) {
      oA += counts[i] * counts[j] * s->lj_a[s->types * i + j];
      oB += counts[i] * counts[j] * s->lj_b[s->types * i + j];
    }
    dA += counts[i] * s->lj_a[s->types * i + i];
    dB += counts[i] * s->lj_b[s->types * i + i];
  }
  free(counts);
  double scale = (double)s->n / (s->n + 1);
  double A = (oA + dA) * scale, B = (oB + dB) * scale;
  if (!(volume > 0) || !(cutoff > 0)) return 0;
  double c3 = cutoff * cutoff * cutoff, c9 = c3 * c3 * c3;
  return (2 * M_PI / volume) * (A / (9 * c9) - B / (3 * c3));
}

int oracle_potential(const oracle_system *s, const double *pos, const oracle_options *o, oracle_terms *out,
                     double *forces) {
  memset(out, 0, sizeof *out);
  double *f = forces;
  if (f) memset(f, 0, sizeof(double) * 3 * s->n);
  int terms = o->terms ? o->terms : ORC_ALL;
  int pme = o->method == 0;
  if (pme && !s->periodic) return -1;
  if (terms & ORC_RESTRAINT) out->restraint = restraint(s, o, pos, f ? f : NULL);
  double *scratch = NULL;
  if (!f) {
    scratch = (double *)calloc(3 * (size_t)s->n, sizeof(double));
    f = scratch;
  }
  if (terms & ORC_BONDED) {
    out->bond = bonds(s, pos, f);
    out->angle = angles(s, pos, f);
    out->torsion = torsions(s, pos, f);
  }
  double alpha = pme ? o->alpha : 0;
  double cutoff = pme ? s->cutoff : INFINITY;
  if (terms & ORC_DIRECT) {
    int32_t *st, *pa;
    excl_set x;
    build_exclusions(s, &x, &st, &pa);
    int done = -1;
    if (o->use_cells && pme) done = direct_cells(s, o, pos, alpha, &x, &out->lj_direct, &out->coulomb_direct, f);
    if (done != 0) direct_all_pairs(s, o, pos, alpha, cutoff, &x, &out->lj_direct, &out->coulomb_direct, f);
    free(st);
    free(pa);
  }
  if (terms & ORC_EXCEPTIONS) exceptions(s, pos, alpha, &out->lj_exception, &out->coulomb_exception, f);
  if (pme) {
    double volume = s->box[0] * s->box[4] * s->box[8];
    if (terms & ORC_RECIPROCAL) out->reciprocal = reciprocal(s, pos, o->alpha, o->grid, f);
    oracle_self_energy(s, o->alpha, volume, &out->self, &out->net);
  }
  out->lennard_jones = out->lj_direct + out->lj_exception;
  out->coulomb_direct_total = out->coulomb_direct + out->coulomb_exception;
  out->coulomb = out->coulomb_direct_total + out->reciprocal + out->self + out->net;
  out->bonded = out->bond + out->angle + out->torsion;
  out->total = out->bonded + out->lennard_jones + out->coulomb + out->restraint;
  free(scratch);
  return 0;
}

/* ---- noise (shared/random.js:24-87) --------------------------------------- */

uint32_t oracle_mix(uint32_t v) {
  v ^= v >> 16;
  v *= 0x7feb352du;
  v ^= v >> 15;
  v *= 0x846ca68bu;
  v ^= v >> 16;
  return v;
}

double oracle_uniform(uint32_t seed, uint32_t a, uint32_t b, uint32_t c) {
  uint32_t st = oracle_mix(seed ^ (a * 0x9e3779b1u));
  st = oracle_mix(st ^ (b * 0x85ebca6bu));
  st = oracle_mix(st ^ (c * 0xc2b2ae35u));
  return ((double)(st >> 9) + 0.5) * (1.0 / 8388608.0);
}

void oracle_gaussian(uint32_t seed, uint32_t atom, uint32_t step, double out[3]) {
  double u1 = oracle_uniform(seed, atom, step, 0), u2 = oracle_uniform(seed, atom, step, 1);
  double u3 = oracle_uniform(seed, atom, step, 2), u4 = oracle_uniform(seed, atom, step, 3);
  double r1 = sqrt(-2 * log(u1)), r2 = sqrt(-2 * log(u3));
  out[0] = r1 * cos(2 * M_PI * u2);
  out[1] = r1 * sin(2 * M_PI * u2);
  out[2] = r2 * cos(2 * M_PI * u4);
}

/* ---- constraints (cpu/integrate/constraints.js:93-222) ------------------- */

static double inv_mass(const oracle_system *s, int a) { return s->masses[a] > 0 ? 1.0 / s->masses[a] : 0.0; }

int oracle_shake(const oracle_system *s, double *p, const double *ref, double *corr, double tol, int32_t maxit) {
  if (s->nconstraint == 0) return 0;
  for (int it = 0; it < maxit; ++it) {
    double worst = 0;
    for (int k = 0; k < s->nconstraint; ++k) {
      int a = s->con_a[k], b = s->con_b[k];
      double t = s->con_d[k];
      double dx = p[3 * a] - p[3 * b], dy = p[3 * a + 1] - p[3 * b + 1], dz = p[3 * a + 2] - p[3 * b + 2];
      double err = dx * dx + dy * dy + dz * dz - t * t;
      double rel = fabs(err) / (t * t);
      if (rel > worst) worst = rel;
      if (rel <= tol) continue;
      double rx = ref[3 * a] - ref[3 * b], ry = ref[3 * a + 1] - ref[3 * b + 1], rz = ref[3 * a + 2] - ref[3 * b + 2];
      double proj = dx * rx + dy * ry + dz * rz;
      double ia = inv_mass(s, a), ib = inv_mass(s, b);
      double den = 2 * (ia + ib) * proj;
      if (fabs(den) < 1e-20) continue;
      double lam = -err / den;
      double sa[3] = {lam * ia * rx, lam * ia * ry, lam * ia * rz}, sb[3] = {lam * ib * rx, lam * ib * ry, lam * ib * rz};
      for (int c = 0; c < 3; ++c) {
        p[3 * a + c] += sa[c];
        p[3 * b + c] -= sb[c];
        if (corr) {
          corr[3 * a + c] += sa[c];
          corr[3 * b + c] -= sb[c];
        }
      }
    }
    if (worst <= tol) return it + 1;
  }
  return -1;
}

int oracle_rattle(const oracle_system *s, const double *p, double *v, double tol, int32_t maxit) {
  if (s->nconstraint == 0) return 0;
  for (int it = 0; it < maxit; ++it) {
    double worst = 0;
    for (int k = 0; k < s->nconstraint; ++k) {
      int a = s->con_a[k], b = s->con_b[k];
      double rx = p[3 * a] - p[3 * b], ry = p[3 * a + 1] - p[3 * b + 1], rz = p[3 * a + 2] - p[3 * b + 2];
      double vx = v[3 * a] - v[3 * b], vy = v[3 * a + 1] - v[3 * b + 1], vz = v[3 * a + 2] - v[3 * b + 2];
      double proj = rx * vx + ry * vy + rz * vz;
      double t = s->con_d[k];
      double rel = fabs(proj) / (t * t);
      if (rel > worst) worst = rel;
      if (rel <= tol) continue;
      double ia = inv_mass(s, a), ib = inv_mass(s, b);
      double den = (ia + ib) * (rx * rx + ry * ry + rz * rz);
      if (den < 1e-20) continue;
      double lam = -proj / den;
      v[3 * a] += lam * ia * rx;
      v[3 * a + 1] += lam * ia * ry;
      v[3 * a + 2] += lam * ia * rz;
      v[3 * b] -= lam * ib * rx;
      v[3 * b + 1] -= lam * ib * ry;
      v[3 * b + 2] -= lam * ib * rz;
    }
    if (worst <= tol) return it + 1;
  }
  return -1;
}

/* SETTLE (Miyamoto & Kollman 1992): the rigid-water position constraint in
 * closed form. The constraint displacements lie along the old bond vectors,
 * so this is the limit SHAKE converges to. */
void oracle_settle(const double old_pos[9], double np[9], double mO, double mH, double dOH, double dHH) {
  const double wohh = mO + 2 * mH;
  const double rc = 0.5 * dHH;
  const double h = sqrt(dOH * dOH - rc * rc);
  const double ra = 2 * mH * h / wohh, rb = h - ra;
  double b0[3], c0[3], com[3], a1[3], b1[3], c1[3];
  for (int k = 0; k < 3; ++k) {
    b0[k] = old_pos[3 + k] - old_pos[k];
    c0[k] = old_pos[6 + k] - old_pos[k];
    com[k] = (mO * np[k] + mH * (np[3 + k] + np[6 + k])) / wohh;
    a1[k] = np[k] - com[k];
    b1[k] = np[3 + k] - com[k];
    c1[k] = np[6 + k] - com[k];
  }
  double ez[3] = {b0[1] * c0[2] - b0[2] * c0[1], b0[2] * c0[0] - b0[0] * c0[2], b0[0] * c0[1] - b0[1] * c0[0]};
  double ex[3] = {a1[1] * ez[2] - a1[2] * ez[1], a1[2] * ez[0] - a1[0] * ez[2], a1[0] * ez[1] - a1[1] * ez[0]};
  double ey[3] = {ez[1] * ex[2] - ez[2] * ex[1], ez[2] * ex[0] - ez[0] * ex[2], ez[0] * ex[1] - ez[1] * ex[0]};
  double lx = 1 / sqrt(ex[0] * ex[0] + ex[1] * ex[1] + ex[2] * ex[2]);
  double ly = 1 / sqrt(ey[0] * ey[0] + ey[1] * ey[1] + ey[2] * ey[2]);
  double lz = 1 / sqrt(ez[0] * ez[0] + ez[1] * ez[1] + ez[2] * ez[2]);
  for (int k = 0; k < 3; ++k) {
    ex[k] *= lx;
    ey[k] *= ly;
    ez[k] *= lz;
  }
#define DOT(u, v) ((u)[0] * (v)[0] + (u)[1] * (v)[1] + (u)[2] * (v)[2])
  double xb0 = DOT(b0, ex), yb0 = DOT(b0, ey), xc0 = DOT(c0, ex), yc0 = DOT(c0, ey);
  double za1 = DOT(a1, ez);
  double xb1 = DOT(b1, ex), yb1 = DOT(b1, ey), zb1 = DOT(b1, ez);
  double xc1 = DOT(c1, ex), yc1 = DOT(c1, ey), zc1 = DOT(c1, ez);
#undef DOT
  double sinphi = za1 / ra;
  double cosphi = sqrt(1 - sinphi * sinphi);
  double sinpsi = (zb1 - zc1) / (2 * rc * cosphi);
  double cospsi = sqrt(1 - sinpsi * sinpsi);
  double ya2 = ra * cosphi, xb2 = -rc * cospsi;
  double t1 = -rb * cosphi, t2 = rc * sinpsi * sinphi;
  double yb2 = t1 - t2, yc2 = t1 + t2;
  double alpa = xb2 * (xb0 - xc0) + yb0 * yb2 + yc0 * yc2;
  double beta = xb2 * (yc0 - yb0) + xb0 * yb2 + xc0 * yc2;
  double gama = xb0 * yb1 - xb1 * yb0 + xc0 * yc1 - xc1 * yc0;
  double al2be2 = alpa * alpa + beta * beta;
  double sinthe = (alpa * gama - beta * sqrt(al2be2 - gama * gama)) / al2be2;
  double costhe = sqrt(1 - sinthe * sinthe);
  double a3[3] = {-ya2 * sinthe, ya2 * costhe, za1};
  double b3[3] = {xb2 * costhe - yb2 * sinthe, xb2 * sinthe + yb2 * costhe, zb1};
  double c3[3] = {-xb2 * costhe - yc2 * sinthe, -xb2 * sinthe + yc2 * costhe, zc1};
  for (int k = 0; k < 3; ++k) {
    np[k] = com[k] + a3[0] * ex[k] + a3[1] * ey[k] + a3[2] * ez[k];
    np[3 + k] = com[k] + b3[0] * ex[k] + b3[1] * ey[k] + b3[2] * ez[k];
    np[6 + k] = com[k] + c3[0] * ex[k] + c3[1] * ey[k] + c3[2] * ez[k];
  }
}

/* ---- the integrator (cpu/integrate/langevin.js:30-114) ------------------- */

static int drift(const oracle_system *s, double *p, double *v, double h, const oracle_integrator *o) {
  int n3 = 3 * s->n;
  double *ref = (double *)malloc(sizeof(double) * n3);
  memcpy(ref, p, sizeof(double) * n3);
  for (int k = 0; k < n3; ++k) p[k] += h * v[k];
  int rc = 0;
  if (s->nconstraint > 0) {
    double *corr = (double *)calloc(n3, sizeof(double));
    if (oracle_shake(s, p, ref, corr, o->tolerance, o->max_iterations) < 0) rc = -1;
    for (int k = 0; k < s->nconstraint; ++k) {
      /* v += correction / h for every constrained atom (once per atom: the
       * corrections of an atom are accumulated, so visiting it twice would
       * double it; mark visited atoms by zeroing their correction). */
      int at[2] = {s->con_a[k], s->con_b[k]};
      for (int q = 0; q < 2; ++q) {
        int a = at[q];
        for (int c = 0; c < 3; ++c) {
          v[3 * a + c] += corr[3 * a + c] / h;
          corr[3 * a + c] = 0;
        }
      }
    }
    free(corr);
  }
  free(ref);
  return rc;
}

int oracle_langevin_step(const oracle_system *s, double *p, double *v, const double *force,
                         const oracle_integrator *o) {
  const double dt = o->timestep, h = 0.5 * dt;
  const double decay = exp(-o->friction * dt), noiseScale = sqrt(1 - decay * decay);
  const double kT = BOLTZMANN * o->temperature;
  int rc = 0;
  for (int a = 0; a < s->n; ++a) {
    double m = s->masses[a];
    if (m <= 0) continue;
    for (int c = 0; c < 3; ++c) v[3 * a + c] += h / m * force[3 * a + c];
  }
  if (oracle_rattle(s, p, v, o->tolerance, o->max_iterations) < 0) rc = -1;
  if (drift(s, p, v, h, o) < 0) rc = -1;
  for (int a = 0; a < s->n; ++a) {
    double m = s->masses[a];
    if (m <= 0) continue;
    double g[3];
    if (o->noise) {
      g[0] = o->noise[3 * a];
      g[1] = o->noise[3 * a + 1];
      g[2] = o->noise[3 * a + 2];
    } else {
      oracle_gaussian(o->seed, (uint32_t)a, (uint32_t)o->step, g);
    }
    double amp = noiseScale * sqrt(kT / m);
    for (int c = 0; c < 3; ++c) v[3 * a + c] = decay * v[3 * a + c] + amp * g[c];
  }
  if (oracle_rattle(s, p, v, o->tolerance, o->max_iterations) < 0) rc = -1;
  if (drift(s, p, v, h, o) < 0) rc = -1;
  return rc;
}

int oracle_langevin_finish(const oracle_system *s, double *p, double *v, const double *force,
                           const oracle_integrator *o) {
  const double h = 0.5 * o->timestep;
  for (int a = 0; a < s->n; ++a) {
    double m = s->masses[a];
    if (m <= 0) continue;
    for (int c = 0; c < 3; ++c) v[3 * a + c] += h / m * force[3 * a + c];
  }
  return oracle_rattle(s, p, v, o->tolerance, o->max_iterations) < 0 ? -1 : 0;
}

double oracle_kinetic(const oracle_system *s, const double *v) {
  double k = 0;
  for (int a = 0; a < s->n; ++a) {
    double m = s->masses[a];
    if (m <= 0) continue;
    k += m * (v[3 * a] * v[3 * a] + v[3 * a + 1] * v[3 * a + 1] + v[3 * a + 2] * v[3 * a + 2]);
  }
  return 0.5 * k;
}

void oracle_velocities(const oracle_system *s, double *v, double temperature, uint32_t seed) {
  const double kT = BOLTZMANN * temperature;
  for (int a = 0; a < s->n; ++a) {
    double m = s->masses[a];
    if (m <= 0) {
      v[3 * a] = v[3 * a + 1] = v[3 * a + 2] = 0;
      continue;
    }
    double g[3];
    oracle_gaussian(seed, (uint32_t)a, 0x5eed, g);
    double sg = sqrt(kT / m);
    for (int c = 0; c < 3; ++c) v[3 * a + c] = sg * g[c];
  }
}

/* ---- barostat (webgpu/kernels/barostat.js:25-162) ------------------------ */

void oracle_scale_molecules(const oracle_system *s, double *p, double scale) {
  for (int m = 0; m < s->nmolecule; ++m) {
    double w[3] = {0, 0, 0}, total = 0;
    for (int k = s->mol_start[m]; k < s->mol_start[m + 1]; ++k) {
      int a = s->mol_atoms[k];
      for (int c = 0; c < 3; ++c) w[c] += s->masses[a] * p[3 * a + c];
      total += s->masses[a];
    }
    if (total > 0) {
      for (int c = 0; c < 3; ++c) w[c] /= total;
    } else {
      int cnt = s->mol_start[m + 1] - s->mol_start[m];
      for (int c = 0; c < 3; ++c) w[c] = 0;
      for (int k = s->mol_start[m]; k < s->mol_start[m + 1]; ++k)
        for (int c = 0; c < 3; ++c) w[c] += p[3 * s->mol_atoms[k] + c];
      for (int c = 0; c < 3; ++c) w[c] /= (cnt > 0 ? cnt : 1);
    }
    for (int k = s->mol_start[m]; k < s->mol_start[m + 1]; ++k) {
      int a = s->mol_atoms[k];
      for (int c = 0; c < 3; ++c) p[3 * a + c] += (scale - 1) * w[c];
    }
  }
}

double oracle_metropolis_weight(double dE, double dV, double ratio, double pressure, double kT, int32_t molecules) {
  return dE + pressure * dV - molecules * kT * log(ratio);
}

/* ---- cell plan (shared/topology/cells.js:25-29,77-181) ------------------ */

/* skinFor (cells.js:77-94) with halfSmallestWidth (:97-101): the measured
 * 0.40 nm anchor scaled linearly by the timestep and by sqrt(T/298) and
 * sqrt(1.008/m_lightest), capped at 0.9 of the room the box leaves beside the
 * cutoff and floored at 0.1 nm. The box widths are its diagonal entries. */
double oracle_skin_for(const oracle_system *s, double timestep, double temperature, const double box[9]) {
  double lightest = INFINITY;
  for (int a = 0; a < s->n; ++a)
    if (s->masses[a] > 0 && s->masses[a] < lightest) lightest = s->masses[a];
  if (!isfinite(lightest)) lightest = 1.008;
  const double scaled = 0.40 * (timestep / 0.002) * sqrt((temperature / 298.0) * (1.008 / lightest));
  double room = INFINITY;
  if (box) {
    double w = fabs(box[0]);
    if (fabs(box[4]) < w) w = fabs(box[4]);
    if (fabs(box[8]) < w) w = fabs(box[8]);
    room = w / 2 - s->cutoff;
  }
  double capped = room > 0 ? room * 0.9 : 0.1;
  double skin = scaled < capped ? scaled : capped;
  return skin > 0.1 ? skin : 0.1;
}

/* axisOffsets (cells.js:25-29): the deduplicated one-deep stencil per axis. */
static int axis_offsets(int cells) { return cells >= 3 ? 3 : cells == 2 ? 2 : 1; }

/* planCellList (cells.js:111-181): floor(width / (cutoff + skin)) cells per
 * axis (at least one), the deduplicated stencil, and the list capacity
 * max(32, ceil(slack * density * 4/3 pi (cutoff + skin)^3)). Returns the
 * capacity; counts, the stencil's entry count and the estimate go out. */
int32_t oracle_plan_cells(const oracle_system *s, double skin, double slack, const double box[9], int32_t counts[3],
                          int32_t *stencil, double *estimated) {
  const double rl = s->cutoff + skin;
  const double widths[3] = {box[0], box[4], box[8]};
  int entries = 1;
  for (int k = 0; k < 3; ++k) {
    int c = (int)floor(fabs(widths[k]) / rl);
    counts[k] = c > 1 ? c : 1;
    entries *= axis_offsets(counts[k]);
  }
  if (stencil) *stencil = entries;
  const double volume = fabs(box[0] * box[4] * box[8]);
  const double estimate = (s->n / volume) * (4.0 / 3.0) * M_PI * rl * rl * rl;
  if (estimated) *estimated = estimate;
  const int cap = (int)ceil(estimate * slack);
  return cap > 32 ? cap : 32;
}

/* ---- minimiser (webgpu/simulation.js:880-979) ------------------------------ */

static double max_atom_force(const oracle_system *s, const double *f) {
  double worst = 0;
  for (int a = 0; a < s->n; ++a) {
    double m = sqrt(f[3 * a] * f[3 * a] + f[3 * a + 1] * f[3 * a + 1] + f[3 * a + 2] * f[3 * a + 2]);
    if (m > worst) worst = m;
  }
  return worst;
}

/* GpuSimulation.minimise (webgpu/simulation.js:880-979): steepest descent in
 * which each trial moves every atom by (max_step / max(worst, 1e-6)) F_i with
 * the displacement capped at max_displacement per atom (displaceShader,
 * webgpu/kernels/integrate.js:653-667), projects the constraints onto the
 * pre-move positions (encodeConstrainPositions :671-688, SHAKE here) and
 * keeps the trial only if the potential went down (max_step x 1.3), else
 * restores the positions and halves max_step, 14 tries per iteration. The
 * largest per-atom force |F_i| decides convergence (forceMagnitudeShader
 * :589-603). restore_forces 0 keeps the rejected trial's forces for the next
 * try, as the reference does; 1 restores them with the positions. */
int oracle_minimise(const oracle_system *s, double *pos, const oracle_options *eo, const oracle_min_options *mo,
                    oracle_min_result *out) {
  const int n3 = 3 * s->n;
  double *f = (double *)calloc(n3, sizeof(double));
  double *fs = (double *)calloc(n3, sizeof(double));
  double *saved = (double *)calloc(n3, sizeof(double));
  oracle_terms t;
  if (!f || !fs || !saved) {
    free(f), free(fs), free(saved);
    return -1;
  }
  double max_step = mo->max_step;
  oracle_potential(s, pos, eo, &t, f);
  double potential = t.total;
  int iterations = 0, rejected = 0, stalled = 0;
  for (int it = 0; it < mo->max_iterations; ++it) {
    iterations = it + 1;
    const double worst = max_atom_force(s, f);
    if (worst <= mo->force_tolerance) break;
    memcpy(saved, pos, sizeof(double) * n3);
    memcpy(fs, f, sizeof(double) * n3);
    int accepted = 0;
    for (int attempt = 0; attempt < 14; ++attempt) {
      const double scale = max_step / (worst > 1e-6 ? worst : 1e-6);
      for (int a = 0; a < s->n; ++a) {
        double d[3] = {scale * f[3 * a], scale * f[3 * a + 1], scale * f[3 * a + 2]};
        double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        double k = len > mo->max_displacement ? mo->max_displacement / len : 1.0;
        for (int c = 0; c < 3; ++c) pos[3 * a + c] = saved[3 * a + c] + k * d[c];
      }
      if (s->nconstraint > 0) oracle_shake(s, pos, saved, NULL, mo->tolerance, 100000);
      oracle_potential(s, pos, eo, &t, f);
      if (t.total < potential) {
        potential = t.total;
        max_step *= 1.3;
        accepted = 1;
        break;
      }
      rejected += 1;
      memcpy(pos, saved, sizeof(double) * n3);
      if (mo->restore_forces) memcpy(f, fs, sizeof(double) * n3);
      max_step *= 0.5;
      if (max_step < 1e-9) break;
    }
    if (!accepted) {
      stalled = 1;
      if (mo->restore_forces) memcpy(f, fs, sizeof(double) * n3);
      break;
    }
  }
  if (out) {
    out->iterations = iterations;
    out->rejected = rejected;
    out->stalled = stalled;
    out->max_force = max_atom_force(s, f);
    out->converged = out->max_force <= mo->force_tolerance;
    out->potential = potential;
    out->max_step = max_step;
  }
  free(f), free(fs), free(saved);
  return 0;
}

/* The brute-force FP32 neighbour sets of oracle_pairs_f32 for large systems:
 * the same pair test (oracle_r2_f32, exclusions, self) over candidates from
 * coarse cells in fractional coordinates whose face-to-face thickness is at
 * least rl + 0.01 nm, so every pair within rl is a candidate (a pair closer
 * than rl differs by less than one cell along every fractional axis). Lists
 * are sorted ascending like the brute force's. box64 the FP64 box rows. */
static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

int32_t oracle_pairs_f32_binned(int32_t n, const float *pos, const double box64[9], const float box9[9],
                                const float inv3[3], float rl2, const int32_t *excl_start,
                                const int32_t *excl_partners, int32_t cap, int32_t *counts, int32_t *lists) {
  const double ax = box64[0], bx = box64[3], by = box64[4], cx = box64[6], cy = box64[7], cz = box64[8];
  const double r[9] = {1 / ax, 0, 0, -bx / (ax * by), 1 / by, 0, (bx * cy - by * cx) / (ax * by * cz), -cy / (by * cz),
                       1 / cz};
  const double rl = sqrt((double)rl2) + 0.01;
  int cells[3];
  for (int k = 0; k < 3; ++k) {
    double face = 1.0 / sqrt(r[k] * r[k] + r[3 + k] * r[3 + k] + r[6 + k] * r[6 + k]);
    cells[k] = (int)floor(face / rl);
    if (cells[k] < 1) cells[k] = 1;
  }
  const int total = cells[0] * cells[1] * cells[2];
  int *cellOf = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  int *start = (int *)calloc(total + 1, sizeof(int));
  int *order = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  if (!cellOf || !start || !order) {
    free(cellOf), free(start), free(order);
    return -1;
  }
  for (int i = 0; i < n; ++i) {
    int c3[3];
    for (int k = 0; k < 3; ++k) {
      double s = pos[3 * i] * r[k] + pos[3 * i + 1] * r[3 + k] + pos[3 * i + 2] * r[6 + k];
      s -= floor(s);
      int ci = (int)(s * cells[k]);
      c3[k] = ci < 0 ? 0 : ci >= cells[k] ? cells[k] - 1 : ci;
    }
    cellOf[i] = (c3[0] * cells[1] + c3[1]) * cells[2] + c3[2];
    start[cellOf[i] + 1] += 1;
  }
  for (int c = 0; c < total; ++c) start[c + 1] += start[c];
  {
    int *cursor = (int *)malloc(sizeof(int) * total);
    memcpy(cursor, start, sizeof(int) * total);
    for (int i = 0; i < n; ++i) order[cursor[cellOf[i]]++] = i;
    free(cursor);
  }
  int32_t widest = 0;
#pragma omp parallel
  {
  int *cand = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  int32_t mine = 0;
#pragma omp for schedule(dynamic, 256)
  for (int i = 0; i < n; ++i) {
    const int c = cellOf[i];
    const int c3[3] = {c / (cells[1] * cells[2]), (c / cells[2]) % cells[1], c % cells[2]};
    int off[3][3], noff[3];
    for (int k = 0; k < 3; ++k) {
      noff[k] = cells[k] >= 3 ? 3 : cells[k];
      for (int o = 0; o < noff[k]; ++o) {
        int d = cells[k] >= 3 ? o - 1 : o;
        off[k][o] = ((c3[k] + d) % cells[k] + cells[k]) % cells[k];
      }
    }
    int m = 0;
    for (int ox = 0; ox < noff[0]; ++ox)
      for (int oy = 0; oy < noff[1]; ++oy)
        for (int oz = 0; oz < noff[2]; ++oz) {
          const int nc = (off[0][ox] * cells[1] + off[1][oy]) * cells[2] + off[2][oz];
          for (int k = start[nc]; k < start[nc + 1]; ++k) {
            const int j = order[k];
            if (j == i) continue;
            if (oracle_r2_f32(pos + 3 * i, pos + 3 * j, box9, inv3) > rl2) continue;
            if (is_excluded(excl_start, excl_partners, i, j)) continue;
            cand[m++] = j;
          }
        }
    qsort(cand, m, sizeof(int), cmp_int);
    if (lists)
      for (int k = 0; k < m && k < cap; ++k) lists[(int64_t)i * cap + k] = cand[k];
    counts[i] = m;
    if (m > mine) mine = m;
  }
#pragma omp critical
  if (mine > widest) widest = mine;
  free(cand);
  }
  free(cellOf), free(start), free(order);
  return widest;
}
