/*
 * rmd_oracle.h - TEST INFRASTRUCTURE ONLY.
 *
 * An FP64 CPU restatement of RainMD's host reference path (cpu/ and the
 * shared/ math it reads), plus an FP32 emulation of the cell binning and the
 * neighbour-list distance test. It is the checker the parity tests, smoke()
 * and the bench's cpu_baseline leg hold the HIP path against; nothing in the
 * product links or calls it.
 *
 * Each function names the reference file:line it restates. Pinning (what
 * independent check holds each function) is listed in the header comment of
 * rmd_oracle.c.
 */
#ifndef RMD_ORACLE_H
#define RMD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_system {
  int32_t n, types;
  const double *charges, *masses;
  const int32_t *atom_types;
  const double *lj_a, *lj_b; /* types*types */
  int32_t nbond;
  const int32_t *bond_a, *bond_b;
  const double *bond_k, *bond_r0;
  int32_t nangle;
  const int32_t *angle_a, *angle_b, *angle_c;
  const double *angle_k, *angle_theta0;
  int32_t ntorsion;
  const int32_t *tor_a, *tor_b, *tor_c, *tor_d;
  const double *tor_k, *tor_n, *tor_phase;
  int32_t nconstraint;
  const int32_t *con_a, *con_b;
  const double *con_d;
  int32_t nexception;
  const int32_t *exc_a, *exc_b;
  const double *exc_qq, *exc_lja, *exc_ljb;
  int32_t nmolecule;
  const int32_t *mol_start, *mol_atoms;
  int32_t periodic;
  double box[9];
  double cutoff;
  int32_t dof;
} oracle_system;

/* Term masks (match the product's RMD_TERM_* values). */
enum { ORC_BONDED = 1, ORC_DIRECT = 2, ORC_EXCEPTIONS = 4, ORC_RECIPROCAL = 8, ORC_RESTRAINT = 16, ORC_ALL = 31 };

typedef struct oracle_options {
  int32_t method;          /* 0 = PME, 1 = NoCutoff                                 */
  double alpha;            /* Ewald alpha (PME)                                     */
  int32_t grid[3];         /* PME grid                                              */
  int32_t terms;           /* ORC_* mask                                            */
  int32_t f32_membership;  /* 1: cutoff membership decided by the FP32 distance test */
  int32_t use_cells;       /* 1: FP64 cell-list direct space (rectangular, >=3 cells) */
  double restraint_k;
  int32_t restraint_count;
  const int32_t *restraint_atoms;
  const double *restraint_reference; /* 3*restraint_count */
} oracle_options;

typedef struct oracle_terms {
  double restraint, bond, angle, torsion;
  double lj_direct, lj_exception, coulomb_direct, coulomb_exception;
  double reciprocal, self, net;
  double lennard_jones, coulomb_direct_total, coulomb, bonded, total;
} oracle_terms;

/* Cody erfc / erf (shared/erf.js:66-107). */
double oracle_erfc(double x);
double oracle_erf(double x);

/* potentialEnergy (cpu/energy.js:25-91) with per-term output; forces (3N,
 * may be NULL) are overwritten. */
int oracle_potential(const oracle_system *s, const double *pos, const oracle_options *o,
                     oracle_terms *out, double *forces);

/* The plain Ewald reciprocal sum over |k| <= kmax per axis: an independent
 * check on the PME grid sum (test/physics.test.js:413-453). */
double oracle_ewald_reciprocal(const oracle_system *s, const double *pos, double alpha, int32_t kmax,
                               double *forces);

/* PME pieces (shared/pme/parameters.js). */
double oracle_alpha_from_tolerance(double cutoff, double tolerance);
void oracle_bspline(double offset, int32_t order, double *weights, double *derivatives);
int oracle_moduli(int32_t size, int32_t order, double *out);

/* Dispersion (shared/dispersion.js:37-82) and self energy (cpu/forces/pme.js:270-286). */
double oracle_dispersion(const oracle_system *s, double volume, double cutoff);
void oracle_self_energy(const oracle_system *s, double alpha, double volume, double *self, double *net);

/* Counter-based noise (shared/random.js:24-87). */
uint32_t oracle_mix(uint32_t v);
double oracle_uniform(uint32_t seed, uint32_t a, uint32_t b, uint32_t c);
void oracle_gaussian(uint32_t seed, uint32_t atom, uint32_t step, double out[3]);

/* One BAOAB step up to the force evaluation (cpu/integrate/langevin.js:30-98)
 * and the closing kick (:101-114). noise may be NULL (generated). Returns 0,
 * or -1 when a constraint did not converge. */
typedef struct oracle_integrator {
  double timestep, temperature, friction, tolerance;
  uint32_t seed;
  int64_t step;
  int32_t max_iterations;
  const double *noise; /* 3N or NULL */
} oracle_integrator;
int oracle_langevin_step(const oracle_system *s, double *pos, double *vel, const double *forces,
                         const oracle_integrator *o);
int oracle_langevin_finish(const oracle_system *s, double *pos, double *vel, const double *forces,
                           const oracle_integrator *o);
/* SHAKE / RATTLE (cpu/integrate/constraints.js:93-222). */
int oracle_shake(const oracle_system *s, double *pos, const double *ref, double *corrections,
                 double tolerance, int32_t max_iterations);
int oracle_rattle(const oracle_system *s, const double *pos, double *vel, double tolerance,
                  int32_t max_iterations);
/* SETTLE for one rigid water (no reference counterpart; Miyamoto & Kollman,
 * J. Comput. Chem. 13, 952 (1992)). Atoms are (O, H1, H2); dOH, dHH the
 * constrained lengths; old/new positions 9 doubles each; new is updated. */
void oracle_settle(const double old_pos[9], double new_pos[9], double mO, double mH, double dOH, double dHH);
/* Kinetic energy and temperature (cpu/integrate/langevin.js:195-209). */
double oracle_kinetic(const oracle_system *s, const double *vel);
/* Initial velocities (langevin.js:211-230). */
void oracle_velocities(const oracle_system *s, double *vel, double temperature, uint32_t seed);

/* Barostat pieces: centroid scaling (webgpu/kernels/barostat.js:25-107) and
 * the Metropolis weight (barostat.js:158-162). */
void oracle_scale_molecules(const oracle_system *s, double *pos, double length_scale);
double oracle_metropolis_weight(double delta_energy, double delta_volume, double volume_ratio,
                                double pressure_kj, double kT, int32_t molecules);

/* skinFor and planCellList (shared/topology/cells.js:77-181). box may be
 * NULL (no cap). oracle_plan_cells returns the neighbour capacity. */
double oracle_skin_for(const oracle_system *s, double timestep, double temperature, const double box[9]);
int32_t oracle_plan_cells(const oracle_system *s, double skin, double slack, const double box[9], int32_t counts[3],
                          int32_t *stencil, double *estimated);

/* Steepest-descent minimiser (webgpu/simulation.js:880-979). */
typedef struct oracle_min_options {
  int32_t max_iterations;
  double force_tolerance, max_step, max_displacement;
  double tolerance;        /* SHAKE tolerance of the constraint projection */
  int32_t restore_forces;  /* 1: a rejected trial restores the forces too  */
} oracle_min_options;
typedef struct oracle_min_result {
  int32_t iterations, rejected, stalled, converged;
  double potential, max_force, max_step;
} oracle_min_result;
int oracle_minimise(const oracle_system *s, double *pos, const oracle_options *energy, const oracle_min_options *o,
                    oracle_min_result *out);

/* FP32 emulation of the cell binning (webgpu/kernels/wgsl-common.js:112-123)
 * and of the neighbour test (neighbours.js:116-139), with this build's fixed
 * operation order. pos is 3N float; recip the 9 f32 reciprocal-box entries;
 * box9/inv3 the f32 box rows and per-axis inverse widths. */
void oracle_cells_f32(int32_t n, const float *pos, const float recip[9], const int32_t counts[3],
                      int32_t *cell_of_atom);
float oracle_r2_f32(const float *pi, const float *pj, const float box9[9], const float inv3[3]);
/* Brute-force FP32 neighbour sets: for each atom i, every j != i, not excluded,
 * with r2_f32 <= rl2. Writes counts[N] and, if lists != NULL, lists[N*cap]
 * sorted ascending. Returns the maximum count. */
int32_t oracle_pairs_f32(int32_t n, const float *pos, const float box9[9], const float inv3[3], float rl2,
                         const int32_t *excl_start, const int32_t *excl_partners, int32_t cap,
                         int32_t *counts, int32_t *lists);

/* oracle_pairs_f32's sets for large systems: the same test over candidates
 * from coarse cells at least rl + 0.01 nm thick. box64 the FP64 box rows.
 * Returns the maximum count, or -1 when it cannot allocate. */
int32_t oracle_pairs_f32_binned(int32_t n, const float *pos, const double box64[9], const float box9[9],
                                const float inv3[3], float rl2, const int32_t *excl_start,
                                const int32_t *excl_partners, int32_t cap, int32_t *counts, int32_t *lists);

#ifdef __cplusplus
}
#endif
#endif
