/*
 * rainmd.h - the C ABI of the MI355X (gfx950) molecular-dynamics step.
 *
 * This is the drop-in boundary for RainMD's GPU-resident MD step. Every entry
 * point below replaces one method (or one host-side helper) that the
 * reference's JavaScript surface exposes; the citation next to each names the
 * reference function it stands in for. A Node.js N-API addon, a Python ctypes
 * binding or a C program can call it: the signatures hold plain pointers,
 * sizes and PODs only.
 *
 * Conventions
 *   - Units are the reference's: nm, ps, kJ/mol, amu, elementary charge, K.
 *   - Every function returns an int status (RMD_OK or a negative RMD_E_*).
 *     A failing call leaves a message retrievable with rmd_last_error() (per
 *     context) or rmd_last_global_error() (for calls without a context).
 *   - Input arrays are copied; output arrays are caller-allocated.
 *   - One context is bound to one HIP device and one stream and is not
 *     re-entrant. Two contexts may live on one device.
 */
#ifndef RAINMD_H
#define RAINMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RMD_ABI_VERSION 1

/* ---- status codes: each maps to one of the reference's error classes ---- */
enum {
  RMD_OK = 0,
  RMD_E_INVALID = -1,          /* plain Error: bad argument / validation     */
  RMD_E_PARSE = -2,            /* parsePrmtop / parseInpcrd refusal           */
  RMD_E_SYSTEM = -3,           /* buildSystem refusal (cutoff > half box ...) */
  RMD_E_DEVICE = -4,           /* HIP error, or no device                     */
  RMD_E_NONFINITE = -5,        /* NonFiniteStateError (shared/finite.js:63)   */
  RMD_E_NEIGHBOUR_OVERFLOW = -6,/* NeighbourOverflowError (neighbours.js:172) */
  RMD_E_STOPPED = -7,          /* SimulationStopped (webgpu/runtime/yield.js:44) */
  RMD_E_BUDGET = -8,           /* GpuBudgetError (webgpu/runtime/allocator.js:141) */
  RMD_E_CONSTRAINT = -9        /* ConstraintFailure (cpu/integrate/langevin.js:164) */
};

/* ======================================================================== */
/* Host side: topology, coordinates and the flattened system                */
/* ======================================================================== */

typedef struct rmd_topology rmd_topology;
typedef struct rmd_coordinates rmd_coordinates;
typedef struct rmd_system rmd_system;

/* Constraint modes of buildSystem (shared/topology/system.js:40-50). */
enum { RMD_CONSTRAINTS_NONE = 0, RMD_CONSTRAINTS_HBONDS = 1, RMD_CONSTRAINTS_ALLBONDS = 2 };

/* Replaces parsePrmtop (shared/amber/prmtop.js:111-385): text of a parm7 file. */
int rmd_topology_parse_prmtop(const char *text, size_t length, rmd_topology **out);
void rmd_topology_free(rmd_topology *topology);

/* Replaces parseInpcrd (shared/amber/inpcrd.js:34-132). expected_atom_count
 * <= 0 skips the NATOM cross-check. */
int rmd_coordinates_parse_inpcrd(const char *text, size_t length, int32_t expected_atom_count,
                                 rmd_coordinates **out);
void rmd_coordinates_free(rmd_coordinates *coordinates);

typedef struct rmd_coordinates_view {
  int32_t atom_count;
  const double *positions;   /* 3N, nm */
  const double *velocities;  /* 3N nm/ps, or NULL */
  int32_t has_box;
  double box_lengths[3];     /* nm */
  double box_angles[3];      /* degrees */
  double time;               /* ps, NaN when absent */
} rmd_coordinates_view;
int rmd_coordinates_view_get(const rmd_coordinates *coordinates, rmd_coordinates_view *view);

/* Replaces boxVectorsFromLengthsAndAngles (shared/amber/inpcrd.js:154-196). */
int rmd_box_vectors_from_lengths_and_angles(const double lengths[3], const double angles_deg[3],
                                            double box_out[9]);

typedef struct rmd_build_options {
  int32_t constraints;       /* RMD_CONSTRAINTS_*; HBonds by default        */
  int32_t rigid_water;       /* 1 by default                                */
  double cutoff;             /* nm; 1.0 by default; 0 = non-periodic use    */
  int32_t has_box_vectors;   /* 1: box_vectors overrides the box below      */
  double box_vectors[9];     /* reduced, row-major                          */
  int32_t has_box;           /* 1: lengths/angles (an inpcrd's box)         */
  double box_lengths[3];
  double box_angles[3];
} rmd_build_options;
void rmd_build_options_default(rmd_build_options *options);

/* Replaces buildSystem (shared/topology/system.js:40-236). The positions of
 * the coordinates (may be NULL) are carried along so a system is
 * self-describing. */
int rmd_system_build(const rmd_topology *topology, const rmd_coordinates *coordinates,
                     const rmd_build_options *options, rmd_system **out);
void rmd_system_free(rmd_system *system);

/* The seeded solvated-peptide builder used by the synthetic workloads: the
 * ACE-ALA-NME solute of ala-wat.prmtop at the box centre and `water_count`
 * TIP3P waters (the same file's WAT residue) on a jittered cubic lattice with
 * uniformly random orientations. No counterpart in the reference (its systems
 * come from tleap); it reuses the reference's own force-field file. */
typedef struct rmd_solvated_options {
  int32_t water_count;
  double box_edge;           /* nm, cubic                                     */
  uint32_t seed;
  int32_t include_solute;    /* 0: pure water box                             */
  double solute_clearance;   /* nm, lattice sites closer to the solute drop   */
  double jitter;             /* nm, uniform +-jitter per axis on the oxygen   */
  double cutoff;             /* nm                                            */
  int32_t constraints;       /* RMD_CONSTRAINTS_*                              */
  int32_t rigid_water;
} rmd_solvated_options;
void rmd_solvated_options_default(rmd_solvated_options *options);
int rmd_system_build_solvated(const char *template_prmtop, size_t template_length,
                              const char *solute_inpcrd, size_t solute_length,
                              const rmd_solvated_options *options, rmd_system **out);

/* The system a later protocol stage runs: the same topology in the live box
 * a barostat left behind, as the page rebuilds it with buildSystem(...,
 * {boxVectors: currentBoxVectors()}) before each stage (web/app.js:354-364,
 * 1103-1124). box is reduced and row-major; the cutoff must stay within half
 * of its smallest width, as in buildSystem (shared/topology/system.js:80-87). */
int rmd_system_rebox(const rmd_system *system, const double box[9], rmd_system **out);
/* soluteHeavyAtoms (shared/topology/restraints.js:46-49): atoms whose residue
 * is neither water nor an ion (the residue classes of
 * shared/analysis/selection.js:9-24,70-77) and whose atomic number is not 1,
 * ascending. Writes at most `capacity` indices; *count is the full count. */
int rmd_system_solute_heavy_atoms(const rmd_system *system, int32_t *out, int32_t capacity, int32_t *count);

/* A flat, read-only view of a built system: every array the kernels and the
 * oracle read. Pointers stay valid until rmd_system_free. */
typedef struct rmd_system_view {
  int32_t atom_count, type_count, molecule_count, residue_count;
  const double *charges, *masses;              /* N                           */
  const int32_t *atom_types, *atomic_numbers;  /* N                           */
  const int32_t *residue_of_atom;              /* N                           */
  const double *lj_a, *lj_b;                   /* T*T, kJ/mol nm^12 / nm^6    */
  int32_t bond_count;                          /* harmonic bonds (unconstrained) */
  const int32_t *bond_a, *bond_b;
  const double *bond_k, *bond_r0;
  int32_t angle_count;
  const int32_t *angle_a, *angle_b, *angle_c;
  const double *angle_k, *angle_theta0;
  int32_t torsion_count;
  const int32_t *torsion_a, *torsion_b, *torsion_c, *torsion_d;
  const double *torsion_k, *torsion_periodicity, *torsion_phase;
  int32_t constraint_count;
  const int32_t *constraint_a, *constraint_b;
  const double *constraint_distance;
  int32_t exception_count;
  const int32_t *exception_a, *exception_b;    /* a < b                        */
  const double *exception_charge_product, *exception_lj_a, *exception_lj_b;
  int32_t exclusion_total;                     /* symmetric, sorted per atom   */
  const int32_t *exclusion_start, *exclusion_partners; /* N+1, exclusion_total */
  const int32_t *molecule_start, *molecule_atoms, *molecule_of_atom; /* M+1, N, N */
  int32_t is_periodic;
  double box[9];
  double cutoff;
  int32_t degrees_of_freedom;
  int32_t constraint_mode, rigid_water;
  const double *positions;                     /* 3N nm (the build's coordinates) or NULL */
  const double *velocities;                    /* 3N nm/ps or NULL              */
  int32_t water_count;                         /* rigid 3-site waters (SETTLE units) */
} rmd_system_view;
int rmd_system_view_get(const rmd_system *system, rmd_system_view *view);

/* Replaces the derived-setup helpers a simulation calls at construction:
 * pmeParameters (shared/pme/parameters.js:227-252), skinFor
 * (shared/topology/cells.js:77-94), planCellList (cells.js:111-181),
 * dispersionCoefficients (shared/dispersion.js:37-66) and ewaldSelfEnergy
 * (cpu/forces/pme.js:270-286). grid[] entries <= 0 are derived; alpha <= 0 is
 * derived. Unlike the reference, radix-7 grids (2^a 3^b 5^c 7^d) are allowed. */
typedef struct rmd_pme_parameters {
  double alpha;
  int32_t grid[3];
  int32_t order;
  double tolerance;
} rmd_pme_parameters;
int rmd_pme_parameters_derive(const double box[9], double cutoff, double tolerance,
                              double forced_alpha, const int32_t forced_grid[3],
                              rmd_pme_parameters *out);
/* B-spline moduli |b(m)|^2 with the Nyquist patch (parameters.js:164-215). */
int rmd_pme_moduli(int32_t grid_size, int32_t order, double *moduli_out);
double rmd_skin_for(const rmd_system *system, double timestep, double temperature);
double rmd_dispersion_energy(const rmd_system *system, double volume);

/* ======================================================================== */
/* Device side: the simulation context (GpuSimulation, webgpu/simulation.js) */
/* ======================================================================== */

typedef struct rmd_ctx rmd_ctx;

/* Constructor options (webgpu/simulation.js:93-188). Fields <= 0 (or < 0 where
 * 0 is meaningful) mean "derive / default", mirroring the `??` defaults. */
typedef struct rmd_options {
  double timestep;               /* 0.002 ps                                   */
  double temperature;            /* 298 K                                      */
  double friction;               /* 1 /ps (negative -> default; 0 allowed)     */
  uint32_t seed;
  double tolerance;              /* 1e-5 relative                              */
  int32_t max_constraint_iterations; /* 150                                    */
  int32_t rebuild_interval;      /* 20; accepted for surface parity (the list
                                    is rebuilt on a device-side skin check)   */
  int32_t adaptive_rebuild;      /* 1                                          */
  int32_t has_pressure;          /* 0: NVT                                     */
  double pressure;               /* bar                                        */
  int32_t barostat_interval;     /* 25 steps                                   */
  int32_t dispersion_correction; /* 1                                          */
  double restraint_force;        /* kJ/mol/nm^2, k*d^2 (no half)               */
  int32_t restraint_count;
  const int32_t *restraint_atoms;       /* restraint_count                     */
  const double *restraint_reference;    /* 3N (only restrained entries read)   */
  double skin;                   /* nm; 0 derives skinFor, < 0 takes the
                                    autotuner's skin for the workload         */
  double neighbour_slack;        /* 2.0                                        */
  int32_t max_neighbours;        /* <= 0 estimates                             */
  double ewald_tolerance;        /* 5e-4                                       */
  double alpha;                  /* <= 0 derives                               */
  int32_t grid[3];               /* <= 0 derives                               */
  int32_t device;                /* HIP device ordinal                         */
  int32_t lanes_per_atom;        /* nonbonded lanes per atom (1..64), 0 = tuned */
  int32_t cell_reach;            /* widest stencil reach the cell plan may take
                                    (1..3; 1 is the reference's plan), 0 = tuned */
  int32_t search_mode;           /* list search: 1 one wave per atom, 2 one wave
                                    per cell, 3 by the plan's cell occupancy,
                                    0 = tuned                                  */
  int32_t pme_bins;              /* PME planes of a single system: 1 bin the
                                    atoms by x plane, 2 every plane scans every
                                    atom, 3 by atom count, 0 = tuned (batch
                                    members always bin)                        */
} rmd_options;
void rmd_options_default(rmd_options *options);

/* new GpuSimulation(device, system, options) (webgpu/simulation.js:83-190). */
int rmd_create(const rmd_system *system, const rmd_options *options, rmd_ctx **out);
/* release() (webgpu/simulation.js:1336-1340). */
void rmd_destroy(rmd_ctx *ctx);
const char *rmd_last_error(const rmd_ctx *ctx);
const char *rmd_last_global_error(void);

/* setState(positions, velocities, {step, suppliedNoise}) (simulation.js:356-369).
 * velocities may be NULL (zero). */
int rmd_set_state(rmd_ctx *ctx, const double *positions, const double *velocities, int64_t step,
                  int32_t supplied_noise);
/* setNoise(noise) (simulation.js:371): 3N normals used when supplied_noise. */
int rmd_set_noise(rmd_ctx *ctx, const float *noise);
/* initialise() (simulation.js:412-422): one force evaluation, fresh list. */
int rmd_initialise(rmd_ctx *ctx);
/* run(count, {blockSize, stopSignal}) (simulation.js:456-527). The barostat
 * move happens between blocks, as in the reference. */
int rmd_run(rmd_ctx *ctx, int64_t steps, int32_t block_size, int64_t *step_out);
/* StopSignal (webgpu/runtime/yield.js:40-51): checked between blocks. */
int rmd_request_stop(rmd_ctx *ctx);

/* energy() (simulation.js:1073-1230). All FP64 host-visible sums. */
typedef struct rmd_energy_terms {
  int64_t step;
  double time;
  double bonded, harmonic_bond, harmonic_angle, periodic_torsion;
  double lennard_jones;      /* direct + exception LJ                         */
  double dispersion;
  double coulomb_direct;     /* direct + exception Coulomb                    */
  double coulomb_reciprocal;
  double coulomb_self;       /* self + net-charge                             */
  double coulomb;
  double position_restraint;
  int32_t restrained_atoms;
  double potential, kinetic, total;
  double temperature, kinetic_for_temperature;
  int32_t kinetic_sampled_in_step; /* 1: "before the closing half kick"       */
  double volume;
  double box[9];
  /* d(energy)/d(ln lambda) when every coordinate and the box scale by lambda:
   * direct-space pairs (sum r dU/dr) and the PME reciprocal sum. Filled by
   * evaluations (0 where not computed); features of the barostat's screen.
   * dscale_molecular: the derivative when molecules are scaled about their
   * centroids with their geometry kept (the barostat's move), every term. */
  double dscale_direct, dscale_reciprocal;
  double dscale_molecular;
} rmd_energy_terms;
int rmd_energy(rmd_ctx *ctx, rmd_energy_terms *out);

/* The context's atom count N (the reference's GpuSimulation.atomCount,
 * webgpu/simulation.js:92), which sizes every per-atom array below; a
 * negative RMD_E_* status for a null context. */
int32_t rmd_atom_count(const rmd_ctx *ctx);

/* readPositions() / readState() (simulation.js:1239-1261). Positions,
 * velocities and forces are 3N float (f32), like the reference's readback. */
int rmd_read_positions(rmd_ctx *ctx, float *positions);
int rmd_read_state(rmd_ctx *ctx, float *positions, float *velocities, float *forces,
                   double box[9], double *volume);

/* scaleVolume(s) (simulation.js:801-811): centroid scaling + box + potential. */
int rmd_scale_volume(rmd_ctx *ctx, double length_scale, double *potential);

/* attemptVolumeChange() (simulation.js:995-1070). */
typedef struct rmd_move {
  int32_t attempted;         /* 0: no barostat (the reference returns null)    */
  int32_t accepted;
  int32_t refused;           /* 1: cutoff would not fit the proposed box       */
  double weight, delta_volume, length_scale;
  double old_potential, new_potential, volume, volume_scale, acceptance_rate;
} rmd_move;
int rmd_attempt_volume_change(rmd_ctx *ctx, rmd_move *out);

/* minimise(options) (simulation.js:880-979). */
typedef struct rmd_min_options {
  int32_t max_iterations;    /* 200 */
  double force_tolerance;    /* 500 kJ/mol/nm */
  double max_step;           /* 0.002 nm */
  double max_displacement;   /* 0.05 nm */
} rmd_min_options;
void rmd_min_options_default(rmd_min_options *options);
typedef struct rmd_min_result {
  int32_t iterations, rejected, stalled, converged;
  double potential, max_force, max_step, energy_resolution;
} rmd_min_result;
int rmd_minimise(rmd_ctx *ctx, const rmd_min_options *options, rmd_min_result *out);

/* diagnostics() (simulation.js:1264-1334). */
typedef struct rmd_diagnostics {
  int32_t max_shake_iterations, shake_hit_iteration_cap, rattle_hit_iteration_cap;
  int32_t neighbour_required, neighbour_capacity, worst_neighbour_required, neighbour_growths;
  double skin, worst_displacement;
  int64_t rebuilds, steps;
  int32_t cell_counts[3];
  double volume;
  int32_t barostat_attempts, barostat_accepted, barostat_refused;
  double volume_scale;
  double constraint_tolerance;
  int32_t pme_grid[3];
  double pme_alpha;
  int32_t lanes_per_atom;
  int32_t water_units, cluster_units, free_units;
  int32_t pme_planes; /* 1: plane PME pipeline, 0: line pipeline */
  int32_t cell_reach;      /* the cell plan's stencil reach (1: the reference's plan) */
  int32_t search_by_cell;  /* 1: the list search runs one wave per cell */
  int64_t steps_since_barostat; /* steps run since the last volume move (a
                                   move is due at a block start once this
                                   reaches the barostat interval) */
  int32_t pme_x_folded; /* 1: force-only evaluations fold the PME x lines into
                           the plane gathers (rmd_set_pme_pipeline mode 4) */
} rmd_diagnostics;
int rmd_diagnostics_get(rmd_ctx *ctx, rmd_diagnostics *out);

/* ---- evaluation entry points used by the parity tests and the bench ---- */

/* Term masks for rmd_evaluate. */
enum {
  RMD_TERM_BONDED = 1, RMD_TERM_DIRECT = 2, RMD_TERM_EXCEPTIONS = 4,
  RMD_TERM_RECIPROCAL = 8, RMD_TERM_RESTRAINT = 16, RMD_TERM_ALL = 31
};
/* One force evaluation at the current positions with a fresh neighbour list,
 * restricted to `terms`. Fills energies (FP64 sums; may be NULL) and forces
 * (3N doubles converted from the fixed-point accumulators; may be NULL).
 * With energies NULL the evaluation runs a step's force-only launches. The
 * reference's counterpart is #evaluatePotential (simulation.js:734-789). */
int rmd_evaluate(rmd_ctx *ctx, int32_t terms, rmd_energy_terms *energies, double *forces);

/* The cell list and neighbour list the device built at the last rebuild.
 * cell_of_atom: N entries. neighbour_count: N. neighbours: N*capacity (the
 * capacity is rmd_diagnostics.neighbour_capacity). Counterparts: the buffers
 * encodeCellList / encodeNeighbours fill (cell-list.js:176-226,
 * neighbours.js:149-169). */
int rmd_read_cells(rmd_ctx *ctx, int32_t *cell_of_atom, int32_t *cell_start, int32_t *cell_atoms);
int rmd_read_neighbours(rmd_ctx *ctx, int32_t *neighbour_count, int32_t *neighbours);

/* Runs exactly one BAOAB step with the current forces and returns, without
 * the closing half kick being folded into anything else. Used to compare one
 * step against the host integrator with supplied noise and forces
 * (tools/gpu/check-integrator.mjs:68-127). */
int rmd_step_once(rmd_ctx *ctx);

/* Verification entry: the forward real-to-complex 3D FFT of a real grid
 * (K0*K1*K2 floats, z fastest) with the context's PME plan and sign
 * convention (exp(+2 pi i m k / K), pme-fft.js:81-144); out receives
 * K0*K1*(K2/2+1) complex values as interleaved floats. */
int rmd_debug_fft_forward(rmd_ctx *ctx, const float *real, float *out);

/* Verification entry: the thermostat's noise for every atom at `step` with
 * the context's seed, drawn by the integrator's own device functions.
 * uniforms (4N floats, may be NULL) receives uniformAt(seed, atom, step, k)
 * for k = 0..3; normals (3N floats) the Box-Muller triple built from them.
 * Counterparts: uniformAt and gaussianTriple (shared/random.js:59-87), whose
 * device arm tools/gpu/check-integrator.mjs:44-66 checks against the host. */
int rmd_debug_noise(rmd_ctx *ctx, int64_t step, float *uniforms, float *normals);

/* The autotuner's shapes compiled into this library for a single system
 * (kind 0) or a replica batch (kind 1) of `atoms` atoms: lanes per atom, the
 * cell plan's widest reach, the search mode, the tuned skin (nm; 0 means the
 * reference's skinFor) and the PME bin choice. Options left at 0 (and a
 * negative skin) take these values. No counterpart in the reference, whose
 * skin and rebuild interval were swept by hand (tools/gpu/bench-skin.mjs). */
typedef struct rmd_tuned_shape {
  int32_t lanes_per_atom, cell_reach, search_mode;
  double skin;
  int32_t pme_bins;
} rmd_tuned_shape;
int rmd_tuned_shape_get(int32_t kind, int32_t atoms, rmd_tuned_shape *out);

/* Tuning and plumbing: the nonbonded lanes per atom (a power of two <= 64),
 * hipGraph capture of step blocks on/off, and the context's HIP stream (for
 * timing on that stream). */
int rmd_set_lanes(rmd_ctx *ctx, int32_t lanes);
int rmd_set_graphs(rmd_ctx *ctx, int32_t enabled);
/* Overlapped force branches (1, the default) or each branch launched on its
 * own, one after another on the context's stream (0). Overlapped, a small
 * system runs its direct space, bonded terms and PME forward planes as one
 * launch; a larger one runs PME and the bonded terms on streams beside the
 * neighbour list and the direct space. Results are identical either way. */
int rmd_set_concurrent_forces(rmd_ctx *ctx, int32_t enabled);
/* PME pipeline: 0 automatic (planes when one x plane of the grid fits in
 * LDS; mode 4 for a single system of at most 32 x planes that does not bin),
 * 1 planes (RMD_E_INVALID when it does not fit), 2 lines, 3 planes with
 * every plane scanning every atom (no atom bins, which large single systems
 * and batch members build by default), 4 as 3 with force-only evaluations
 * folding the x lines into the plane gathers (one launch fewer per
 * evaluation; RMD_E_INVALID beyond 32 x planes or for a batch member).
 * Results are identical for 1 and 3; mode 4's forces differ from theirs by
 * FP32 rounding (evaluations with energies are those of mode 3). */
int rmd_set_pme_pipeline(rmd_ctx *ctx, int32_t mode);
int rmd_get_stream(rmd_ctx *ctx, void **stream);
/* Average launch duration (ms) of one kernel family at the current state,
 * timed with HIP events on the context's stream over `reps` launches:
 * 0 nonbonded, 1 PME chain, 2 integrator, 3 bonded/exceptions, 4 rebuild. */
int rmd_profile_kernel(rmd_ctx *ctx, int32_t kernel, int32_t reps, double *avg_ms);
/* Average durations (ms) of the PME chain's three stages at the current
 * state, each over reps back-to-back launches of that stage on the context's
 * stream (HIP events around them): stage_ms[0] the charge spread with the
 * forward transforms it carries, [1] the remaining transforms and the
 * reciprocal convolution (0 where force-only evaluations fold them into the
 * gathers, rmd_set_pme_pipeline mode 4), [2] the inverse transforms and the
 * force gather (webgpu/kernels/pme.js:137-336). */
int rmd_profile_pme_stages(rmd_ctx *ctx, int32_t reps, double *stage_ms);
/* Measurement calibration: `reps` launches of a plain device-to-device copy
 * of `bytes` bytes (16-byte vector loads and stores) on the context's stream;
 * its hardware byte counters scale the HBM traffic read for other kernels. */
int rmd_copy_probe(rmd_ctx *ctx, int64_t bytes, int32_t reps, double *avg_ms);

/* ======================================================================== */
/* The barostat's learned screen (no counterpart in the reference)          */
/* ======================================================================== */
/* A 16 -> 64 -> 64 -> 1 tanh MLP (RMD_SURROGATE_PARAMS FP32 parameters:
 * W1[16][64], b1[64], W2[64][64], b2[64], W3[64], b3, input-major) whose
 * layer products run as BF16 MFMA with FP32 accumulation on the GPU, for
 * inference and for training (Adam, mean-squared error). The reference's
 * barostat evaluates every move exactly (webgpu/simulation.js:995-1070);
 * this screen lets a move be turned down before its exact evaluation with
 * a delayed-acceptance second stage that keeps the NPT ensemble exact. */
#define RMD_SURROGATE_INPUTS 16
#define RMD_SURROGATE_PARAMS 5313
typedef struct rmd_surrogate rmd_surrogate;
/* Parameters drawn from the counter hash with `seed` (uniform, scaled by
 * sqrt(6 / (fan_in + fan_out))), biases zero. */
int rmd_surrogate_create(int32_t device, uint32_t seed, rmd_surrogate **out);
void rmd_surrogate_free(rmd_surrogate *net);
int rmd_surrogate_get_params(rmd_surrogate *net, float *params);
int rmd_surrogate_set_params(rmd_surrogate *net, const float *params);
/* y[k] = f(x[k]), x: n x 16. */
int rmd_surrogate_forward(rmd_surrogate *net, const float *x, int32_t n, float *y);
/* rmd_surrogate_forward that allocates nothing: no counterpart in the
 * reference. For callers that cannot allocate device memory (code under
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). The
 * inputs and outputs are staged through one static array per device (the
 * device holds it until the process ends), so n is at most
 * RMD_SURROGATE_NOALLOC_SAMPLES (larger n is refused with RMD_E_INVALID),
 * and calls that overlap on one device (two networks on two host threads)
 * race on that array. Outputs equal rmd_surrogate_forward's. */
#define RMD_SURROGATE_NOALLOC_SAMPLES 16384
int rmd_surrogate_forward_noalloc(rmd_surrogate *net, const float *x, int32_t n, float *y);
/* loss = mean_k (f(x_k) - t_k)^2 and its gradient (RMD_SURROGATE_PARAMS). */
int rmd_surrogate_gradient(rmd_surrogate *net, const float *x, const float *t, int32_t n, float *loss, float *grad);
/* `steps` Adam steps (mini-batches of 64 drawn from the n samples);
 * returns the mean mini-batch loss of the last 16 steps. */
int rmd_surrogate_train(rmd_surrogate *net, const float *x, const float *t, int32_t n, int32_t steps,
                        float learning_rate, float *loss);

/* The screen on a simulation's barostat. Phase 0 runs plain moves and keeps
 * (features, exact energy change) samples; after `training_moves` of them
 * the network is trained (`training_steps` Adam steps) and frozen, and
 * phase 1 screens: stage one accepts a move with min(1, exp(-w~/kT)) from
 * the predicted weight w~ (no evaluation when it fails); a survivor is
 * evaluated exactly and accepted with
 * min(1, exp(-w/kT) * a1(reverse) / a1(forward)). */
typedef struct rmd_screen_options {
  int32_t enabled;
  int32_t training_moves;   /* 2000 */
  int32_t training_steps;   /* 4000 */
  float learning_rate;      /* 1e-3 */
  uint32_t seed;            /* network initialisation */
} rmd_screen_options;
void rmd_screen_options_default(rmd_screen_options *options);
int rmd_screen_configure(rmd_ctx *ctx, const rmd_screen_options *options);
typedef struct rmd_screen_stats {
  int32_t phase;                 /* 0 collecting samples, 1 screening */
  int64_t samples;               /* (features, exact change) pairs kept */
  int64_t moves, exact_evaluations, screened_out, second_stage_rejections, accepted;
  double training_loss;          /* normalised mean-squared error at the end of training */
  double residual_rms;           /* kJ/mol: rms of (exact - predicted) energy change on the screened moves */
  double estimate_rms;           /* kJ/mol: rms of (exact - first-order estimate) on the same moves */
} rmd_screen_stats;
int rmd_screen_stats_get(rmd_ctx *ctx, rmd_screen_stats *out);
/* The kept samples: up to `capacity` rows of 16 raw features and the exact
 * energy change residual target (kJ/mol); *count receives the number kept. */
int rmd_screen_samples(rmd_ctx *ctx, float *features, float *targets, int32_t capacity, int32_t *count);

/* Device and build identity. */
typedef struct rmd_device_info {
  int32_t device_count;
  char name[128];
  char arch[64];
  int32_t compute_units;
  int64_t global_memory;
} rmd_device_info;
int rmd_device_info_get(int32_t device, rmd_device_info *out);
int32_t rmd_abi_version(void);

/* ======================================================================== */
/* Replica batches                                                          */
/* ======================================================================== */

/* Many independent systems stepped together (uneven replicas, each with its
 * own box, PME grid, seed and barostat): the "rmd_batch_* variants" a host
 * needs for replica workloads. The reference steps one GpuSimulation per
 * system (webgpu/simulation.js:82); a batch is R member contexts stepped by
 * batched launches, each replica computing exactly what its own context
 * would alone. Every replica takes the same options except its seed
 * (seeds[r], or options->seed when seeds is NULL), needs the same Ewald
 * alpha, at most 16384 atoms and 4096 cells, and a PME grid whose planes fit
 * in LDS; a batch holds at most RMD_BATCH_MAX_REPLICAS replicas.
 * options->lanes_per_atom picks the direct space's lanes per atom for every
 * member (8, 16, 32 or 64; 0 = 16). */
#define RMD_BATCH_MAX_REPLICAS 4096
typedef struct rmd_batch rmd_batch;
int rmd_batch_create(const rmd_system *const *systems, int32_t count, const rmd_options *options,
                     const uint32_t *seeds, rmd_batch **out);
void rmd_batch_destroy(rmd_batch *batch);
const char *rmd_batch_last_error(const rmd_batch *batch);
/* Replica r's own context, owned by the batch (rmd_destroy leaves it alone):
 * every single-system call works on it - setState, minimise, energy,
 * readState, diagnostics, and a run of that replica on its own. NULL when r
 * is out of range. */
rmd_ctx *rmd_batch_replica(rmd_batch *batch, int32_t replica);
typedef struct rmd_batch_info {
  int32_t replicas;
  int64_t atoms;                 /* over all replicas                          */
  int64_t blocks_run;            /* step blocks run by rmd_batch_run           */
  int64_t step_workgroups;       /* workgroups launched per step, all kernels  */
  int64_t nonbonded_workgroups;  /* of the batched direct-space launch         */
} rmd_batch_info;
int rmd_batch_info_get(rmd_batch *batch, rmd_batch_info *out);
/* initialise() of every replica. */
int rmd_batch_initialise(rmd_batch *batch);
/* One evaluation of the given RMD_TERM_* terms of every replica at its
 * current positions (the counterpart of rmd_evaluate); each replica's forces
 * are then readable through its own context. */
int rmd_batch_evaluate(rmd_batch *batch, int32_t terms);
/* run(count, {blockSize}) of every replica at once (webgpu/simulation.js:
 * 456-527); each replica's barostat moves when its own interval is due, and
 * the moves of all due replicas are evaluated together. *step_out is replica
 * 0's step. Errors name the replica. */
int rmd_batch_run(rmd_batch *batch, int64_t steps, int32_t block, int64_t *step_out);
int rmd_batch_request_stop(rmd_batch *batch);
int rmd_batch_set_graphs(rmd_batch *batch, int32_t enabled);
int rmd_batch_set_concurrent_forces(rmd_batch *batch, int32_t enabled);
/* Average duration of the batched direct-space launch (profiling). */
int rmd_batch_profile_nonbonded(rmd_batch *batch, int32_t reps, double *avg_ms);
/* The learned delayed-acceptance screen for a batch's barostat: one network
 * shared by every replica (features and targets per molecule and kT), trained
 * on the first training_moves exact moves of all replicas (both directions of
 * each), then screening every replica's moves with the single system's two
 * stages. No counterpart in the reference. The stats count the screening
 * phase only: moves decided by the network, exact evaluations, moves turned
 * down without one, second-stage rejections, acceptances. */
int rmd_batch_screen_configure(rmd_batch *batch, const rmd_screen_options *options);
int rmd_batch_screen_stats_get(rmd_batch *batch, rmd_screen_stats *out);

#ifdef __cplusplus
}
#endif

#endif /* RAINMD_H */
