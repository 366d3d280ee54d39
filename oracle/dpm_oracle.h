/*
 * dpm_oracle.h -- CPU oracle for the deformable-ring (DPM2D) vertex engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a host restatement, in plain C, of the
 * reference's force, neighbour, integrator and FIRE arithmetic.  Only the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark may load it,
 * and only as the checker.  The product library never links or calls it.
 *
 * Layout follows the reference (AoS): pos/vel/force are double[numVertices*2],
 * index arrays are long (int64), exactly as in include/DPM2D.h:65-177 of the
 * reference.  Every function cites the reference file:line it restates.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   geometry/topology   closed forms (regular n-gon), minimum image straddling L/2
 *   shape force         central finite differences of the shape energy, zero force
 *                       at rest shape, sum of forces = 0 per ring
 *   pair potentials     closed-form values at d = sigma, 2^(1/6) sigma, cut-offs,
 *                       finite differences, C1 continuity
 *   vertex-vertex       all-pairs brute force == list form, Newton's third law
 *   smooth              finite differences of the smooth energy away from branch
 *                       switches, hand-built vertex-near-segment cases
 *   neighbour list      brute-force all-pairs sets
 *   integrators         NVE time reversibility and O(dt^2) energy drift
 *   FIRE                convergence to |F|max < 1e-12 on a harmonic packing
 *   RNG                 Philox4x32-10 known-answer vectors (Random123)
 *   per-ring costs      (an addition) finite differences with distinct costs per ring;
 *                       uniform per-ring costs equal the global ones bit for bit
 *   Langevin            per-ring noise pinned through the shared Philox stream and
 *                       <K> against T; the reference's own Thrust/time(0) streams are
 *                       not reproducible even by the reference: parity unpinned by design
 * The set-up draws (drand48 order of setPolySizeDistribution, setRandomParticles,
 * initVerticesOnParticles) are restated in oracle/setup_oracle.py.
 */
#ifndef DPM_ORACLE_H
#define DPM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* enum orders follow simControlStruct, include/DPM2D.h:22-32 of the reference */
enum { ORC_HARMONIC = 0, ORC_LJ = 1, ORC_ADHESIVE = 2, ORC_WCA = 3 };
enum { ORC_VERTEX_VERTEX = 0, ORC_VERTEX_SMOOTH = 1, ORC_CELL_SMOOTH = 2, ORC_ALL = 3 };
enum { ORC_MONO_HARMONIC = 0, ORC_MONO_FENE = 1 };
enum { ORC_PASSIVE = 0, ORC_ACTIVE = 1, ORC_ACTIVE_SHAPE = 2 };
enum { ORC_SIM_NONE = 0, ORC_SIM_LANGEVIN = 1, ORC_SIM_NVE = 2, ORC_SIM_NVE_RESCALE = 3 };
enum { ORC_NOISE_LANGEVIN = 0, ORC_NOISE_BAOAB = 1, ORC_NOISE_BROWNIAN = 2, ORC_NOISE_DRIVEN_BROWNIAN = 3 };
enum { ORC_GEOM_NORMAL = 0, ORC_GEOM_FIXED_BOX = 1, ORC_GEOM_FIXED_SIDES = 2, ORC_GEOM_ROUND_BOX = 3 };

/* field ids shared by orc_set_field / orc_get_field */
enum {
  ORC_F_POS = 0, ORC_F_VEL, ORC_F_FORCE, ORC_F_ENERGY, ORC_F_RAD, ORC_F_L0,
  ORC_F_THETA0, ORC_F_THETA, ORC_F_LENGTH, ORC_F_A0, ORC_F_AREA, ORC_F_PERIMETER,
  ORC_F_PPOS, ORC_F_PENERGY, ORC_F_PANGLE, ORC_F_LASTPOS, ORC_F_BOX, ORC_F_WALLFORCE,
  ORC_F_PVEL, ORC_F_PANGVEL, ORC_F_PFORCE, ORC_F_PTORQUE, ORC_F_VTORQUE, ORC_F_PRAD,
  ORC_F_COUNT
};

typedef struct orc_state orc_state;

/* lifecycle: numVertexInParticleList is copied; vertices of ring p are contiguous */
orc_state* orc_create(long numParticles, const long* numVertexInParticleList);
void orc_destroy(orc_state* s);
long orc_num_vertices(const orc_state* s);
long orc_field_size(const orc_state* s, int field);
int orc_set_field(orc_state* s, int field, const double* data, long n);
int orc_get_field(const orc_state* s, int field, double* data, long n);

/* parameters (DPM2D.cu:1875-1960) */
void orc_set_types(orc_state* s, int potentialType, int interactionType, int monomerType, int particleType);
void orc_set_ring_costs(orc_state* s, const double* ka, const double* kl, const double* kb);
void orc_set_energy_costs(orc_state* s, double ea, double el, double eb, double ec);
void orc_set_attraction_constants(orc_state* s, double l1, double l2);
void orc_set_lj_cutoff(orc_state* s, double LJcutoff);
void orc_set_fene_constants(orc_state* s, double stiff, double ext);
void orc_set_time_step(orc_state* s, double dt);
void orc_set_self_propulsion(orc_state* s, double driving, double taup);
void orc_set_rng(orc_state* s, uint64_t seed, uint32_t systemId);

/* geometry (cudaKernel.cuh) */
double orc_pbc_distance(double x1, double x2, double size);
long orc_next_id(const orc_state* s, long vertexId, long particleId);
long orc_previous_id(const orc_state* s, long vertexId, long particleId);
void orc_calc_particle_shape(orc_state* s);
void orc_calc_particle_positions(orc_state* s);

/* forces */
void orc_calc_shape_force_energy(orc_state* s);
double orc_pair_grad_multiple_and_energy(const orc_state* s, double distance, double radSum, double* epot);
void orc_calc_vertex_interaction(orc_state* s);
void orc_calc_all_to_all_interaction(orc_state* s);
void orc_calc_smooth_interaction(orc_state* s);
void orc_add_self_propulsion(orc_state* s);
void orc_calc_force_energy(orc_state* s);

/* neighbours (DPM2D.cu:1300-1423, 3479-3507; cudaKernel.cuh:1735-1763, 1906-1914) */
double orc_set_displacement_cutoff(orc_state* s, double cutoff, double size);
void orc_calc_neighbor_list(orc_state* s, double cutDistance);
int orc_check_neighbors(orc_state* s);
long orc_neighbor_list_size(const orc_state* s);
long orc_get_neighbors(const orc_state* s, long* list, long* counts, long cap);
long orc_get_update_count(const orc_state* s);
void orc_reset_update_count(orc_state* s);
void orc_reset_last_positions(orc_state* s);
double orc_get_cut_distance(const orc_state* s);
/* row build of calcNeighborList: 1 = cell list (default; the same rows), 0 = all-pairs loop */
void orc_set_list_method(orc_state* s, int method);
/* OpenMP threads of the per-vertex loops (default 1: serial, deterministic order) */
void orc_set_threads(orc_state* s, int threads);

/* integrators (Simulator.cu:23-89, 206-238; DPM2D.cu:3713-3817) */
void orc_set_noise_type(orc_state* s, int noiseType);
void orc_set_geometry(orc_state* s, int geometryType);
void orc_set_wall_energy_scale(orc_state* s, double ew);
void orc_set_box_radius(orc_state* s, double boxRadius);
double orc_get_box_radius(orc_state* s);
void orc_calc_wall_forces(orc_state* s);
/* observables */
void orc_calc_contacts(orc_state* s);
void orc_calc_geometric_contacts(orc_state* s, double gap);
long orc_contact_list_size(const orc_state* s);
void orc_get_contacts(const orc_state* s, long* list, long* counts);
double orc_get_hexatic_order(orc_state* s);
void orc_set_initial_positions(orc_state* s);
double orc_get_vertex_msd(orc_state* s);
double orc_get_particle_msd(orc_state* s);
double orc_get_vertex_isf(orc_state* s);
double orc_get_particle_isf(orc_state* s, double waveNumber);
void orc_calc_per_particle_stress(orc_state* s, double* out4np);
void orc_calc_stress_tensor(orc_state* s, double* out4);
double orc_get_pressure(orc_state* s);
void orc_init_langevin(orc_state* s, double T, double gamma, int readState, int zeroMomentum);
void orc_set_active_tension(orc_state* s, double l0Tau, double l0Diff);
void orc_update_rest_length(orc_state* s);
void orc_init_nve(orc_state* s, double T, int readState, int zeroMomentum);
void orc_init_nve_rescale(orc_state* s, double T);
void orc_inject_kinetic_energy(orc_state* s, int zeroMomentum);
void orc_langevin_loop(orc_state* s, long steps);
void orc_nve_loop(orc_state* s, long steps);
void orc_nve_rescale_loop(orc_state* s, long steps);

/* FIRE (FIRE.cu:29-41, 300-411; DPM2D.cu:3660-3675) */
void orc_init_fire(orc_state* s, const double* params7, long minStep, long numStep);
void orc_fire_loop(orc_state* s, long iterations);
/* out[0..4] = fire_dt, a, numStep, last P, updateCount */
void orc_get_fire_state(const orc_state* s, double* out);
long orc_fire_minimize(orc_state* s, double tol, long maxIter);

/* observables (DPM2D.cu:739-1238) */
double orc_get_potential_energy(orc_state* s);
double orc_get_kinetic_energy(const orc_state* s);
double orc_get_temperature(const orc_state* s);
double orc_get_max_unbalanced_force(const orc_state* s);
double orc_get_total_force_magnitude(const orc_state* s);
double orc_get_phi(orc_state* s);
double orc_get_preferred_phi(const orc_state* s);
double orc_get_mean_particle_size(const orc_state* s);
double orc_get_mean_vertex_radius(const orc_state* s);
void orc_adjust_temperature(orc_state* s, double targetTemp);
double orc_adjust_kinetic_energy(orc_state* s, double prevEtot);
void orc_scale_vertices(orc_state* s, double scale);
void orc_scale_box(orc_state* s, double scale);
/* rigid rings */
void orc_calc_rigid_force_energy(orc_state* s);
void orc_init_rigid(orc_state* s, double T, double gamma, int langevin, int readState);
void orc_rigid_loop(orc_state* s, long steps);
void orc_init_rigid_fire(orc_state* s, const double* p, long minStep, long numStep);
void orc_rigid_fire_loop(orc_state* s, long iterations);
double orc_get_rigid_max_unbalanced_force(const orc_state* s);
double orc_get_rigid_kinetic_energy(const orc_state* s);
void orc_get_particle_shapes(orc_state* s, double* shape3);

/* RNG: Philox4x32-10 (Salmon et al., SC'11) and the Box-Muller pair built on it */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_gauss_pair(uint64_t seed, uint32_t systemId, uint32_t tag, uint64_t counter,
                    uint32_t index, double* z0, double* z1);

#ifdef __cplusplus
}
#endif
#endif
