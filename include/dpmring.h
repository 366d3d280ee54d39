/*
 * dpmring.h -- C ABI of the MI355X (gfx950) DPM2D vertex engine, libdpmring.so.
 *
 * The reference (farceri/cudaRing) has no FFI: its seam is the DPM2D / FIRE /
 * SimInterface C++ class API over raw thrust::device_vector pointers and
 * __constant__ device globals.  This ABI is what those classes (re-created in
 * cudaring_amd/host/) and the Python/torch wrapper call.  Each entry point
 * names the reference method it stands in for.
 *
 * Conventions
 *   - every call returns an int status (DPM_OK = 0, negative on error) and never
 *     throws; dpm_last_error() returns the message of the last failure;
 *   - a handle is a "set" of B >= 1 independent systems (packings) stored back to
 *     back on one device; single-system use is B = 1.  The reference's
 *     one-DPM2D-per-process limit (device __constant__ state) does not apply;
 *   - host-side 2-vectors (positions, velocities, forces, ring centres) use the
 *     reference's interleaved layout (x0, y0, x1, y1, ...); on the device they are
 *     stored as structure-of-arrays FP64;
 *   - vertex and ring indices are global over the set, in system order;
 *   - `stream` arguments are hipStream_t passed as void*; NULL = default stream;
 *   - a handle is not thread-safe; distinct handles are independent.
 */
#ifndef DPMRING_H
#define DPMRING_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DPM_OK 0
#define DPM_ERR_ARG (-1)
#define DPM_ERR_HIP (-2)
#define DPM_ERR_OOM (-3)
#define DPM_ERR_UNSUPPORTED (-4)
#define DPM_ERR_OVERFLOW (-5)
#define DPM_ERR_STATE (-6)

typedef struct dpm_set_s* dpm_set_t;

/* field ids for dpm_set_field / dpm_get_field (sizes: V = vertices, R = rings, B = systems) */
enum {
  DPM_FIELD_POS = 0,       /* 2V  setVertexPositions / getVertexPositions, DPM2D.h:357-358     */
  DPM_FIELD_VEL = 1,       /* 2V  setVertexVelocities / getVertexVelocities, DPM2D.h:363-364    */
  DPM_FIELD_FORCE = 2,     /* 2V  setVertexForces / getVertexForces, DPM2D.h:365-366            */
  DPM_FIELD_ENERGY = 3,    /* V   d_energy (per-vertex shape + vertex-vertex energy)            */
  DPM_FIELD_RAD = 4,       /* V   setVertexRadii / getVertexRadii, DPM2D.h:283-284              */
  DPM_FIELD_L0 = 5,        /* V   setRestLengths / getRestLengths, DPM2D.h:287-288              */
  DPM_FIELD_THETA0 = 6,    /* V   setRestAngles / getRestAngles, DPM2D.h:289-290                */
  DPM_FIELD_THETA = 7,     /* V   getSegmentAngles, DPM2D.h:292                                  */
  DPM_FIELD_LENGTH = 8,    /* V   getSegmentLengths, DPM2D.h:291                                 */
  DPM_FIELD_A0 = 9,        /* R   setRestAreas / getRestAreas, DPM2D.h:285-286                  */
  DPM_FIELD_AREA = 10,     /* R   getAreas, DPM2D.h:294                                          */
  DPM_FIELD_PERIMETER = 11,/* R   getPerimeters, DPM2D.h:295                                     */
  DPM_FIELD_PPOS = 12,     /* 2R  getParticlePositions (ring centres), DPM2D.h:318               */
  DPM_FIELD_PENERGY = 13,  /* R   getParticleEnergies (smooth interaction energy), DPM2D.h:325   */
  DPM_FIELD_PANGLE = 14,   /* R   setParticleAngles / getParticleAngles, DPM2D.h:326-327         */
  DPM_FIELD_LASTPOS = 15,  /* 2V  d_lastPos (neighbour-list reference positions)               */
  DPM_FIELD_BOX = 16,      /* 2B  setBoxSize / getBoxSize, DPM2D.h:275-276 (per system)        */
  DPM_FIELD_WALLFORCE = 17,/* 2V  d_wallForce (force on the walls per vertex), DPM2D.h:146       */
  DPM_FIELD_PVEL = 18,     /* 2R  setParticleVelocities / getParticleVelocities, DPM2D.h:339-340  */
  DPM_FIELD_PANGVEL = 19,  /* R   getParticleAngularVelocities (rigid rings), DPM2D.h:351         */
  DPM_FIELD_PFORCE = 20,   /* 2R  getParticleForces (rigid ring force), DPM2D.h:343               */
  DPM_FIELD_PTORQUE = 21,  /* R   getParticleTorques (ring torque / moment of inertia), DPM2D.h:354 */
  DPM_FIELD_VTORQUE = 22,  /* V   getVertexTorques (about the ring centre), DPM2D.h:373           */
  DPM_FIELD_PRAD = 23,     /* R   setParticleRadii / getParticleRadii, DPM2D.h:327-328           */
  DPM_FIELD_COUNT = 24
};

/* potential / interaction / monomer / particle enums keep simControlStruct's order, DPM2D.h:22-32 */
enum { DPM_POT_HARMONIC = 0, DPM_POT_LJ = 1, DPM_POT_ADHESIVE = 2, DPM_POT_WCA = 3 };
enum { DPM_INT_VERTEX_VERTEX = 0, DPM_INT_VERTEX_SMOOTH = 1, DPM_INT_ALL = 3 };
enum { DPM_MONO_HARMONIC = 0, DPM_MONO_FENE = 1 };
enum { DPM_PART_PASSIVE = 0, DPM_PART_ACTIVE = 1, DPM_PART_ACTIVE_SHAPE = 2 };
enum { DPM_NOISE_LANGEVIN = 0, DPM_NOISE_BAOAB = 1, DPM_NOISE_BROWNIAN = 2, DPM_NOISE_DRIVEN_BROWNIAN = 3 };
enum { DPM_GEOM_NORMAL = 0, DPM_GEOM_FIXED_BOX = 1, DPM_GEOM_FIXED_SIDES = 2, DPM_GEOM_ROUND_BOX = 3 };

/* ---- lifecycle ------------------------------------------------------------ */
/* Replaces DPM2D(long nParticles, long dim, long nVertexPerParticle) (DPM2D.h:47) followed by
 * setNumVertexInParticleList (DPM2D.h:267) and initParticleIdList (DPM2D.h:219), for B systems:
 * ringsPerSystem[B], numVertexInParticleList[sum(ringsPerSystem)] (n >= 3).  A ring of more than
 * 64 vertices is cut into slices of 64 (one tile each) with its ring sums taken by an extra launch
 * per step; sets holding such rings support passive particles only, and no rigid-ring, adjoint
 * or one-workgroup-per-system FIRE path (those calls return DPM_ERR_UNSUPPORTED; FIRE runs the
 * multi-launch kernels). */
int dpm_create(dpm_set_t* out, int device, long numSystems, const long* ringsPerSystem,
               const long* numVertexInParticleList);
int dpm_destroy(dpm_set_t h);
int dpm_set_stream(dpm_set_t h, void* stream);
const char* dpm_last_error(dpm_set_t h);
/* getNumParticles / getNumVertices (DPM2D.h:259-263), summed over the set */
int dpm_get_sizes(dpm_set_t h, long* numSystems, long* numParticles, long* numVertices);
int dpm_synchronize(dpm_set_t h);

/* ---- state ---------------------------------------------------------------- */
/* the host_vector setters and getters of DPM2D.h:267-366 (see the field ids above) */
int dpm_set_field(dpm_set_t h, int field, const double* host, long n);
int dpm_get_field(dpm_set_t h, int field, double* host, long n);
/* device pointers of the SoA storage, for zero-copy use by a caller on the same stream:
 * component 0 = x (or the scalar field), 1 = y.  Positions point at the current buffer. */
int dpm_device_ptr(dpm_set_t h, int field, int component, void** ptr);

/* ---- parameters ----------------------------------------------------------- */
/* setPotentialType / setInteractionType / setMonomerType / setParticleType, DPM2D.h:225-250 */
int dpm_set_types(dpm_set_t h, int potential, int interaction, int monomer, int particle);
/* setGeometryType, DPM2D.h:228 / DPM2D.cu:338-358 (simControlStruct::geometryEnum order): walls
 * (WCA with energy scale ew) at x = 0, Lx and y = 0, Ly (fixedBox), at y = 0, Ly with x periodic
 * (fixedSides), or on the circle of radius boxRadius about the origin (roundBox); distances lose
 * their periodic image along walled directions (boxDistance, cudaKernel.cuh:76-94) */
int dpm_set_geometry(dpm_set_t h, int geometry);
/* setBoxEnergyCost(ew), DPM2D.h:537 / DPM2D.cu:1911-1914 (default 1e3, DPM2D.cu:70) */
int dpm_set_box_energy_cost(dpm_set_t h, double ew);
/* setBoxRadius / getBoxRadius, DPM2D.h:278-280 / DPM2D.cu:627-664 (roundBox) */
int dpm_set_box_radius(dpm_set_t h, double boxRadius);
int dpm_get_box_radius(dpm_set_t h, double* boxRadius);
/* setEnergyCosts(ea, el, eb, ec), DPM2D.h:525 */
int dpm_set_energy_costs(dpm_set_t h, double ea, double el, double eb, double ec);
/* Per-ring shape costs ka[R], kl[R], kb[R] replacing the global ea, el, eb of setEnergyCosts in the
 * area, perimeter and bending terms; all three NULL returns to the globals.  No reference
 * counterpart: the reference's costs are device-global scalars (cudaKernel.cuh:39-42). */
int dpm_set_ring_costs(dpm_set_t h, const double* ka, const double* kl, const double* kb);
/* setAttractionConstants(l1, l2), DPM2D.h:535 */
int dpm_set_attraction_constants(dpm_set_t h, double l1, double l2);
/* setLJcutoff, DPM2D.h:539 (energy and force shifts use the current ec) */
int dpm_set_lj_cutoff(dpm_set_t h, double LJcutoff);
/* setFENEconstants, DPM2D.h:541 */
int dpm_set_fene_constants(dpm_set_t h, double stiff, double ext);
/* setTimeStep, DPM2D.h:545 */
int dpm_set_time_step(dpm_set_t h, double dt);
/* setSelfPropulsionParams(driving, taup), DPM2D.h:527 */
int dpm_set_self_propulsion(dpm_set_t h, double driving, double taup);
/* setActiveTensionParams(l0Tau, l0Diff), DPM2D.h:531 / DPM2D.cu:1898-1903: the rest lengths of the
 * activeShape particle type become the current l0; each force call then moves
 * l0 += ((l0 - l0Rest) / l0Tau + sqrt(2 l0Diff / dt) xi_ring) dt (updateRestLength, DPM2D.cu:2037-2056) */
int dpm_set_active_tension(dpm_set_t h, double l0Tau, double l0Diff);
/* seed of the Philox streams that replace thrust::default_random_engine + lrand48
 * (defs.h:30-79, DPM2D.cu:43); resets the per-stream call counters */
int dpm_set_seed(dpm_set_t h, uint64_t seed);

/* ---- neighbours ----------------------------------------------------------- */
/* setDisplacementCutoff(cutoff, size) -> cutDistance, DPM2D.h:447 / DPM2D.cu:1300-1329 */
int dpm_set_displacement_cutoff(dpm_set_t h, double cutoff, double size, double* cutDistance);
/* calcNeighborList(cutDistance), DPM2D.h:640 / DPM2D.cu:3479-3507 (built with a cell list;
 * the list holds the same sets as the reference's all-pairs build) */
int dpm_calc_neighbor_list(dpm_set_t h, double cutDistance);
/* Cell side of the neighbour search as a multiple (>= 1) of its smallest value, the list radius
 * plus the 2 cutoff / 3 a vertex may move between binning and rebuild; a tuning knob (the
 * reference builds its lists without cells, DPM2D.cu:3479-3507, so no counterpart).  The rows
 * are the same sets for every scale. */
int dpm_set_cell_scale(dpm_set_t h, double scale);
/* checkNeighbors(), DPM2D.h:459 / DPM2D.cu:1408-1423: device-side displacement test and rebuild */
int dpm_check_neighbors(dpm_set_t h);
/* getNeighbors + getMaxNeighborList (DPM2D.h:411-412): rows of `cap` ascending global ids,
 * -1 padded, plus counts[V]; returns the device row capacity in *width */
int dpm_get_neighbor_list(dpm_set_t h, long* list, long* counts, long cap, long* width);
/* resetLastPositions (DPM2D.h:359), getUpdateCount / resetUpdateCount (DPM2D.h:449-450) */
int dpm_reset_last_positions(dpm_set_t h);
int dpm_get_update_count(dpm_set_t h, long* perSystem);
int dpm_reset_update_count(dpm_set_t h);

/* ---- forces --------------------------------------------------------------- */
/* calcParticleShape + calcParticlePositions, DPM2D.h:303-304 */
int dpm_calc_particle_shape(dpm_set_t h);
/* calcForceEnergy(), DPM2D.h:547 / DPM2D.cu:1962-2104 */
int dpm_calc_force_energy(dpm_set_t h);
/* getInteractionForces (DPM2D.h:575; written to forces.dat by savePacking, FileIO.h:424): the
 * forces of the vertex interaction alone at the current positions, AoS 2V.  Stored forces and
 * energies are left as they were.  (The reference's vertexVertex branch writes per-vertex
 * energies past the end of a per-ring array, DPM2D.cu:2221-2251; this returns the forces it means.) */
int dpm_get_interaction_forces(dpm_set_t h, double* host, long n);

/* ---- observables ---------------------------------------------------------- */
/* setInitialPositions + setInitialParticlePositions, DPM2D.cu:921-930: the MSD / ISF reference
 * positions (vertices and ring centres); the Langevin, NVE and NVE-rescale inits also store them */
int dpm_set_initial_positions(dpm_set_t h);
/* getVertexMSD / getParticleMSD, DPM2D.h:441-443 / DPM2D.cu:1284-1298 (per system) */
int dpm_get_vertex_msd(dpm_set_t h, double* perSystem);
int dpm_get_particle_msd(dpm_set_t h, double* perSystem);
/* getVertexISF / getParticleISF(waveNumber), DPM2D.h:473-475 / DPM2D.cu:1486-1503 (per system) */
int dpm_get_vertex_isf(dpm_set_t h, double* perSystem);
int dpm_get_particle_isf(dpm_set_t h, double waveNumber, double* perSystem);
/* calcContacts / getContacts, DPM2D.h:389, 632 / DPM2D.cu:1019-1024, 3408-3428: per ring the rings
 * owning an entry of its vertices' rows, in order of first appearance; rows of `cap` (-1 padded),
 * counts[R], device row width in *width */
int dpm_calc_contacts(dpm_set_t h);
/* calcGeometricContacts, DPM2D.h:634 / DPM2D.cu:3430-3450: rings within one radius sum of a
 * vertex (gap 1), rows of 8 (4 * nDim), contacts beyond 8 dropped; read with dpm_get_contacts */
int dpm_calc_geometric_contacts(dpm_set_t h);
int dpm_get_contacts(dpm_set_t h, long* list, long* counts, long cap, long* width);
/* getParticleHexaticOrder, DPM2D.h:477 / DPM2D.cu:1505-1510 (from the last dpm_calc_contacts) */
int dpm_get_hexatic_order(dpm_set_t h, double* perSystem);
/* getPerParticleStressTensor / getStressTensor / getPressure, DPM2D.h:375-379 / DPM2D.cu:962-983,
 * 2181-2198 (harmonic contacts): 4 per ring, 4 per system, 1 per system */
int dpm_get_per_particle_stress(dpm_set_t h, double* perRing4);
int dpm_get_stress_tensor(dpm_set_t h, double* perSystem4);
int dpm_get_pressure(dpm_set_t h, double* perSystem);

/* ---- integrators ---------------------------------------------------------- */
/* setNoiseType, DPM2D.h:237 / DPM2D.cu:410-423: the integrator of langevinLoop, in
 * simControlStruct::noiseEnum order (0 langevin, 1 baoab, 2 brownian, 3 drivenBrownian;
 * Simulator.cu:23-30, 112-144, 147-179, 182-203) */
int dpm_set_noise_type(dpm_set_t h, int noiseType);
/* initLangevin(T, gamma, readState) / langevinLoop(), DPM2D.h:680-682, DPM2D.cu:3713-3753.
 * zeroMomentum removes the mean velocity after injection (an addition; the reference does not). */
int dpm_init_langevin(dpm_set_t h, double T, double gamma, int readState, int zeroMomentum);
int dpm_langevin_loop(dpm_set_t h, long steps);
/* The stages of Langevin::integrate (Simulator.cu:23-95) for callers that drive them one by
 * one (SimInterface, Simulator.h:47-52): injectKineticEnergy (per-ring Gaussian velocities of
 * amplitude sqrt(T), nothing else), updatePosition (x += timeStep v), updateVelocity
 * (v += timeStep F), updateThermalVel (F += noise xi - gamma v with one draw per ring, the draw the
 * fused step of the same call index uses; needs dpm_init_langevin with Langevin noise) and
 * conserveMomentum (the mean vertex velocity of each system removed).  In that order with
 * checkNeighbors and calcForceEnergy, they reproduce one dpm_langevin_loop step. */
int dpm_inject_kinetic_energy(dpm_set_t h, double T);
int dpm_update_position(dpm_set_t h, double timeStep);
int dpm_update_velocity(dpm_set_t h, double timeStep);
int dpm_update_thermal_vel(dpm_set_t h);
int dpm_conserve_momentum(dpm_set_t h);
/* initNVE(T, readState) / NVELoop(), DPM2D.h:684-686 */
int dpm_init_nve(dpm_set_t h, double T, int readState, int zeroMomentum);
int dpm_nve_loop(dpm_set_t h, long steps);
/* initNVERescale(T) / NVERescaleLoop(), DPM2D.h:688-690 */
int dpm_init_nve_rescale(dpm_set_t h, double T);
int dpm_nve_rescale_loop(dpm_set_t h, long steps);
/* adjustTemperature(T), DPM2D.h:429 / DPM2D.cu:1185-1200 (per system) */
int dpm_adjust_temperature(dpm_set_t h, double T);
/* adjustKineticEnergy(prevEtot), DPM2D.h:427 / DPM2D.cu:1161-1183, per system: with K and U the
 * current kinetic and potential energies and dE = U + K - prevEtot[s], the velocities scale by
 * sqrt((K - dE) / K) when K > dE and are left alone otherwise; scaleOut[s] (may be NULL) receives
 * the scale, or 0 where no adjustment was made */
int dpm_adjust_kinetic_energy(dpm_set_t h, const double* prevEtot, double* scaleOut);

/* ---- FIRE ----------------------------------------------------------------- */
/* initFIRE(params{a_start, f_dec, f_inc, f_a, dt, dt_max, a}, minStep, numStep, numDOF),
 * DPM2D.h:665; the state is per system */
int dpm_init_fire(dpm_set_t h, const double* params7, long minStep, long numStep);
/* vertexFIRELoop() x iterations, DPM2D.h:669 / FIRE.cu:397-411 (all systems advance) */
int dpm_fire_loop(dpm_set_t h, long iterations);
/* Kernel of dpm_fire_minimize: persistent = 1 (default) runs one workgroup per system that
 * iterates with workgroup barriers only (wavesPerSystem 1..16, 0 = one per tile of the largest
 * system up to 16); persistent = 0 replays graphs of the split FIRE step kernels with per-system
 * freezing.  No reference counterpart (launch configuration, chosen by the autotuner). */
int dpm_set_fire_kernel(dpm_set_t h, int persistent, int wavesPerSystem);
/* FIRE state per system: out[5*s + {0..4}] = fire_dt, a, numStep, last P = v.F, iterations */
int dpm_get_fire_state(dpm_set_t h, double* out);
/* The compressFIRE.cpp:121-131 loop for every system at once: iterate each system until
 * getMaxUnbalancedForce() <= tol or maxIter; converged systems stop moving.  Each call tests the
 * stopping rule of every system afresh and continues the FIRE state (dt, alpha, N+) where the
 * last call left it; maxIter bounds the iterations since dpm_init_fire, whose per-system count
 * iterations[s] receives.  Runs without host synchronisation between iterations. */
int dpm_fire_minimize(dpm_set_t h, double tol, long maxIter, long* iterations);

/* ---- observables (per system) --------------------------------------------- */
/* getPotentialEnergy / getKineticEnergy / getTemperature, DPM2D.h:417-420 */
int dpm_get_potential_energy(dpm_set_t h, double* perSystem);
int dpm_get_kinetic_energy(dpm_set_t h, double* perSystem);
int dpm_get_temperature(dpm_set_t h, double* perSystem);
/* getMaxUnbalancedForce / getTotalForceMagnitude, DPM2D.h:408-409 */
int dpm_get_max_unbalanced_force(dpm_set_t h, double* perSystem);
int dpm_get_total_force_magnitude(dpm_set_t h, double* perSystem);
/* getPhi / getPreferredPhi, DPM2D.h:439-440 */
int dpm_get_phi(dpm_set_t h, double* perSystem);
int dpm_get_preferred_phi(dpm_set_t h, double* perSystem);
/* getParticleShapes (A, P, P^2/4piA per ring), DPM2D.h:296 */
int dpm_get_particle_shapes(dpm_set_t h, double* shape3);
/* scaleVertices(scale) per system, DPM2D.h:492 / DPM2D.cu:1802-1811 */
int dpm_scale_vertices(dpm_set_t h, const double* scalePerSystem);
/* scaleBox(scale) per system, DPM2D.h:498 / DPM2D.cu:1693-1719: ring centres and the box divide
 * by scale and each ring moves with its centre (translateVertices, DPM2D.cu:1834-1839); in the
 * round box the radius divides instead, with one scale for every system.  The neighbour list
 * is not rebuilt (call dpm_calc_neighbors, as the reference's drivers do). */
int dpm_scale_box(dpm_set_t h, const double* scalePerSystem);

/* ---- rigid rings (shapeEnum::rigid) ------------------------------------------------------ *
 * Each ring moves as a rigid body: its centre (DPM_FIELD_PPOS), velocity (PVEL), angle (PANGLE)
 * and angular velocity (PANGVEL) are driven by the sum of its vertices' interaction forces
 * (PFORCE) and their torque about the centre divided by the moment of inertia sum |r|^2 (PTORQUE,
 * an angular acceleration).  No shape forces, walls or activity; vertexVertex or vertexSmooth,
 * periodic boxes. */
/* calcRigidForceEnergy, DPM2D.h:567 / DPM2D.cu:2128-2142 */
int dpm_calc_rigid_force_energy(dpm_set_t h);
/* initRigidLangevin(T, gamma, readState) / rigidLangevinLoop(), DPM2D.h:693-695 /
 * DPM2D.cu:3823-3847; Simulator.cu:241-313 */
int dpm_init_rigid_langevin(dpm_set_t h, double T, double gamma, int readState);
int dpm_rigid_langevin_loop(dpm_set_t h, long steps);
/* initRigidNVE(T, readState) / rigidNVELoop(), DPM2D.h:697-699 / DPM2D.cu:3850-3868;
 * Simulator.cu:316-371 */
int dpm_init_rigid_nve(dpm_set_t h, double T, int readState);
int dpm_rigid_nve_loop(dpm_set_t h, long steps);
/* initRigidFIRE(params7, minStep, numStep, ...) / rigidFIRELoop(), DPM2D.h:675-677 /
 * DPM2D.cu:3700-3710; FIRE.cu:173-298 (iterations per call) */
int dpm_init_rigid_fire(dpm_set_t h, const double* params7, long minStep, long numStep);
int dpm_rigid_fire_loop(dpm_set_t h, long iterations);
/* rigid FIRE scalars per system: dt, a, numStep, last power v.F, iterations (FIRE.h:30-40);
 * no reference counterpart (the reference keeps them in its FIRE object) */
int dpm_get_rigid_fire_state(dpm_set_t h, double* out5PerSystem);
/* getRigidMaxUnbalancedForce, DPM2D.h:660 / DPM2D.cu:3643-3652 (per system) */
int dpm_get_rigid_max_unbalanced_force(dpm_set_t h, double* perSystem);
/* getRigidKineticEnergy, DPM2D.h:425 / DPM2D.cu:1149-1159 (per system; needs DPM_FIELD_PRAD) */
int dpm_get_rigid_kinetic_energy(dpm_set_t h, double* perSystem);

/* ---- launch configuration and measurement (no reference counterpart) ------- */
/* Wavefronts per ring tile in the step kernel (1..8): the gather over a tile's neighbour rows
 * is split across them.  Chosen by the autotuner (cudaring_amd/autotune.py). */
int dpm_set_launch_config(dpm_set_t h, int wavesPerTile);
/* Average duration of the fused Langevin step kernel (ring shape + interaction + thermostat +
 * kicks + drift) over `reps` back-to-back launches on the handle's stream, timed with hipEvents
 * on that stream.  Velocities are restored afterwards; *meanNeighbors = mean list length. */
int dpm_time_step_kernel(dpm_set_t h, long reps, double* avgMicroseconds, double* meanNeighbors);

/* ---- adjoint (no reference counterpart: differentiation is an addition) ---- */
/* Supported for the vertexVertex and vertexSmooth interactions with harmonic monomers in a
 * periodic box, all four potentials (vertexSmooth: exact on the branch the positions select, the
 * interaction being piecewise smooth); passive particles, and for every entry point but the FIRE
 * pair also self-propelled ones (active: their driving force depends on neither positions nor
 * parameters).  Vectors are device arrays owned
 * by the caller and used on the handle's stream; a "2V SoA" vector holds the V x components
 * followed by the V y components. */
/* Hessian-vector product of the potential energy at the current positions: out = H u
 * (ux, uy, outx, outy: device arrays of V). */
int dpm_force_hvp(dpm_set_t h, const double* ux, const double* uy, double* outx, double* outy);
/* Gradient of the potential energy at the current shape variables (after calcForceEnergy)
 * with respect to the per-ring costs, gring = [dU/dka(R) | dU/dkl(R) | dU/dkb(R)], and to the
 * rest lengths, gl0[V]; either may be NULL. */
int dpm_energy_param_grad(dpm_set_t h, double* gring, double* gl0);
/* dpm_energy_param_grad without any device allocation, for callers that cannot allocate (code
 * under HIP graph capture, tight serving or optimisation loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Its per-vertex partials live in a fixed
 * library buffer of 4 x DPM_NOALLOC_MAX_VERTICES doubles per GPU (8 MiB, held until the process
 * ends); sets with more vertices get DPM_ERR_ARG.  Calls on one GPU share that buffer: calls that
 * overlap on one GPU (other streams or threads) race.  Same arguments and results as
 * dpm_energy_param_grad. */
#define DPM_NOALLOC_MAX_VERTICES 262144
int dpm_energy_param_grad_noalloc(dpm_set_t h, double* gring, double* gl0);
/* NVELoop() x steps (DPM2D.h:686) from the current state, keeping a checkpoint of positions,
 * velocities and forces (self-propelled rings: also ring angles and RNG call counters, so that the
 * recomputed segments draw the same angle increments) every `every` steps for dpm_nve_backward. */
int dpm_nve_forward(dpm_set_t h, long steps, long every);
/* Reverse pass of the last dpm_nve_forward: given ax = dL/dx_N and av = dL/dv_N (2V SoA), returns
 * gx0 = dL/dx_0, gv0 = dL/dv_0 (2V SoA), gring = dL/d(ka, kl, kb) per ring (3R) and
 * gl0 = dL/dl0 (V) through the dynamics; outputs may be NULL.  Segments between checkpoints are
 * recomputed; the set is left in the state the forward pass ended in. */
int dpm_nve_backward(dpm_set_t h, const double* ax, const double* av, double* gx0, double* gv0, double* gring,
                     double* gl0);

/* vertexFIRELoop() x iterations (DPM2D.h:669, FIRE.cu:338-411) from the current state after
 * dpm_init_fire, every system advancing (no stopping rule), recording for dpm_fire_backward the
 * state before every iteration (positions, velocities, forces) and each system's FIRE scalars
 * after its parameter update.  The tape holds (iterations + 1) x 6 V doubles. */
int dpm_fire_forward(dpm_set_t h, long iterations);
/* Reverse pass of the last dpm_fire_forward with its branch sequence held fixed (each system's
 * dt, alpha and restart of every iteration, as recorded): given ax = dL/dx_N and av = dL/dv_N
 * (2V SoA), returns gx0 = dL/dx_0, gv0 = dL/dv_0 (2V SoA), gring = dL/d(ka, kl, kb) per ring (3R)
 * and gl0 = dL/dl0 (V); outputs may be NULL.  One Hessian-vector product per recorded position;
 * the set keeps the state the forward pass ended in. */
int dpm_fire_backward(dpm_set_t h, const double* ax, const double* av, double* gx0, double* gv0, double* gring,
                      double* gl0);

#ifdef __cplusplus
}
#endif
#endif /* DPMRING_H */
