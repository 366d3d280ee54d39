// DPM2D.cpp -- DPM2D over libdpmring.so (one system per object).
//
// Set-up (size distributions, random particles, soft-disk FIRE, vertices on particles) runs on
// the host before a device set exists, as the reference's element-wise host loops do
// (DPM2D.cu:1522-1691).  The first call that needs the device creates a one-system set and
// uploads the host state; from then on the device holds the state and getters copy it back.
#include "DPM2D.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <iomanip>
#include <iostream>
#include <numeric>
#include <stdexcept>

#include "FIRE.h"
#include "Simulator.h"

using std::cout;
using std::endl;

namespace {
// minimum image, round half away from zero (cudaKernel.cuh:70-74)
inline double pbc(double d, double L) { return d - L * std::round(d / L); }
}  // namespace

DPM2D::DPM2D(long nParticles, long dim, long nVertexPerParticle) {
  if (dim != 2) throw std::invalid_argument("DPM2D: only nDim = 2 is supported");
  if (nParticles < 1) throw std::invalid_argument("DPM2D: nParticles must be >= 1");
  if (nVertexPerParticle < 3) throw std::invalid_argument("DPM2D: vertices per particle must be >= 3");
  srand48(time(0));  // as DPM2D.cu:43; setSeed() makes a run reproducible
  seed = (uint64_t)time(0);
  nDim = dim;
  numParticles = nParticles;
  numVertexPerParticle = nVertexPerParticle;
  numVertices = numParticles * numVertexPerParticle;
  simControl.simulationType = simControlStruct::simulationEnum::gpu;
  simControl.geometryType = simControlStruct::geometryEnum::normal;
  simControl.particleType = simControlStruct::particleEnum::passive;
  simControl.shapeType = simControlStruct::shapeEnum::deformable;
  simControl.noiseType = simControlStruct::noiseEnum::langevin;
  simControl.potentialType = simControlStruct::potentialEnum::harmonic;
  simControl.interactionType = simControlStruct::interactionEnum::vertexVertex;
  simControl.neighborType = simControlStruct::neighborEnum::neighbor;
  simControl.monomerType = simControlStruct::monomerEnum::harmonic;
  // default constants of DPM2D.cu:63-71
  dt = 1e-3;
  rho0 = 1;
  calA0 = 1;
  ea = 1e5;
  el = 20;
  eb = 10;
  ec = 1;
  driving = 0;
  taup = 1;
  l1 = 0;
  l2 = 0;
  LJcutoff = 0;
  LJecut = 0;
  LJfshift = 0;
  stiff = 0;
  extSq = 0;
  cutDistance = 1;
  cutoff = 0;
  boxSize.assign(2, 1.0);
  std::vector<long> n(numParticles, numVertexPerParticle);
  setNumVertexInParticleList(n);
  particlePos.assign(2 * numParticles, 0.0);
  particleVel.assign(2 * numParticles, 0.0);
  particleForce.assign(2 * numParticles, 0.0);
  particleRad.assign(numParticles, 0.0);
  particleEnergy.assign(numParticles, 0.0);
  particleLastPos.assign(2 * numParticles, 0.0);
}

DPM2D::~DPM2D() {
  delete fire_;
  delete sim_;
  if (h_) dpm_destroy(h_);
}

void DPM2D::check(int rc, const char* what) {
  if (rc == DPM_OK) return;
  std::string msg = std::string("DPM2D::") + what + " failed (" + std::to_string(rc) + ")";
  if (h_) msg += ": " + std::string(dpm_last_error(h_));
  throw std::runtime_error(msg);
}

// ---------------------------------------------------------------- device set
dpm_set_t DPM2D::handle() {
  if (h_) return h_;
  long rings = numParticles;
  dpm_set_t h = nullptr;
  int rc = dpm_create(&h, 0, 1, &rings, numVertexInParticleList.data());
  if (rc != DPM_OK) throw std::runtime_error("DPM2D: dpm_create failed (" + std::to_string(rc) + ")");
  h_ = h;
  set_field(DPM_FIELD_BOX, boxSize);
  set_field(DPM_FIELD_RAD, hRad);
  set_field(DPM_FIELD_L0, hL0);
  set_field(DPM_FIELD_THETA0, hTheta0);
  set_field(DPM_FIELD_A0, hA0);
  set_field(DPM_FIELD_POS, hPos);
  set_field(DPM_FIELD_VEL, hVel);
  push_params();
  return h_;
}

void DPM2D::release() {
  if (!h_) return;
  hPos = field(DPM_FIELD_POS, 2 * numVertices);
  hVel = field(DPM_FIELD_VEL, 2 * numVertices);
  hRad = field(DPM_FIELD_RAD, numVertices);
  hL0 = field(DPM_FIELD_L0, numVertices);
  hTheta0 = field(DPM_FIELD_THETA0, numVertices);
  hA0 = field(DPM_FIELD_A0, numParticles);
  dpm_destroy(h_);
  h_ = nullptr;
  if (fire_) fire_->deviceReady = false;
}

void DPM2D::push_params() {
  if (!h_) return;
  int pot = (int)simControl.potentialType;
  int inter;
  switch (simControl.interactionType) {
    case simControlStruct::interactionEnum::vertexVertex: inter = DPM_INT_VERTEX_VERTEX; break;
    case simControlStruct::interactionEnum::vertexSmooth: inter = DPM_INT_VERTEX_SMOOTH; break;
    case simControlStruct::interactionEnum::all: inter = DPM_INT_ALL; break;
    default: throw std::invalid_argument("DPM2D: cellSmooth interaction is not provided");
  }
  const int part = (int)simControl.particleType;  // passive, active, activeShape = DPM_PART_*
  check(dpm_set_types(h_, pot, inter, (int)simControl.monomerType, part), "setTypes");
  check(dpm_set_noise_type(h_, (int)simControl.noiseType), "setNoiseType");
  check(dpm_set_geometry(h_, (int)simControl.geometryType), "setGeometryType");
  check(dpm_set_box_energy_cost(h_, ew), "setBoxEnergyCost");
  if (boxRadius > 0) check(dpm_set_box_radius(h_, boxRadius), "setBoxRadius");
  check(dpm_set_energy_costs(h_, ea, el, eb, ec), "setEnergyCosts");
  check(dpm_set_attraction_constants(h_, l1, l2), "setAttractionConstants");
  if (LJcutoff > 0) check(dpm_set_lj_cutoff(h_, LJcutoff), "setLJcutoff");
  if (extSq > 0) check(dpm_set_fene_constants(h_, stiff, std::sqrt(extSq)), "setFENEconstants");
  check(dpm_set_time_step(h_, dt), "setTimeStep");
  check(dpm_set_self_propulsion(h_, driving, taup), "setSelfPropulsionParams");
  check(dpm_set_seed(h_, seed), "setSeed");
  if (cutoff > 0) {
    double cd;
    check(dpm_set_displacement_cutoff(h_, cutoffArg, sizeArg, &cd), "setDisplacementCutoff");
  }
}

std::vector<double> DPM2D::field(int f, long n) {
  std::vector<double> out(n);
  check(dpm_get_field(handle(), f, out.data(), n), "getField");
  return out;
}

void DPM2D::set_field(int f, const std::vector<double>& v) {
  check(dpm_set_field(handle(), f, v.data(), (long)v.size()), "setField");
}

// ---------------------------------------------------------------- types
void DPM2D::setSimulationType(simControlStruct::simulationEnum t) {
  if (t != simControlStruct::simulationEnum::gpu)
    throw std::invalid_argument("DPM2D: this engine runs on the GPU only (simulationType gpu)");
  simControl.simulationType = t;
}
void DPM2D::setGeometryType(simControlStruct::geometryEnum t) {
  simControl.geometryType = t;  // normal, fixedBox, fixedSides, roundBox = DPM_GEOM_*
  push_params();
}
void DPM2D::setBoxEnergyCost(double ew_) {
  ew = ew_;
  push_params();
}
void DPM2D::setBoxRadius(double boxRadius_) {
  boxRadius = boxRadius_;
  push_params();
}
void DPM2D::setParticleType(simControlStruct::particleEnum t) {
  simControl.particleType = t;
  push_params();
}
void DPM2D::setNoiseType(simControlStruct::noiseEnum t) {
  simControl.noiseType = t;  // langevin, baoab, brownian, drivenBrownian = DPM_NOISE_*
  push_params();
}
// setActiveTensionParams, DPM2D.cu:1898-1903: the rest lengths become the current l0
void DPM2D::setActiveTensionParams(double l0Tau_, double l0Diff_) {
  l0Tau = l0Tau_;
  l0Diff = l0Diff_;
  check(dpm_set_active_tension(handle(), l0Tau, l0Diff), "setActiveTensionParams");
}
void DPM2D::getActiveTensionParams(double& l0Tau_, double& l0Diff_) {
  l0Tau_ = l0Tau;
  l0Diff_ = l0Diff;
}
void DPM2D::setPotentialType(simControlStruct::potentialEnum t) {
  simControl.potentialType = t;
  push_params();
}
void DPM2D::setInteractionType(simControlStruct::interactionEnum t) {
  simControl.interactionType = t;
  push_params();
}
void DPM2D::setNeighborType(simControlStruct::neighborEnum t) {
  // The device list is always a Verlet list built through cell bins; it holds the sets of
  // the reference's neighbor mode, which is also what its cell mode computes.
  if (t == simControlStruct::neighborEnum::allToAll)
    throw std::invalid_argument("DPM2D: use interactionType all for the all-to-all sum");
  simControl.neighborType = t;
}
void DPM2D::setMonomerType(simControlStruct::monomerEnum t) {
  simControl.monomerType = t;
  push_params();
}

// ---------------------------------------------------------------- topology
void DPM2D::setNumVertexInParticleList(std::vector<long>& list) {
  if ((long)list.size() != numParticles)
    throw std::invalid_argument("DPM2D::setNumVertexInParticleList: size must equal numParticles");
  for (long n : list)
    if (n < 3) throw std::invalid_argument("DPM2D: ring sizes must be >= 3");
  long nv = std::accumulate(list.begin(), list.end(), 0L);
  bool same = (list == numVertexInParticleList);
  if (!same) release();
  numVertexInParticleList = list;
  numVertices = nv;
  initParticleIdList();
  if ((long)hRad.size() != nv) {
    hRad.assign(nv, 0.0);
    hL0.assign(nv, 0.0);
    hTheta0.assign(nv, 0.0);
    hPos.assign(2 * nv, 0.0);
    hVel.assign(2 * nv, 0.0);
  }
  if ((long)hA0.size() != numParticles) hA0.assign(numParticles, 0.0);
}

void DPM2D::initParticleIdList() {
  particleIdList.assign(numVertices, 0);
  firstVertexInParticleId.assign(numParticles, 0);
  long first = 0;
  for (long p = 0; p < numParticles; p++) {
    firstVertexInParticleId[p] = first;
    for (long i = 0; i < numVertexInParticleList[p]; i++) particleIdList[first + i] = p;
    first += numVertexInParticleList[p];
  }
}

// ---------------------------------------------------------------- state
static void need(const std::vector<double>& v, long n, const char* what) {
  if ((long)v.size() != n) throw std::invalid_argument(std::string("DPM2D::") + what + ": wrong size");
}

void DPM2D::setBoxSize(std::vector<double>& b) {
  need(b, 2, "setBoxSize");
  boxSize = b;
  if (h_) set_field(DPM_FIELD_BOX, b);
}
#define DPM_HOST_FIELD(SET, GET, FIELD, HOSTVEC, N)                  \
  void DPM2D::SET(std::vector<double>& v) {                          \
    need(v, N, #SET);                                                \
    if (h_) set_field(FIELD, v); else HOSTVEC = v;                   \
  }                                                                  \
  std::vector<double> DPM2D::GET() { return h_ ? field(FIELD, N) : HOSTVEC; }
DPM_HOST_FIELD(setVertexRadii, getVertexRadii, DPM_FIELD_RAD, hRad, numVertices)
DPM_HOST_FIELD(setRestAreas, getRestAreas, DPM_FIELD_A0, hA0, numParticles)
DPM_HOST_FIELD(setRestLengths, getRestLengths, DPM_FIELD_L0, hL0, numVertices)
DPM_HOST_FIELD(setRestAngles, getRestAngles, DPM_FIELD_THETA0, hTheta0, numVertices)
DPM_HOST_FIELD(setVertexPositions, getVertexPositions, DPM_FIELD_POS, hPos, 2 * numVertices)
DPM_HOST_FIELD(setVertexVelocities, getVertexVelocities, DPM_FIELD_VEL, hVel, 2 * numVertices)
#undef DPM_HOST_FIELD

void DPM2D::setVertexForces(std::vector<double>& f) {
  need(f, 2 * numVertices, "setVertexForces");
  set_field(DPM_FIELD_FORCE, f);
}
std::vector<double> DPM2D::getVertexForces() { return field(DPM_FIELD_FORCE, 2 * numVertices); }
std::vector<double> DPM2D::getSegmentLengths() { return field(DPM_FIELD_LENGTH, numVertices); }
std::vector<double> DPM2D::getSegmentAngles() { return field(DPM_FIELD_THETA, numVertices); }
std::vector<double> DPM2D::getAreas() { return field(DPM_FIELD_AREA, numParticles); }
std::vector<double> DPM2D::getPerimeters() { return field(DPM_FIELD_PERIMETER, numParticles); }
std::vector<double> DPM2D::getParticlePositions() { return field(DPM_FIELD_PPOS, 2 * numParticles); }
std::vector<double> DPM2D::getParticleEnergies() { return field(DPM_FIELD_PENERGY, numParticles); }
std::vector<double> DPM2D::getVertexEnergies() { return field(DPM_FIELD_ENERGY, numVertices); }
void DPM2D::setParticleAngles(std::vector<double>& a) {
  need(a, numParticles, "setParticleAngles");
  set_field(DPM_FIELD_PANGLE, a);
}
std::vector<double> DPM2D::getParticleAngles() { return field(DPM_FIELD_PANGLE, numParticles); }
std::vector<double> DPM2D::getParticleShapes() {
  // (area, perimeter, perimeter^2 / (4 pi area)) per ring, DPM2D.cu:739-756
  std::vector<double> s3(3 * numParticles);
  check(dpm_get_particle_shapes(handle(), s3.data()), "getParticleShapes");
  return s3;
}
std::vector<double> DPM2D::getInteractionForces() {
  std::vector<double> f(2 * numVertices);
  check(dpm_get_interaction_forces(handle(), f.data(), (long)f.size()), "getInteractionForces");
  return f;
}

void DPM2D::setLengthScale() {
  std::vector<double> a0 = getRestAreas();
  rho0 = std::sqrt(std::accumulate(a0.begin(), a0.end(), 0.0) / numParticles);  // DPM2D.cu:591-595
}

// ---------------------------------------------------------------- parameters
void DPM2D::setEnergyCosts(double ea_, double el_, double eb_, double ec_) {
  ea = ea_;
  el = el_;
  eb = eb_;
  ec = ec_;
  if (h_) check(dpm_set_energy_costs(h_, ea, el, eb, ec), "setEnergyCosts");
}
void DPM2D::setSelfPropulsionParams(double driving_, double taup_) {
  driving = driving_;
  taup = taup_;
  if (h_) check(dpm_set_self_propulsion(h_, driving, taup), "setSelfPropulsionParams");
}
void DPM2D::setAttractionConstants(double l1_, double l2_) {
  l1 = l1_;
  l2 = l2_;
  if (h_) check(dpm_set_attraction_constants(h_, l1, l2), "setAttractionConstants");
}
void DPM2D::setLJcutoff(double c) {
  LJcutoff = c;
  // shifts as DPM2D.cu:1923-1932 (printed only; the library keeps its own copy)
  double r6 = std::pow(1.0 / c, 6);
  LJecut = 4 * ec * (r6 * r6 - r6);
  LJfshift = 24 * ec * (2 * r6 * r6 - r6) / c;
  if (h_) check(dpm_set_lj_cutoff(h_, c), "setLJcutoff");
}
void DPM2D::setFENEconstants(double stiff_, double ext_) {
  stiff = stiff_;
  extSq = ext_ * ext_;
  if (h_) check(dpm_set_fene_constants(h_, stiff, ext_), "setFENEconstants");
}
double DPM2D::setTimeScale(double dt_) {
  const double l00 = getRestLengths()[0];
  const double ta = rho0 / std::sqrt(ea);
  const double tl = (rho0 * l00) / std::sqrt(ea * el);
  const double tb = (rho0 * l00) / std::sqrt(ea * eb);
  double tmin = 1e08;
  if (ta < tmin) tmin = ta;
  if (tl < tmin) tmin = tl;
  if (tb < tmin) tmin = tb;
  return setTimeStep(tmin * dt_);
}

double DPM2D::setTimeStep(double dt_) {
  dt = dt_;
  if (h_) check(dpm_set_time_step(h_, dt), "setTimeStep");
  return dt;
}
void DPM2D::setSeed(uint64_t s) {
  seed = s;
  srand48((long)s);
  if (h_) check(dpm_set_seed(h_, s), "setSeed");
}

// ---------------------------------------------------------------- observables
static double one(int (*fn)(dpm_set_t, double*), DPM2D* d, const char* what) {
  double v = 0;
  d->check(fn(d->handle(), &v), what);
  return v;
}
double DPM2D::getTotalForceMagnitude() { return one(dpm_get_total_force_magnitude, this, "getTotalForceMagnitude"); }
double DPM2D::getMaxUnbalancedForce() { return one(dpm_get_max_unbalanced_force, this, "getMaxUnbalancedForce"); }
double DPM2D::getPotentialEnergy() { return one(dpm_get_potential_energy, this, "getPotentialEnergy"); }
double DPM2D::getKineticEnergy() { return one(dpm_get_kinetic_energy, this, "getKineticEnergy"); }
double DPM2D::getTemperature() { return one(dpm_get_temperature, this, "getTemperature"); }
double DPM2D::getPhi() { return one(dpm_get_phi, this, "getPhi"); }

// ---------------------------------------------------------------- observables
void DPM2D::setInitialPositions() { check(dpm_set_initial_positions(handle()), "setInitialPositions"); }
double DPM2D::getVertexMSD() { return one(dpm_get_vertex_msd, this, "getVertexMSD"); }
double DPM2D::getParticleMSD() { return one(dpm_get_particle_msd, this, "getParticleMSD"); }
double DPM2D::getVertexISF() { return one(dpm_get_vertex_isf, this, "getVertexISF"); }
double DPM2D::getParticleISF(double waveNumber) {
  double v = 0;
  check(dpm_get_particle_isf(handle(), waveNumber, &v), "getParticleISF");
  return v;
}
double DPM2D::getParticleHexaticOrder() { return one(dpm_get_hexatic_order, this, "getParticleHexaticOrder"); }
double DPM2D::getPressure() { return one(dpm_get_pressure, this, "getPressure"); }
void DPM2D::calcContacts() { check(dpm_calc_contacts(handle()), "calcContacts"); }
void DPM2D::calcGeometricContacts() {
  check(dpm_calc_geometric_contacts(handle()), "calcGeometricContacts");
  contactListSize = 8;
}
std::vector<long> DPM2D::getContacts() {
  calcContacts();
  long w = 0;
  check(dpm_get_contacts(handle(), nullptr, nullptr, 0, &w), "getContacts");
  contactListSize = w > 0 ? w : 1;
  std::vector<long> rows((size_t)numParticles * contactListSize);
  check(dpm_get_contacts(handle(), rows.data(), nullptr, contactListSize, &w), "getContacts");
  return rows;
}
std::vector<double> DPM2D::getPerParticleStressTensor() {
  std::vector<double> out(4 * numParticles);
  check(dpm_get_per_particle_stress(handle(), out.data()), "getPerParticleStressTensor");
  return out;
}
std::vector<double> DPM2D::getStressTensor() {
  std::vector<double> out(4);
  check(dpm_get_stress_tensor(handle(), out.data()), "getStressTensor");
  return out;
}

double DPM2D::getPreferredPhi() {
  if (h_) return one(dpm_get_preferred_phi, this, "getPreferredPhi");
  // host form for set-up before a device set exists (DPM2D.cu:1222-1238, cudaKernel.cuh:1636-1642)
  double a = std::accumulate(hA0.begin(), hA0.end(), 0.0), v = 0;
  for (long p = 0; p < numParticles; p++) {
    double r = hRad[firstVertexInParticleId[p]];
    v += r * r * (0.5 * numVertexInParticleList[p] - 1);
  }
  if (simControl.geometryType == simControlStruct::geometryEnum::roundBox) return (a + PI * v) / (PI * boxRadius * boxRadius);
  return (a + PI * v) / (boxSize[0] * boxSize[1]);
}
double DPM2D::getMeanParticleSize() {
  std::vector<double> a0 = getRestAreas();
  return 2 * std::sqrt(std::accumulate(a0.begin(), a0.end(), 0.0) / (PI * numParticles));
}
double DPM2D::getMeanVertexRadius() {
  std::vector<double> r = getVertexRadii();
  return std::accumulate(r.begin(), r.end(), 0.0) / numVertices;
}
double DPM2D::getMinVertexRadius() {
  std::vector<double> r = getVertexRadii();
  return std::min(1.0, *std::min_element(r.begin(), r.end()));  // reduce with init 1 (DPM2D.cu:774-776)
}
// per-vertex kinetic energy split into ring centre-of-mass and shape parts, using the ring's
// mean vertex velocity (cudaKernel.cuh:1615-1627 sums n-1 vertices into an uninitialised sum;
// this is the intended mean)
double DPM2D::getParticleKineticEnergy() {
  std::vector<double> v = getVertexVelocities();
  double k = 0;
  for (long p = 0; p < numParticles; p++) {
    double sx = 0, sy = 0;
    long f = firstVertexInParticleId[p], n = numVertexInParticleList[p];
    for (long i = 0; i < n; i++) {
      sx += v[2 * (f + i)];
      sy += v[2 * (f + i) + 1];
    }
    sx /= n;
    sy /= n;
    k += 0.5 * (sx * sx + sy * sy);
  }
  return k;
}
double DPM2D::getShapeKineticEnergy() { return getKineticEnergy() - getParticleKineticEnergy(); }
void DPM2D::adjustTemperature(double T) { check(dpm_adjust_temperature(handle(), T), "adjustTemperature"); }
void DPM2D::adjustKineticEnergy(double prevEtot) {
  // DPM2D.cu:1161-1183, on the device
  double scale = 0;
  check(dpm_adjust_kinetic_energy(handle(), &prevEtot, &scale), "adjustKineticEnergy");
  if (scale > 0) {
    cout << "DPM2D::adjustKineticEnergy:: scale: " << scale << endl;
  } else {
    cout << "DPM2D::adjustKineticEnergy:: kinetic energy is less then change in total energy - no adjustment is made" << endl;
  }
}

// ---------------------------------------------------------------- neighbours
double DPM2D::setDisplacementCutoff(double cutoff_, double size_) {
  cutoffArg = cutoff_;
  sizeArg = size_;
  cutoff = cutoff_ * size_;
  check(dpm_set_displacement_cutoff(handle(), cutoff_, size_, &cutDistance), "setDisplacementCutoff");
  cout << "DPM2D::setDisplacementCutoff - cutDistance: " << cutDistance << " cutoff: " << cutoff << endl;
  return cutDistance;
}
void DPM2D::calcNeighbors(double c) { calcNeighborList(c); }
void DPM2D::calcNeighborList(double c) {
  cutDistance = c;
  check(dpm_calc_neighbor_list(handle(), c), "calcNeighborList");
}
void DPM2D::checkNeighbors() { check(dpm_check_neighbors(handle()), "checkNeighbors"); }
void DPM2D::resetLastPositions() { check(dpm_reset_last_positions(handle()), "resetLastPositions"); }
void DPM2D::resetUpdateCount() { check(dpm_reset_update_count(handle()), "resetUpdateCount"); }
long DPM2D::getUpdateCount() {
  long n = 0;
  check(dpm_get_update_count(handle(), &n), "getUpdateCount");
  return n;
}
std::vector<long> DPM2D::getNeighbors() {
  long width = 0;
  check(dpm_get_neighbor_list(handle(), nullptr, nullptr, 0, &width), "getNeighbors");
  std::vector<long> list(numVertices * std::max(width, 1L)), counts(numVertices);
  check(dpm_get_neighbor_list(handle(), list.data(), counts.data(), std::max(width, 1L), &width), "getNeighbors");
  neighborListSize = width;
  return list;
}
std::vector<long> DPM2D::getMaxNeighborList() {
  long width = 0;
  check(dpm_get_neighbor_list(handle(), nullptr, nullptr, 0, &width), "getMaxNeighborList");
  std::vector<long> list(numVertices * std::max(width, 1L)), counts(numVertices);
  check(dpm_get_neighbor_list(handle(), list.data(), counts.data(), std::max(width, 1L), &width),
        "getMaxNeighborList");
  return counts;
}

// ---------------------------------------------------------------- forces
void DPM2D::calcParticleShape() { check(dpm_calc_particle_shape(handle()), "calcParticleShape"); }
void DPM2D::calcForceEnergy() { check(dpm_calc_force_energy(handle()), "calcForceEnergy"); }

// ---------------------------------------------------------------- initialisation
void DPM2D::setMonoSizeDistribution() {
  // DPM2D.cu:1522-1541
  std::vector<long> n(numParticles, numVertexPerParticle);
  setNumVertexInParticleList(n);
  std::vector<double> a0(numParticles), l0(numVertices), th(numVertices), rad(numVertices);
  for (long p = 0; p < numParticles; p++) {
    double nv = (double)numVertexInParticleList[p];
    a0[p] = 1;
    double c = nv * std::tan(PI / nv) / PI;
    for (long i = 0; i < numVertexInParticleList[p]; i++) {
      long v = firstVertexInParticleId[p] + i;
      l0[v] = 2. * std::sqrt(PI * c * a0[p]) / nv;
      th[v] = 2. * PI / nv;
      rad[v] = 0.5 * l0[v];
    }
  }
  setRestAreas(a0);
  setRestLengths(l0);
  setRestAngles(th);
  setVertexRadii(rad);
}

void DPM2D::sizeRings(std::vector<long>& n, double minN) {
  setNumVertexInParticleList(n);
  std::vector<double> a0(numParticles), l0(numVertices), th(numVertices), rad(numVertices);
  for (long p = 0; p < numParticles; p++) {
    double nv = (double)numVertexInParticleList[p];
    a0[p] = (nv / minN) * (nv / minN);
    double c = calA0 * nv * std::tan(PI / nv) / PI;
    for (long i = 0; i < numVertexInParticleList[p]; i++) {
      long v = firstVertexInParticleId[p] + i;
      l0[v] = 2. * std::sqrt(PI * c * a0[p]) / nv;
      th[v] = 2. * PI / nv;
      rad[v] = 0.5 * l0[v];
    }
  }
  setRestAreas(a0);
  setRestLengths(l0);
  setRestAngles(th);
  setVertexRadii(rad);
}

// DPM2D.cu:1590-1598
void DPM2D::setSinusoidalRestAngles(double thetaA, double thetaK) {
  std::vector<double> th(numVertices);
  for (long p = 0; p < numParticles; p++) {
    const double thetaR = 2. * PI / (double)numVertexInParticleList[p];
    for (long i = 0; i < numVertexInParticleList[p]; i++)
      th[firstVertexInParticleId[p] + i] = thetaA * thetaR * std::cos(thetaR * thetaK * (double)i);
  }
  setRestAngles(th);
}

void DPM2D::setPolySizeDistribution(double calA0_, double polyDispersity) {
  // DPM2D.cu:1545-1588: one Box-Muller draw per ring from drand48, floor, clamp below
  calA0 = calA0_;
  double minN = (double)numVertexPerParticle;
  std::vector<long> n(numParticles);
  for (long p = 0; p < numParticles; p++) {
    double r1 = drand48();
    double r2 = drand48();
    double g = std::sqrt(-2. * std::log(r1)) * std::cos(2. * PI * r2);
    double k = std::floor(polyDispersity * numVertexPerParticle * g + numVertexPerParticle);
    if (k < minN) k = minN;
    n[p] = (long)k;
  }
  sizeRings(n, minN);
  cout << "DPM2D::setPolySizeDistribution: numVertices: " << numVertices << endl;
}

void DPM2D::setBiSizeDistribution(double calA0_, long nSmall, long nLarge, double fracLarge) {
  calA0 = calA0_;
  long nl = std::lround(fracLarge * numParticles);
  std::vector<long> n(numParticles, nSmall);
  for (long p = 0; p < nl; p++) n[p] = nLarge;
  for (long p = numParticles - 1; p > 0; p--) {  // Fisher-Yates with drand48
    long q = (long)(drand48() * (p + 1));
    std::swap(n[p], n[q]);
  }
  sizeRings(n, (double)std::min(nSmall, nLarge));
  cout << "DPM2D::setBiSizeDistribution: numVertices: " << numVertices << endl;
}

void DPM2D::setRandomParticles(double phi0, double extraRad) {
  // DPM2D.cu:1601-1630 (unit box, then rest sizes scaled to reach phi0)
  std::vector<double> box = {1.0, 1.0};
  setBoxSize(box);
  double scale = std::sqrt(getPreferredPhi() / phi0);
  std::vector<double> a0 = getRestAreas(), l0 = getRestLengths(), rad = getVertexRadii();
  double areaSum = 0;
  particlePos.assign(2 * numParticles, 0.0);
  particleRad.assign(numParticles, 0.0);
  for (long p = 0; p < numParticles; p++) {
    a0[p] /= (scale * scale);
    for (long d = 0; d < 2; d++) particlePos[2 * p + d] = boxSize[d] * drand48();
    double nv = (double)numVertexInParticleList[p];
    particleRad[p] = extraRad * std::sqrt((2. * a0[p]) / (nv * std::sin(2. * PI / nv)));
    areaSum += PI * particleRad[p] * particleRad[p];
  }
  for (long v = 0; v < numVertices; v++) {
    l0[v] /= scale;
    rad[v] /= scale;
  }
  setRestAreas(a0);
  setRestLengths(l0);
  setVertexRadii(rad);
  particleVel.assign(2 * numParticles, 0.0);
  particleForce.assign(2 * numParticles, 0.0);
  particleEnergy.assign(numParticles, 0.0);
  particleLastPos = particlePos;
  setLengthScaleToOne();
  cout << "DPM2D::setRandomParticles: particle packing fraction: " << getPreferredPhi() << " "
       << areaSum / (boxSize[0] * boxSize[1]) << endl;
}

// DPM2D.cu:1631-1653: the box (lx, ly) grown so that the rest shapes fill phi0, disks drawn in
// 0.95 of it (x then y per disk)
void DPM2D::setScaledRandomParticles(double phi0, double extraRad, double lx, double ly) {
  std::vector<double> box = {lx, ly};
  setBoxSize(box);
  double scale = std::sqrt(getPreferredPhi() / phi0);
  box = {lx * scale, ly * scale};
  setBoxSize(box);
  std::vector<double> a0 = getRestAreas();
  double areaSum = 0;
  particlePos.assign(2 * numParticles, 0.0);
  particleRad.assign(numParticles, 0.0);
  for (long p = 0; p < numParticles; p++) {
    for (long d = 0; d < 2; d++) particlePos[2 * p + d] = 0.95 * boxSize[d] * drand48();
    double nv = (double)numVertexInParticleList[p];
    particleRad[p] = extraRad * std::sqrt((2. * a0[p]) / (nv * std::sin(2. * PI / nv)));
    areaSum += PI * particleRad[p] * particleRad[p];
  }
  particleVel.assign(2 * numParticles, 0.0);
  particleForce.assign(2 * numParticles, 0.0);
  particleEnergy.assign(numParticles, 0.0);
  particleLastPos = particlePos;
  setLengthScaleToOne();
  cout << "DPM2D::setScaledRandomParticles: packing fraction: " << getPreferredPhi() << " "
       << areaSum / (boxSize[0] * boxSize[1]) << endl;
}

// DPM2D.cu:1655-1676: the round box (centred on the origin) grown to phi0, disks drawn at radius
// 0.95 R sqrt(U) and angle 2 pi U - pi
void DPM2D::setRoundScaledRandomParticles(double phi0, double extraRad, double boxRadius_) {
  setBoxRadius(boxRadius_);
  double scale = std::sqrt(getPreferredPhi() / phi0);
  setBoxRadius(boxRadius_ * scale);
  std::vector<double> a0 = getRestAreas();
  double areaSum = 0;
  particlePos.assign(2 * numParticles, 0.0);
  particleRad.assign(numParticles, 0.0);
  for (long p = 0; p < numParticles; p++) {
    double r = 0.95 * boxRadius * std::sqrt(drand48());
    double th = 2 * PI * drand48() - PI;
    particlePos[2 * p] = r * std::cos(th);
    particlePos[2 * p + 1] = r * std::sin(th);
    double nv = (double)numVertexInParticleList[p];
    particleRad[p] = extraRad * std::sqrt((2. * a0[p]) / (nv * std::sin(2. * PI / nv)));
    areaSum += PI * particleRad[p] * particleRad[p];
  }
  particleVel.assign(2 * numParticles, 0.0);
  particleForce.assign(2 * numParticles, 0.0);
  particleEnergy.assign(numParticles, 0.0);
  particleLastPos = particlePos;
  setLengthScaleToOne();
  cout << "DPM2D::setRoundScaledRandomParticles: packing fraction: " << getPreferredPhi() << " "
       << areaSum / (PI * boxRadius * boxRadius) << endl;
}

// scalePacking, DPM2D.cu:1728-1763: every length (disks, vertices, rest sizes, box) divided by scale
void DPM2D::scalePacking(double scale) {
  for (auto& x : particlePos) x /= scale;
  for (auto& r : particleRad) r /= scale;
  std::vector<double> pos = getVertexPositions(), rad = getVertexRadii(), l0 = getRestLengths(), a0 = getRestAreas();
  for (auto& x : pos) x /= scale;
  for (auto& r : rad) r /= scale;
  for (auto& l : l0) l /= scale;
  for (auto& a : a0) a /= (scale * scale);
  setVertexPositions(pos);
  setVertexRadii(rad);
  setRestLengths(l0);
  setRestAreas(a0);
  if (simControl.geometryType == simControlStruct::geometryEnum::roundBox) {
    setBoxRadius(boxRadius / scale);
  } else {
    std::vector<double> box = {boxSize[0] / scale, boxSize[1] / scale};
    setBoxSize(box);
  }
}

// scaleParticlePacking, DPM2D.cu:1765-1800: disks, vertex radii, rest sizes and box divided by the
// mean disk radius (vertex positions are not, they are placed on the disks afterwards)
void DPM2D::scaleParticlePacking() {
  double scale = getMeanParticleSigma();
  for (auto& r : particleRad) r /= scale;
  for (auto& x : particlePos) x /= scale;
  std::vector<double> rad = getVertexRadii(), l0 = getRestLengths(), a0 = getRestAreas();
  for (auto& r : rad) r /= scale;
  for (auto& l : l0) l /= scale;
  for (auto& a : a0) a /= (scale * scale);
  setVertexRadii(rad);
  setRestLengths(l0);
  setRestAreas(a0);
  if (simControl.geometryType == simControlStruct::geometryEnum::roundBox) {
    setBoxRadius(boxRadius / scale);
  } else {
    std::vector<double> box = {boxSize[0] / scale, boxSize[1] / scale};
    setBoxSize(box);
  }
}

void DPM2D::initVerticesOnParticles() {
  // DPM2D.cu:1678-1691: the angle uses the global vertex id, so ring p starts rotated by
  // 2 pi firstVertex / n; jitter 1e-3 l0 U(0,1), x draw before y draw
  std::vector<double> a0 = getRestAreas(), l0 = getRestLengths();
  std::vector<double> pos(2 * numVertices), vel(2 * numVertices, 0.0);
  for (long v = 0; v < numVertices; v++) {
    long p = particleIdList[v];
    double nv = (double)numVertexInParticleList[p];
    double R = std::sqrt((2. * a0[p]) / (nv * std::sin(2. * PI / nv)));
    pos[2 * v] = R * std::cos((2. * PI * v) / nv) + particlePos[2 * p] + 1e-03 * l0[v] * drand48();
    pos[2 * v + 1] = R * std::sin((2. * PI * v) / nv) + particlePos[2 * p + 1] + 1e-03 * l0[v] * drand48();
  }
  setVertexPositions(pos);
  setVertexVelocities(vel);
}

void DPM2D::scaleVertices(double scale) {
  check(dpm_scale_vertices(handle(), &scale), "scaleVertices");
}

// getDeformableWaveNumber, DPM2D.cu:1453-1462
double DPM2D::getDeformableWaveNumber() {
  if (simControl.geometryType == simControlStruct::geometryEnum::roundBox)
    return PI / (2. * std::sqrt(boxRadius * boxRadius * getPreferredPhi() / numParticles));
  return PI / (2. * std::sqrt(boxSize[0] * boxSize[1] * getPreferredPhi() / (PI * numParticles)));
}

// scaleBox, DPM2D.cu:1693-1719 (the library moves each ring with its centre)
void DPM2D::scaleBox(double scale) {
  check(dpm_scale_box(handle(), &scale), "scaleBox");
  if (simControl.geometryType == simControlStruct::geometryEnum::roundBox) {
    boxRadius /= scale;
  } else {
    for (auto& L : boxSize) L /= scale;
  }
}

// ---------------------------------------------------------------- soft disks (set-up only)
// Host restatement of the soft-disk path the compression driver runs before vertices exist:
// kernelCalcParticleNeighborList (cudaKernel.cuh:1765-1787, list width grown as a power of two,
// DPM2D.cu:3521-3541) and kernelCalcParticleInteraction (cudaKernel.cuh:1185-1217).  The
// neighbour test uses twice the particle's own radius as the pair's radius sum, as the
// reference's extractOtherParticle does (cudaKernel.cuh:289-296).
void DPM2D::calcParticleNeighborList(double cutDistance_) {
  std::vector<long> count(numParticles, 0);
  std::vector<std::vector<long>> rows(numParticles);
  for (long i = 0; i < numParticles; i++) {
    for (long j = 0; j < numParticles; j++) {
      if (j == i) continue;
      double dx = boxd(particlePos[2 * i] - particlePos[2 * j], 0);
      double dy = boxd(particlePos[2 * i + 1] - particlePos[2 * j + 1], 1);
      double radSum = 2 * particleRad[i];
      double overlap = 1 - std::sqrt(dx * dx + dy * dy) / radSum;
      if (-overlap < cutDistance_) rows[i].push_back(j);
    }
    count[i] = (long)rows[i].size();
  }
  long maxN = count.empty() ? 0 : *std::max_element(count.begin(), count.end());
  if (maxN > partNeighborListSize) partNeighborListSize = (long)std::pow(2, std::ceil(std::log2((double)maxN)));
  partMaxNeighborList = count;
  partNeighborList.assign(numParticles * std::max(partNeighborListSize, 1L), -1L);
  for (long i = 0; i < numParticles; i++)
    for (long k = 0; k < count[i]; k++) partNeighborList[i * partNeighborListSize + k] = rows[i][k];
}

void DPM2D::calcParticleForceEnergy() {
  bool wca = simControl.potentialType == simControlStruct::potentialEnum::wca;
  if (!wca && simControl.potentialType != simControlStruct::potentialEnum::harmonic)
    throw std::invalid_argument("DPM2D: the soft-disk set-up supports the harmonic and WCA potentials");
  particleForce.assign(2 * numParticles, 0.0);
  particleEnergy.assign(numParticles, 0.0);
  for (long i = 0; i < numParticles; i++) {
    for (long k = 0; k < partMaxNeighborList[i]; k++) {
      long j = partNeighborList[i * partNeighborListSize + k];
      if (j < 0 || j == i) continue;
      double dx = boxd(particlePos[2 * i] - particlePos[2 * j], 0);
      double dy = boxd(particlePos[2 * i + 1] - particlePos[2 * j + 1], 1);
      double d = std::sqrt(dx * dx + dy * dy), radSum = particleRad[i] + particleRad[j];
      if (!wca) {
        double ov = 1 - d / radSum;  // calcContactInteraction, cudaKernel.cuh:531-544
        if (ov > 0) {
          double g = ec * ov / radSum;
          particleForce[2 * i] += g * dx / d;
          particleForce[2 * i + 1] += g * dy / d;
          particleEnergy[i] += (0.5 * ec * ov * ov) * 0.5;
        }
      } else if (d <= WCAcut * radSum) {  // calcWCAInteraction, cudaKernel.cuh:653-668
        double r = radSum / d, r6 = std::pow(r, 6), r12 = r6 * r6;
        double g = 24 * ec * (2 * r12 - r6) / d;
        particleForce[2 * i] += g * dx / d;
        particleForce[2 * i + 1] += g * dy / d;
        particleEnergy[i] += 0.5 * ec * (4 * (r12 - r6) + 1);
      }
    }
  }
  if (simControl.geometryType != simControlStruct::geometryEnum::normal) calcParticleWallInteraction();
}

// boxDistance (cudaKernel.cuh:76-94): the minimum image along periodic directions only
double DPM2D::boxd(double d, int dim) const {
  using G = simControlStruct::geometryEnum;
  const G g = simControl.geometryType;
  if (g == G::normal || (g == G::fixedSides && dim == 0)) return pbc(d, boxSize[dim]);
  return d;
}

// calcParticleFixedWallInteraction, DPM2D.cu:3590-3609 (kernelCalcParticleSquareWallInteraction,
// kernelCalcParticleSidesInteraction2D, kernelCalcParticleRoundWallInteraction with
// calcWallWCAInteraction, cudaKernel.cuh:1221-1330): WCA of energy scale ew against the
// particle's image on the wall, radius sum = the particle's radius
void DPM2D::calcParticleWallInteraction() {
  using G = simControlStruct::geometryEnum;
  auto wall = [&](long p, double wx, double wy) {
    double x = particlePos[2 * p], y = particlePos[2 * p + 1], radSum = particleRad[p];
    double dx = x - wx, dy = y - wy, d = std::sqrt(dx * dx + dy * dy);
    if (d < WCAcut * radSum) {
      double r = radSum / d, r6 = std::pow(r, 6), r12 = r6 * r6;
      double g = 24 * ew * (2 * r12 - r6) / d;
      particleForce[2 * p] += g * dx / d;
      particleForce[2 * p + 1] += g * dy / d;
      particleEnergy[p] += ew * (4 * (r12 - r6) + 1);
    }
  };
  const G g = simControl.geometryType;
  for (long p = 0; p < numParticles; p++) {
    double x = particlePos[2 * p], y = particlePos[2 * p + 1], rc = WCAcut * particleRad[p];
    if (g == G::fixedBox) {
      if (x < rc) wall(p, 0, y);
      else if (boxSize[0] - x < rc) wall(p, boxSize[0], y);
    }
    if (g == G::fixedBox || g == G::fixedSides) {
      if (y < rc) wall(p, x, 0);
      else if (boxSize[1] - y < rc) wall(p, x, boxSize[1]);
    }
    if (g == G::roundBox) {
      double R = std::sqrt(x * x + y * y), th = std::atan2(y, x);
      if (boxRadius - R < rc) wall(p, boxRadius * std::cos(th), boxRadius * std::sin(th));
    }
  }
}

double DPM2D::getParticleMaxUnbalancedForce() {
  double m = -1;
  for (double f : particleForce) m = std::max(m, f * f);
  return std::sqrt(m);
}
double DPM2D::getParticleMaxDisplacement() {
  double m = -1;
  for (long p = 0; p < numParticles; p++) {
    double dx = particlePos[2 * p] - particleLastPos[2 * p], dy = particlePos[2 * p + 1] - particleLastPos[2 * p + 1];
    m = std::max(m, std::sqrt(dx * dx + dy * dy));
  }
  return m;
}
double DPM2D::getParticlePotentialEnergy() { return std::accumulate(particleEnergy.begin(), particleEnergy.end(), 0.0); }
double DPM2D::getParticlePhi() {
  double s = 0;
  for (double r : particleRad) s += r * r;
  if (simControl.geometryType == simControlStruct::geometryEnum::roundBox) return s / (boxRadius * boxRadius);
  return s * PI / (boxSize[0] * boxSize[1]);
}
double DPM2D::getMeanParticleSigma() { return std::accumulate(particleRad.begin(), particleRad.end(), 0.0) / numParticles; }
double DPM2D::getMinParticleSigma() { return std::min(1.0, *std::min_element(particleRad.begin(), particleRad.end())); }

// ---------------------------------------------------------------- minimisers and integrators
void DPM2D::initFIRE(std::vector<double>& p, long minStep_, long numStep_, long numDOF_) {
  // DPM2D.cu:3660-3675
  delete fire_;
  fire_ = new FIRE(this);
  if (p.size() == 7) {
    fire_->initMinimizer(p[0], p[1], p[2], p[3], p[4], p[5], p[6], minStep_, numStep_, numDOF_);
    resetParticleLastPositions();
  } else {
    cout << "DPM2D::initFIRE: wrong number of FIRE parameters, must be 7" << endl;
  }
}
void DPM2D::setParticleMassFIRE() { fire_->setParticleMass(); }
void DPM2D::setTimeStepFIRE(double timeStep) { fire_->setFIRETimeStep(timeStep); }
void DPM2D::particleFIRELoop() { fire_->minimizerParticleLoop(); }
void DPM2D::vertexFIRELoop() { fire_->minimizerVertexLoop(); }
void DPM2D::vertexFIRELoop(long iterations) { fire_->minimizerVertexLoop(iterations); }
long DPM2D::vertexFIREMinimize(double tol, long maxIter) { return fire_->minimize(tol, maxIter); }

void DPM2D::initLangevin(double Temp, double gamma, bool readState) {
  // DPM2D.cu:3713-3758 (langevin noise type)
  delete sim_;
  sim_ = new Langevin(this, SimConfig(Temp));
  sim_->gamma = gamma;
  sim_->noise = std::sqrt(2. * Temp * gamma / dt);
  check(dpm_init_langevin(handle(), Temp, gamma, readState ? 1 : 0, 0), "initLangevin");
  cout << "DPM2D::initLangevin:: Langevin integrator, current temperature: " << std::setprecision(12)
       << getTemperature() << endl;
}
void DPM2D::langevinLoop() { sim_->integrate(); }
void DPM2D::langevinLoop(long steps) { check(dpm_langevin_loop(handle(), steps), "langevinLoop"); }
void DPM2D::initNVE(double Temp, bool readState) {
  delete sim_;
  sim_ = new NVE(this, SimConfig(Temp));
  check(dpm_init_nve(handle(), Temp, readState ? 1 : 0, 0), "initNVE");
  cout << "DPM2D::initNVE:: current temperature: " << std::setprecision(12) << getTemperature() << endl;
}
void DPM2D::NVELoop() { sim_->integrate(); }
void DPM2D::NVELoop(long steps) { check(dpm_nve_loop(handle(), steps), "NVELoop"); }
void DPM2D::initNVERescale(double Temp) {
  delete sim_;
  sim_ = new NVERescale(this, SimConfig(Temp));
  check(dpm_init_nve_rescale(handle(), Temp), "initNVERescale");
  cout << "DPM2D::initNVERescale:: current temperature: " << std::setprecision(12) << getTemperature() << endl;
}
void DPM2D::NVERescaleLoop() { sim_->integrate(); }

// ---------------------------------------------------------------- rigid rings
// setShapeType, DPM2D.cu:391-402
void DPM2D::setShapeType(simControlStruct::shapeEnum t) { simControl.shapeType = t; }
void DPM2D::setParticleRadii(std::vector<double>& r) {
  need(r, numParticles, "setParticleRadii");
  particleRad = r;
  set_field(DPM_FIELD_PRAD, r);
}
std::vector<double> DPM2D::getParticleRadii() { return field(DPM_FIELD_PRAD, numParticles); }
void DPM2D::setParticlePositions(std::vector<double>& p) {
  need(p, 2 * numParticles, "setParticlePositions");
  set_field(DPM_FIELD_PPOS, p);
}
void DPM2D::setParticleVelocities(std::vector<double>& v) {
  need(v, 2 * numParticles, "setParticleVelocities");
  set_field(DPM_FIELD_PVEL, v);
}
std::vector<double> DPM2D::getParticleVelocities() { return field(DPM_FIELD_PVEL, 2 * numParticles); }
std::vector<double> DPM2D::getParticleForces() { return field(DPM_FIELD_PFORCE, 2 * numParticles); }
std::vector<double> DPM2D::getParticleAngularVelocities() { return field(DPM_FIELD_PANGVEL, numParticles); }
void DPM2D::setParticleAngularVelocities(std::vector<double>& w) {
  need(w, numParticles, "setParticleAngularVelocities");
  set_field(DPM_FIELD_PANGVEL, w);
}
std::vector<double> DPM2D::getParticleTorques() { return field(DPM_FIELD_PTORQUE, numParticles); }
std::vector<double> DPM2D::getVertexTorques() { return field(DPM_FIELD_VTORQUE, numVertices); }

void DPM2D::calcRigidForceEnergy() { check(dpm_calc_rigid_force_energy(handle()), "calcRigidForceEnergy"); }
double DPM2D::getRigidMaxUnbalancedForce() {
  return one(dpm_get_rigid_max_unbalanced_force, this, "getRigidMaxUnbalancedForce");
}
double DPM2D::getRigidKineticEnergy() { return one(dpm_get_rigid_kinetic_energy, this, "getRigidKineticEnergy"); }
// getParticleTemperature, DPM2D.cu:1112-1114
double DPM2D::getParticleTemperature() {
  std::vector<double> v = getParticleVelocities();
  double k = 0;
  for (double x : v) k += x * x;
  return 2 * (0.5 * k) / (numParticles * nDim);
}
// getRigidPhi, DPM2D.cu:1240-1250: rest areas over the box (or the round box)
double DPM2D::getRigidPhi() {
  std::vector<double> a0 = getRestAreas();
  double a = std::accumulate(a0.begin(), a0.end(), 0.0);
  if (simControl.geometryType == simControlStruct::geometryEnum::roundBox) return a / (PI * boxRadius * boxRadius);
  return a / (boxSize[0] * boxSize[1]);
}
// getRigidWaveNumber, DPM2D.cu:1464-1466
double DPM2D::getRigidWaveNumber() {
  return PI / (2. * std::sqrt(boxSize[0] * boxSize[1] * getPreferredPhi() / (PI * numParticles)));
}
// initRigidLangevin / rigidLangevinLoop, DPM2D.cu:3823-3848
void DPM2D::initRigidLangevin(double Temp, double gamma, bool readState) {
  check(dpm_init_rigid_langevin(handle(), Temp, gamma, readState ? 1 : 0), "initRigidLangevin");
  cout << "DPM2D::initRigidLangevin:: current temperature: " << std::setprecision(12) << getParticleTemperature()
       << endl;
}
void DPM2D::rigidLangevinLoop() { rigidLangevinLoop(1); }
void DPM2D::rigidLangevinLoop(long steps) { check(dpm_rigid_langevin_loop(handle(), steps), "rigidLangevinLoop"); }
// initRigidNVE / rigidNVELoop, DPM2D.cu:3850-3872
void DPM2D::initRigidNVE(double Temp, bool readState) {
  check(dpm_init_rigid_nve(handle(), Temp, readState ? 1 : 0), "initRigidNVE");
  cout << "DPM2D::initRigidNVE:: current temperature: " << std::setprecision(12) << getParticleTemperature()
       << endl;
}
void DPM2D::rigidNVELoop() { rigidNVELoop(1); }
void DPM2D::rigidNVELoop(long steps) { check(dpm_rigid_nve_loop(handle(), steps), "rigidNVELoop"); }
// initRigidFIRE / rigidFIRELoop, DPM2D.cu:3700-3710 (numDOF and cutDist are kept by the
// reference's FIRE object only for its own bookkeeping)
void DPM2D::initRigidFIRE(std::vector<double>& p, long minStep_, long numStep_, long numDOF_, double cutDist_) {
  (void)numDOF_;
  (void)cutDist_;
  if (p.size() != 7) {
    cout << "DPM2D::initRigidFIRE: wrong number of FIRE parameters, must be 7" << endl;
    return;
  }
  check(dpm_init_rigid_fire(handle(), p.data(), minStep_, numStep_), "initRigidFIRE");
}
void DPM2D::rigidFIRELoop() { rigidFIRELoop(1); }
void DPM2D::rigidFIRELoop(long iterations) { check(dpm_rigid_fire_loop(handle(), iterations), "rigidFIRELoop"); }

// ---------------------------------------------------------------- two-ring interaction test
// setTwoParticleTest, DPM2D.cu:3229-3267: two equal rings (setMonoSizeDistribution) in an
// lx x ly box at (0.65 lx, y0 ly) and (0.35 lx, y1 ly); ring 1 moves along x at vel1
void DPM2D::setTwoParticleTest(double lx, double ly, double y0, double y1, double vel1) {
  setMonoSizeDistribution();
  std::vector<double> a0 = getRestAreas();
  particleRad.assign(numParticles, 0.0);
  for (long p = 0; p < numParticles; p++) {
    double nv = (double)numVertexInParticleList[p];
    particleRad[p] = std::sqrt((2. * a0[p]) / (nv * std::sin(2. * PI / nv)));
  }
  std::vector<double> box = {lx, ly};
  setBoxSize(box);
  particlePos.assign(2 * numParticles, 0.0);
  particlePos[0] = lx * 0.65;
  particlePos[1] = ly * y0;
  particlePos[2] = lx * 0.35;
  particlePos[3] = ly * y1;
  initVerticesOnParticles();
  setInitialPositions();
  std::vector<double> vel(2 * numVertices, 0.0);
  if (simControl.shapeType == simControlStruct::shapeEnum::rigid) {
    std::vector<double> pv(2 * numParticles, 0.0);
    pv[2] = vel1;
    setParticleVelocities(pv);
  } else {
    for (long v = firstVertexInParticleId[1]; v < firstVertexInParticleId[1] + numVertexInParticleList[1]; v++)
      vel[2 * v] = vel1;
  }
  setVertexVelocities(vel);
  setLengthScaleToOne();
}

// testInteraction / testDeformableInteraction, DPM2D.cu:3269-3315, 3374-3383: velocity Verlet of
// ring 1's vertices only (ring 0 stays put), the list check and the force call between the
// half kicks.  A test harness: the half kicks run on the host through the field copies.
void DPM2D::testInteraction(double timeStep) {
  if (simControl.shapeType == simControlStruct::shapeEnum::rigid)
    throw std::invalid_argument("DPM2D::testInteraction: deformable rings only");
  const long v0 = firstVertexInParticleId[1], v1 = v0 + numVertexInParticleList[1];
  std::vector<double> pos = getVertexPositions(), vel = getVertexVelocities(), f = getVertexForces();
  for (long v = v0; v < v1; v++)
    for (int d = 0; d < 2; d++) {
      vel[2 * v + d] += 0.5 * timeStep * f[2 * v + d];
      pos[2 * v + d] += timeStep * vel[2 * v + d];
    }
  setVertexPositions(pos);
  checkNeighbors();
  calcForceEnergy();
  f = getVertexForces();
  for (long v = v0; v < v1; v++)
    for (int d = 0; d < 2; d++) vel[2 * v + d] += 0.5 * timeStep * f[2 * v + d];
  setVertexVelocities(vel);
}
