// DPM2D.h -- the deformable-ring system class over libdpmring.so.
//
// Keeps the class name, simControlStruct enums and the hot-path method names and
// signatures of include/DPM2D.h in farceri/cudaRing (line numbers cited per group),
// with std::vector<T> for thrust::host_vector<T>.  Device state lives in the library
// (one dpm_set_t of one system); getters copy to the host like the reference's.
#ifndef DPM_HOST_DPM2D_H
#define DPM_HOST_DPM2D_H

#include <string>
#include <vector>

#include "../../include/dpmring.h"
#include "defs.h"

// DPM2D.h:22-32
struct simControlStruct {
  enum class simulationEnum { gpu, cpu, omp } simulationType;
  enum class geometryEnum { normal, fixedBox, fixedSides, roundBox } geometryType;
  enum class particleEnum { passive, active, activeShape } particleType;
  enum class shapeEnum { deformable, rigid } shapeType;
  enum class noiseEnum { langevin, baoab, brownian, drivenBrownian } noiseType;
  enum class potentialEnum { harmonic, lennardJones, adhesive, wca } potentialType;
  enum class interactionEnum { vertexVertex, vertexSmooth, cellSmooth, all } interactionType;
  enum class neighborEnum { neighbor, cell, allToAll } neighborType;
  enum class monomerEnum { harmonic, FENE } monomerType;
};

class FIRE;
class SimInterface;

class DPM2D {
 public:
  DPM2D(long nParticles, long dim, long nVertexPerParticle);  // DPM2D.h:47
  ~DPM2D();
  DPM2D(const DPM2D&) = delete;
  DPM2D& operator=(const DPM2D&) = delete;

  FIRE* fire_ = nullptr;
  SimInterface* sim_ = nullptr;
  simControlStruct simControl;

  long nDim, numParticles, numVertexPerParticle, numVertices;
  double dt, rho0, calA0, ea, el, eb, ec, driving, taup, l1, l2, LJcutoff, LJecut, LJfshift, stiff, extSq;
  double l0Tau = 1, l0Diff = 0;  // activeShape rest-length dynamics
  double ew = 1e3, boxRadius = 0;  // wall energy scale (DPM2D.cu:70), round-box radius
  double cutoff, cutDistance;
  uint64_t seed;

  // ---- library handle and status
  dpm_set_t handle();      // the library set (created on first use of the topology)
  void check(int rc, const char* what);

  // ---- types (DPM2D.h:225-250)
  void setSimulationType(simControlStruct::simulationEnum t);
  void setGeometryType(simControlStruct::geometryEnum t);
  void setParticleType(simControlStruct::particleEnum t);
  void setNoiseType(simControlStruct::noiseEnum t);
  void setPotentialType(simControlStruct::potentialEnum t);
  void setInteractionType(simControlStruct::interactionEnum t);
  void setNeighborType(simControlStruct::neighborEnum t);
  void setMonomerType(simControlStruct::monomerEnum t);
  simControlStruct::potentialEnum getPotentialType() { return simControl.potentialType; }
  simControlStruct::interactionEnum getInteractionType() { return simControl.interactionType; }

  // ---- sizes and topology (DPM2D.h:255-273, 219)
  long getNDim() { return nDim; }
  long getNumParticles() { return numParticles; }
  long getNumVertices() { return numVertices; }
  long getNumVertexPerParticle() { return numVertexPerParticle; }
  void setNumVertices(long n) { numVertices = n; }
  void setNumVertexInParticleList(std::vector<long>& numVertexInParticleList_);
  std::vector<long> getNumVertexInParticleList() { return numVertexInParticleList; }
  void initParticleIdList();
  std::vector<long> getParticleIdList() { return particleIdList; }
  std::vector<long> getFirstVertexInParticleId() { return firstVertexInParticleId; }

  // ---- box and shape variables (DPM2D.h:275-366)
  void setBoxSize(std::vector<double>& boxSize_);
  std::vector<double> getBoxSize() { return boxSize; }
  void setVertexRadii(std::vector<double>& rad_);
  std::vector<double> getVertexRadii();
  void setRestAreas(std::vector<double>& a0_);
  std::vector<double> getRestAreas();
  void setRestLengths(std::vector<double>& l0_);
  std::vector<double> getRestLengths();
  void setRestAngles(std::vector<double>& theta0_);
  std::vector<double> getRestAngles();
  std::vector<double> getSegmentLengths();
  std::vector<double> getSegmentAngles();
  std::vector<double> getAreas();
  std::vector<double> getPerimeters();
  std::vector<double> getParticleShapes();        // 3 per ring: area, perimeter, shape parameter
  std::vector<double> getInteractionForces();     // DPM2D.h:575
  void setVertexPositions(std::vector<double>& pos_);
  std::vector<double> getVertexPositions();
  void setVertexVelocities(std::vector<double>& vel_);
  std::vector<double> getVertexVelocities();
  void setVertexForces(std::vector<double>& force_);
  std::vector<double> getVertexForces();
  std::vector<double> getParticlePositions();
  std::vector<double> getParticleEnergies();
  std::vector<double> getVertexEnergies();
  void setParticleAngles(std::vector<double>& angle_);
  std::vector<double> getParticleAngles();
  void setLengthScale();
  void setLengthScaleToOne() { rho0 = 1; }

  // ---- parameters (DPM2D.h:525-545)
  void setEnergyCosts(double ea_, double el_, double eb_, double ec_);
  void setSelfPropulsionParams(double driving_, double taup_);
  void getSelfPropulsionParams(double& driving_, double& taup_) { driving_ = driving; taup_ = taup; }
  void setActiveTensionParams(double l0Tau_, double l0Diff_);
  void setBoxEnergyCost(double ew_);
  void setBoxRadius(double boxRadius_);
  double getBoxRadius() const { return boxRadius; }
  void getActiveTensionParams(double& l0Tau_, double& l0Diff_);
  void setAttractionConstants(double l1_, double l2_);
  void setLJcutoff(double LJcutoff_);
  void setFENEconstants(double stiff_, double ext_);
  double setTimeStep(double dt_);
  // setTimeScale, DPM2D.cu:1941-1954: dt_ times the smallest of the area, segment and bending
  // time scales (rho0 / sqrt(ea), rho0 l0 / sqrt(ea el), rho0 l0 / sqrt(ea eb), l0 of vertex 0)
  double setTimeScale(double dt_);
  void setSeed(uint64_t seed_);

  // ---- observables (DPM2D.h:408-440, 296-300)
  double getTotalForceMagnitude();
  double getMaxUnbalancedForce();
  double getPotentialEnergy();
  double getKineticEnergy();
  double getEnergy() { return getPotentialEnergy() + getKineticEnergy(); }
  double getTemperature();
  double getShapeKineticEnergy();
  double getParticleKineticEnergy();
  double getShapeCOMEnergyRatio() { return getShapeKineticEnergy() / getParticleKineticEnergy(); }
  double getPhi();
  // observables (DPM2D.h:375-477 of the reference)
  void setInitialPositions();
  double getVertexMSD();
  double getParticleMSD();
  double getVertexISF();
  double getParticleISF(double waveNumber);
  double getDeformableWaveNumber();  // DPM2D.cu:1453-1462
  double getParticleHexaticOrder();
  double getPressure();
  void calcContacts();
  void calcGeometricContacts();  // DPM2D.h:634
  std::vector<long> getContacts();  // numParticles x contactListSize, -1 padded
  std::vector<double> getPerParticleStressTensor();
  std::vector<double> getStressTensor();
  long contactListSize = 0;
  double getPreferredPhi();
  double getMeanParticleSize();
  double getMeanVertexRadius();
  double getMinVertexRadius();
  double getVertexRadius() { return getVertexRadii()[0]; }
  void adjustTemperature(double targetTemp);
  void adjustKineticEnergy(double prevEtot);

  // ---- neighbours (DPM2D.h:447-459, 638-640)
  double setDisplacementCutoff(double cutoff_, double size_);
  void calcNeighbors(double cutDistance_);
  void calcNeighborList(double cutDistance_);
  void checkNeighbors();
  void resetLastPositions();
  void resetUpdateCount();
  long getUpdateCount();
  std::vector<long> getNeighbors();
  std::vector<long> getMaxNeighborList();

  // ---- forces (DPM2D.h:547, 303-304)
  void calcParticleShape();
  void calcParticlePositions() { calcParticleShape(); }
  void calcForceEnergy();

  // ---- initialisation and compression (DPM2D.h:482-495)
  void setMonoSizeDistribution();
  void setPolySizeDistribution(double calA0_, double polyDispersity);
  // rest angle of vertex i of an n-vertex ring: thetaA (2 pi / n) cos((2 pi / n) thetaK i)
  void setSinusoidalRestAngles(double thetaA, double thetaK);
  // 50/50 (or `fracLarge`) mixture of nSmall- and nLarge-vertex rings; the reference leaves
  // setBiSizeDistribution as a stub (DPM2D.h:484).  Sizing as setPolySizeDistribution.
  void setBiSizeDistribution(double calA0_, long nSmall, long nLarge, double fracLarge);
  void setRandomParticles(double phi0, double extraRad);
  void setScaledRandomParticles(double phi0, double extraRad, double lx, double ly);
  void setRoundScaledRandomParticles(double phi0, double extraRad, double boxRadius_);
  void scalePacking(double scale);
  void scaleParticlePacking();
  void initVerticesOnParticles();
  void scaleVertices(double scale);
  void scaleBox(double scale);

  // ---- soft-disk (particle) set-up path, host side (FIRE.cu:43-171, DPM2D.cu:3521-3600)
  std::vector<double> particlePos, particleRad, particleVel, particleForce, particleEnergy, particleLastPos;
  std::vector<long> partNeighborList, partMaxNeighborList;
  long partNeighborListSize = 0;
  void calcParticleNeighborList(double cutDistance_);
  void calcParticleForceEnergy();
  void calcParticleWallInteraction();
  double boxd(double d, int dim) const;  // box distance along dim (periodic directions only)
  double getParticleMaxUnbalancedForce();
  double getParticleMaxDisplacement();
  void resetParticleLastPositions() { particleLastPos = particlePos; }
  double getParticlePotentialEnergy();
  double getParticlePhi();
  double getMeanParticleSigma();
  double getMinParticleSigma();
  void setParticleMassFIRE();
  void particleFIRELoop();

  // ---- integrators and FIRE (DPM2D.h:665-690)
  void initFIRE(std::vector<double>& FIREparams, long minStep, long numStep, long numDOF);
  void setTimeStepFIRE(double timeStep);
  void vertexFIRELoop();
  void vertexFIRELoop(long iterations);
  // vertexFIRELoop until getMaxUnbalancedForce() <= tol or maxIter, without a host round trip
  // per iteration (the compressFIRE.cpp:121-131 loop); returns the iterations taken
  long vertexFIREMinimize(double tol, long maxIter);
  void initLangevin(double Temp, double gamma, bool readState);
  void langevinLoop();
  void langevinLoop(long steps);  // several steps in one call (one host synchronisation)
  void initNVE(double Temp, bool readState);
  void NVELoop();
  void NVELoop(long steps);
  void initNVERescale(double Temp);
  void NVERescaleLoop();

  // ---- rigid rings (DPM2D.h:327-373, 425-469, 567, 660-699)
  void setShapeType(simControlStruct::shapeEnum t);
  void setParticleRadii(std::vector<double>& particleRad_);
  std::vector<double> getParticleRadii();
  void setParticlePositions(std::vector<double>& particlePos_);
  void setParticleVelocities(std::vector<double>& particleVel_);
  std::vector<double> getParticleVelocities();
  std::vector<double> getParticleForces();
  void setParticleAngularVelocities(std::vector<double>& particleAngvel_);
  std::vector<double> getParticleAngularVelocities();
  std::vector<double> getParticleTorques();
  std::vector<double> getVertexTorques();
  void calcRigidForceEnergy();
  double getRigidMaxUnbalancedForce();
  double getRigidKineticEnergy();
  double getParticleTemperature();
  double getRigidPhi();
  double getRigidWaveNumber();
  void initRigidLangevin(double Temp, double gamma, bool readState);
  void rigidLangevinLoop();
  void rigidLangevinLoop(long steps);
  void initRigidNVE(double Temp, bool readState);
  void rigidNVELoop();
  void rigidNVELoop(long steps);
  void initRigidFIRE(std::vector<double>& FIREparams, long minStep, long numStep, long numDOF, double cutDist);
  void rigidFIRELoop();
  void rigidFIRELoop(long iterations);

  // ---- two-ring interaction test (DPM2D.h:613-627)
  void setTwoParticleTest(double lx, double ly, double y0, double y1, double vel1);
  void testInteraction(double timeStep);

  long neighborListSize = 0;

 private:
  dpm_set_t h_ = nullptr;
  std::vector<long> numVertexInParticleList, firstVertexInParticleId, particleIdList;
  std::vector<double> boxSize;
  // host copies that hold the state while no device set exists (set-up phase)
  std::vector<double> hPos, hVel, hRad, hL0, hTheta0, hA0;
  double cutoffArg = 0, sizeArg = 0;  // last setDisplacementCutoff arguments
  void release();                     // copy the device state back and drop the set
  void push_params();
  void sizeRings(std::vector<long>& n, double minN);
  std::vector<double> field(int f, long n);
  void set_field(int f, const std::vector<double>& v);
};

#endif
