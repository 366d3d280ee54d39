// runRigid -- rigid-ring dynamics or minimisation of a saved packing (the flows of runRigidNVE /
// compressRigidFIRE in farceri/cudaRing, whose sources no longer build against their DPM2D).
//
// usage: runRigid inDir outDir mode timeStep Tinject damping maxStep numParticles seed
//                 [potential=harmonic|wca|lj] [interaction=vertex|smooth]
//   mode: nve | nvt | fire.  timeStep and damping are in units of sigma / sqrt(ec) with sigma the
//   mean particle size; the skin is setDisplacementCutoff(0.5, 2 <r_v>).  The rings keep the shapes
//   they have in inDir: a packing saved by savePacking (its particleRad.dat gives the lever arm of
//   the rotational kinetic energy) or a rigid packing saved by saveRigidPacking.  nve / nvt:
//   maxStep steps with the energy log (step, time, U/N, K/N, E/N, phi) every maxStep/100 steps;
//   fire: maxStep FIRE iterations (or until the largest ring force or torque is below 1e-12),
//   logged the same way.  The rigid packing (saveRigidPacking) and boxSize.dat go to outDir.
#include <sys/stat.h>

#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>

#include "../host/DPM2D.h"
#include "../host/FileIO.h"

int main(int argc, char** argv) {
  if (argc < 10) {
    std::cerr << "usage: runRigid inDir outDir nve|nvt|fire timeStep Tinject damping maxStep numParticles seed "
                 "[potential] [interaction]" << std::endl;
    return 2;
  }
  std::string inDir = argv[1], outDir = argv[2], mode = argv[3];
  if (inDir.back() != '/') inDir += "/";
  if (outDir.back() != '/') outDir += "/";
  double timeStep = std::atof(argv[4]), Tinject = std::atof(argv[5]), damping = std::atof(argv[6]);
  long maxStep = std::atol(argv[7]), numParticles = std::atol(argv[8]);
  unsigned long seed = std::strtoul(argv[9], nullptr, 10);
  std::string pot = argc > 10 ? argv[10] : "harmonic", inter = argc > 11 ? argv[11] : "vertex";
  if (mode != "nve" && mode != "nvt" && mode != "fire") {
    std::cerr << "runRigid: mode must be nve, nvt or fire" << std::endl;
    return 2;
  }
  const double ec = 1, LJcut = 1.5, cutoff = 0.5;
  mkdir(outDir.c_str(), 0755);

  DPM2D dpm(numParticles, 2, 32);
  dpm.setSeed(seed);
  dpm.setShapeType(simControlStruct::shapeEnum::rigid);
  if (pot == "wca") dpm.setPotentialType(simControlStruct::potentialEnum::wca);
  else if (pot == "lj") {
    dpm.setPotentialType(simControlStruct::potentialEnum::lennardJones);
    dpm.setLJcutoff(LJcut);
  }
  if (inter == "smooth") dpm.setInteractionType(simControlStruct::interactionEnum::vertexSmooth);
  ioDPMFile io(&dpm);
  struct stat st;
  if (stat((inDir + "particleAngles.dat").c_str(), &st) == 0) {
    io.readRigidPackingFromDirectory(inDir, numParticles, 2);  // a packing saved by saveRigidPacking
  } else {
    io.readPackingFromDirectory(inDir, numParticles, 2);
    std::vector<double> prad = io.read1DFile(inDir + "particleRad.dat", numParticles);
    dpm.setParticleRadii(prad);
  }
  dpm.setEnergyCosts(0, 0, 0, ec);
  double sigma = dpm.getMeanParticleSize(), timeUnit = sigma / std::sqrt(ec);
  timeStep = dpm.setTimeStep(timeStep * timeUnit);
  std::cout << "Units - time: " << timeUnit << " space: " << sigma << " time step: " << timeStep << std::endl;
  double size = 2 * dpm.getMeanVertexRadius();
  double cutDistance = dpm.setDisplacementCutoff(cutoff, size);
  dpm.calcNeighbors(cutDistance);
  dpm.calcRigidForceEnergy();
  if (mode == "nve") dpm.initRigidNVE(Tinject, false);
  else if (mode == "nvt") dpm.initRigidLangevin(Tinject, damping / timeUnit, false);
  else {
    std::vector<double> fp = {0.2, 0.5, 1.1, 0.99, timeStep, 10 * timeStep, 0.2};
    dpm.initRigidFIRE(fp, 20, 0, numParticles, cutDistance);
  }
  // the inits clear the ring torques: evaluate them at the start state
  dpm.calcRigidForceEnergy();
  dpm.resetUpdateCount();
  std::ofstream log((outDir + "energy.dat").c_str());
  long saveFreq = std::max(1L, maxStep / 100), step = 0;
  while (step < maxStep) {
    long n = std::min(saveFreq, maxStep - step);
    if (mode == "nve") dpm.rigidNVELoop(n);
    else if (mode == "nvt") dpm.rigidLangevinLoop(n);
    else dpm.rigidFIRELoop(n);
    step += n;
    double u = dpm.getPotentialEnergy() / numParticles, k = dpm.getRigidKineticEnergy() / numParticles;
    log << step << "\t" << step * timeStep << "\t" << std::setprecision(precision) << u << "\t" << k << "\t" << u + k
        << "\t" << dpm.getRigidPhi() << std::endl;
    if (mode == "fire" && dpm.getRigidMaxUnbalancedForce() < 1e-12) break;
  }
  log.close();
  std::cout << mode << ": steps: " << step << " U/N: " << dpm.getPotentialEnergy() / numParticles
            << " T: " << dpm.getParticleTemperature() << " max force/torque: " << dpm.getRigidMaxUnbalancedForce()
            << " updates: " << dpm.getUpdateCount() << std::endl;
  io.saveRigidPacking(outDir);
  // saveRigidPacking writes no box (FileIO.h:437-458) and readRigidPackingFromDirectory needs one
  io.save1DFile(outDir + "boxSize.dat", dpm.getBoxSize());
  return 0;
}
