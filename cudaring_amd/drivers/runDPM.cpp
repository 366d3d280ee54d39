// runDPM -- run NVE, Langevin NVT or active Langevin dynamics of a saved packing
// (the flow of runNVE.cpp / runNVT.cpp / runActive.cpp in farceri/cudaRing).
//
// usage: runDPM inDir outDir mode timeStep Tinject damping maxStep numParticles seed
//               [potential=harmonic|wca|lj|adhesive] [interaction=vertex|smooth] [driving] [taup]
//   mode: nve | nvt | active.  timeStep and damping are in units of sigma / sqrt(ec) with sigma
//   the mean particle size (runNVT.cpp:119-126); the skin is setDisplacementCutoff(0.5, 2 <r_v>).
// Writes energy.dat (every maxStep/100 steps) and the final packing to outDir.
#include <sys/stat.h>

#include <cmath>
#include <cstdlib>
#include <iostream>
#include <string>

#include "../host/DPM2D.h"
#include "../host/FileIO.h"

int main(int argc, char** argv) {
  if (argc < 10) {
    std::cerr << "usage: runDPM inDir outDir nve|nvt|active timeStep Tinject damping maxStep numParticles seed "
                 "[potential] [interaction] [driving] [taup]" << std::endl;
    return 2;
  }
  std::string inDir = argv[1], outDir = argv[2], mode = argv[3];
  if (inDir.back() != '/') inDir += "/";
  if (outDir.back() != '/') outDir += "/";
  double timeStep = std::atof(argv[4]), Tinject = std::atof(argv[5]), damping = std::atof(argv[6]);
  long maxStep = std::atol(argv[7]), numParticles = std::atol(argv[8]);
  unsigned long seed = std::strtoul(argv[9], nullptr, 10);
  std::string pot = argc > 10 ? argv[10] : "harmonic", inter = argc > 11 ? argv[11] : "vertex";
  double driving = argc > 12 ? std::atof(argv[12]) : 1.0, taup = argc > 13 ? std::atof(argv[13]) : 100.0;
  const double ea = 1e5, el = 20, eb = 10, ec = 1, LJcut = 1.5, cutoff = 0.5;
  mkdir(outDir.c_str(), 0755);

  DPM2D dpm(numParticles, 2, 32);
  dpm.setSeed(seed);
  if (pot == "wca") dpm.setPotentialType(simControlStruct::potentialEnum::wca);
  else if (pot == "lj") {
    dpm.setPotentialType(simControlStruct::potentialEnum::lennardJones);
    dpm.setLJcutoff(LJcut);
  } else if (pot == "adhesive") {
    dpm.setPotentialType(simControlStruct::potentialEnum::adhesive);
    dpm.setAttractionConstants(0.1, 0.2);
  }
  if (inter == "smooth") dpm.setInteractionType(simControlStruct::interactionEnum::vertexSmooth);
  ioDPMFile io(&dpm);
  io.readPackingFromDirectory(inDir, numParticles, 2);
  dpm.setEnergyCosts(ea, el, eb, ec);
  long numVertices = dpm.getNumVertices();
  double sigma = dpm.getMeanParticleSize(), timeUnit = sigma / std::sqrt(ec);
  timeStep = dpm.setTimeStep(timeStep * timeUnit);
  std::cout << "Units - time: " << timeUnit << " space: " << sigma << " time step: " << timeStep << std::endl;
  if (mode == "active") {
    dpm.setParticleType(simControlStruct::particleEnum::active);
    dpm.setSelfPropulsionParams(driving, taup * timeUnit);
  }
  double size = 2 * dpm.getMeanVertexRadius();
  double cutDistance = dpm.setDisplacementCutoff(cutoff, size);
  dpm.calcNeighbors(cutDistance);
  dpm.calcForceEnergy();
  if (mode == "nve") dpm.initNVE(Tinject, false);
  else dpm.initLangevin(Tinject, damping / timeUnit, false);
  dpm.resetUpdateCount();
  io.openEnergyFile(outDir + "energy.dat");
  long saveFreq = std::max(1L, maxStep / 100), step = 0;
  while (step < maxStep) {
    long n = std::min(saveFreq, maxStep - step);
    if (mode == "nve") dpm.NVELoop(n);
    else dpm.langevinLoop(n);
    step += n;
    io.saveEnergy(step - 1, timeStep, numParticles, numVertices);
  }
  io.closeEnergyFile();
  std::cout << mode << ": steps: " << step << " E/N: " << dpm.getEnergy() / numVertices
            << " T: " << dpm.getTemperature() << " phi: " << dpm.getPhi() << " updates: " << dpm.getUpdateCount()
            << std::endl;
  io.savePacking(outDir);
  return 0;
}
