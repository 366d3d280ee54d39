// runNVT -- Langevin (or Brownian / driven Brownian) dynamics of a saved packing, with the
// command line and output layout of runNVT.cpp in farceri/cudaRing:
//
//   runNVT inDir timeStep Tinject damping maxStep initialStep numParticles noiseType potType mode
//
// noiseType: langevin (default) | brownian | driven; potType: harmonic (default) | wca | lj;
// vertex-segment ("smooth") interaction; shape costs ea = 1e5, el = 20, eb = 10, ec = 1;
// timeStep and damping in units of sigma / sqrt(ec) with sigma the mean particle size; skin
// setDisplacementCutoff(0.5, 2 <r_v>).  The sample directory is
// <dynamics>/T<Tinject>-damping<damping>/ (run_common.h for the modes).  An optional 11th
// argument seeds the random streams (the reference seeds from the clock).
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <string>

#include "../host/DPM2D.h"
#include "../host/FileIO.h"
#include "run_common.h"

int main(int argc, char** argv) {
  if (argc < 11) {
    std::cerr << "usage: runNVT inDir timeStep Tinject damping maxStep initialStep numParticles noiseType potType "
                 "mode [seed]" << std::endl;
    return 2;
  }
  std::string inDir = argv[1], noiseType = argv[8], potType = argv[9], mode = argv[10];
  double timeStep = std::atof(argv[2]), Tinject = std::atof(argv[3]), damping = std::atof(argv[4]);
  long maxStep = (long)std::atof(argv[5]), initialStep = (long)std::atof(argv[6]), numParticles = std::atol(argv[7]);
  const long nDim = 2, numVertexPerParticle = 32;
  const double cutoff = 0.5, ea = 1e05, el = 20, eb = 10, ec = 1, LJcut = 1.5;
  std::string whichDynamics = "langevin";

  DPM2D dpm(numParticles, nDim, numVertexPerParticle);
  if (argc > 11) dpm.setSeed(std::strtoull(argv[11], nullptr, 10));
  if (noiseType == "brownian") {
    dpm.setNoiseType(simControlStruct::noiseEnum::brownian);
    whichDynamics = "brownian";
  } else if (noiseType == "driven") {
    dpm.setNoiseType(simControlStruct::noiseEnum::drivenBrownian);
    whichDynamics = "driven";
  } else {
    std::cout << "Setting default langevin noise" << std::endl;
  }
  if (potType == "lj") {
    dpm.setPotentialType(simControlStruct::potentialEnum::lennardJones);
    whichDynamics += "-lj/";
    dpm.setLJcutoff(LJcut);
  } else if (potType == "wca") {
    dpm.setPotentialType(simControlStruct::potentialEnum::wca);
    whichDynamics += "-wca/";
  } else {
    std::cout << "Setting default harmonic potential" << std::endl;
  }
  dpm.setInteractionType(simControlStruct::interactionEnum::vertexSmooth);
  dpm.setNeighborType(simControlStruct::neighborEnum::neighbor);
  // with the harmonic potential whichDynamics has no trailing '/', so the sample directory sits
  // directly in inDir, as runNVT.cpp:87 makes it
  const std::string dirSample = whichDynamics + "T" + argv[3] + "-damping" + argv[4] + "/";
  RunPlan plan = plan_run(inDir, mode, whichDynamics, dirSample, "../", false);
  ioDPMFile io(&dpm);
  io.readPackingFromDirectory(plan.inDir, numParticles, nDim);
  dpm.setEnergyCosts(ea, el, eb, ec);
  if (plan.readState) io.readState(plan.inDir, numParticles, dpm.getNumVertices(), nDim);
  const long numVertices = dpm.getNumVertices();
  io.openEnergyFile(plan.outDir + "energy.dat");

  const double sigma = dpm.getMeanParticleSize(), timeUnit = sigma / std::sqrt(ec);
  timeStep = dpm.setTimeStep(timeStep * timeUnit);
  std::cout << "Units - time: " << timeUnit << " space: " << sigma << " time step: " << timeStep << std::endl;
  std::cout << "Thermostat - damping: " << damping << " Tinject: " << Tinject << std::endl;
  damping /= timeUnit;
  io.saveDriveParams(plan.outDir, damping);
  dpm.initLangevin(Tinject, damping, plan.readState);
  const double size = 2 * dpm.getMeanVertexRadius();
  dpm.calcNeighbors(dpm.setDisplacementCutoff(cutoff, size));
  dpm.calcForceEnergy();
  dpm.resetUpdateCount();
  const double waveQ = dpm.getDeformableWaveNumber();

  RunLoop loop{maxStep, initialStep, timeStep, numVertices, "NVT"};
  loop.advance = [&](long n) { dpm.langevinLoop(n); };
  loop.progress = [&] {
    std::cout << " Ks/Kc: " << dpm.getShapeCOMEnergyRatio() << " ISF: " << dpm.getParticleISF(waveQ);
  };
  loop.run(dpm, io, plan);
  io.closeEnergyFile();
  return 0;
}
