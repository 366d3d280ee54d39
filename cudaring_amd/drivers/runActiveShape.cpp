// runActiveShape -- rings with fluctuating rest lengths (activeShape: Ornstein-Uhlenbeck rest
// lengths, Langevin thermostat, WCA vertex-segment interaction) from a saved packing, with the
// command line and output layout of runActiveShape.cpp in farceri/cudaRing:
//
//   runActiveShape inDir timeStep Tinject damping l0Tau l0Diff maxStep initialStep numParticles
//                  potType noiseType mode
//
// timeStep, damping and l0Tau in units of sigma / sqrt(ec), l0Diff in sigma^2 / that time unit;
// shape costs ea = 1e5, el = 20, eb = 10, ec = 1; skin setDisplacementCutoff(0.5, 2 <r_v>).
// potType and noiseType are read as there and leave the WCA potential and the Langevin noise in
// place (the reference only prints the drive differently for drivenBrownian).  Reads the state in
// every mode; the sample directory is shape/l0Tau<l0Tau>-l0Diff<l0Diff>/ and a "change" run writes
// it beside the input directory (one level up, runActiveShape.cpp:88-90).  An optional 13th
// argument seeds the random streams.
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <string>

#include "../host/DPM2D.h"
#include "../host/FileIO.h"
#include "run_common.h"

int main(int argc, char** argv) {
  if (argc < 13) {
    std::cerr << "usage: runActiveShape inDir timeStep Tinject damping l0Tau l0Diff maxStep initialStep "
                 "numParticles potType noiseType mode [seed]" << std::endl;
    return 2;
  }
  std::string inDir = argv[1], noiseType = argv[11], mode = argv[12];
  double timeStep = std::atof(argv[2]), Tinject = std::atof(argv[3]), damping = std::atof(argv[4]);
  double l0Tau = std::atof(argv[5]), l0Diff = std::atof(argv[6]);
  long maxStep = (long)std::atof(argv[7]), initialStep = (long)std::atof(argv[8]), numParticles = std::atol(argv[9]);
  const long nDim = 2, numVertexPerParticle = 32;
  const double cutoff = 0.5, ea = 1e05, el = 20, eb = 10, ec = 1;
  const std::string whichDynamics = "shape/";

  DPM2D dpm(numParticles, nDim, numVertexPerParticle);
  if (argc > 13) dpm.setSeed(std::strtoull(argv[13], nullptr, 10));
  dpm.setParticleType(simControlStruct::particleEnum::activeShape);
  dpm.setNoiseType(simControlStruct::noiseEnum::langevin);
  dpm.setPotentialType(simControlStruct::potentialEnum::wca);
  dpm.setInteractionType(simControlStruct::interactionEnum::vertexSmooth);
  dpm.setNeighborType(simControlStruct::neighborEnum::neighbor);
  const std::string dirSample = whichDynamics + "l0Tau" + argv[5] + "-l0Diff" + argv[6] + "/";
  RunPlan plan = plan_run(inDir, mode, whichDynamics, dirSample, "../", true);
  ioDPMFile io(&dpm);
  io.readPackingFromDirectory(plan.inDir, numParticles, nDim);
  dpm.setEnergyCosts(ea, el, eb, ec);
  if (plan.readState) io.readState(plan.inDir, numParticles, dpm.getNumVertices(), nDim);
  const long numVertices = dpm.getNumVertices();
  io.openEnergyFile(plan.outDir + "energy.dat");

  const double sigma = dpm.getMeanParticleSize(), timeUnit = sigma / std::sqrt(ec);
  const double diffUnit = sigma * sigma / timeUnit;
  timeStep = dpm.setTimeStep(timeStep * timeUnit);
  std::cout << "Units - time: " << timeUnit << " space: " << sigma << " time step: " << timeStep << std::endl;
  if (noiseType == "drivenBrownian") {
    std::cout << "Driven Shape - damping: " << damping << " l0Tau: " << l0Tau << " l0Diff: " << l0Diff << std::endl;
  } else {
    std::cout << "Thermostat - damping: " << damping << " Tinject: " << Tinject << std::endl;
    std::cout << "Active Shape - l0Tau: " << l0Tau << " l0Diff: " << l0Diff << std::endl;
  }
  damping /= timeUnit;
  l0Tau *= timeUnit;
  l0Diff *= diffUnit;
  dpm.setActiveTensionParams(l0Tau, l0Diff);
  io.saveDriveParams(plan.outDir, damping);
  dpm.initLangevin(Tinject, damping, plan.readState);
  const double size = 2 * dpm.getMeanVertexRadius();
  dpm.calcNeighbors(dpm.setDisplacementCutoff(cutoff, size));
  dpm.calcForceEnergy();
  dpm.resetUpdateCount();
  const double waveQ = dpm.getDeformableWaveNumber();

  RunLoop loop{maxStep, initialStep, timeStep, numVertices, "Active Shape"};
  loop.advance = [&](long n) { dpm.langevinLoop(n); };
  loop.progress = [&] {
    std::cout << " Ks/Kc: " << dpm.getShapeCOMEnergyRatio() << " ISF: " << dpm.getParticleISF(waveQ);
  };
  loop.run(dpm, io, plan);
  io.closeEnergyFile();
  return 0;
}
