// runNVE -- NVE dynamics of a saved packing, with the command line and output layout of
// runNVE.cpp in farceri/cudaRing:
//
//   runNVE inDir timeStep Tinject maxStep initialStep numParticles potType intType (unused) mode
//
// potType: harmonic (default) | wca | lj; intType: smooth selects the vertex-segment
// interaction, anything else the vertex-vertex one; shape costs ea = 1e5, el = 20, eb = 10,
// ec = 1; timeStep in units of sigma / sqrt(ec); skin setDisplacementCutoff(0.5, 2 <r_v>).  Reads
// the state (velocities) in every mode, as the reference does; a "change" run rescales to Tinject
// after every energy-log line.  The sample directory is <nve[-pot]>/T<Tinject>/.  An optional
// 11th argument seeds the random streams.
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <string>

#include "../host/DPM2D.h"
#include "../host/FileIO.h"
#include "run_common.h"

int main(int argc, char** argv) {
  if (argc < 11) {
    std::cerr << "usage: runNVE inDir timeStep Tinject maxStep initialStep numParticles potType intType unused mode "
                 "[seed]" << std::endl;
    return 2;
  }
  std::string inDir = argv[1], potType = argv[7], intType = argv[8], mode = argv[10];
  double timeStep = std::atof(argv[2]), Tinject = std::atof(argv[3]);
  long maxStep = (long)std::atof(argv[4]), initialStep = (long)std::atof(argv[5]), numParticles = std::atol(argv[6]);
  const long nDim = 2, numVertexPerParticle = 32;
  const double cutoff = 0.5, ea = 1e05, el = 20, eb = 10, ec = 1, LJcut = 1.5;
  std::string whichDynamics;

  DPM2D dpm(numParticles, nDim, numVertexPerParticle);
  if (argc > 11) dpm.setSeed(std::strtoull(argv[11], nullptr, 10));
  if (potType == "lj") {
    dpm.setPotentialType(simControlStruct::potentialEnum::lennardJones);
    whichDynamics = "nve-lj/";
    dpm.setLJcutoff(LJcut);
  } else if (potType == "wca") {
    dpm.setPotentialType(simControlStruct::potentialEnum::wca);
    whichDynamics = "nve-wca/";
  } else {
    std::cout << "Setting default harmonic potential" << std::endl;
    whichDynamics = "nve/";
  }
  if (intType == "smooth") {
    dpm.setInteractionType(simControlStruct::interactionEnum::vertexSmooth);
    dpm.setNeighborType(simControlStruct::neighborEnum::neighbor);
  }
  const std::string dirSample = whichDynamics + "T" + argv[3] + "/";
  RunPlan plan = plan_run(inDir, mode, whichDynamics, dirSample, "../../", true);
  ioDPMFile io(&dpm);
  io.readPackingFromDirectory(plan.inDir, numParticles, nDim);
  dpm.setEnergyCosts(ea, el, eb, ec);
  if (plan.readState) io.readState(plan.inDir, numParticles, dpm.getNumVertices(), nDim);
  io.openEnergyFile(plan.outDir + "energy.dat");

  const double sigma = dpm.getMeanParticleSize(), timeUnit = sigma / std::sqrt(ec);
  timeStep = dpm.setTimeStep(timeStep * timeUnit);
  std::cout << "Time step: " << timeStep << " sigma: " << sigma << " Tinject: " << Tinject << std::endl;
  const long numVertices = dpm.getNumVertices();
  dpm.initNVE(Tinject, plan.readState);
  if (plan.makeNewDir) dpm.adjustTemperature(Tinject);
  const double size = 2 * dpm.getMeanVertexRadius();
  dpm.calcNeighbors(dpm.setDisplacementCutoff(cutoff, size));
  dpm.calcForceEnergy();
  dpm.resetUpdateCount();
  const double waveQ = dpm.getDeformableWaveNumber();

  RunLoop loop{maxStep, initialStep, timeStep, numVertices, "NVE"};
  loop.advance = [&](long n) { dpm.NVELoop(n); };
  loop.progress = [&] { std::cout << " ISF: " << dpm.getParticleISF(waveQ); };
  if (plan.makeNewDir) loop.onEnergy = [&] { dpm.adjustTemperature(Tinject); };
  loop.run(dpm, io, plan);
  io.closeEnergyFile();
  return 0;
}
