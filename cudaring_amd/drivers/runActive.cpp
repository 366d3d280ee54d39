// runActive -- active Brownian rings (self-propulsion + Langevin thermostat, WCA vertex-segment
// interaction) from a saved packing, with the command line and output layout of runActive.cpp
// in farceri/cudaRing:
//
//   runActive inDir timeStep Tinject damping taup driving maxStep initialStep numParticles mode
//
// timeStep, damping and taup in units of sigma / sqrt(ec), driving in ec / sigma; shape costs
// ea = 1e5, el = 20, eb = 10, ec = 1; skin setDisplacementCutoff(0.5, 2 <r_v>).  Reads the state
// in every mode, as the reference does.  The sample directory is active/tp<taup>-f0<driving>/;
// state snapshots also hold the ring contacts.  An optional 11th argument seeds the random streams.
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <string>

#include "../host/DPM2D.h"
#include "../host/FileIO.h"
#include "run_common.h"

int main(int argc, char** argv) {
  if (argc < 11) {
    std::cerr << "usage: runActive inDir timeStep Tinject damping taup driving maxStep initialStep numParticles mode "
                 "[seed]" << std::endl;
    return 2;
  }
  std::string inDir = argv[1], mode = argv[10];
  double timeStep = std::atof(argv[2]), Tinject = std::atof(argv[3]), damping = std::atof(argv[4]);
  double tp = std::atof(argv[5]), driving = std::atof(argv[6]);
  long maxStep = (long)std::atof(argv[7]), initialStep = (long)std::atof(argv[8]), numParticles = std::atol(argv[9]);
  const long nDim = 2, numVertexPerParticle = 32;
  const double cutoff = 0.5, ea = 1e05, el = 20, eb = 10, ec = 1;
  const std::string whichDynamics = "active/";

  DPM2D dpm(numParticles, nDim, numVertexPerParticle);
  if (argc > 11) dpm.setSeed(std::strtoull(argv[11], nullptr, 10));
  dpm.setParticleType(simControlStruct::particleEnum::active);
  dpm.setNoiseType(simControlStruct::noiseEnum::langevin);
  dpm.setPotentialType(simControlStruct::potentialEnum::wca);
  dpm.setInteractionType(simControlStruct::interactionEnum::vertexSmooth);
  dpm.setNeighborType(simControlStruct::neighborEnum::neighbor);
  const std::string dirSample = whichDynamics + "tp" + argv[5] + "-f0" + argv[6] + "/";
  RunPlan plan = plan_run(inDir, mode, whichDynamics, dirSample, "../../", true);
  ioDPMFile io(&dpm);
  io.readPackingFromDirectory(plan.inDir, numParticles, nDim);
  dpm.setEnergyCosts(ea, el, eb, ec);
  if (plan.readState) io.readState(plan.inDir, numParticles, dpm.getNumVertices(), nDim);
  const long numVertices = dpm.getNumVertices();
  io.openEnergyFile(plan.outDir + "energy.dat");

  const double sigma = dpm.getMeanParticleSize(), timeUnit = sigma / std::sqrt(ec), forceUnit = ec / sigma;
  timeStep = dpm.setTimeStep(timeStep * timeUnit);
  std::cout << "Units - time: " << timeUnit << " space: " << sigma << " time step: " << timeStep << std::endl;
  std::cout << "Thermostat - damping: " << damping << " Tinject: " << Tinject
            << " magnitude:" << std::sqrt(2. * Tinject * damping / timeStep) << std::endl;
  std::cout << "Activity - taup: " << tp << " driving: " << driving
            << " Pe = 3 v_0 tau_p / sigma: " << 3 * (driving / damping) * tp / sigma << std::endl;
  damping /= timeUnit;
  tp *= timeUnit;
  driving *= forceUnit;
  dpm.setSelfPropulsionParams(driving, tp);
  io.saveDriveParams(plan.outDir, damping);
  dpm.initLangevin(Tinject, damping, plan.readState);
  const double size = 2 * dpm.getMeanVertexRadius();
  dpm.calcNeighbors(dpm.setDisplacementCutoff(cutoff, size));
  dpm.calcForceEnergy();
  dpm.resetUpdateCount();
  const double waveQ = dpm.getDeformableWaveNumber();

  RunLoop loop{maxStep, initialStep, timeStep, numVertices, "Active Motility"};
  loop.saveContacts = true;
  loop.advance = [&](long n) { dpm.langevinLoop(n); };
  loop.progress = [&] {
    std::cout << " Ks/Kc: " << dpm.getShapeCOMEnergyRatio() << " ISF: " << dpm.getParticleISF(waveQ);
  };
  loop.run(dpm, io, plan);
  io.closeEnergyFile();
  return 0;
}
