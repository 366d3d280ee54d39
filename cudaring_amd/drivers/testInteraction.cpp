// testInteraction -- two rings, one moving into the other (testInteraction.cpp in
// farceri/cudaRing, which no longer builds against its DPM2D): setTwoParticleTest(2.8, 2.8, 0.28,
// y1, 0.2) with 20-vertex rings, vertex-segment interaction, costs (1e5, 20, 10, 1), then
// testInteraction steps (velocity Verlet of ring 1 only).
//
// usage: testInteraction outDir timeStep maxStep [potential=lj|wca|harmonic] [seed]
//   lj: Lennard-Jones with cutoff 1.5 and y1 = 0.69; otherwise y1 = 0.65.  Writes the start and
//   end packings (outDir/start/, outDir/end/) and the energy log every maxStep / 100 steps.
#include <sys/stat.h>

#include <cmath>
#include <cstdlib>
#include <iostream>
#include <string>

#include "../host/DPM2D.h"
#include "../host/FileIO.h"

int main(int argc, char** argv) {
  if (argc < 4) {
    std::cerr << "usage: testInteraction outDir timeStep maxStep [lj|wca|harmonic] [seed]" << std::endl;
    return 2;
  }
  std::string outDir = argv[1];
  if (outDir.back() != '/') outDir += "/";
  double timeStep = std::atof(argv[2]);
  long maxStep = std::atol(argv[3]);
  std::string pot = argc > 4 ? argv[4] : "lj";
  const long numParticles = 2, numVertexPerParticle = 20;
  const double LJcutoff = 1.5, cutoff = 0.5, lx = 2.8, ly = 2.8, vel1 = 0.2, y0 = 0.28;
  double y1 = 0.65;
  mkdir(outDir.c_str(), 0755);

  DPM2D dp(numParticles, 2, numVertexPerParticle);
  if (argc > 5) dp.setSeed(std::strtoul(argv[5], nullptr, 10));
  dp.setInteractionType(simControlStruct::interactionEnum::vertexSmooth);
  dp.setEnergyCosts(1e5, 20, 10, 1);
  if (pot == "lj") {
    dp.setPotentialType(simControlStruct::potentialEnum::lennardJones);
    dp.setLJcutoff(LJcutoff);
    y1 = 0.69;
  } else if (pot == "wca") {
    dp.setPotentialType(simControlStruct::potentialEnum::wca);
  }
  ioDPMFile io(&dp);
  dp.setTwoParticleTest(lx, ly, y0, y1, vel1);
  long numVertices = dp.getNumVertices();
  timeStep = dp.setTimeStep(timeStep);
  std::cout << "initial velocity on particle 1: " << vel1 << " time step: " << timeStep << std::endl;
  double size = 2 * dp.getVertexRadius();
  double cutDistance = dp.setDisplacementCutoff(cutoff, size);
  dp.calcNeighbors(cutDistance);
  dp.calcForceEnergy();
  dp.resetUpdateCount();
  dp.resetLastPositions();
  std::string startDir = outDir + "start/", endDir = outDir + "end/";
  mkdir(startDir.c_str(), 0755);
  mkdir(endDir.c_str(), 0755);
  io.savePacking(startDir);
  io.openEnergyFile(outDir + "energy.dat");
  long saveFreq = std::max(1L, maxStep / 100), checkFreq = std::max(1L, maxStep / 10);
  for (long step = 0; step < maxStep; step++) {
    dp.testInteraction(timeStep);
    if (step % saveFreq == 0) io.saveEnergy(step, timeStep, numParticles, numVertices);
    if (step % checkFreq == 0) {
      double epot = dp.getPotentialEnergy() / numVertices, ekin = dp.getKineticEnergy() / numVertices;
      std::cout << "NVE: current step: " << step << " U: " << epot << " K: " << ekin << " Energy: " << epot + ekin
                << " updates: " << dp.getUpdateCount() << std::endl;
    }
  }
  io.closeEnergyFile();
  io.savePacking(endDir);
  return 0;
}
