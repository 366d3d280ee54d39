// compressNVE -- quasistatic compression of a saved packing by NVE runs between box shrinks
// (the flow of compressNVE.cpp:143-231 in farceri/cudaRing, from a packing on disk).
//
// usage: compressNVE inDir outDir timeStep Tinject maxStep numParticles phiTarget
//                    [potential=wca|lj|harmonic] [box=periodic|square|sides] [deltaPhi=2e-3] [dyn=nve|scalevel]
//   At each packing fraction: neighbours and forces, the energy injected by the last box shrink
//   removed (adjustKineticEnergy to the previous total energy), maxStep NVE steps with the energy
//   log every maxStep/100 steps, the packing saved under outDir/<phi>/, then
//   scaleBox(sqrt((phi + deltaPhi) / phi)) until phi reaches phiTarget.  timeStep is in units of
//   sigma / sqrt(ec) with sigma the mean particle size; costs are compressNVE.cpp:38's
//   (1e5, 20, 10, 1) with walls of energy scale ew = 10.  The packing's rings must lie inside the
//   box for the walled geometries.
#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

#include "../host/DPM2D.h"
#include "../host/FileIO.h"

int main(int argc, char** argv) {
  if (argc < 8) {
    std::cerr << "usage: compressNVE inDir outDir timeStep Tinject maxStep numParticles phiTarget "
                 "[potential] [box] [deltaPhi] [nve|scalevel]" << std::endl;
    return 2;
  }
  std::string inDir = argv[1], outDir = argv[2];
  if (inDir.back() != '/') inDir += "/";
  if (outDir.back() != '/') outDir += "/";
  double timeStep = std::atof(argv[3]), Tinject = std::atof(argv[4]);
  long maxStep = std::atol(argv[5]), numParticles = std::atol(argv[6]);
  double phiTh = std::atof(argv[7]);
  std::string pot = argc > 8 ? argv[8] : "wca", boxType = argc > 9 ? argv[9] : "periodic";
  double deltaPhi = argc > 10 ? std::atof(argv[10]) : 2e-3;
  std::string dynType = argc > 11 ? argv[11] : "nve";
  const double ea = 1e5, el = 20, eb = 10, ec = 1, ew = 10, LJcut = 1.5, cutoff = 0.5;
  const long maxSearchStep = 1500;
  mkdir(outDir.c_str(), 0755);

  DPM2D dpm(numParticles, 2, 32);
  if (boxType == "square") dpm.setGeometryType(simControlStruct::geometryEnum::fixedBox);
  else if (boxType == "sides") dpm.setGeometryType(simControlStruct::geometryEnum::fixedSides);
  else if (boxType != "periodic") {
    std::cerr << "compressNVE: box must be periodic, square or sides" << std::endl;
    return 2;
  }
  if (boxType != "periodic") dpm.setBoxEnergyCost(ew);
  ioDPMFile io(&dpm);
  io.readPackingFromDirectory(inDir, numParticles, 2);
  if (boxType != "periodic") {
    std::vector<double> pos = dpm.getVertexPositions(), L = dpm.getBoxSize();
    for (size_t i = 0; i < pos.size(); i++) {
      const double x = pos[i], Lw = L[i % 2];
      const bool walled = boxType == "square" || i % 2 == 1;
      if (walled && (x < 0 || x > Lw)) {
        std::cerr << "compressNVE: vertex " << i / 2 << " lies outside the walled box" << std::endl;
        return 2;
      }
    }
  }
  if (pot == "wca") dpm.setPotentialType(simControlStruct::potentialEnum::wca);
  else if (pot == "lj") {
    dpm.setPotentialType(simControlStruct::potentialEnum::lennardJones);
    dpm.setLJcutoff(LJcut);
  } else dpm.setPotentialType(simControlStruct::potentialEnum::harmonic);
  dpm.setInteractionType(simControlStruct::interactionEnum::vertexSmooth);
  dpm.setNeighborType(simControlStruct::neighborEnum::neighbor);
  dpm.setEnergyCosts(ea, el, eb, ec);
  long numVertices = dpm.getNumVertices();

  dpm.calcParticleShape();
  double currentPhi = dpm.getPhi(), prevEnergy = 0;
  std::cout << "CURRENT PHI: " << currentPhi << " preferred phi: " << dpm.getPreferredPhi() << std::endl;
  double sigma = dpm.getMeanParticleSize(), size = 2 * dpm.getMeanVertexRadius();
  double timeUnit = sigma / std::sqrt(ec);
  timeStep = dpm.setTimeStep(timeStep * timeUnit);
  std::cout << "Time step: " << timeStep << " sigma: " << sigma << " size: " << size << " Tinject: " << Tinject
            << std::endl;
  if (dynType == "scalevel") dpm.initNVERescale(Tinject);
  else dpm.initNVE(Tinject, false);
  long saveFreq = std::max(1L, maxStep / 100);
  for (long searchStep = 0; searchStep < maxSearchStep; searchStep++) {
    char phiName[32];
    std::snprintf(phiName, sizeof phiName, "%.3f", std::round(dpm.getPhi() * 1000) / 1000);
    std::string currentDir = outDir + phiName + "/";
    mkdir(currentDir.c_str(), 0755);
    io.openEnergyFile(currentDir + "energy.dat");
    double cutDistance = dpm.setDisplacementCutoff(cutoff, size);
    dpm.calcNeighbors(cutDistance);
    dpm.calcForceEnergy();
    dpm.resetUpdateCount();
    dpm.setInitialPositions();
    double waveQ = dpm.getDeformableWaveNumber();
    if (searchStep != 0 && dynType == "nve") {
      std::cout << "Energy after compression - E/N: " << dpm.getEnergy() / numVertices << std::endl;
      dpm.adjustKineticEnergy(prevEnergy);
      dpm.calcForceEnergy();
      std::cout << "Energy after adjustment - E/N: " << dpm.getEnergy() / numVertices << std::endl;
    }
    // the log line of step k follows the (k + 1)-th step, for k a multiple of saveFreq
    long step = 0;
    while (step < maxStep) {
      long n = step == 0 ? 1 : std::min(saveFreq, maxStep - step);
      if (dynType == "scalevel") {
        for (long i = 0; i < n; i++) dpm.NVERescaleLoop();
      } else {
        dpm.NVELoop(n);
      }
      if ((step + n - 1) % saveFreq == 0) io.saveEnergy(step + n - 1, timeStep, numParticles, numVertices);
      step += n;
    }
    std::cout << "NVE: steps: " << step << " E/N: " << dpm.getEnergy() / numVertices << " T: " << dpm.getTemperature()
              << " ISF: " << dpm.getParticleISF(waveQ) << " phi: " << dpm.getPhi()
              << " updates: " << dpm.getUpdateCount() << std::endl;
    prevEnergy = dpm.getEnergy();
    io.savePacking(currentDir);
    io.closeEnergyFile();
    if (currentPhi >= phiTh) {
      std::cout << "\nTarget density met, current phi: " << currentPhi << std::endl;
      break;
    }
    double scaleFactor = std::sqrt((currentPhi + deltaPhi) / currentPhi);
    dpm.scaleBox(scaleFactor);
    currentPhi = dpm.getPhi();
    std::vector<double> L = dpm.getBoxSize();
    std::cout << "\nNEW PACKING FRACTION: " << currentPhi << " preferred: " << dpm.getPreferredPhi()
              << " Lx: " << L[0] << " Ly: " << L[1] << " scale: " << scaleFactor << std::endl;
  }
  return 0;
}
