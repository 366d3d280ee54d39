// compressNVE -- quasistatic compression by NVE runs between box shrinks (compressNVE.cpp in
// farceri/cudaRing, same arguments and directory layout).
//
// usage: compressNVE outDir inDir timeStep Tinject shape0 numParticles maxStep lx ly potType boxType
//                    [phiTarget=0.96] [deltaPhi=2e-3] [seed] [dyn=nve|scalevel]
//   inDir "0": a new packing -- polydisperse sizes (calA0 = shape0, polydispersity 0.1), soft disks
//   at phi 0.1 in an lx x ly box (boxType "circle": a round box of radius lx), soft-disk FIRE,
//   vertices on the disks, lengths in units of the mean particle size; written to
//   outDir/<N>-box<lx><ly>-A<shape0>/{fire,initial}/ (circle: outDir/circle/<N>-A<shape0>/).
//   Otherwise the packing and velocities of outDir/inDir/ are read.
//   boxType: square (fixedBox), sides (fixedSides), circle (roundBox), anything else periodic;
//   walls have energy scale ew = 10.  potType: wca, lj or harmonic, vertex-segment interaction.
//   Then at each packing fraction: neighbours and forces, the energy injected by the last box
//   shrink removed (adjustKineticEnergy to the previous total energy), maxStep NVE steps with the
//   energy log every maxStep/100 steps, the packing saved under <dir>/<phi>/, then
//   scaleBox(sqrt((phi + deltaPhi) / phi)) until phi reaches phiTarget.  timeStep is in units of
//   sigma / sqrt(ec), sigma the mean particle size; costs (1e5, 20, 10, 1) as compressNVE.cpp:38.
//   maxStep 0 stops after the set-up (no device is used).
#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

#include "../host/DPM2D.h"
#include "../host/FIRE.h"
#include "../host/FileIO.h"

int main(int argc, char** argv) {
  if (argc < 12) {
    std::cerr << "usage: compressNVE outDir inDir timeStep Tinject shape0 numParticles maxStep lx ly potType "
                 "boxType [phiTarget] [deltaPhi] [seed] [nve|scalevel]" << std::endl;
    return 2;
  }
  std::string outDir = argv[1], inDir = argv[2];
  if (outDir.back() != '/') outDir += "/";
  double timeStep = std::atof(argv[3]), Tinject = std::atof(argv[4]), shape0 = std::atof(argv[5]);
  long numParticles = std::atol(argv[6]), maxStep = std::atol(argv[7]);
  double lx = std::atof(argv[8]), ly = std::atof(argv[9]);
  std::string potType = argv[10], boxType = argv[11];
  double phiTh = argc > 12 ? std::atof(argv[12]) : 0.96, deltaPhi = argc > 13 ? std::atof(argv[13]) : 2e-3;
  unsigned long seed = argc > 14 ? std::strtoul(argv[14], nullptr, 10) : 0;
  std::string dynType = argc > 15 ? argv[15] : "nve";
  const long maxSearchStep = 1500, maxIterations = 5000000, minStep = 20;
  const double polydispersity = 0.1, cutoff = 0.5, LJcut = 1.5, forceTolerance = 1e-12, FIREStep = 1e-2;
  const double ea = 1e5, el = 20, eb = 10, ec = 1, ew = 10;
  double currentPhi = 0.1, prevEnergy = 0;
  mkdir(outDir.c_str(), 0755);

  DPM2D dpm(numParticles, 2, 32);
  if (argc > 14) dpm.setSeed(seed);
  if (boxType == "square") dpm.setGeometryType(simControlStruct::geometryEnum::fixedBox);
  else if (boxType == "sides") dpm.setGeometryType(simControlStruct::geometryEnum::fixedSides);
  else if (boxType == "circle") dpm.setGeometryType(simControlStruct::geometryEnum::roundBox);
  else dpm.setGeometryType(simControlStruct::geometryEnum::normal);
  if (boxType == "square" || boxType == "sides" || boxType == "circle") dpm.setBoxEnergyCost(ew);
  ioDPMFile io(&dpm);

  if (inDir != "0") {
    std::cout << "Reading initial configuration from " << inDir << std::endl;
    inDir = outDir + inDir + "/";
    io.readPackingFromDirectory(inDir, numParticles, 2);
    io.readState(inDir, numParticles, dpm.getNumVertices(), 2);
  } else {
    char tag[128];
    if (boxType == "circle") {
      outDir += "circle/";
      mkdir(outDir.c_str(), 0755);
      std::snprintf(tag, sizeof tag, "%s-A%s/", argv[6], argv[5]);
    } else {
      std::snprintf(tag, sizeof tag, "%s-box%s%s-A%s/", argv[6], argv[8], argv[9], argv[5]);
    }
    outDir += tag;
    mkdir(outDir.c_str(), 0755);
    // soft disks with the harmonic contact, relaxed by FIRE (compressNVE.cpp:70-103)
    dpm.setPotentialType(simControlStruct::potentialEnum::harmonic);
    dpm.setInteractionType(simControlStruct::interactionEnum::vertexVertex);
    dpm.setPolySizeDistribution(shape0, polydispersity);
    if (boxType == "circle") dpm.setRoundScaledRandomParticles(currentPhi, 1.5, lx);
    else dpm.setScaledRandomParticles(currentPhi, 1.5, lx, ly);
    dpm.scaleParticlePacking();
    dpm.setEnergyCosts(0, 0, 0, ec);
    std::vector<double> fp = {0.2, 0.5, 1.1, 0.99, FIREStep, 10 * FIREStep, 0.2};
    dpm.initFIRE(fp, minStep, 0, numParticles);
    dpm.setParticleMassFIRE();
    double size = 2 * dpm.getMinParticleSigma();
    // setDisplacementCutoff(0.5, size) of the harmonic contact (DPM2D.cu:1300-1329): list radius
    // 1 + 0.5 particle diameters, refresh after a displacement of 0.5 size / 3
    double cutDistance = 1 + cutoff, dispCut = cutoff * size;
    dpm.calcParticleNeighborList(cutDistance);
    dpm.calcParticleForceEnergy();
    dpm.resetParticleLastPositions();
    long iteration = 0;
    while (dpm.getParticleMaxUnbalancedForce() > forceTolerance && iteration != maxIterations) {
      dpm.particleFIRELoop();
      if (3 * dpm.getParticleMaxDisplacement() > dispCut) {
        dpm.calcParticleNeighborList(cutDistance);
        dpm.resetParticleLastPositions();
      }
      iteration += 1;
    }
    std::cout << "\nFIRE: iteration: " << iteration << " maxUnbalancedForce: " << std::setprecision(precision)
              << dpm.getParticleMaxUnbalancedForce() << " energy: " << dpm.getParticlePotentialEnergy() << std::endl;
    std::string fireDir = outDir + "fire/";
    mkdir(fireDir.c_str(), 0755);
    io.saveParticlePacking(fireDir);
    dpm.initVerticesOnParticles();
    dpm.scalePacking(dpm.getMeanParticleSize());
    if (maxStep <= 0) {  // set-up only: host files, no device
      std::string dp = outDir + "initial/";
      mkdir(dp.c_str(), 0755);
      io.save1DIndexFile(dp + "numVertexInParticleList.dat", dpm.getNumVertexInParticleList());
      io.save1DFile(dp + "radii.dat", dpm.getVertexRadii());
      io.save2DFile(dp + "positions.dat", dpm.getVertexPositions(), 2);
      io.save1DFile(dp + "restAreas.dat", dpm.getRestAreas());
      io.save1DFile(dp + "restLengths.dat", dpm.getRestLengths());
      io.save1DFile(dp + "restAngles.dat", dpm.getRestAngles());
      io.save1DFile(dp + "particleRad.dat", dpm.particleRad);
      if (boxType == "circle") io.save0DFile(dp + "boxSize.dat", dpm.getBoxRadius());
      else io.save1DFile(dp + "boxSize.dat", dpm.getBoxSize());
      return 0;
    }
    std::string initDir = outDir + "initial/";
    mkdir(initDir.c_str(), 0755);
    io.savePacking(initDir);
  }
  if (maxStep <= 0) return 0;

  if (potType == "wca") dpm.setPotentialType(simControlStruct::potentialEnum::wca);
  else if (potType == "lj") {
    dpm.setPotentialType(simControlStruct::potentialEnum::lennardJones);
    dpm.setLJcutoff(LJcut);
  } else dpm.setPotentialType(simControlStruct::potentialEnum::harmonic);
  dpm.setInteractionType(simControlStruct::interactionEnum::vertexSmooth);
  dpm.setNeighborType(simControlStruct::neighborEnum::neighbor);
  dpm.setEnergyCosts(ea, el, eb, ec);
  long numVertices = dpm.getNumVertices();

  // quasistatic isothermal compression (compressNVE.cpp:143-231)
  dpm.calcParticleShape();
  currentPhi = dpm.getPhi();
  std::cout << "CURRENT PHI: " << currentPhi << " preferred phi: " << dpm.getPreferredPhi() << std::endl;
  double sigma = dpm.getMeanParticleSize(), size = 2 * dpm.getMeanVertexRadius();
  double timeUnit = sigma / std::sqrt(ec);
  timeStep = dpm.setTimeStep(timeStep * timeUnit);
  std::cout << "Time step: " << timeStep << " sigma: " << sigma << " size: " << size << " Tinject: " << Tinject
            << std::endl;
  if (dynType == "scalevel") dpm.initNVERescale(Tinject);
  else dpm.initNVE(Tinject, inDir != "0");
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
    std::cout << "\nNEW PACKING FRACTION: " << currentPhi << " preferred: " << dpm.getPreferredPhi();
    if (boxType == "circle") {
      std::cout << " boxR: " << dpm.getBoxRadius();
    } else {
      std::vector<double> L = dpm.getBoxSize();
      std::cout << " Lx: " << L[0] << " Ly: " << L[1];
    }
    std::cout << " scale: " << scaleFactor << std::endl;
  }
  return 0;
}
