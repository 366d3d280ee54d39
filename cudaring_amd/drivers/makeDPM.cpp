// makeDPM -- make one minimised deformable-ring packing (the flow of makeDPM.cpp in
// farceri/cudaRing, with the set-up calls of its current API, compressFIRE.cpp:47-90):
// polydisperse ring sizes, sinusoidal rest angles, random soft disks at phi0, soft-disk FIRE,
// vertices on the disks, then vertex FIRE until the total force magnitude is below the
// tolerance, and the packing saved in outDir.
//
// usage: makeDPM outDir [numParticles=32] [numVertexPerParticle=32] [seed=0] [maxIter=1000000]
// The first argument is the reference's whole command line; the options change the constants it
// fixes (32 rings of 32 vertices, polydispersity 0.1, phi0 0.4, calA0 1.01, costs 1e3, 1, 1e-2, 1,
// thetaA 1, thetaK 0, cutDistance 2, force tolerance 1e-10, FIRE dt0 5e-2).
#include <sys/stat.h>

#include <cstdio>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "../host/DPM2D.h"
#include "../host/FIRE.h"
#include "../host/FileIO.h"

int main(int argc, char** argv) {
  if (argc < 2) {
    std::cerr << "usage: makeDPM outDir [numParticles=] [numVertexPerParticle=] [seed=] [maxIter=]" << std::endl;
    return 2;
  }
  std::string outDir = argv[1];
  if (outDir.back() != '/') outDir += "/";
  long numParticles = 32, nvp = 32, maxIterations = 1000000;
  unsigned long seed = 0;
  for (int i = 2; i < argc; i++) {
    std::string a = argv[i];
    const size_t eq = a.find('=');
    const std::string key = a.substr(0, eq), val = eq == std::string::npos ? "" : a.substr(eq + 1);
    if (key == "numParticles") numParticles = std::atol(val.c_str());
    else if (key == "numVertexPerParticle") nvp = std::atol(val.c_str());
    else if (key == "seed") seed = std::strtoul(val.c_str(), nullptr, 10);
    else if (key == "maxIter") maxIterations = std::atol(val.c_str());
    else {
      std::cerr << "makeDPM: unknown option " << a << std::endl;
      return 2;
    }
  }
  const long minStep = 20;
  const double polydispersity = 0.1, phi0 = 0.4, cutDistance = 2., forceTolerance = 1e-10;
  const double ea = 1e3, el = 1, eb = 1e-2, ec = 1, dt0 = 5e-2, calA0 = 1.01, thetaA = 1., thetaK = 0.;
  mkdir(outDir.c_str(), 0755);

  DPM2D dpm(numParticles, 2, nvp);
  dpm.setSeed(seed);
  dpm.setEnergyCosts(ea, el, eb, ec);
  ioDPMFile io(&dpm);
  dpm.setPolySizeDistribution(calA0, polydispersity);
  dpm.setSinusoidalRestAngles(thetaA, thetaK);
  dpm.setRandomParticles(phi0, 1.0);

  // ---- soft disks
  std::vector<double> pf = {0.2, 0.5, 1.1, 0.99, dt0, 10 * dt0, 0.2};
  dpm.initFIRE(pf, minStep, 0, numParticles);
  dpm.setParticleMassFIRE();
  dpm.calcParticleNeighborList(cutDistance);
  dpm.calcParticleForceEnergy();
  double cutoff = (1 + cutDistance) * dpm.getMinParticleSigma();
  dpm.resetParticleLastPositions();
  long it = 0;
  while (dpm.getParticleMaxUnbalancedForce() > forceTolerance && it != maxIterations) {
    dpm.particleFIRELoop();
    if (3 * dpm.getParticleMaxDisplacement() > cutoff) {
      dpm.calcParticleNeighborList(cutDistance);
      dpm.resetParticleLastPositions();
    }
    it++;
  }
  std::cout << "FIRE: iteration: " << it << " maxUnbalancedForce: " << std::setprecision(precision)
            << dpm.getParticleMaxUnbalancedForce() << " energy: " << dpm.getParticlePotentialEnergy() << std::endl;
  dpm.initVerticesOnParticles();
  std::cout << "current packing fraction: " << std::setprecision(precision) << dpm.getPhi() << std::endl;

  // ---- deformable rings
  dpm.calcNeighborList(cutDistance);
  dpm.calcForceEnergy();
  const double timeStep = dpm.setTimeScale(dt0);
  std::cout << "fire time step: " << timeStep << std::endl;
  std::vector<double> vf = {0.2, 0.5, 1.1, 0.99, timeStep, 10 * timeStep, 0.2};
  dpm.initFIRE(vf, minStep, 0, dpm.getNumVertices());
  dpm.resetLastPositions();
  dpm.setDisplacementCutoff((1 + cutDistance) * dpm.getMeanVertexRadius(), cutDistance);
  // makeDPM.cpp:76-87: one FIRE iteration, then the total force magnitude, per loop pass
  it = 0;
  double forceCheck = dpm.getTotalForceMagnitude();
  const long printFreq = 100;
  while (forceCheck > forceTolerance && it != maxIterations) {
    dpm.vertexFIRELoop();
    forceCheck = dpm.getTotalForceMagnitude();
    if (it % printFreq == 0)
      std::cout << "FIRE: iteration: " << it << " maxUnbalancedForce: " << std::setprecision(precision) << forceCheck
                << " energy: " << dpm.getPotentialEnergy() << std::endl;
    it++;
  }
  std::cout << "FIRE: iteration: " << it << " maxUnbalancedForce: " << std::setprecision(precision)
            << dpm.getMaxUnbalancedForce() << " totalForce: " << dpm.getTotalForceMagnitude()
            << " energy: " << dpm.getPotentialEnergy() << std::endl;
  io.savePacking(outDir);
  return 0;
}
