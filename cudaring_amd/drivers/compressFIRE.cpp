// compressFIRE -- make a deformable-ring packing and compress it with FIRE
// (the flow of compressFIRE.cpp in farceri/cudaRing: polydisperse sizes, random soft disks,
// soft-disk FIRE, vertices on the disks, then vertex FIRE at growing packing fraction).
//
// usage: compressFIRE outDir numParticles numVertexPerParticle calA0 polydispersity phiTarget seed
//                     [deltaPhi=2e-3] [phi0=0.2] [costs=ea,el,eb,ec] [maxIter=5000000] [bi=nSmall,nLarge]
//   phiTarget <= 0 stops after the set-up (no device is used) and writes the set-up files.
// Each compression step: scaleVertices(sqrt((phi + deltaPhi) / phi)), then FIRE until the
// largest force component is below 1e-10 / (mean particle size), as compressFIRE.cpp:104-131.
// The default costs are compressFIRE.cpp:30's (1e3, 1, 0, 1).
#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "../host/DPM2D.h"
#include "../host/FIRE.h"
#include "../host/FileIO.h"

int main(int argc, char** argv) {
  if (argc < 8) {
    std::cerr << "usage: compressFIRE outDir numParticles numVertexPerParticle calA0 polydispersity "
                 "phiTarget seed [deltaPhi=] [phi0=] [costs=ea,el,eb,ec] [maxIter=] [bi=nSmall,nLarge]" << std::endl;
    return 2;
  }
  std::string outDir = argv[1];
  if (outDir.back() != '/') outDir += "/";
  long numParticles = std::atol(argv[2]), nvp = std::atol(argv[3]);
  double calA0 = std::atof(argv[4]), poly = std::atof(argv[5]), phiTarget = std::atof(argv[6]);
  unsigned long seed = std::strtoul(argv[7], nullptr, 10);
  double deltaPhi = 2e-3, phi0 = 0.2, ea = 1e3, el = 1, eb = 0, ec = 1;
  long nSmall = 0, nLarge = 0, maxIterations = 5000000;
  for (int i = 8; i < argc; i++) {
    std::string a = argv[i];
    const size_t eq = a.find('=');
    const std::string key = a.substr(0, eq), val = eq == std::string::npos ? "" : a.substr(eq + 1);
    if (key == "deltaPhi") deltaPhi = std::atof(val.c_str());
    else if (key == "phi0") phi0 = std::atof(val.c_str());
    else if (key == "maxIter") maxIterations = std::atol(val.c_str());
    else if (key == "costs") std::sscanf(val.c_str(), "%lf,%lf,%lf,%lf", &ea, &el, &eb, &ec);
    else if (key == "bi") std::sscanf(val.c_str(), "%ld,%ld", &nSmall, &nLarge);
    else {
      std::cerr << "compressFIRE: unknown option " << a << std::endl;
      return 2;
    }
  }
  const long minStep = 20;
  const double cutDistance = 1, dt = 1e-2, forceTolerance0 = 1e-10;
  mkdir(outDir.c_str(), 0755);

  DPM2D dpm(numParticles, 2, nvp);
  dpm.setSeed(seed);
  dpm.setEnergyCosts(ea, el, eb, ec);
  ioDPMFile io(&dpm);

  // ---- soft disks
  if (nSmall > 0) dpm.setBiSizeDistribution(calA0, nSmall, nLarge, 0.5);
  else dpm.setPolySizeDistribution(calA0, poly);
  dpm.setRandomParticles(phi0, 1.1);
  double sigma = dpm.getMeanParticleSize();
  double step = dpm.setTimeStep(dt * 10 * sigma);
  std::vector<double> pf = {0.2, 0.5, 1.1, 0.99, step, 10 * step, 0.2};
  dpm.initFIRE(pf, minStep, 0, numParticles);
  dpm.setParticleMassFIRE();
  dpm.calcParticleNeighborList(cutDistance);
  dpm.calcParticleForceEnergy();
  double cutoff = (1 + cutDistance) * dpm.getMinParticleSigma();
  dpm.resetParticleLastPositions();
  double tol = forceTolerance0 / dpm.getMeanParticleSigma();
  long it = 0;
  while (dpm.getParticleMaxUnbalancedForce() > tol && it != maxIterations) {
    dpm.particleFIRELoop();
    if (3 * dpm.getParticleMaxDisplacement() > cutoff) {
      dpm.calcParticleNeighborList(cutDistance);
      dpm.resetParticleLastPositions();
    }
    it++;
  }
  std::cout << "FIRE: iteration: " << it << " maxUnbalancedForce: " << std::setprecision(precision)
            << dpm.getParticleMaxUnbalancedForce() << " energy: " << dpm.getParticlePotentialEnergy() << std::endl;
  std::string sp = outDir + "sp/";
  mkdir(sp.c_str(), 0755);
  io.saveParticlePacking(sp);
  dpm.initVerticesOnParticles();
  if (phiTarget <= 0) {  // set-up only: host files, no device
    std::string dp = outDir + "setup/";
    mkdir(dp.c_str(), 0755);
    io.save1DIndexFile(dp + "numVertexInParticleList.dat", dpm.getNumVertexInParticleList());
    io.save1DFile(dp + "radii.dat", dpm.getVertexRadii());
    io.save2DFile(dp + "positions.dat", dpm.getVertexPositions(), 2);
    io.save1DFile(dp + "restAreas.dat", dpm.getRestAreas());
    io.save1DFile(dp + "restLengths.dat", dpm.getRestLengths());
    io.save1DFile(dp + "restAngles.dat", dpm.getRestAngles());
    io.save1DFile(dp + "boxSize.dat", dpm.getBoxSize());
    io.save1DFile(dp + "particleRad.dat", dpm.particleRad);
    return 0;
  }

  // ---- vertex compression
  dpm.calcParticleShape();
  double phi = dpm.getPhi();
  std::cout << "current packing fraction: " << std::setprecision(precision) << phi << std::endl;
  while (true) {
    sigma = dpm.getMeanParticleSize();
    step = dpm.setTimeStep(dt * sigma);
    std::vector<double> vf = {0.2, 0.5, 1.1, 0.99, step, 10 * step, 0.2};
    dpm.initFIRE(vf, minStep, 0, dpm.getNumVertices());
    dpm.calcNeighborList(cutDistance);
    dpm.calcForceEnergy();
    dpm.resetLastPositions();
    dpm.setDisplacementCutoff((1 + cutDistance) * dpm.getMeanVertexRadius(), cutDistance);
    tol = forceTolerance0 / dpm.getMeanParticleSize();
    it = dpm.vertexFIREMinimize(tol, maxIterations);
    std::cout << "FIRE: iteration: " << it << " maxUnbalancedForce: " << std::setprecision(precision)
              << dpm.getMaxUnbalancedForce() << " energy: " << dpm.getPotentialEnergy() << " phi: " << phi << std::endl;
    if (phi + 0.5 * deltaPhi >= phiTarget) break;
    double next = std::min(phi + deltaPhi, phiTarget);
    dpm.scaleVertices(std::sqrt(next / phi));
    phi = dpm.getPhi();
  }
  std::string dp = outDir + "dp/";
  mkdir(dp.c_str(), 0755);
  io.savePacking(dp);
  std::cout << "saved " << dp << " phi " << dpm.getPhi() << std::endl;
  return 0;
}
