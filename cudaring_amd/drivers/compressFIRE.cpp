// compressFIRE -- make a deformable-ring packing and compress it with FIRE
// (the flow of compressFIRE.cpp in farceri/cudaRing: polydisperse sizes, random soft disks,
// soft-disk FIRE, vertices on the disks, then vertex FIRE at growing packing fraction).
//
// usage: compressFIRE outDir [seed=] [numParticles=256] [numVertexPerParticle=32] [phi0=0.2]
//                     [maxSearchStep=500] [maxStep=10000] [maxIter=5000000] [costs=1e3,1,0,1]
//   the reference's command line (compressFIRE.cpp:24-37: the output directory, its constants as
//   defaults) and its jamming search (compressFIRE.cpp:95-193), see reference_flow below;
// usage: compressFIRE outDir numParticles numVertexPerParticle calA0 polydispersity phiTarget seed
//                     [deltaPhi=2e-3] [phi0=0.2] [costs=ea,el,eb,ec] [maxIter=5000000] [bi=nSmall,nLarge]
//   compression to a target packing fraction; phiTarget <= 0 stops after the set-up (no device is
//   used) and writes the set-up files.  Each compression step: scaleVertices(sqrt((phi + deltaPhi)
//   / phi)), then FIRE until the largest force component is below 1e-10 / (mean particle size),
//   as compressFIRE.cpp:104-131.  The default costs are compressFIRE.cpp:30's (1e3, 1, 0, 1).
#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "../host/DPM2D.h"
#include "../host/FIRE.h"
#include "../host/FileIO.h"

// Soft-disk set-up of compressFIRE.cpp:51-86: particle FIRE until the largest particle force is
// below 1e-10 / (mean particle sigma), rebuilding the particle list when a disk has moved more than
// cutoff / 3; returns the iterations.
static long soft_disk_fire(DPM2D& dpm, long numParticles, long minStep, double dt, double cutDistance,
                           double forceTolerance0, long maxIterations) {
  double sigma = dpm.getMeanParticleSize();
  double step = dpm.setTimeStep(dt * 10 * sigma);
  std::cout << "FIRE: timeStep: " << step << std::endl;
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
  return it;
}

// The reference driver's flow with its constants (compressFIRE.cpp:26-37): 256 rings of 32
// vertices, calA0 1.1, polydispersity 0.2, soft disks at phi 0.2, costs (1e3, 1, 0, 1); then up
// to maxSearchStep steps of [vertex FIRE to 1e-10 / sigma (at most maxIter iterations); save the
// packing in outDir/<phi>/; under-jammed (force below the tolerance): keep a copy and grow by
// deltaPhi; over-jammed (FIRE stopped above twice the tolerance): halve deltaPhi and grow the kept
// copy by it; maxStep Langevin steps at T 1e-2, damping sqrt(10) / sigma; scaleVertices].  A
// jammed packing (force between one and two tolerances) ends the search: the reference prints it
// and loops on the same packing without a further step (compressFIRE.cpp:142-145).
//
// With those costs the vertex FIRE is unstable once FIRE has grown its step to dt_max = 0.1 sigma:
// the ring-area mode has curvature ea * 4 pi calA0 / (n A0), so lambda * dt_max^2 =
// 0.176 ea / n (5.5 for ea 1e3 and n 32, above the leapfrog limit of 4) whatever the box size, and
// the zig-zag modes of the perimeter springs and of the overlapping neighbour vertices add
// 4 (el + ec) n^2 / (100 pi^2 calA0) each (3.8 at n 32).  The oracle's FIRE diverges the same way
// from the same start, so the costs can be given (costs=ea,el,eb,ec) for a stable search.
static int reference_flow(const std::string& outDir, unsigned long seed, long numParticles, long nvp, double phi0,
                          long maxSearchStep, long maxStep, long maxIterations, const double* costs) {
  const long minStep = 20, printFreq = 100000;
  const double polydispersity = 0.2, phiTh = 0.2, calA0 = 1.1, cutDistance = 1;
  const double forceTolerance0 = 1e-10, iod = 10, dt = 1e-2, Tinject = 1e-2;
  const double ea = costs[0], el = costs[1], eb = costs[2], ec = costs[3];
  double deltaPhi = 2e-3;
  mkdir(outDir.c_str(), 0755);
  DPM2D dpm(numParticles, 2, nvp);
  dpm.setSeed(seed);
  dpm.setEnergyCosts(ea, el, eb, ec);
  ioDPMFile io(&dpm);
  dpm.setPolySizeDistribution(calA0, polydispersity);
  dpm.setRandomParticles(phi0, 1.1);
  soft_disk_fire(dpm, numParticles, minStep, dt, cutDistance, forceTolerance0, maxIterations);
  std::string dir = outDir + "sp/";
  mkdir(dir.c_str(), 0755);
  io.saveParticlePacking(dir);
  dpm.initVerticesOnParticles();
  // forces.dat holds the interaction forces, which need the vertex list (the reference writes its
  // force buffer, never filled at this point); the first search step rebuilds it anyway
  dpm.calcNeighborList(cutDistance);
  dir = outDir + "dp/";
  mkdir(dir.c_str(), 0755);
  io.savePacking(dir);
  dpm.calcParticleShape();
  std::cout << "current packing fraction: " << std::setprecision(precision) << dpm.getPhi() << std::endl;
  const long numVertices = dpm.getNumVertices();
  std::vector<double> positions, radii, lengths, areas;
  double currentPhi = dpm.getPhi(), previousPhi = currentPhi, scaleFactor = 1;
  for (long searchStep = 0; searchStep < maxSearchStep;) {
    double sigma = dpm.getMeanParticleSize();
    double step = dpm.setTimeStep(dt * sigma);
    std::cout << "\nFIRE: timeStep: " << step;
    std::vector<double> vf = {0.2, 0.5, 1.1, 0.99, step, 10 * step, 0.2};
    dpm.initFIRE(vf, minStep, 0, numVertices);
    dpm.calcNeighborList(cutDistance);
    dpm.calcForceEnergy();
    dpm.resetLastPositions();
    dpm.setDisplacementCutoff((1 + cutDistance) * dpm.getMeanVertexRadius(), cutDistance);
    const double tol = forceTolerance0 / dpm.getMeanParticleSize();
    std::cout << " force tollerance: " << tol << std::endl;
    // the reference's loop checks the force after every iteration and prints every printFreq
    long it = 0;
    double forceCheck = dpm.getMaxUnbalancedForce();
    while (forceCheck > tol && it != maxIterations) {
      const long chunk = std::min(printFreq - it % printFreq, maxIterations - it);
      it += dpm.vertexFIREMinimize(tol, chunk);
      forceCheck = dpm.getMaxUnbalancedForce();
      if (it % printFreq == 0 && it != 0 && forceCheck > tol)
        std::cout << "FIRE: iteration: " << it << " maxUnbalancedForce: " << std::setprecision(precision)
                  << forceCheck << " energy: " << dpm.getPotentialEnergy() << std::endl;
    }
    std::cout << "FIRE: iteration: " << it << " maxUnbalancedForce: " << std::setprecision(precision) << forceCheck
              << " energy: " << dpm.getPotentialEnergy() << std::endl;
    dir = outDir + std::to_string(dpm.getPhi()) + "/";
    mkdir(dir.c_str(), 0755);
    io.savePacking(dir);
    const bool jam = forceCheck < 2.0 * tol && forceCheck > tol;
    const bool overJam = forceCheck > 2.0 * tol;
    const bool underJam = forceCheck < tol;
    if (jam) {
      std::cout << "Compression step: " << searchStep << " Found jammed configuration, packing fraction: "
                << currentPhi << std::endl;
      break;
    }
    if (underJam) {
      scaleFactor = std::sqrt((currentPhi + deltaPhi) / currentPhi);
      std::cout << "Compression step: " << searchStep << " Found underjammed configuration, scaleFactor: "
                << scaleFactor << std::endl;
      previousPhi = currentPhi;
      positions = dpm.getVertexPositions();
      radii = dpm.getVertexRadii();
      lengths = dpm.getRestLengths();
      areas = dpm.getRestAreas();
    } else if (overJam) {
      deltaPhi *= 0.5;
      scaleFactor = std::sqrt((previousPhi + deltaPhi) / previousPhi);
      std::cout << "Compression step: " << searchStep << " Found overjammed configuration, scaleFactor: "
                << scaleFactor << std::endl;
      if (!positions.empty()) {  // none kept yet when the first step is over-jammed: compress it anyway
        dpm.setVertexPositions(positions);
        dpm.setVertexRadii(radii);
        dpm.setRestLengths(lengths);
        dpm.setRestAreas(areas);
      }
    }
    if (currentPhi > phiTh) {
      sigma = dpm.getMeanParticleSize();
      const double damping = std::sqrt(iod) / sigma;
      const double timeStep = dpm.setTimeStep(dt / damping);
      std::cout << "Thermalization, timeStep: " << timeStep << std::endl;
      dpm.initLangevin(Tinject, damping, true);
      const long u0 = dpm.getUpdateCount();
      dpm.langevinLoop(maxStep);  // the list is rebuilt on the device at each skin crossing
      std::cout << "NVT: current step: 0 E/N: " << (dpm.getPotentialEnergy() + dpm.getKineticEnergy()) / numVertices
                << " T: " << dpm.getTemperature() << " phi: " << dpm.getPhi()
                << " updates: " << dpm.getUpdateCount() - u0 << std::endl;
    }
    dpm.scaleVertices(scaleFactor);
    currentPhi = dpm.getPhi();
    std::cout << "new packing fraction: " << currentPhi << std::endl;
    searchStep++;
  }
  return 0;
}

int main(int argc, char** argv) {
  if (argc >= 2 && (argc == 2 || std::string(argv[2]).find('=') != std::string::npos)) {
    std::string outDir = argv[1];
    if (outDir.back() != '/') outDir += "/";
    unsigned long seed = 0;
    long numParticles = 256, nvp = 32, maxSearchStep = 500, maxStep = 10000, maxIterations = 5000000;
    double phi0 = 0.2, costs[4] = {1e3, 1, 0, 1};
    for (int i = 2; i < argc; i++) {
      std::string a = argv[i];
      const size_t eq = a.find('=');
      const std::string key = a.substr(0, eq), val = eq == std::string::npos ? "" : a.substr(eq + 1);
      if (key == "seed") seed = std::strtoul(val.c_str(), nullptr, 10);
      else if (key == "numParticles") numParticles = std::atol(val.c_str());
      else if (key == "numVertexPerParticle") nvp = std::atol(val.c_str());
      else if (key == "phi0") phi0 = std::atof(val.c_str());
      else if (key == "maxSearchStep") maxSearchStep = std::atol(val.c_str());
      else if (key == "maxStep") maxStep = std::atol(val.c_str());
      else if (key == "maxIter") maxIterations = std::atol(val.c_str());
      else if (key == "costs" &&
               std::sscanf(val.c_str(), "%lf,%lf,%lf,%lf", &costs[0], &costs[1], &costs[2], &costs[3]) == 4) {
      } else {
        std::cerr << "compressFIRE: unknown option " << a << std::endl;
        return 2;
      }
    }
    return reference_flow(outDir, seed, numParticles, nvp, phi0, maxSearchStep, maxStep, maxIterations, costs);
  }
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
  soft_disk_fire(dpm, numParticles, minStep, dt, cutDistance, forceTolerance0, maxIterations);
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
    const double sigma = dpm.getMeanParticleSize();
    const double step = dpm.setTimeStep(dt * sigma);
    std::vector<double> vf = {0.2, 0.5, 1.1, 0.99, step, 10 * step, 0.2};
    dpm.initFIRE(vf, minStep, 0, dpm.getNumVertices());
    dpm.calcNeighborList(cutDistance);
    dpm.calcForceEnergy();
    dpm.resetLastPositions();
    dpm.setDisplacementCutoff((1 + cutDistance) * dpm.getMeanVertexRadius(), cutDistance);
    const double tol = forceTolerance0 / dpm.getMeanParticleSize();
    const long it = dpm.vertexFIREMinimize(tol, maxIterations);
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
