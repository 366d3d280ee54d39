// FileIO.h -- ioDPMFile: the packing / state / energy-log text format of include/FileIO.h in
// farceri/cudaRing (deformable rings, periodic box).
//
// Format: one value per line for 1-D files, "x\ty\t" per line for 2-D files, 14 significant
// digits (setprecision(precision), defs.h), params.dat as "key\tvalue" at the stream default
// precision.  Files are read with sscanf("%lf") per line like the reference.  Open failures print
// to stderr and exit(1) as the reference does (FileIO.h:39-53).
#ifndef DPM_HOST_FILEIO_H
#define DPM_HOST_FILEIO_H

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <numeric>
#include <string>
#include <vector>

#include "DPM2D.h"
#include "defs.h"

class ioDPMFile {
 public:
  std::ifstream inputFile;
  std::ofstream outputFile;
  std::ofstream energyFile;
  DPM2D* dpm_ = nullptr;

  ioDPMFile() = default;
  explicit ioDPMFile(DPM2D* dpmPtr) : dpm_(dpmPtr) {}

  // ---- files
  void openInputFile(const std::string& name) {
    inputFile = std::ifstream(name.c_str());
    if (!inputFile.is_open()) {
      std::cerr << "ioDPMFile::openInputFile: error: could not open input file " << name << std::endl;
      std::exit(1);
    }
  }
  void openOutputFile(const std::string& name) {
    outputFile = std::ofstream(name.c_str());
    if (!outputFile.is_open()) {
      std::cerr << "ioDPMFile::openOutputFile: error: could not open output file " << name << std::endl;
      std::exit(1);
    }
  }
  void openEnergyFile(const std::string& name) {
    energyFile = std::ofstream(name.c_str());
    if (!energyFile.is_open()) {
      std::cerr << "ioDPMFile::openEnergyFile: error: could not open energy file " << name << std::endl;
      std::exit(1);
    }
  }
  void closeEnergyFile() { energyFile.close(); }

  // Energy log (FileIO.h:58-70): step+1, time, U/Nv, K/Nv, E/Nv, shape/COM kinetic energy ratio,
  // hexatic order (from the ring contacts), phi.
  void saveEnergy(long step, double timeStep, long numParticles, long numVertices) {
    (void)numParticles;
    double epot = dpm_->getPotentialEnergy() / numVertices;
    double ekin = dpm_->getKineticEnergy() / numVertices;
    energyFile << step + 1 << "\t" << (step + 1) * timeStep << "\t";
    energyFile << std::setprecision(precision) << epot << "\t" << ekin << "\t" << epot + ekin << "\t";
    energyFile << dpm_->getShapeCOMEnergyRatio() << "\t" << dpm_->getParticleHexaticOrder() << "\t";
    energyFile << dpm_->getPhi() << std::endl;
  }

  // the deformable-ring columns of saveEnergy (FileIO.h:71-82), as the run drivers call it
  void saveDeformableEnergy(long step, double timeStep, long numVertices) {
    saveEnergy(step, timeStep, dpm_->getNumParticles(), numVertices);
  }

  // ---- scalar, 1-D and 2-D files
  std::vector<long> read1DIndexFile(const std::string& name, long rows) {
    std::vector<long> data;
    openInputFile(name);
    std::string line;
    for (long r = 0; r < rows; r++) {
      long v = 0;
      std::getline(inputFile, line);
      std::sscanf(line.c_str(), "%ld", &v);
      data.push_back(v);
    }
    inputFile.close();
    return data;
  }
  void save1DIndexFile(const std::string& name, const std::vector<long>& data) {
    openOutputFile(name);
    for (long v : data) outputFile << v << std::endl;
    outputFile.close();
  }
  double read0DFile(const std::string& name) {
    double v = 0;
    openInputFile(name);
    std::string line;
    std::getline(inputFile, line);
    std::sscanf(line.c_str(), "%lf", &v);
    inputFile.close();
    return v;
  }
  void save0DFile(const std::string& name, double v) {
    openOutputFile(name);
    outputFile << std::setprecision(precision) << v << std::endl;
    outputFile.close();
  }
  std::vector<double> read1DFile(const std::string& name, long rows) {
    std::vector<double> data;
    openInputFile(name);
    std::string line;
    for (long r = 0; r < rows; r++) {
      double v = 0;
      std::getline(inputFile, line);
      std::sscanf(line.c_str(), "%lf", &v);
      data.push_back(v);
    }
    inputFile.close();
    return data;
  }
  void save1DFile(const std::string& name, const std::vector<double>& data) {
    openOutputFile(name);
    for (double v : data) outputFile << std::setprecision(precision) << v << std::endl;
    outputFile.close();
  }
  std::vector<double> read2DFile(const std::string& name, long rows) {
    std::vector<double> data;
    openInputFile(name);
    std::string line;
    for (long r = 0; r < rows; r++) {
      double a = 0, b = 0;
      std::getline(inputFile, line);
      std::sscanf(line.c_str(), "%lf %lf", &a, &b);
      data.push_back(a);
      data.push_back(b);
    }
    inputFile.close();
    return data;
  }
  void save2DFile(const std::string& name, const std::vector<double>& data, long cols) {
    openOutputFile(name);
    long rows = (long)data.size() / cols;
    for (long r = 0; r < rows; r++) {
      for (long c = 0; c < cols; c++) outputFile << std::setprecision(precision) << data[r * cols + c] << "\t";
      outputFile << std::endl;
    }
    outputFile.close();
  }

  // integer rows of `cols` entries, tab-separated (save2DIndexFile, FileIO.h:224-238)
  void save2DIndexFile(const std::string& name, const std::vector<long>& data, long cols) {
    openOutputFile(name);
    long rows = cols > 0 ? (long)data.size() / cols : 0;
    for (long r = 0; r < rows; r++) {
      for (long c = 0; c < cols; c++) outputFile << data[r * cols + c] << "\t";
      outputFile << std::endl;
    }
    outputFile.close();
  }

  // neighbour rows and ring contacts (saveNeighbors / saveInitialNeighbors, FileIO.h:531-550)
  void saveNeighbors(const std::string& dir) {
    save2DIndexFile(dir + "neighborList.dat", dpm_->getNeighbors(), dpm_->neighborListSize);
  }
  void saveInitialNeighbors(const std::string& dir, const std::string& file1 = "initialNeighbors.dat",
                            const std::string& file2 = "initialContacts.dat") {
    save2DIndexFile(dir + file1, dpm_->getNeighbors(), dpm_->neighborListSize);
    save2DIndexFile(dir + file2, dpm_->getContacts(), dpm_->contactListSize);
  }
  // ring contact rows (saveContacts, FileIO.h:527-529)
  void saveContacts(const std::string& dir) {
    save2DIndexFile(dir + "contacts.dat", dpm_->getContacts(), dpm_->contactListSize);
  }
  void saveConfiguration(const std::string& dir) {
    savePacking(dir);
    saveNeighbors(dir);
  }

  // driveParams.dat: damping, and taup / f0 of active rings or l0Tau / l0Diff of activeShape
  // rings (saveDriveParams, FileIO.h:557-574)
  void saveDriveParams(const std::string& dir, double damping) {
    std::ofstream p((dir + "driveParams.dat").c_str());
    p << "damping" << "\t" << damping << std::endl;
    if (dpm_->simControl.particleType == simControlStruct::particleEnum::active) {
      double driving, taup;
      dpm_->getSelfPropulsionParams(driving, taup);
      p << "taup" << "\t" << taup << std::endl;
      p << "f0" << "\t" << driving << std::endl;
    } else if (dpm_->simControl.particleType == simControlStruct::particleEnum::activeShape) {
      double l0Tau, l0Diff;
      dpm_->getActiveTensionParams(l0Tau, l0Diff);
      p << "l0Tau" << "\t" << l0Tau << std::endl;
      p << "l0Diff" << "\t" << l0Diff << std::endl;
    }
    p.close();
  }

  // ---- packings (FileIO.h:270-326, 383-435)
  void saveParticlePacking(const std::string& dir) {
    std::ofstream p((dir + "params.dat").c_str());
    p << "numParticles" << "\t" << dpm_->getNumParticles() << std::endl;
    p << "dt" << "\t" << dpm_->dt << std::endl;
    p << "phi" << "\t" << dpm_->getParticlePhi() << std::endl;
    p << "energy" << "\t" << dpm_->getParticlePotentialEnergy() / dpm_->getNumParticles() << std::endl;
    p.close();
    save1DFile(dir + "boxSize.dat", dpm_->getBoxSize());
    save2DFile(dir + "particlePos.dat", dpm_->particlePos, 2);
    save1DFile(dir + "particleRad.dat", dpm_->particleRad);
  }

  void savePacking(const std::string& dir) {
    std::ofstream p((dir + "params.dat").c_str());
    p << "numParticles" << "\t" << dpm_->getNumParticles() << std::endl;
    p << "ea" << "\t" << dpm_->ea << std::endl;
    p << "el" << "\t" << dpm_->el << std::endl;
    p << "eb" << "\t" << dpm_->eb << std::endl;
    p << "ec" << "\t" << dpm_->ec << std::endl;
    p << "dt" << "\t" << dpm_->dt << std::endl;
    p << "phi" << "\t" << dpm_->getPhi() << std::endl;
    p << "phi0" << "\t" << dpm_->getPreferredPhi() << std::endl;
    p << "epot" << "\t" << dpm_->getPotentialEnergy() / dpm_->getNumParticles() << std::endl;
    p << "temperature" << "\t" << dpm_->getTemperature() << std::endl;
    p.close();
    save1DIndexFile(dir + "numVertexInParticleList.dat", dpm_->getNumVertexInParticleList());
    save1DFile(dir + "radii.dat", dpm_->getVertexRadii());
    save2DFile(dir + "positions.dat", dpm_->getVertexPositions(), 2);
    save2DFile(dir + "forces.dat", dpm_->getInteractionForces(), 2);
    save1DFile(dir + "restAreas.dat", dpm_->getRestAreas());
    save1DFile(dir + "restLengths.dat", dpm_->getRestLengths());
    save1DFile(dir + "restAngles.dat", dpm_->getRestAngles());
    save2DFile(dir + "velocities.dat", dpm_->getVertexVelocities(), 2);
    save1DFile(dir + "particleRad.dat", dpm_->particleRad);
    save2DFile(dir + "shapes.dat", dpm_->getParticleShapes(), 3);
    if (dpm_->simControl.geometryType == simControlStruct::geometryEnum::roundBox)
      save0DFile(dir + "boxSize.dat", dpm_->getBoxRadius());  // FileIO.h:389-396
    else
      save1DFile(dir + "boxSize.dat", dpm_->getBoxSize());
  }

  void readPackingFromDirectory(const std::string& dir, long numParticles_, long nDim_) {
    std::vector<long> n = read1DIndexFile(dir + "numVertexInParticleList.dat", numParticles_);
    dpm_->setNumVertexInParticleList(n);
    long nv = std::accumulate(n.begin(), n.end(), 0L);
    std::cout << "readFromDirectory:: numVertices: " << nv << " on device: " << dpm_->getNumVertices() << std::endl;
    std::vector<double> pos = read2DFile(dir + "positions.dat", nv);
    dpm_->setVertexPositions(pos);
    std::vector<double> rad = read1DFile(dir + "radii.dat", nv);
    dpm_->setVertexRadii(rad);
    std::vector<double> a0 = read1DFile(dir + "restAreas.dat", numParticles_);
    dpm_->setRestAreas(a0);
    std::vector<double> l0 = read1DFile(dir + "restLengths.dat", nv);
    dpm_->setRestLengths(l0);
    std::vector<double> th0 = read1DFile(dir + "restAngles.dat", nv);
    dpm_->setRestAngles(th0);
    dpm_->setLengthScale();
    if (dpm_->simControl.geometryType == simControlStruct::geometryEnum::roundBox) {
      dpm_->setBoxRadius(read0DFile(dir + "boxSize.dat"));  // the round box's radius (FileIO.h:311-316)
      std::cout << "FileIO::readPackingFromDirectory: preferred phi: " << dpm_->getPreferredPhi()
                << " boxRadius: " << dpm_->getBoxRadius() << std::endl;
    } else {
      std::vector<double> box = read1DFile(dir + "boxSize.dat", nDim_);
      dpm_->setBoxSize(box);
      std::cout << "FileIO::readPackingFromDirectory: preferred phi: " << dpm_->getPreferredPhi()
                << " box-Lx: " << box[0] << ", Ly: " << box[1] << std::endl;
    }
    dpm_->calcParticleShape();
  }

  // ---- states (FileIO.h:460-526)
  void readState(const std::string& dir, long numParticles_, long numVertices_, long nDim_) {
    (void)numParticles_;
    (void)nDim_;
    std::vector<double> vel = read2DFile(dir + "velocities.dat", numVertices_);
    dpm_->setVertexVelocities(vel);
  }
  void readDeformableActiveState(const std::string& dir, long numParticles_, long numVertices_, long nDim_) {
    readState(dir, numParticles_, numVertices_, nDim_);
    std::vector<double> ang = read1DFile(dir + "particleAngles.dat", numParticles_);
    dpm_->setParticleAngles(ang);
  }
  void saveState(const std::string& dir) {
    save2DFile(dir + "positions.dat", dpm_->getVertexPositions(), 2);
    save2DFile(dir + "velocities.dat", dpm_->getVertexVelocities(), 2);
    save2DFile(dir + "shapes.dat", dpm_->getParticleShapes(), 3);
    save2DFile(dir + "particlePos.dat", dpm_->getParticlePositions(), 2);
  }
  void saveDeformableActiveState(const std::string& dir) {
    save2DFile(dir + "positions.dat", dpm_->getVertexPositions(), 2);
    save2DFile(dir + "velocities.dat", dpm_->getVertexVelocities(), 2);
    save1DFile(dir + "particleAngles.dat", dpm_->getParticleAngles());
  }

  // ---- rigid rings: packing read (FileIO.h:328-381), packing save (437-458), state save (517-525)
  void readRigidPackingFromDirectory(const std::string& dir, long numParticles_, long nDim_) {
    std::vector<long> n = read1DIndexFile(dir + "numVertexInParticleList.dat", numParticles_);
    dpm_->setNumVertexInParticleList(n);
    long nv = std::accumulate(n.begin(), n.end(), 0L);
    std::cout << "readRigidPackingFromDirectory:: numVertices: " << nv << " on device: " << dpm_->getNumVertices()
              << std::endl;
    std::vector<double> pos = read2DFile(dir + "positions.dat", nv);
    dpm_->setVertexPositions(pos);
    std::vector<double> rad = read1DFile(dir + "radii.dat", nv);
    dpm_->setVertexRadii(rad);
    std::vector<double> a0 = read1DFile(dir + "restAreas.dat", numParticles_);
    dpm_->setRestAreas(a0);
    std::vector<double> prad = read1DFile(dir + "particleRad.dat", numParticles_);
    dpm_->setParticleRadii(prad);
    std::vector<double> ppos = read2DFile(dir + "particlePos.dat", numParticles_);
    dpm_->setParticlePositions(ppos);
    std::vector<double> ang = read1DFile(dir + "particleAngles.dat", numParticles_);
    dpm_->setParticleAngles(ang);
    dpm_->setLengthScaleToOne();
    if (dpm_->simControl.geometryType == simControlStruct::geometryEnum::roundBox) {
      dpm_->setBoxRadius(read0DFile(dir + "boxSize.dat"));
    } else {
      std::vector<double> box = read1DFile(dir + "boxSize.dat", nDim_);
      dpm_->setBoxSize(box);
    }
  }
  void saveRigidPacking(const std::string& dir) {
    std::ofstream p((dir + "params.dat").c_str());
    p << "numParticles" << "\t" << dpm_->getNumParticles() << std::endl;
    p << "phi" << "\t" << dpm_->getPhi() << std::endl;
    p << "epot" << "\t" << dpm_->getPotentialEnergy() << std::endl;
    p << "temperature" << "\t" << dpm_->getTemperature() << std::endl;
    p.close();
    save1DIndexFile(dir + "numVertexInParticleList.dat", dpm_->getNumVertexInParticleList());
    save1DFile(dir + "radii.dat", dpm_->getVertexRadii());
    save2DFile(dir + "forces.dat", dpm_->getVertexForces(), 2);
    save2DFile(dir + "positions.dat", dpm_->getVertexPositions(), 2);
    save1DFile(dir + "restAreas.dat", dpm_->getRestAreas());
    save1DFile(dir + "particleRad.dat", dpm_->getParticleRadii());
    save2DFile(dir + "particlePos.dat", dpm_->getParticlePositions(), 2);
    save2DFile(dir + "particleVel.dat", dpm_->getParticleVelocities(), 2);
    save1DFile(dir + "particleAngles.dat", dpm_->getParticleAngles());
    save1DFile(dir + "particleAngvel.dat", dpm_->getParticleAngularVelocities());
  }
  void saveRigidState(const std::string& dir) {
    save2DFile(dir + "positions.dat", dpm_->getVertexPositions(), 2);
    save2DFile(dir + "velocities.dat", dpm_->getVertexVelocities(), 2);
    save2DFile(dir + "particlePos.dat", dpm_->getParticlePositions(), 2);
    save2DFile(dir + "particleVel.dat", dpm_->getParticleVelocities(), 2);
    save1DFile(dir + "particleAngles.dat", dpm_->getParticleAngles());
    save1DFile(dir + "particleAngvel.dat", dpm_->getParticleAngularVelocities());
  }

  // ---- cell format (FileIO.h:611-704): NUMCL / PACKF / BOXSZ / STRSS header, then per ring
  // "CINFO n . . a0 area p0" and per vertex "VINFO . . x y rad l0 theta0 fx fy"
  void readPackingFromCellFormat(const std::string& name, long skiplines) {
    openInputFile(name);
    std::string line;
    for (long l = 0; l < skiplines; l++) std::getline(inputFile, line);
    long np = 0;
    double phi = 0, lx = 0, ly = 0, s[3];
    std::getline(inputFile, line);
    std::sscanf(line.c_str(), "NUMCL %ld", &np);
    if (np < 1) {
      std::cerr << "FileIO::readFromCellFormat: error: numParticles = " << np << ". Ending here." << std::endl;
      std::exit(1);
    }
    std::getline(inputFile, line);
    std::sscanf(line.c_str(), "PACKF %lf", &phi);
    std::getline(inputFile, line);
    std::sscanf(line.c_str(), "BOXSZ %lf %lf", &lx, &ly);
    std::getline(inputFile, line);
    std::sscanf(line.c_str(), "STRSS %lf %lf %lf", &s[0], &s[1], &s[2]);
    std::vector<long> n;
    std::vector<double> pos, rad, a0, l0, th0, box = {lx, ly};
    for (long p = 0; p < np; p++) {
      long k = 0;
      double a = 0, area = 0, p0 = 0;
      std::getline(inputFile, line);
      std::sscanf(line.c_str(), "CINFO %ld %*d %*d %lf %lf %lf", &k, &a, &area, &p0);
      n.push_back(k);
      a0.push_back(a);
      for (long i = 0; i < k; i++) {
        double x = 0, y = 0, r = 0, l = 0, t = 0, fx = 0, fy = 0;
        std::getline(inputFile, line);
        std::sscanf(line.c_str(), "VINFO %*d %*d %lf %lf %lf %lf %lf %lf %lf", &x, &y, &r, &l, &t, &fx, &fy);
        pos.push_back(x);
        pos.push_back(y);
        rad.push_back(r);
        l0.push_back(l);
        th0.push_back(t);
      }
    }
    inputFile.close();
    dpm_->setNumVertexInParticleList(n);
    dpm_->setBoxSize(box);
    dpm_->setVertexPositions(pos);
    dpm_->setVertexRadii(rad);
    dpm_->setRestAreas(a0);
    dpm_->setRestLengths(l0);
    dpm_->setRestAngles(th0);
    dpm_->setLengthScale();
    std::cout << "FileIO::readFromCellFormat: preferred phi: " << dpm_->getPreferredPhi() << std::endl;
  }
};

#endif
