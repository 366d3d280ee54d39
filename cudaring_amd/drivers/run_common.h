// run_common.h -- the directory plan and the step loop shared by the runNVE, runNVT, runActive and
// runActiveShape drivers, whose command lines and output layout are those of runNVE.cpp,
// runNVT.cpp, runActive.cpp and runActiveShape.cpp in farceri/cudaRing:
//   * the mode argument: "change" reads the packing and its state and writes a new directory
//     beside it, "over" runs again in the sample directory, "run" / "runlog" write the trajectory
//     to its dynamics/ or dynamics-log/ subdirectory, anything else starts a new sample directory;
//   * the energy log every maxStep/1000 steps, a progress line and the packing every maxStep/10,
//     state snapshots t<step>/ every maxStep/100 steps (log-spaced within each tenth for runlog).
// Steps between two of these events run as one device loop call.  The reference divides by
// these intervals, so a maxStep below 1000 makes it divide by zero; here they are at least 1.
#ifndef DPM_DRIVERS_RUN_COMMON_H
#define DPM_DRIVERS_RUN_COMMON_H

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <filesystem>
#include <functional>
#include <iostream>
#include <string>

#include "../host/DPM2D.h"
#include "../host/FileIO.h"

struct RunPlan {
  std::string inDir, outDir;
  bool readState = true, makeNewDir = false, logSave = false, linSave = true;
};

// The directory logic of the run drivers (runNVT.cpp:43-115): `up` is the path from the input
// directory to the parent of a "change" run's new sample directory ("../" or "../../").
inline RunPlan plan_run(std::string inDir, const std::string& mode, const std::string& whichDynamics,
                        const std::string& dirSample, const std::string& up, bool readStateByDefault) {
  namespace fs = std::filesystem;
  RunPlan p;
  p.readState = readStateByDefault;
  bool sameDir = false, dynamics = false;
  if (mode == "change") {
    p.makeNewDir = true;
    std::cout << "Change mode: make new directory and run dynamics with different parameters" << std::endl;
  } else if (mode == "over") {
    sameDir = true;
    std::cout << "Over mode: run dynamics with same parameters and save in same directory" << std::endl;
  } else if (mode == "run") {
    sameDir = dynamics = true;
    std::cout << "Run mode: make DYNAMICS directory and run dynamics with same parameters" << std::endl;
  } else if (mode == "runlog") {
    sameDir = dynamics = p.logSave = true;
    p.linSave = false;
    std::cout << "Run Log mode: make DYNAMICS-LOG directory, run dynamics with same parameters and save "
                 "log-spaced trajectories" << std::endl;
  } else {
    std::cout << "Default mode: make new directory path and run initial dynamics" << std::endl;
  }
  if (sameDir) {
    p.readState = true;
    inDir += dirSample;
    p.outDir = inDir;
    if (dynamics) {
      p.outDir += p.logSave ? "dynamics-log/" : "dynamics/";
      if (fs::exists(p.outDir)) inDir = p.outDir;
      else fs::create_directory(p.outDir);
    }
  } else {
    if (p.makeNewDir) {
      p.readState = true;
      p.outDir = inDir + up + dirSample;
    } else {
      if (!fs::exists(inDir + whichDynamics)) fs::create_directory(inDir + whichDynamics);
      p.outDir = inDir + dirSample;
    }
    fs::create_directory(p.outDir);
  }
  p.inDir = inDir;
  std::cout << "inDir: " << p.inDir << std::endl << "outDir: " << p.outDir << std::endl;
  return p;
}

// The loop of the run drivers (runNVT.cpp:148-196): after the step with index `step` (from 0),
// the energy log when step % saveEnergyFreq == 0, then the progress line and the packing when
// step % checkPointFreq == 0; the state snapshots by the linear or the log schedule.
// `advance(n)` runs n steps, `onEnergy` runs after each energy-log line, `progress` prints the
// quantities of the driver's progress line.
struct RunLoop {
  long maxStep, initialStep;
  double timeStep;
  long numVertices;
  std::string label;
  bool saveContacts = false;
  std::function<void(long)> advance;
  std::function<void()> onEnergy = [] {};
  std::function<void()> progress = [] {};

  void run(DPM2D& dpm, ioDPMFile& io, const RunPlan& p) {
    namespace fs = std::filesystem;
    const long checkPointFreq = std::max(1L, maxStep / 10), linFreq = std::max(1L, checkPointFreq / 10);
    const long saveEnergyFreq = std::max(1L, linFreq / 10);
    long multiple = 1, saveFreq = 1, done = 0;
    auto t0 = std::chrono::steady_clock::now();
    io.savePacking(p.outDir);
    io.saveInitialNeighbors(p.outDir);
    for (long step = 0; step < maxStep; step++) {
      const bool energy = step % saveEnergyFreq == 0;
      bool snapshot = p.linSave && step % linFreq == 0;
      if (p.logSave) {  // runNVT.cpp:171-183, evaluated for every step as there
        if (step > multiple * checkPointFreq) {
          saveFreq = 1;
          multiple += 1;
        }
        if ((step - (multiple - 1) * checkPointFreq) > saveFreq * 10) saveFreq *= 10;
        if (((step - (multiple - 1) * checkPointFreq) % saveFreq) == 0) snapshot = true;
      }
      if (!energy && !snapshot) continue;
      advance(step + 1 - done);  // every step up to this one
      done = step + 1;
      if (energy) {
        io.saveDeformableEnergy(step, timeStep, numVertices);
        if (step % checkPointFreq == 0) {
          std::cout << label << ": current step: " << step << " E/N: " << dpm.getEnergy() / numVertices
                    << " T: " << dpm.getTemperature();
          progress();
          std::cout << " phi: " << dpm.getPhi();
          const long updates = dpm.getUpdateCount();
          if (step != 0 && updates > 0)
            std::cout << " number of updates: " << updates << " frequency " << checkPointFreq / updates << std::endl;
          else
            std::cout << " no updates" << std::endl;
          dpm.resetUpdateCount();
          io.savePacking(p.outDir);
        }
        onEnergy();
      }
      if (snapshot) {
        const std::string dir = p.outDir + "/t" + std::to_string(initialStep + step) + "/";
        fs::create_directory(dir);
        io.saveState(dir);
        if (saveContacts) io.saveContacts(dir);
      }
    }
    if (done < maxStep) advance(maxStep - done);
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::printf("Elapsed time: %f ms, %f s.\n", ms, ms / 1000);
    io.savePacking(p.outDir);
  }
};

#endif
