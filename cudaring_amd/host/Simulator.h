// Simulator.h -- integrator classes of a DPM2D (SimConfig / SimInterface / Langevin / NVE /
// NVERescale of include/Simulator.h:16-113 in farceri/cudaRing).
//
// One integrate() is one fused device step (kick, drift, neighbour check, forces, thermostat,
// kick) of libdpmring.so.  The reference's separate stages (updatePosition, updateVelocity,
// updateThermalVel, conserveMomentum, injectKineticEnergy) are also callable one by one, each a
// device call of its own: in Langevin::integrate's order (with DPM2D::checkNeighbors and
// calcForceEnergy between) they reproduce the fused step.  integrate() itself does not call the
// virtual stages, so a subclass that overrides one changes only stage-wise use.
#ifndef DPM_HOST_SIMULATOR_H
#define DPM_HOST_SIMULATOR_H

#include "defs.h"

class DPM2D;

class SimConfig {
 public:
  double Tinject = 0;
  SimConfig() = default;
  SimConfig(double Tin) : Tinject(Tin) {}
};

class SimInterface {
 public:
  DPM2D* dpm_ = nullptr;
  SimConfig config;
  double noise = 0;
  double gamma = 1;
  long firstIndex = 10;

  SimInterface() = default;
  SimInterface(DPM2D* dpmPtr, SimConfig config) : dpm_(dpmPtr), config(config) {}
  virtual ~SimInterface() = default;

  virtual void injectKineticEnergy() = 0;
  virtual void updatePosition(double timeStep) = 0;
  virtual void updateVelocity(double timeStep) = 0;
  virtual void updateThermalVel() = 0;
  virtual void conserveMomentum() = 0;
  virtual void integrate() = 0;
};

class Langevin : public SimInterface {
 public:
  Langevin() = default;
  Langevin(DPM2D* dpmPtr, SimConfig config) : SimInterface(dpmPtr, config) {}
  void injectKineticEnergy() override;
  void updatePosition(double timeStep) override;
  void updateVelocity(double timeStep) override;
  void updateThermalVel() override;
  void conserveMomentum() override;
  void integrate() override;
};

class NVE : public Langevin {
 public:
  NVE() = default;
  NVE(DPM2D* dpmPtr, SimConfig config) : Langevin(dpmPtr, config) {}
  void integrate() override;
};

class NVERescale : public Langevin {
 public:
  NVERescale() = default;
  NVERescale(DPM2D* dpmPtr, SimConfig config) : Langevin(dpmPtr, config) {}
  void injectKineticEnergy() override;
  void integrate() override;
};

#endif
