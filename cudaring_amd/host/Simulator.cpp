// Simulator.cpp -- integrators over libdpmring.so (Simulator.cu:23-89, 206-238 of the reference).
#include "Simulator.h"

#include <stdexcept>

#include "DPM2D.h"

static void fused(const char* what) {
  throw std::logic_error(std::string(what) +
                         ": this stage is fused into integrate() (one device step) and is not callable alone");
}

// per-ring Gaussian velocities of amplitude sqrt(T) (Simulator.cu:32-52)
void Langevin::injectKineticEnergy() {
  dpm_->check(dpm_init_langevin(dpm_->handle(), config.Tinject, gamma, 0, 0), "injectKineticEnergy");
}
void Langevin::updatePosition(double) { fused("Langevin::updatePosition"); }
void Langevin::updateVelocity(double) { fused("Langevin::updateVelocity"); }
void Langevin::updateThermalVel() { fused("Langevin::updateThermalVel"); }
void Langevin::conserveMomentum() { fused("Langevin::conserveMomentum"); }
void Langevin::integrate() { dpm_->check(dpm_langevin_loop(dpm_->handle(), 1), "Langevin::integrate"); }

void NVE::injectKineticEnergy() {
  dpm_->check(dpm_init_nve(dpm_->handle(), config.Tinject, 0, 0), "injectKineticEnergy");
}
void NVE::integrate() { dpm_->check(dpm_nve_loop(dpm_->handle(), 1), "NVE::integrate"); }

void NVERescale::injectKineticEnergy() {
  dpm_->check(dpm_init_nve_rescale(dpm_->handle(), config.Tinject), "injectKineticEnergy");
}
void NVERescale::integrate() { dpm_->check(dpm_nve_rescale_loop(dpm_->handle(), 1), "NVERescale::integrate"); }
