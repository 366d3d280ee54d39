// Simulator.cpp -- integrators over libdpmring.so (Simulator.cu:23-89, 206-238 of the reference).
#include "Simulator.h"

#include "DPM2D.h"

// per-ring Gaussian velocities of amplitude sqrt(T) (Simulator.cu:32-52); velocities only
void Langevin::injectKineticEnergy() {
  dpm_->check(dpm_inject_kinetic_energy(dpm_->handle(), config.Tinject), "injectKineticEnergy");
}
// kernelUpdateVertexPos / kernelUpdateVertexVel (Simulator.cu:85-95)
void Langevin::updatePosition(double timeStep) {
  dpm_->check(dpm_update_position(dpm_->handle(), timeStep), "Langevin::updatePosition");
}
void Langevin::updateVelocity(double timeStep) {
  dpm_->check(dpm_update_velocity(dpm_->handle(), timeStep), "Langevin::updateVelocity");
}
// F += noise xi - gamma v, one draw per ring (Simulator.cu:54-77)
void Langevin::updateThermalVel() {
  dpm_->check(dpm_update_thermal_vel(dpm_->handle()), "Langevin::updateThermalVel");
}
// the mean vertex velocity removed (Simulator.cu:97-110)
void Langevin::conserveMomentum() {
  dpm_->check(dpm_conserve_momentum(dpm_->handle()), "Langevin::conserveMomentum");
}
// the stages above in Langevin::integrate's order, fused into one device step
void Langevin::integrate() { dpm_->check(dpm_langevin_loop(dpm_->handle(), 1), "Langevin::integrate"); }

void NVE::integrate() { dpm_->check(dpm_nve_loop(dpm_->handle(), 1), "NVE::integrate"); }

// velocities rescaled to the temperature Tinject (Simulator.cu:224-238)
void NVERescale::injectKineticEnergy() {
  dpm_->check(dpm_adjust_temperature(dpm_->handle(), config.Tinject), "NVERescale::injectKineticEnergy");
}
void NVERescale::integrate() { dpm_->check(dpm_nve_rescale_loop(dpm_->handle(), 1), "NVERescale::integrate"); }
