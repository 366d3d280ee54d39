// FIRE.cpp -- FIRE over libdpmring.so (vertices) and on the host (soft-disk set-up).
#include "FIRE.h"

#include <algorithm>
#include <cmath>
#include <iostream>

#include "DPM2D.h"

void FIRE::initMinimizer(double a_start_, double f_dec_, double f_inc_, double f_a_, double fire_dt_,
                         double fire_dt_max_, double a_, long minStep_, long numStep_, long numDOF_) {
  (void)numDOF_;  // sizes scratch vectors in the reference; the library sizes its own
  a_start = a_start_;
  f_dec = f_dec_;
  f_inc = f_inc_;
  f_a = f_a_;
  fire_dt = fire_dt_;
  fire_dt_max = fire_dt_max_;
  a = a_;
  minStep = minStep_;
  numStep = numStep_;
  deviceReady = false;  // pushed to the device at the first vertex iteration
}

// ---------------------------------------------------------------- particle minimiser
void FIRE::updateParticlePositionAndVelocity() {
  // FIRE.cu:44-71: half kick with mass pi / r^2, drift; zero velocity if the force sums to 0
  std::vector<double>& pos = dpm_->particlePos;
  std::vector<double>& vel = dpm_->particleVel;
  const std::vector<double>& f = dpm_->particleForce;
  for (long p = 0; p < dpm_->numParticles; p++) {
    double total = 0;
    for (long d = 0; d < 2; d++) {
      vel[2 * p + d] += 0.5 * fire_dt * f[2 * p + d] / d_mass[2 * p + d];
      pos[2 * p + d] += fire_dt * vel[2 * p + d];
      total += f[2 * p + d];
    }
    if (total == 0) vel[2 * p] = vel[2 * p + 1] = 0;
  }
}

void FIRE::updateParticleVelocity() {
  // FIRE.cu:74-99
  std::vector<double>& vel = dpm_->particleVel;
  const std::vector<double>& f = dpm_->particleForce;
  for (long p = 0; p < dpm_->numParticles; p++) {
    double total = 0;
    for (long d = 0; d < 2; d++) {
      vel[2 * p + d] += 0.5 * fire_dt * f[2 * p + d] / d_mass[2 * p + d];
      total += f[2 * p + d];
    }
    if (total == 0) vel[2 * p] = vel[2 * p + 1] = 0;
  }
}

void FIRE::bendParticleVelocityTowardsForce() {
  // FIRE.cu:102-141
  std::vector<double>& vel = dpm_->particleVel;
  const std::vector<double>& f = dpm_->particleForce;
  double vDotF = 0;
  for (size_t i = 0; i < vel.size(); i++) vDotF += vel[i] * f[i];
  if (vDotF < 0) {
    std::fill(vel.begin(), vel.end(), 0.0);
    numStep = 0;
    fire_dt = std::max(fire_dt * f_dec, fire_dt_max / 10);
    a = a_start;
  } else if (numStep > minStep) {
    fire_dt = std::min(fire_dt * f_inc, fire_dt_max);
    a *= f_a;
  }
  double vv = 0, ff = 0;
  for (size_t i = 0; i < vel.size(); i++) {
    vv += vel[i] * vel[i];
    ff += f[i] * f[i];
  }
  if (ff == 0) {
    std::cout << "FIRE::bendVelocityTowardsForceFIRE: forceNormSquared is zero" << std::endl;
    std::fill(vel.begin(), vel.end(), 0.0);
  } else {
    double ratio = std::sqrt(vv / ff);
    for (size_t i = 0; i < vel.size(); i++) vel[i] = (1 - a) * vel[i] + a * f[i] * ratio;
  }
}

void FIRE::setParticleMass() {
  // DPM2D.cu:3677-3685
  d_mass.resize(2 * dpm_->numParticles);
  for (long p = 0; p < dpm_->numParticles; p++)
    for (long d = 0; d < 2; d++) d_mass[2 * p + d] = PI / (dpm_->particleRad[p] * dpm_->particleRad[p]);
}

void FIRE::setFIRETimeStep(double dt) {
  fire_dt = dt;
  fire_dt_max = 10 * dt;
  deviceReady = false;
}

void FIRE::minimizerParticleLoop() {
  // FIRE.cu:159-171.  The reference's checkParticleNeighbors here refreshes the vertex list
  // (DPM2D.cu:1425-1433), which does not affect the disks; the driver refreshes the particle
  // list after each iteration.
  if ((long)d_mass.size() != 2 * dpm_->numParticles) setParticleMass();
  updateParticlePositionAndVelocity();
  dpm_->calcParticleForceEnergy();
  updateParticleVelocity();
  bendParticleVelocityTowardsForce();
  numStep++;
}

// ---------------------------------------------------------------- vertex minimiser
void FIRE::ensureDevice() {
  if (deviceReady) return;
  double p7[7] = {a_start, f_dec, f_inc, f_a, fire_dt, fire_dt_max, a};
  dpm_->check(dpm_init_fire(dpm_->handle(), p7, minStep, numStep), "initFIRE");
  deviceReady = true;
}

void FIRE::minimizerVertexLoop() { minimizerVertexLoop(1); }

void FIRE::minimizerVertexLoop(long iterations) {
  ensureDevice();
  dpm_->check(dpm_fire_loop(dpm_->handle(), iterations), "vertexFIRELoop");
}

long FIRE::minimize(double tol, long maxIter) {
  ensureDevice();
  long it = 0;
  dpm_->check(dpm_fire_minimize(dpm_->handle(), tol, maxIter, &it), "minimizeFIRE");
  return it;
}

void FIRE::refresh() {
  if (!deviceReady) return;
  double s[5];
  dpm_->check(dpm_get_fire_state(dpm_->handle(), s), "getFIREState");
  fire_dt = s[0];
  a = s[1];
  numStep = (long)s[2];
}
