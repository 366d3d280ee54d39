// FIRE.h -- FIRE minimiser of a DPM2D (fields and methods of include/FIRE.h:20-76 in
// farceri/cudaRing).  The vertex minimiser runs in libdpmring.so (FIRE state per system on the
// device); the soft-disk (particle) minimiser used during set-up runs on the host.
#ifndef DPM_HOST_FIRE_H
#define DPM_HOST_FIRE_H

#include <vector>

#include "defs.h"

class DPM2D;

class FIRE {
 public:
  DPM2D* dpm_;

  FIRE() = default;
  explicit FIRE(DPM2D* dpmPtr) : dpm_(dpmPtr) {}
  ~FIRE() = default;

  double a_start = 0, a = 0, f_dec = 0, f_inc = 0, f_a = 0, fire_dt_max = 0, fire_dt = 0;
  double cutDistance = 0;
  long minStep = 0, numStep = 0;
  std::vector<double> d_mass;  // per particle degree of freedom (particle minimiser)
  bool deviceReady = false;    // the library's FIRE state matches the fields above

  // FIRE.cu:29-41
  void initMinimizer(double a_start_, double f_dec_, double f_inc_, double f_a_, double fire_dt_,
                     double dt_max_, double a_, long minStep_, long numStep_, long numDOF_);

  // ---- particle minimiser (FIRE.cu:43-171), host side
  void updateParticlePositionAndVelocity();
  void updateParticleVelocity();
  void bendParticleVelocityTowardsForce();
  void setParticleMass();
  void setFIRETimeStep(double timeStep);
  void minimizerParticleLoop();

  // ---- vertex minimiser (FIRE.cu:300-411), on the device.  The reference's sub-steps
  // (updateVertexPosition/Velocity, checkFIREParams, bendVertexVelocityTowardsForce) are fused
  // into the iteration and are not callable one by one.
  void minimizerVertexLoop();
  void minimizerVertexLoop(long iterations);
  long minimize(double tol, long maxIter);
  // copy fire_dt, a, numStep of the device state into the fields above
  void refresh();

 private:
  void ensureDevice();
};

#endif
