// dpm_device.h -- device-side arithmetic of the DPM2D vertex engine (gfx950).
//
// Layout: structure-of-arrays FP64 per vertex (x, y, vx, vy, fx, fy, rad, l0,
// theta0 ...), int32 topology, per-ring and per-system arrays.  A "set" holds
// B independent systems (packings) stored back to back; every ring lies inside
// one system and every vertex inside one ring.
//
// The pair and shape formulas restate include/cudaKernel.cuh of
// farceri/cudaRing (line numbers cited per function).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dpm {

constexpr double kPI = 3.141592653589793238462643383279502884197;  // defs.h:14
constexpr double kWCAcut = 1.122462048;                            // defs.h:17

enum Potential : int { POT_HARMONIC = 0, POT_LJ = 1, POT_ADHESIVE = 2, POT_WCA = 3 };
enum Interaction : int { INT_VV = 0, INT_SMOOTH = 1, INT_CELL_SMOOTH = 2, INT_ALL = 3 };
enum Monomer : int { MONO_HARMONIC = 0, MONO_FENE = 1 };
enum ParticleKind : int { PART_PASSIVE = 0, PART_ACTIVE = 1, PART_ACTIVE_SHAPE = 2 };
enum Geometry : int { GEOM_NORMAL = 0, GEOM_FIXED_BOX = 1, GEOM_FIXED_SIDES = 2, GEOM_ROUND_BOX = 3 };

// Philox stream tags (counter word 3, high byte)
enum RngTag : uint32_t { TAG_THERMAL = 1, TAG_INJECT = 2, TAG_ACTIVE = 3, TAG_SHAPE = 4 };

struct Params {
  int potential, interaction, monomer, particle, geometry;
  double ew, boxRadius;  // wall energy scale (setWallEnergyScale), round-box radius (setBoxRadius)
  double ea, el, eb, ec, l1, l2, LJcutoff, LJecut, LJfshift, stiff, extSq;
  double dt;
  double driving, taup;
  double gamma, noise;         // of the noise type (initLangevin, DPM2D.cu:3713-3753)
  double l0Tau, l0Diff;        // activeShape rest-length dynamics (setActiveTensionParams)
  double cutDistance, cutoff;  // neighbour list radius factor and rebuild threshold
  uint64_t seed;
};

// Device view of a set.  Positions are double buffered (x/y current, xn/yn next).
struct View {
  int nv, np, nsys, W;
  double *x, *y, *xn, *yn;
  double *vx, *vy, *fx, *fy;
  double *rad, *l0, *th0, *th, *len, *ev, *es;
  double* l0r;  // activeShape rest lengths, or null
  double *wfx, *wfy;  // force on the walls per vertex (getWallForce), written with walls
  double *lx, *ly;
  int *vring, *vsys;
  int4* vinfo;  // per vertex in one load: ring, ring's first vertex, ring size, system
  int *rfirst, *rnv, *rsys;
  double *a0, *area, *perim, *cx, *cy, *pe, *ang;
  double *bx, *by, *ibx, *iby;  // per-system box and its reciprocal
  int *svoff, *sroff;  // per-system vertex / ring offsets (nsys + 1)
  double *mbx, *mby;   // per-system image multipliers (Box::mx, Box::my)
  double* kr;          // per-ring shape costs [ka(np) | kl(np) | kb(np)], or null for the globals
  int *nbr, *ncnt;     // ELL neighbour list nbr[k * nv + v], count[v]
  int2* nbrp;          // beside each entry u: (prev(u), prev(prev(u))) in u's ring
  int *tile_v0, *tile_v1, *tile_sys;  // ring-aligned tiles (<= 64 vertices, one system)
  int* stile;          // per-system first tile (nsys + 1)
  int ntiles;
  const struct Rigid* rg;  // rigid-ring arrays (device copy of the pointers; one kernel argument)
};

// Rigid rings, per ring: velocity, angular velocity, force, torque / moment of inertia (an
// angular acceleration), moment of inertia, particle radius, last centre and angle; per vertex:
// torque about the ring centre.  Reached through View::rg so that the step kernel's argument
// block stays the size it has without them.
struct Rigid {
  double *pvx, *pvy, *pw, *pfx, *pfy, *ptq, *moi, *prad, *plx, *ply, *pla, *vtq;
};

// Box of one system with its reciprocal: the minimum image multiplies by 1/L instead of dividing
// and rounds to nearest-even (one instruction) instead of half away from zero; both differ from
// the reference's round(d / L) only when d is an exact odd multiple of half the box, where the
// minimum image itself is ambiguous (and far beyond every interaction range).
struct Box {
  double Lx, Ly, iLx, iLy;
  double mx, my;  // image multipliers: 1/L along a periodic direction, 0 along a walled one
};

// ------------------------------------------------------------ geometry ---
// pbcDistance, cudaKernel.cuh:70-74 (round = half away from zero)
__device__ __forceinline__ double pbc(double delta, double size) {
  return delta - size * round(delta / size);
}
// boxDistance, cudaKernel.cuh:76-94: with a zero multiplier (a walled direction) the difference
// is returned as it is
__device__ __forceinline__ double pbx(double delta, const Box& b) { return delta - b.Lx * rint(delta * b.mx); }
__device__ __forceinline__ double pby(double delta, const Box& b) { return delta - b.Ly * rint(delta * b.my); }

// ----------------------------------------------------------- Philox RNG ---
// Philox4x32-10 (Salmon et al., SC'11): 10 rounds, 9 key bumps.
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; r++) {
    if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint32_t hi0 = __umulhi(0xD2511F53u, c[0]), lo0 = 0xD2511F53u * c[0];
    uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]), lo1 = 0xCD9E8D57u * c[2];
    uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
  }
}

__device__ __forceinline__ double u01(uint32_t hi, uint32_t lo) {
  uint64_t v = ((uint64_t)hi << 32) | lo;
  return ((double)(v >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// Gaussian pair for (system, tag, call counter, local ring index): Box-Muller on Philox.
__device__ __forceinline__ void gauss_pair(uint64_t seed, uint32_t sys, uint32_t tag, uint64_t counter,
                                           uint32_t index, double& z0, double& z1) {
  uint32_t c[4] = {index, (uint32_t)counter, (uint32_t)(counter >> 32), (tag << 24) | (sys & 0xFFFFFFu)};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  double uA = u01(c[0], c[1]);
  double uB = u01(c[2], c[3]);
  double r = sqrt(-2.0 * log(uA));
  double a = 2.0 * kPI * uB;
  z0 = r * cos(a);
  z1 = r * sin(a);
}

// ---------------------------------------------------------- pair terms ---
// Directed pair on `this` from `other` at separation (dx, dy) = this - other:
// calcContactInteraction 531-544, calcLJInteraction 562-583,
// calcAdhesiveInteraction 604-626, calcWCAInteraction 653-668.
// Adds to (fx, fy) and returns the directed-visit energy.
template <int POT>
__device__ __forceinline__ double pair_force(const Params& P, double dx, double dy, double radSum,
                                             double& fx, double& fy) {
  // the force formulas multiply by 1/distance (one division) where the reference divides
  double distance = sqrt(dx * dx + dy * dy);
  if constexpr (POT == POT_HARMONIC) {
    double overlap = 1 - distance / radSum;
    if (overlap > 0) {
      double gm = P.ec * overlap / radSum;
      const double id = 1.0 / distance;
      fx += gm * dx * id;
      fy += gm * dy * id;
      return (0.5 * P.ec * overlap * overlap) * 0.5;
    }
    return 0.0;
  } else if constexpr (POT == POT_LJ) {
    if (distance <= (P.LJcutoff * radSum)) {
      const double id = 1.0 / distance;
      double ratio = radSum * id;
      double r2 = ratio * ratio;
      double ratio6 = r2 * r2 * r2;
      double ratio12 = ratio6 * ratio6;
      double forceShift = P.LJfshift / radSum;
      double gm = 24 * P.ec * (2 * ratio12 - ratio6) * id - forceShift;
      fx += gm * dx * id;
      fy += gm * dy * id;
      return 0.5 * (4 * P.ec * (ratio12 - ratio6) - P.LJecut - fabs(forceShift) * (distance - P.LJcutoff * radSum));
    }
    return 0.0;
  } else if constexpr (POT == POT_ADHESIVE) {
    double overlap = 1 - distance / radSum;
    if (distance < (1 + P.l2) * radSum) {
      double gm, epot;
      if (distance < (1 + P.l1) * radSum) {
        gm = P.ec * overlap / radSum;
        epot = 0.5 * P.ec * (overlap * overlap - P.l1 * P.l2) * 0.5;
      } else {
        gm = -(P.ec * P.l1 / (P.l2 - P.l1)) * (overlap + P.l2) / radSum;
        epot = -(0.5 * (P.ec * P.l1 / (P.l2 - P.l1)) * (overlap + P.l2) * (overlap + P.l2)) * 0.5;
      }
      const double id = 1.0 / distance;
      fx += gm * dx * id;
      fy += gm * dy * id;
      return epot;
    }
    return 0.0;
  } else {  // WCA
    if (distance <= (kWCAcut * radSum)) {
      const double id = 1.0 / distance;
      double ratio = radSum * id;
      double r2 = ratio * ratio;
      double ratio6 = r2 * r2 * r2;
      double ratio12 = ratio6 * ratio6;
      double gm = 24 * P.ec * (2 * ratio12 - ratio6) * id;
      fx += gm * dx * id;
      fy += gm * dy * id;
      return 0.5 * P.ec * (4 * (ratio12 - ratio6) + 1);
    }
    return 0.0;
  }
}

// calcGradMultipleAndEnergy, cudaKernel.cuh:806-858: scalar gradient multiple and energy.
template <int POT>
__device__ __forceinline__ double grad_multiple(const Params& P, double distance, double radSum, double& epot) {
  if constexpr (POT == POT_HARMONIC) {
    double overlap = 1 - distance / radSum;
    if (overlap > 0) {
      epot = 0.5 * P.ec * overlap * overlap * 0.5;
      return P.ec * overlap / radSum;
    }
    return 0;
  } else if constexpr (POT == POT_LJ) {
    if (distance <= (P.LJcutoff * radSum)) {
      const double id = 1.0 / distance;
      double ratio = radSum * id;
      double r2 = ratio * ratio;
      double ratio6 = r2 * r2 * r2;
      double ratio12 = ratio6 * ratio6;
      double forceShift = P.LJfshift / radSum;
      epot = 0.5 * (4 * P.ec * (ratio12 - ratio6) - P.LJecut - fabs(forceShift) * (distance - P.LJcutoff * radSum));
      return 24 * P.ec * (2 * ratio12 - ratio6) * id - forceShift;
    }
    return 0;
  } else if constexpr (POT == POT_ADHESIVE) {
    double overlap = 1 - distance / radSum;
    if (distance < (1 + P.l1) * radSum) {
      epot = 0.5 * P.ec * (overlap * overlap - P.l1 * P.l2) * 0.5;
      return P.ec * overlap / radSum;
    } else if (distance < (1 + P.l2) * radSum) {
      epot = -(0.5 * (P.ec * P.l1 / (P.l2 - P.l1)) * (overlap + P.l2) * (overlap + P.l2)) * 0.5;
      return -(P.ec * P.l1 / (P.l2 - P.l1)) * (overlap + P.l2) / radSum;
    }
    return 0;
  } else {
    if (distance <= (kWCAcut * radSum)) {
      const double id = 1.0 / distance;
      double ratio = radSum * id;
      double r2 = ratio * ratio;
      double ratio6 = r2 * r2 * r2;
      double ratio12 = ratio6 * ratio6;
      epot = 0.5 * P.ec * (4 * (ratio12 - ratio6) + 1);
      return 4 * P.ec * (12 * ratio12 - 6 * ratio6) * id;
    }
    return 0;
  }
}

// ------------------------------------------------------------------ walls ---
// calcWallWCAInteraction, cudaKernel.cuh:1221-1237: WCA (energy scale ew) between a vertex of
// radius r and its image (wx, wy) on the wall; the wall takes the opposite force.
__device__ __forceinline__ double wall_wca(const Params& P, const Box& B, double x, double y, double r, double wx,
                                           double wy, double& fx, double& fy, double& wfx, double& wfy) {
  const double dx = pbx(x - wx, B), dy = pby(y - wy, B);
  const double distance = sqrt(dx * dx + dy * dy);
  if (distance < (kWCAcut * r)) {
    const double ratio = r / distance;
    const double ratio6 = pow(ratio, 6);
    const double ratio12 = ratio6 * ratio6;
    const double gm = 24 * P.ew * (2 * ratio12 - ratio6) / distance;
    fx += gm * dx / distance;
    fy += gm * dy / distance;
    wfx -= gm * dx / distance;
    wfy -= gm * dy / distance;
    return P.ew * (4 * (ratio12 - ratio6) + 1);
  }
  return 0.0;
}

// kernelCalcVertexSquareWallInteraction (fixedBox), kernelCalcVertexSidesInteraction2D
// (fixedSides) and kernelCalcVertexRoundWallInteraction (roundBox, centred on the origin),
// cudaKernel.cuh:1324-1410.  Adds to (fx, fy), returns the wall energy, writes the wall force.
__device__ __forceinline__ double wall_forces(const Params& P, const Box& B, double x, double y, double r, double& fx,
                                              double& fy, double& wfx, double& wfy) {
  wfx = wfy = 0.0;
  double e = 0.0;
  const double rc = kWCAcut * r;
  if (P.geometry == GEOM_FIXED_BOX) {
    if (x < rc) e += wall_wca(P, B, x, y, r, 0.0, y, fx, fy, wfx, wfy);
    else if ((B.Lx - x) < rc) e += wall_wca(P, B, x, y, r, B.Lx, y, fx, fy, wfx, wfy);
  }
  if (P.geometry == GEOM_FIXED_BOX || P.geometry == GEOM_FIXED_SIDES) {
    if (y < rc) e += wall_wca(P, B, x, y, r, x, 0.0, fx, fy, wfx, wfy);
    else if ((B.Ly - y) < rc) e += wall_wca(P, B, x, y, r, x, B.Ly, fx, fy, wfx, wfy);
  }
  if (P.geometry == GEOM_ROUND_BOX) {
    const double rr = sqrt(x * x + y * y), th = atan2(y, x);
    if ((P.boxRadius - rr) < rc)
      e += wall_wca(P, B, x, y, r, P.boxRadius * cos(th), P.boxRadius * sin(th), fx, fy, wfx, wfy);
  }
  return e;
}

// ------------------------------------------------- smooth vertex-segment ---
// One evaluation of the smooth interaction of vertex i with the segment
// (prev(j), j) of its listed neighbour j, kernelCalcSmoothInteraction
// cudaKernel.cuh:922-975 with getProjection 861-864, getProjectionPos 866-871,
// calcCross 873-875, calcVertexSegmentInteraction 877-897 and
// calcVertexVertexInteraction 899-912.  Instead of scattering to three
// vertices with atomics, the caller asks for the share of one role:
//   ROLE_THIS  -> force on i,  ROLE_OTHER -> force on j,  ROLE_PREV -> force on prev(j).
// The energy e (added by the reference to both rings) is returned for ROLE_THIS
// and ROLE_OTHER.  Inputs: ti = r_i, tj = r_j, tp = r_prev(j), tq = r_prev(prev(j)).
enum { ROLE_THIS = 0, ROLE_OTHER = 1, ROLE_PREV = 2 };

// Shared by both: with sg = r_j - r_p, ip = r_i - r_p, ps = r_q - r_p and jt = r_j - r_i (each one
// minimum image), every difference the reference forms is one of these or its exact negative
// (the minimum image is odd), the sign cross/|cross| is a comparison, and the divisions by the
// segment length are multiplications by its reciprocal.  The two projections, which select the
// branches, are formed exactly as the reference forms them.
template <int POT, int ROLE>
__device__ __forceinline__ double smooth_pair(const Params& P, const Box& B, double tix, double tiy, double tjx,
                                              double tjy, double tpx, double tpy, double tqx, double tqy,
                                              double radSum, double& fx, double& fy) {
  // segment = other - previous, length, projection of i on (prev, other)
  const double sgx = pbx(tjx - tpx, B), sgy = pby(tjy - tpy, B);
  const double length = sqrt(sgx * sgx + sgy * sgy);
  const double ipx = pbx(tix - tpx, B), ipy = pby(tiy - tpy, B);  // this - previous
  const double proj = (ipx * sgx + ipy * sgy) / (length * length);
  // previous segment = secondPrevious - previous; previous - secondPrevious = -ps
  const double psx = pbx(tqx - tpx, B), psy = pby(tqy - tpy, B);
  const double previousLength = sqrt(psx * psx + psy * psy);
  const double previousProj =
      (pbx(tix - tqx, B) * -psx + pby(tiy - tqy, B) * -psy) / (previousLength * previousLength);
  double energy = 0.0;
  if (proj >= 0 && proj < 1) {
    const double qx = tpx + proj * sgx, qy = tpy + proj * sgy;  // projection point (proj in [0, 1))
    const double dqx = pbx(tix - qx, B), dqy = pby(tiy - qy, B);
    double epot = 0.0;
    const double gm = grad_multiple<POT>(P, sqrt(dqx * dqx + dqy * dqy), radSum, epot);
    if (gm != 0) {
      // calcCross: (prev - other)_x (other - this)_y - (other - this)_x (prev - other)_y
      const double jtx = pbx(tjx - tix, B), jty = pby(tjy - tiy, B);
      const double cross = -sgx * jty - jtx * -sgy;
      const double absCross = fabs(cross);
      const double sign = cross > 0 ? 1.0 : (cross < 0 ? -1.0 : 0.0);
      const double il = 1.0 / length;
      if constexpr (ROLE == ROLE_THIS) {
        fx += gm * sign * -sgy * il;
        fy += gm * sign * sgx * il;
      } else if constexpr (ROLE == ROLE_OTHER) {
        const double c = absCross * il * il;
        fx += gm * (sign * ipy + c * -sgx) * il;
        fy += gm * (sign * -ipx + c * -sgy) * il;
      } else {
        const double c = absCross * il * il;
        fx += gm * (sign * jty - c * -sgx) * il;
        fy += gm * (sign * -jtx - c * -sgy) * il;
      }
      energy += epot;
    }
    if (previousProj >= 0 && previousProj < 1) {
      // concave: subtract the pair (prev(j), i); calcVertexVertexInteraction(previousPos, thisPos)
      const double dx = -ipx, dy = -ipy;
      const double d = sqrt(dx * dx + dy * dy);
      double ep = 0.0;
      const double g = grad_multiple<POT>(P, d, radSum, ep);
      if (g != 0) {
        if constexpr (ROLE == ROLE_THIS) { fx += g * dx * (1.0 / d); fy += g * dy * (1.0 / d); }
        if constexpr (ROLE == ROLE_PREV) { fx += -g * dx * (1.0 / d); fy += -g * dy * (1.0 / d); }
        energy -= ep;
      }
    }
  } else if (proj < 0) {
    if (previousProj >= 1) {
      // convex corner: add the pair (i, prev(j)); calcVertexVertexInteraction(thisPos, previousPos)
      const double dx = ipx, dy = ipy;
      const double d = sqrt(dx * dx + dy * dy);
      double ep = 0.0;
      const double g = grad_multiple<POT>(P, d, radSum, ep);
      if (g != 0) {
        if constexpr (ROLE == ROLE_THIS) { fx += g * dx * (1.0 / d); fy += g * dy * (1.0 / d); }
        if constexpr (ROLE == ROLE_PREV) { fx += -g * dx * (1.0 / d); fy += -g * dy * (1.0 / d); }
        energy += ep;
      }
    }
  }
  return energy;
}

// The same evaluation for the pair (i, segment (prev(j), j)) returning both the share of j
// (ROLE_OTHER) and the share of prev(j) (ROLE_PREV) in one pass; the caller owns j, so both
// receivers lie in its own ring.  Returns the energy e added to both rings.
template <int POT>
__device__ __forceinline__ double smooth_pair_op(const Params& P, const Box& B, double tix, double tiy, double tjx,
                                                 double tjy, double tpx, double tpy, double tqx, double tqy,
                                                 double radSum, double& ox, double& oy, double& px, double& py) {
  const double sgx = pbx(tjx - tpx, B), sgy = pby(tjy - tpy, B);
  const double length = sqrt(sgx * sgx + sgy * sgy);
  const double ipx = pbx(tix - tpx, B), ipy = pby(tiy - tpy, B);
  const double proj = (ipx * sgx + ipy * sgy) / (length * length);
  const double psx = pbx(tqx - tpx, B), psy = pby(tqy - tpy, B);
  const double previousLength = sqrt(psx * psx + psy * psy);
  const double previousProj =
      (pbx(tix - tqx, B) * -psx + pby(tiy - tqy, B) * -psy) / (previousLength * previousLength);
  double energy = 0.0;
  if (proj >= 0 && proj < 1) {
    const double qx = tpx + proj * sgx, qy = tpy + proj * sgy;
    const double dqx = pbx(tix - qx, B), dqy = pby(tiy - qy, B);
    double epot = 0.0;
    const double gm = grad_multiple<POT>(P, sqrt(dqx * dqx + dqy * dqy), radSum, epot);
    if (gm != 0) {
      const double jtx = pbx(tjx - tix, B), jty = pby(tjy - tiy, B);
      const double cross = -sgx * jty - jtx * -sgy;
      const double absCross = fabs(cross);
      const double sign = cross > 0 ? 1.0 : (cross < 0 ? -1.0 : 0.0);
      const double il = 1.0 / length;
      const double c = absCross * il * il;
      ox += gm * (sign * ipy + c * -sgx) * il;
      oy += gm * (sign * -ipx + c * -sgy) * il;
      px += gm * (sign * jty - c * -sgx) * il;
      py += gm * (sign * -jtx - c * -sgy) * il;
      energy += epot;
    }
    if (previousProj >= 0 && previousProj < 1) {
      const double dx = -ipx, dy = -ipy;
      const double d = sqrt(dx * dx + dy * dy);
      double ep = 0.0;
      const double g = grad_multiple<POT>(P, d, radSum, ep);
      if (g != 0) {
        const double id = 1.0 / d;
        px += -g * dx * id;
        py += -g * dy * id;
        energy -= ep;
      }
    }
  } else if (proj < 0) {
    if (previousProj >= 1) {
      const double dx = ipx, dy = ipy;
      const double d = sqrt(dx * dx + dy * dy);
      double ep = 0.0;
      const double g = grad_multiple<POT>(P, d, radSum, ep);
      if (g != 0) {
        const double id = 1.0 / d;
        px += -g * dx * id;
        py += -g * dy * id;
        energy += ep;
      }
    }
  }
  return energy;
}

}  // namespace dpm
