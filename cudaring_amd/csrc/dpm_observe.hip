// dpm_observe.hip -- observables of a set: particle contacts, hexatic order, mean-square
// displacement and intermediate scattering function against stored initial positions, and the
// per-particle contact stress.  One thread per ring (or vertex); they run outside the step loop.
#include "dpm_internal.h"
#include "dpm_rows.h"

namespace dpm {

// kernelCalcContacts, cudaKernel.cuh:1791-1822: the rings owning any entry of the rows of ring
// p's vertices, in order of first appearance, into a -1 padded row of `width`; `overflow` is set
// when a ring has more.  With gap >= 0, kernelCalcGeometricContacts (1823-1861): only entries
// within -(1 - d / radSum) < gap count, and contacts beyond the row width are dropped.
__global__ void k_contacts(View V, int* list, int* count, int width, int* overflow, double gap) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= V.np) return;
  int* row = list + (long)p * width;
  for (int q = 0; q < width; q++) row[q] = -1;
  int added = 0;
  const int f = V.rfirst[p], n = V.rnv[p];
  for (int v = f; v < f + n; v++) {
    const int c = V.ncnt[v];
    for (int k = 0; k < c; k++) {
      const int o = V.nbr[(long)k * V.nv + v];
      if (o == v || o < 0) continue;
      if (gap >= 0) {
        const Box B = box_of(V, V.rsys[p]);
        const double dx = pbx(V.x[v] - V.x[o], B), dy = pby(V.y[v] - V.y[o], B);
        if (!(-(1 - sqrt(dx * dx + dy * dy) / (V.rad[v] + V.rad[o])) < gap)) continue;
        if (added == width) continue;
      }
      const int q = V.vring[o];
      bool isNew = true;
      for (int t = 0; t < added; t++)
        if (row[t] == q) isNew = false;
      if (!isNew) continue;
      if (added == width) {
        atomicOr(overflow, 1);
        count[p] = added;
        return;
      }
      row[added++] = q;
    }
  }
  count[p] = added;
}

// kernelCalcHexaticOrderParameter, cudaKernel.cuh:1951-1976: |<exp(6 i theta)>| over the
// contacts of ring p (directions between ring centres), 0 without contacts
__global__ void k_hexatic(View V, const int* list, const int* count, int width, double* psi6) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= V.np) return;
  const Box B = box_of(V, V.rsys[p]);
  const int n = count[p];
  double re = 0, im = 0;
  if (n > 0) {
    for (int q = 0; q < n; q++) {
      const int o = list[(long)p * width + q];
      if (o == p || o < 0) continue;
      const double a = atan2(pby(V.cy[p] - V.cy[o], B), pbx(V.cx[p] - V.cx[o], B));
      re += cos(6.0 * a);
      im += sin(6.0 * a);
    }
    re /= n;
    im /= n;
  }
  psi6[p] = hypot(re, im);
}

// Displacement of n points (vertices, or ring centres with rings = 1) from their initial
// positions: d2[i] = |dr|^2 (kernelCalcVertexDistanceSq / kernelCalcParticleDistanceSq,
// cudaKernel.cuh:1888-1925) and sf[i] = sin(k |dr|) / (k |dr|) (kernelCalc*ScatteringFunction,
// 1935-1949), with k = waveNum, or pi / (2 r) with r the radius of the system's first vertex
// when waveNum <= 0 (getVertexISF, DPM2D.cu:1486-1494).
__global__ void k_displacement(View V, const double* x, const double* y, const double* x0, const double* y0,
                               int n, int rings, double waveNum, double* d2, double* sf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = rings ? V.rsys[i] : V.vsys[i];
  const Box B = box_of(V, s);
  const double dx = pbx(x[i] - x0[i], B), dy = pby(y[i] - y0[i], B);
  const double dd = dx * dx + dy * dy;
  const double k = waveNum > 0 ? waveNum : kPI / (2 * V.rad[V.svoff[s]]);
  const double d = sqrt(dd);
  d2[i] = dd;
  sf[i] = sin(k * d) / (k * d);
}

// kernelCalcPerParticleStressTensor, cudaKernel.cuh:1525-1565 (harmonic contacts): for every
// overlapping listed pair of ring p's vertices, lever (x) f with f = ec overlap / radSum *
// delta / (distance radSum), delta = other - this; the lever starts at the vertex position
// relative to the ring centre and, as in the reference, gains delta / 2 at each overlap of the
// same vertex.  out[4p..4p+3] = (xx, xy, yx, yy).
__global__ void k_particle_stress(View V, Params P, double* out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= V.np) return;
  const Box B = box_of(V, V.rsys[p]);
  double sxx = 0, sxy = 0, syx = 0, syy = 0;
  const int f = V.rfirst[p], n = V.rnv[p];
  for (int v = f; v < f + n; v++) {
    double relx = pbx(V.x[v] - V.cx[p], B), rely = pby(V.y[v] - V.cy[p], B);
    const int c = V.ncnt[v];
    for (int k = 0; k < c; k++) {
      const int o = V.nbr[(long)k * V.nv + v];
      if (o == v || o < 0) continue;
      const double radSum = V.rad[v] + V.rad[o];
      const double ddx = pbx(V.x[o] - V.x[v], B), ddy = pby(V.y[o] - V.y[v], B);
      const double distance = sqrt(ddx * ddx + ddy * ddy);
      const double overlap = 1 - distance / radSum;
      if (overlap > 0) {
        const double gm = P.ec * overlap / radSum;
        const double fx = gm * ddx / (distance * radSum), fy = gm * ddy / (distance * radSum);
        relx += ddx * 0.5;
        rely += ddy * 0.5;
        sxx += relx * fx;
        syy += rely * fy;
        sxy += relx * fy;
        syx += rely * fx;
      }
    }
  }
  out[4 * p] = sxx;
  out[4 * p + 1] = sxy;
  out[4 * p + 2] = syx;
  out[4 * p + 3] = syy;
}

static inline unsigned grid_o(long n) { return (unsigned)((n + 255) / 256); }

hipError_t launch_contacts(dpm_set_s* S, int* list, int* count, int width, int* overflow, double gap) {
  k_contacts<<<grid_o(S->np), 256, 0, S->stream>>>(S->V, list, count, width, overflow, gap);
  return hipGetLastError();
}

hipError_t launch_hexatic(dpm_set_s* S, const int* list, const int* count, int width, double* psi6) {
  k_hexatic<<<grid_o(S->np), 256, 0, S->stream>>>(S->V, list, count, width, psi6);
  return hipGetLastError();
}

hipError_t launch_displacement(dpm_set_s* S, const double* x, const double* y, const double* x0, const double* y0,
                               long n, int rings, double waveNum, double* d2, double* sf) {
  k_displacement<<<grid_o(n), 256, 0, S->stream>>>(S->V, x, y, x0, y0, (int)n, rings, waveNum, d2, sf);
  return hipGetLastError();
}

hipError_t launch_particle_stress(dpm_set_s* S, double* out) {
  k_particle_stress<<<grid_o(S->np), 256, 0, S->stream>>>(S->V, S->P, out);
  return hipGetLastError();
}

}  // namespace dpm
