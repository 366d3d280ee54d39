// dpm_adjoint.hip -- Hessian-vector products and the reverse pass of velocity Verlet.
//
// No reference counterpart: differentiation is an addition.  The Hessian of the potential
// energy is never formed.  H w is the directional derivative of the force field, -dF(x + e w)/de
// at e = 0, evaluated in forward mode: every position carries a tangent, and the shape and pair
// force formulas of the step kernel (calcAreaForceEnergy, calcPerimeterForceEnergy,
// calcBendingForceEnergy, cudaKernel.cuh:429-471; the pair potentials 531-668 and 806-858) are
// written once over a dual-number type.  Since F = -grad U holds exactly for these formulas, the
// result is the exact Hessian-vector product (to rounding), not a finite difference.
//
// Reverse pass of one NVE step (v' = v + dt/2 F(x), x1 = x + dt v', v1 = v' + dt/2 F(x1)):
// the two half kicks that meet at x_n use the same adjoint vector w_n, so each position of the
// trajectory costs one H w product:
//     ax -= c dt H(x_n) w,  g_theta += c dt (dF/dtheta)^T w,  w <- w + dt ax,
// with c = 1/2 at the two ends of the trajectory and 1 inside.
#include "dpm_internal.h"
#include "dpm_rows.h"
#include "dpm_tile.h"

namespace dpm {

// ------------------------------------------------------------- dual numbers ---
struct Dual {
  double v, t;  // value and tangent
};
__device__ __forceinline__ Dual mk(double v, double t) { return Dual{v, t}; }
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return Dual{a.v + b.v, a.t + b.t}; }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return Dual{a.v - b.v, a.t - b.t}; }
__device__ __forceinline__ Dual operator-(Dual a) { return Dual{-a.v, -a.t}; }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return Dual{a.v * b.v, a.t * b.v + a.v * b.t}; }
__device__ __forceinline__ Dual operator*(double a, Dual b) { return Dual{a * b.v, a * b.t}; }
__device__ __forceinline__ Dual operator*(Dual a, double b) { return Dual{a.v * b, a.t * b}; }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return Dual{q, (a.t - q * b.t) / b.v};
}
__device__ __forceinline__ Dual operator/(Dual a, double b) { return Dual{a.v / b, a.t / b}; }
__device__ __forceinline__ Dual operator/(double a, Dual b) {
  const double q = a / b.v;
  return Dual{q, -q * b.t / b.v};
}
__device__ __forceinline__ Dual operator+(Dual a, double b) { return Dual{a.v + b, a.t}; }
__device__ __forceinline__ Dual operator-(Dual a, double b) { return Dual{a.v - b, a.t}; }
__device__ __forceinline__ Dual operator-(double a, Dual b) { return Dual{a - b.v, -b.t}; }
__device__ __forceinline__ Dual dsqrt(Dual a) {
  const double r = sqrt(a.v);
  return Dual{r, 0.5 * a.t / r};
}
// atan2(y, x): d = (x dy - y dx) / (x^2 + y^2)
__device__ __forceinline__ Dual datan2(Dual y, Dual x) {
  return Dual{atan2(y.v, x.v), (x.v * y.t - y.v * x.t) / (x.v * x.v + y.v * y.v)};
}

struct D2 {
  Dual x, y;
};
__device__ __forceinline__ D2 operator+(D2 a, D2 b) { return D2{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ D2 operator-(D2 a, D2 b) { return D2{a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ D2 operator*(Dual s, D2 a) { return D2{s * a.x, s * a.y}; }
__device__ __forceinline__ D2 operator*(double s, D2 a) { return D2{s * a.x, s * a.y}; }
__device__ __forceinline__ Dual dot(D2 a, D2 b) { return a.x * b.x + a.y * b.y; }
__device__ __forceinline__ Dual cross(D2 a, D2 b) { return a.x * b.y - a.y * b.x; }

// ------------------------------------------------------------ pair forces ---
// Gradient multiple g(d) of the four potentials (grad_multiple in dpm_device.h), on a dual
// distance; the force on i is g(d) (r_i - r_j) / d.
template <int POT>
__device__ __forceinline__ bool gm_dual(const Params& P, Dual d, double radSum, Dual& g) {
  if constexpr (POT == POT_HARMONIC) {
    const Dual overlap = 1.0 - d / radSum;
    if (overlap.v > 0) {
      g = P.ec * overlap / radSum;
      return true;
    }
    return false;
  } else if constexpr (POT == POT_LJ || POT == POT_WCA) {
    const double cut = POT == POT_LJ ? P.LJcutoff : kWCAcut;
    if (d.v <= cut * radSum) {
      const Dual ratio = radSum / d;
      const Dual r2 = ratio * ratio;
      const Dual r6 = r2 * r2 * r2;
      const Dual r12 = r6 * r6;
      g = 24 * P.ec * (2.0 * r12 - r6) / d;
      if constexpr (POT == POT_LJ) g = g - P.LJfshift / radSum;
      return true;
    }
    return false;
  } else {
    const Dual overlap = 1.0 - d / radSum;
    if (d.v < (1 + P.l1) * radSum) {
      g = P.ec * overlap / radSum;
      return true;
    } else if (d.v < (1 + P.l2) * radSum) {
      g = -(P.ec * P.l1 / (P.l2 - P.l1)) * (overlap + P.l2) / radSum;
      return true;
    }
    return false;
  }
}

// --------------------------------------------------------- displacement flag ---
// Verlet-list validity at the positions in V.x: flags systems whose vertices moved more than a
// third of the skin since the rows were built (checkDisplacement, DPM2D.cu:1390-1406), and
// clears the other parity's cell counts for the next use.
__global__ void k_adj_flag(View V, Params P, int* flag, int* any, int* cellOther, int ncells) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  for (int c = t; c < ncells; c += gridDim.x * blockDim.x) cellOther[c] = 0;
  if (t >= V.nv) return;
  const int s = V.vsys[t];
  const Box B = box_of(V, s);
  const double ddx = pbx(V.x[t] - V.lx[t], B), ddy = pby(V.y[t] - V.ly[t], B);
  if (3 * sqrt(ddx * ddx + ddy * ddy) > P.cutoff) {
    atomicOr(&flag[s], 1);
    atomicOr(any, 1);
  }
}

// One Verlet list for a whole recomputed segment: the rows built at the reference positions
// (lastPos) stay valid while every vertex of every recorded slot is within the displacement
// bound of checkDisplacement (3 |x - x_ref| <= cutoff); `bad` is set otherwise.
__global__ void k_adj_seg_check(View V, Params P, const double* seg, int slots, int* bad) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V.nv) return;
  const Box B = box_of(V, V.vsys[v]);
  const double rx = V.lx[v], ry = V.ly[v];
  int out = 0;
  for (int n = 0; n < slots; n++) {
    const double* sx = seg + (long)n * 2 * V.nv;
    const double dx = pbx(sx[v] - rx, B), dy = pby(sx[V.nv + v] - ry, B);
    if (3 * sqrt(dx * dx + dy * dy) > P.cutoff) out = 1;
  }
  if (out) atomicOr(bad, 1);
}

// ------------------------------------------------------------ H w per tile ---
// One workgroup per ring-aligned tile, as the step kernel: wave 0 evaluates the shape force and
// its tangent with the ring walks done as segmented wave scans (the unwrapped position of a
// vertex relative to the ring's first vertex is a scan of minimum-image segments; its tangent is
// w_v - w_first), while waves 1..nw-1 share the pair gather, of which only the tangent is needed.
// AdjArgs (dpm_internal.h): mode 0 writes H w; mode 1 applies the reverse-pass update.
template <int POT>
__global__ __launch_bounds__(WAVE* MAXSPLIT) void k_adj_hvp(View V, Params P, CellGrid G, AdjArgs A) {
  extern __shared__ double red[];  // red[((w - 1) * 2 + q) * WAVE + lane], sized by the launch
  const int w = threadIdx.x / WAVE;
  const int nw = blockDim.x / WAVE;
  Lane L;
  lane_ctx(V, A.xcd ? xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x, threadIdx.x & (WAVE - 1), L);
  const Box B = box_of(V, L.s);
  const double x = V.x[L.v], y = V.y[L.v];
  const double wvx = A.wx[L.v], wvy = A.wy[L.v];
  if (A.cellOther)
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < A.ncells; c += gridDim.x * blockDim.x)
      A.cellOther[c] = 0;
  if (A.flag) {
    const bool rebuild = A.flag[L.s] != 0;
    if (rebuild && w == 0 && L.act) build_row(V, P, G, A.count, G.slots, L.v, L.s, B, A.err, A.maxcnt);
    if (rebuild) __syncthreads();
  }

  // pair tangent: d/de of sum_j g(d) (r_v - r_j) / d along w (kernelCalcVertexInteraction,
  // cudaKernel.cuh:690-722)
  double tx = 0, ty = 0;
  const int g0 = nw > 1 ? 1 : 0;
  if (w >= g0 && L.act) {
    const double radv = V.rad[L.v];
    const int c = V.ncnt[L.v];
    for (int k = w - g0; k < c; k += nw - g0) {
      const int j = V.nbr[(long)k * V.nv + L.v];
      const Dual dx = mk(pbx(x - V.x[j], B), wvx - A.wx[j]);
      const Dual dy = mk(pby(y - V.y[j], B), wvy - A.wy[j]);
      const Dual d = dsqrt(dx * dx + dy * dy);
      Dual g;
      if (gm_dual<POT>(P, d, radv + V.rad[j], g)) {
        const Dual q = g / d;
        tx += (q * dx).t;
        ty += (q * dy).t;
      }
    }
  }
  if (nw > 1 && w > 0) {
    red[(w - 1) * 2 * WAVE + L.lane] = tx;
    red[((w - 1) * 2 + 1) * WAVE + L.lane] = ty;
  }

  // ---- wave 0: shape force of the lane's vertex and its tangent
  D2 Fa{}, Fp{}, Fb{};
  double wnx = 0, wny = 0, nSx = 0, nSy = 0, nl = 1, l0v = 1, ka = 0, kl = 0, kb = 0;
  if (w == 0) {
    const double xn = shfl_d(x, L.nl), yn = shfl_d(y, L.nl);
    const double xp = shfl_d(x, L.pl), yp = shfl_d(y, L.pl);
    wnx = shfl_d(wvx, L.nl);
    wny = shfl_d(wvy, L.nl);
    const double wpx = shfl_d(wvx, L.pl), wpy = shfl_d(wvy, L.pl);
    const double wfx = shfl_d(wvx, L.segS), wfy = shfl_d(wvy, L.segS);
    // segments (v, next) and (prev, v) of the reference's unwrapped walk
    const D2 nS = {mk(pbx(xn - x, B), wnx - wvx), mk(pby(yn - y, B), wny - wvy)};
    const D2 pS = {mk(-pbx(xp - x, B), wvx - wpx), mk(-pby(yp - y, B), wvy - wpy)};
    // area (calcAreaForceEnergy, cudaKernel.cuh:429-435): shoelace over the unwrapped walk
    const double Ix = seg_scan(L.act ? nS.x.v : 0.0, L.lane, L.segS);
    const double Iy = seg_scan(L.act ? nS.y.v : 0.0, L.lane, L.segS);
    const D2 u = {mk(Ix - nS.x.v, wvx - wfx), mk(Iy - nS.y.v, wvy - wfy)};
    const Dual shoe = cross(u, nS);
    const double Sv = seg_sum(L.act ? shoe.v : 0.0, L.lane, L.segS, L.segE);
    const double St = seg_sum(L.act ? shoe.t : 0.0, L.lane, L.segS, L.segE);
    const Dual area = mk(0.5 * fabs(Sv), (Sv < 0 ? -0.5 : 0.5) * St);
    const double a0 = V.a0[L.r];
    const Dual gA = 0.5 * ((area / a0) - 1.0) / a0;
    Fa = D2{gA * (-(pS.y + nS.y)), gA * (pS.x + nS.x)};
    // perimeter (calcPerimeterForceEnergy 437-447)
    const Dual nLen = dsqrt(dot(nS, nS)), pLen = dsqrt(dot(pS, pS));
    l0v = V.l0[L.v];
    const double l0p = shfl_d(l0v, L.pl);
    Fp = (((nLen / l0v) - 1.0) / (l0v * nLen)) * nS - (((pLen / l0p) - 1.0) / (l0p * pLen)) * pS;
    // bending (calcBendingForceEnergy 457-471; calcAngle 423-427)
    const Dual D = datan2(cross(nS, pS), dot(nS, pS)) - V.th0[L.v];
    const Dual Dp = mk(shfl_d(D.v, L.pl), shfl_d(D.t, L.pl));
    const Dual Dn = mk(shfl_d(D.v, L.nl), shfl_d(D.t, L.nl));
    const Dual preVar = (D - Dp) / dot(pS, pS), nextVar = (D - Dn) / dot(nS, nS);
    Fb = D2{preVar * pS.y + nextVar * nS.y, -(preVar * pS.x) - nextVar * nS.x};
    ka = V.kr ? V.kr[L.r] : P.ea;
    kl = V.kr ? V.kr[V.np + L.r] : P.el;
    kb = V.kr ? V.kr[2 * V.np + L.r] : P.eb;
    nSx = nS.x.v;
    nSy = nS.y.v;
    nl = nLen.v;
  }
  if (nw > 1) {
    __syncthreads();
    if (w > 0) return;
    for (int q = 1; q < nw; q++) {
      tx += red[(q - 1) * 2 * WAVE + L.lane];
      ty += red[((q - 1) * 2 + 1) * WAVE + L.lane];
    }
  }
  if (!L.act) return;
  // Verlet-list validity at the positions processed next (checkDisplacement, DPM2D.cu:1390-1406);
  // a row rebuilt above was rebuilt by this lane, so lastPos is this lane's own write
  if (A.px) {
    const double ddx = pbx(A.px[L.v] - V.lx[L.v], B), ddy = pby(A.py[L.v] - V.ly[L.v], B);
    if (3 * sqrt(ddx * ddx + ddy * ddy) > P.cutoff) {
      atomicOr(&A.flagOut[L.s], 1);
      atomicOr(A.anyOut, 1);
    }
  }
  const double hwx = -(ka * Fa.x.t + kl * Fp.x.t + kb * Fb.x.t + tx);  // H w = -dF
  const double hwy = -(ka * Fa.y.t + kl * Fp.y.t + kb * Fb.y.t + ty);
  if (A.mode == 0) {
    A.hx[L.v] = hwx;
    A.hy[L.v] = hwy;
    return;
  }
  // reverse-pass update and parameter partials (values of the unit-cost forces)
  const long nv = V.nv;
  const double cd = A.cd;
  A.g[L.v] += cd * (wvx * Fa.x.v + wvy * Fa.y.v);
  A.g[nv + L.v] += cd * (wvx * Fp.x.v + wvy * Fp.y.v);
  A.g[2 * nv + L.v] += cd * (wvx * Fb.x.v + wvy * Fb.y.v);
  // rest length l0_v of the segment (v, next): its force is +f on v and -f on next with
  // f = kl (1 / l0^2 - 1 / (l0 |n|)) n, so d/dl0 contracts with (w_v - w_next)
  const double dg = -2.0 / (l0v * l0v * l0v) + 1.0 / (l0v * l0v * nl);
  A.g[3 * nv + L.v] += cd * kl * dg * ((wvx - wnx) * nSx + (wvy - wny) * nSy);
  const double ax = A.ax[L.v] - cd * hwx, ay = A.ay[L.v] - cd * hwy;
  A.ax[L.v] = ax;
  A.ay[L.v] = ay;
  A.nwx[L.v] = A.update ? wvx + A.dt * ax : wvx;
  A.nwy[L.v] = A.update ? wvy + A.dt * ay : wvy;
}

// ------------------------------------------------- energy parameter partials ---
// Per vertex, at the current shape variables (area, len, theta from the last force call):
// out[v] = dU/dka share (1/n of the ring's), out[nv+v] = dU/dkl share (segment v),
// out[2nv+v] = dU/dkb share (angle v), out[3nv+v] = dU/dl0_v.
__global__ void k_energy_param_partials(View V, Params P, double* out) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V.nv) return;
  const int r = V.vring[v];
  const long nv = V.nv;
  const double dA = V.area[r] / V.a0[r] - 1.0;
  const double dl = V.len[v] / V.l0[v] - 1.0;
  const double D = V.th[v] - V.th0[v];
  const double kl = V.kr ? V.kr[V.np + r] : P.el;
  out[v] = 0.5 * dA * dA / V.rnv[r];
  out[nv + v] = 0.5 * dl * dl;
  out[2 * nv + v] = 0.5 * D * D;
  out[3 * nv + v] = kl * dl * (-V.len[v] / (V.l0[v] * V.l0[v]));
}

// per-ring sums of three per-vertex arrays in vertex order (deterministic):
// out[q * np + r] = sum over the ring's vertices of in[q * nv + v], q = 0, 1, 2
__global__ void k_ring_sum3(View V, const double* in, double* out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= V.np) return;
  const int f = V.rfirst[r], n = V.rnv[r];
  for (int q = 0; q < 3; q++) {
    double acc = 0;
    for (int i = 0; i < n; i++) acc += in[(long)q * V.nv + f + i];
    out[(long)q * V.np + r] = acc;
  }
}

__global__ void k_fill_d(double* p, double value, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = value;
}

__global__ void k_copy_d(double* dst, const double* src, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// ------------------------------------------------------------------ launchers ---
static inline unsigned grid_n(long n) { return (unsigned)((n + 255) / 256); }

hipError_t launch_adj_flag(dpm_set_s* S, const View& Vn, int parity) {
  const long n = S->nv > S->G.ncells ? S->nv : S->G.ncells;
  k_adj_flag<<<grid_n(n), 256, 0, S->stream>>>(Vn, S->P, S->d_flag[parity], S->d_any + parity,
                                               S->G.count[1 - parity], S->G.ncells);
  return hipGetLastError();
}

hipError_t launch_adj_hvp(dpm_set_s* S, const View& Vn, AdjArgs A) {
  A.err = S->d_err;
  A.maxcnt = S->d_maxcnt;
  A.xcd = S->xcdMap;
  const unsigned g = (unsigned)S->V.ntiles, b = WAVE * S->split;
  const size_t sh = (size_t)(S->split - 1) * 2 * WAVE * sizeof(double);
  hipStream_t st = S->stream;
  switch (S->P.potential) {
    case POT_HARMONIC: k_adj_hvp<POT_HARMONIC><<<g, b, sh, st>>>(Vn, S->P, S->G, A); break;
    case POT_LJ: k_adj_hvp<POT_LJ><<<g, b, sh, st>>>(Vn, S->P, S->G, A); break;
    case POT_ADHESIVE: k_adj_hvp<POT_ADHESIVE><<<g, b, sh, st>>>(Vn, S->P, S->G, A); break;
    default: k_adj_hvp<POT_WCA><<<g, b, sh, st>>>(Vn, S->P, S->G, A); break;
  }
  return hipGetLastError();
}

hipError_t launch_adj_seg_check(dpm_set_s* S, const double* seg, int slots, int* bad) {
  k_adj_seg_check<<<grid_n(S->nv), 256, 0, S->stream>>>(S->V, S->P, seg, slots, bad);
  return hipGetLastError();
}

hipError_t launch_energy_param_partials(dpm_set_s* S, double* out) {
  k_energy_param_partials<<<grid_n(S->nv), 256, 0, S->stream>>>(S->V, S->P, out);
  return hipGetLastError();
}

hipError_t launch_ring_sum3(dpm_set_s* S, const double* in, double* out) {
  k_ring_sum3<<<grid_n(S->np), 256, 0, S->stream>>>(S->V, in, out);
  return hipGetLastError();
}

hipError_t launch_fill_d(double* p, double value, long n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  k_fill_d<<<grid_n(n), 256, 0, st>>>(p, value, n);
  return hipGetLastError();
}

hipError_t launch_copy_d(double* dst, const double* src, long n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  k_copy_d<<<grid_n(n), 256, 0, st>>>(dst, src, n);
  return hipGetLastError();
}

}  // namespace dpm
