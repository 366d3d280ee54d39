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

// ------------------------------------------------------- ring areas (dual) ---
// Per ring: signed shoelace sum over the minimum-image walk from the first vertex, with tangents
// w; area = |S| / 2 and its directional derivative.
__global__ void k_adj_area(View V, const double* wx, const double* wy, double* A, double* dA) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= V.np) return;
  const int f = V.rfirst[r], n = V.rnv[r];
  const int s = V.rsys[r];
  const Box B = box_of(V, s);
  Dual ux = mk(0, wx[f]), uy = mk(0, wy[f]);
  Dual S = mk(0, 0);
  for (int i = 0; i < n; i++) {
    const int a = f + i, b = (i + 1 < n) ? a + 1 : f;
    const Dual dx = mk(pbx(V.x[b] - V.x[a], B), wx[b] - wx[a]);
    const Dual dy = mk(pby(V.y[b] - V.y[a], B), wy[b] - wy[a]);
    S = S + (ux * dy - dx * uy);
    ux = ux + dx;
    uy = uy + dy;
  }
  const double sg = S.v < 0 ? -1.0 : 1.0;
  A[r] = 0.5 * fabs(S.v);
  dA[r] = 0.5 * sg * S.t;
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

// the two passes above in one launch: thread t tests vertex t, clears cell t, and walks ring t
__global__ void k_adj_prep(View V, Params P, const double* wx, const double* wy, double* A, double* dA, int* flag,
                           int* any, int* cellOther, int ncells) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  for (int c = t; c < ncells; c += gridDim.x * blockDim.x) cellOther[c] = 0;
  if (t < V.nv) {
    const int s = V.vsys[t];
    const Box B = box_of(V, s);
    const double ddx = pbx(V.x[t] - V.lx[t], B), ddy = pby(V.y[t] - V.ly[t], B);
    if (3 * sqrt(ddx * ddx + ddy * ddy) > P.cutoff) {
      atomicOr(&flag[s], 1);
      atomicOr(any, 1);
    }
  }
  if (t < V.np) {
    const int f = V.rfirst[t], n = V.rnv[t];
    const Box B = box_of(V, V.rsys[t]);
    Dual ux = mk(0, wx[f]), uy = mk(0, wy[f]);
    Dual S = mk(0, 0);
    for (int i = 0; i < n; i++) {
      const int a = f + i, b = (i + 1 < n) ? a + 1 : f;
      const Dual dx = mk(pbx(V.x[b] - V.x[a], B), wx[b] - wx[a]);
      const Dual dy = mk(pby(V.y[b] - V.y[a], B), wy[b] - wy[a]);
      S = S + (ux * dy - dx * uy);
      ux = ux + dx;
      uy = uy + dy;
    }
    const double sg = S.v < 0 ? -1.0 : 1.0;
    A[t] = 0.5 * fabs(S.v);
    dA[t] = 0.5 * sg * S.t;
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

// ------------------------------------------------------------ H w per vertex ---
// AdjArgs (dpm_internal.h): mode 0 writes H w; mode 1 applies the reverse-pass update.
template <int POT>
__global__ void k_adj_hvp(View V, Params P, CellGrid G, AdjArgs A) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V.nv) return;
  const int r = V.vring[v], s = V.vsys[v];
  const Box B = box_of(V, s);
  if (A.flag && A.flag[s]) build_row(V, P, G, A.count, v, s, B, A.err, A.maxcnt);

  const int f = V.rfirst[r], n = V.rnv[r];
  const int im1 = (v == f) ? f + n - 1 : v - 1;
  const int im2 = (im1 == f) ? f + n - 1 : im1 - 1;
  const int ip1 = (v == f + n - 1) ? f : v + 1;
  const int ip2 = (ip1 == f + n - 1) ? f : ip1 + 1;
  const double x0 = V.x[v], y0 = V.y[v];
  // ring neighbourhood unwrapped by minimum-image hops from v, with tangents w
  const D2 R0 = {mk(0, A.wx[v]), mk(0, A.wy[v])};
  const double hm1x = -pbx(x0 - V.x[im1], B), hm1y = -pby(y0 - V.y[im1], B);
  const D2 Rm1 = {mk(hm1x, A.wx[im1]), mk(hm1y, A.wy[im1])};
  const D2 Rm2 = {mk(hm1x - pbx(V.x[im1] - V.x[im2], B), A.wx[im2]),
                  mk(hm1y - pby(V.y[im1] - V.y[im2], B), A.wy[im2])};
  const double hp1x = pbx(V.x[ip1] - x0, B), hp1y = pby(V.y[ip1] - y0, B);
  const D2 Rp1 = {mk(hp1x, A.wx[ip1]), mk(hp1y, A.wy[ip1])};
  const D2 Rp2 = {mk(hp1x + pbx(V.x[ip2] - V.x[ip1], B), A.wx[ip2]),
                  mk(hp1y + pby(V.y[ip2] - V.y[ip1], B), A.wy[ip2])};

  const double ka = V.kr ? V.kr[r] : P.ea;
  const double kl = V.kr ? V.kr[V.np + r] : P.el;
  const double kb = V.kr ? V.kr[2 * V.np + r] : P.eb;
  const double a0 = V.a0[r];
  // area term per unit ka: the ring's area and its derivative along w, from the precomputed
  // pass or walked here by this thread (n <= 64 minimum-image hops)
  Dual area;
  if (A.A) {
    area = mk(A.A[r], A.dA[r]);
  } else {
    Dual ux = mk(0, A.wx[f]), uy = mk(0, A.wy[f]), S2 = mk(0, 0);
    for (int i = 0; i < n; i++) {
      const int a = f + i, b = (i + 1 < n) ? a + 1 : f;
      const Dual dx = mk(pbx(V.x[b] - V.x[a], B), A.wx[b] - A.wx[a]);
      const Dual dy = mk(pby(V.y[b] - V.y[a], B), A.wy[b] - A.wy[a]);
      S2 = S2 + (ux * dy - dx * uy);
      ux = ux + dx;
      uy = uy + dy;
    }
    area = mk(0.5 * fabs(S2.v), (S2.v < 0 ? -0.5 : 0.5) * S2.t);
  }
  const Dual gA = 0.5 * ((area / a0) - 1.0) / a0;
  const D2 Fa = {gA * (Rm1.y - Rp1.y), gA * (Rp1.x - Rm1.x)};
  // perimeter term per unit kl
  const D2 nS = Rp1 - R0, pS = R0 - Rm1;
  const Dual nLen = dsqrt(dot(nS, nS)), pLen = dsqrt(dot(pS, pS));
  const double l0v = V.l0[v], l0p = V.l0[im1];
  const Dual nDelta = nLen / l0v - 1.0, pDelta = pLen / l0p - 1.0;
  const D2 Fp = (nDelta / (l0v * nLen)) * nS - (pDelta / (l0p * pLen)) * pS;
  // bending term per unit kb: turning angles of im1, v, ip1 (calcAngle, cudaKernel.cuh:423-427)
  const D2 pm = Rm1 - Rm2, np1 = Rp2 - Rp1;
  const Dual thm = datan2(cross(pS, pm), dot(pS, pm));
  const Dual th = datan2(cross(nS, pS), dot(nS, pS));
  const Dual thp = datan2(cross(np1, nS), dot(np1, nS));
  const Dual Dm = thm - V.th0[im1], D = th - V.th0[v], Dp = thp - V.th0[ip1];
  const Dual preVar = (D - Dm) / dot(pS, pS), nextVar = (D - Dp) / dot(nS, nS);
  const D2 Fb = {preVar * pS.y + nextVar * nS.y, -(preVar * pS.x) - nextVar * nS.x};

  Dual fx = ka * Fa.x + kl * Fp.x + kb * Fb.x;
  Dual fy = ka * Fa.y + kl * Fp.y + kb * Fb.y;
  // vertex-vertex pairs over the row (kernelCalcVertexInteraction, cudaKernel.cuh:690-722)
  const double radv = V.rad[v];
  const int c = V.ncnt[v];
  for (int k = 0; k < c; k++) {
    const int j = V.nbr[(long)k * V.nv + v];
    const Dual dx = mk(pbx(x0 - V.x[j], B), A.wx[v] - A.wx[j]);
    const Dual dy = mk(pby(y0 - V.y[j], B), A.wy[v] - A.wy[j]);
    const Dual d = dsqrt(dx * dx + dy * dy);
    Dual g;
    if (gm_dual<POT>(P, d, radv + V.rad[j], g)) {
      fx = fx + g * dx / d;
      fy = fy + g * dy / d;
    }
  }
  const double hwx = -fx.t, hwy = -fy.t;  // H w = -dF
  if (A.mode == 0) {
    A.hx[v] = hwx;
    A.hy[v] = hwy;
    return;
  }
  // reverse-pass update and parameter partials (values of the unit-cost forces)
  const double wvx = A.wx[v], wvy = A.wy[v];
  const long nv = V.nv;
  A.g[v] += A.cd * (wvx * Fa.x.v + wvy * Fa.y.v);
  A.g[nv + v] += A.cd * (wvx * Fp.x.v + wvy * Fp.y.v);
  A.g[2 * nv + v] += A.cd * (wvx * Fb.x.v + wvy * Fb.y.v);
  // rest length l0_v of the segment (v, ip1): its force is +f on v and -f on ip1 with
  // f = kl (1 / l0^2 - 1 / (l0 |n|)) n, so d/dl0 contracts with (w_v - w_ip1)
  const double nl = nLen.v;
  const double dg = -2.0 / (l0v * l0v * l0v) + 1.0 / (l0v * l0v * nl);
  A.g[3 * nv + v] += A.cd * kl * dg * ((wvx - A.wx[ip1]) * nS.x.v + (wvy - A.wy[ip1]) * nS.y.v);
  const double ax = A.ax[v] - A.cd * hwx, ay = A.ay[v] - A.cd * hwy;
  A.ax[v] = ax;
  A.ay[v] = ay;
  if (A.update) {
    A.nwx[v] = wvx + A.dt * ax;
    A.nwy[v] = wvy + A.dt * ay;
  } else {
    A.nwx[v] = wvx;
    A.nwy[v] = wvy;
  }
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

hipError_t launch_adj_area(dpm_set_s* S, const View& Vn, const double* wx, const double* wy, double* A, double* dA) {
  k_adj_area<<<grid_n(S->np), 256, 0, S->stream>>>(Vn, wx, wy, A, dA);
  return hipGetLastError();
}

hipError_t launch_adj_prep(dpm_set_s* S, const View& Vn, int parity, const double* wx, const double* wy, double* A,
                           double* dA) {
  long n = S->nv > S->G.ncells ? S->nv : S->G.ncells;
  if (S->np > n) n = S->np;
  k_adj_prep<<<grid_n(n), 256, 0, S->stream>>>(Vn, S->P, wx, wy, A, dA, S->d_flag[parity], S->d_any + parity,
                                               S->G.count[1 - parity], S->G.ncells);
  return hipGetLastError();
}

hipError_t launch_adj_flag(dpm_set_s* S, const View& Vn, int parity) {
  const long n = S->nv > S->G.ncells ? S->nv : S->G.ncells;
  k_adj_flag<<<grid_n(n), 256, 0, S->stream>>>(Vn, S->P, S->d_flag[parity], S->d_any + parity,
                                               S->G.count[1 - parity], S->G.ncells);
  return hipGetLastError();
}

hipError_t launch_adj_hvp(dpm_set_s* S, const View& Vn, AdjArgs A) {
  A.err = S->d_err;
  A.maxcnt = S->d_maxcnt;
  const unsigned gr = grid_n(S->nv);
  switch (S->P.potential) {
    case POT_HARMONIC: k_adj_hvp<POT_HARMONIC><<<gr, 256, 0, S->stream>>>(Vn, S->P, S->G, A); break;
    case POT_LJ: k_adj_hvp<POT_LJ><<<gr, 256, 0, S->stream>>>(Vn, S->P, S->G, A); break;
    case POT_ADHESIVE: k_adj_hvp<POT_ADHESIVE><<<gr, 256, 0, S->stream>>>(Vn, S->P, S->G, A); break;
    default: k_adj_hvp<POT_WCA><<<gr, 256, 0, S->stream>>>(Vn, S->P, S->G, A); break;
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
