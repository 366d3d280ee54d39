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
// The vertex-segment interaction (vertexSmooth) is differentiated through the same branches as
// the force: its forces are piecewise smooth in the positions, and the tangent is that of the
// branch the positions select (exact away from the branch switches).
//
// Reverse pass of one NVE step (v' = v + dt/2 F(x), x1 = x + dt v', v1 = v' + dt/2 F(x1)):
// the two half kicks that meet at x_n use the same adjoint vector w_n, so each position of the
// trajectory costs one H w product:
//     ax -= c dt H(x_n) w,  g_theta += c dt (dF/dtheta)^T w,  w <- w + dt ax,
// with c = 1/2 at the two ends of the trajectory and 1 inside.
#include <algorithm>

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

// ------------------------------------------------------ vertex-segment forces ---
// minimum image of a dual difference: the image shift is constant wherever it is differentiable
__device__ __forceinline__ Dual pbxd(Dual d, const Box& B) { return Dual{pbx(d.v, B), d.t}; }
__device__ __forceinline__ Dual pbyd(Dual d, const Box& B) { return Dual{pby(d.v, B), d.t}; }

// The vertex-segment interaction of vertex i with the segment (p, j), p = prev(j), q = prev(p),
// over dual positions: the branches of smooth_pair / smooth_pair_op (kernelCalcSmoothInteraction,
// cudaKernel.cuh:922-975) taken on the values, the forces on i (fi), on j (fj) and on p (fp)
// accumulated as duals, so their tangents are the directional derivatives along the positions'
// tangents.
template <int POT>
__device__ __forceinline__ void smooth_dual(const Params& P, const Box& B, D2 ti, D2 tj, D2 tp, D2 tq, double radSum,
                                            D2& fi, D2& fj, D2& fp) {
  const D2 sg = {pbxd(tj.x - tp.x, B), pbyd(tj.y - tp.y, B)};
  const Dual len2 = dot(sg, sg);
  const D2 ip = {pbxd(ti.x - tp.x, B), pbyd(ti.y - tp.y, B)};
  const Dual proj = dot(ip, sg) / len2;
  const D2 ps = {pbxd(tq.x - tp.x, B), pbyd(tq.y - tp.y, B)};
  const D2 iq = {pbxd(ti.x - tq.x, B), pbyd(ti.y - tq.y, B)};
  const double previousProj = (iq.x.v * -ps.x.v + iq.y.v * -ps.y.v) / (ps.x.v * ps.x.v + ps.y.v * ps.y.v);
  auto pair = [&](D2 d, double s) {  // vertex pair along d = r_i - r_p, counted with sign s on i
    const Dual dist = dsqrt(dot(d, d));
    Dual g;
    if (gm_dual<POT>(P, dist, radSum, g)) {
      const Dual k = s * (g / dist);
      fi = fi + k * d;
      fp = fp - k * d;
    }
  };
  if (proj.v >= 0 && proj.v < 1) {
    const D2 q = {tp.x + proj * sg.x, tp.y + proj * sg.y};
    const D2 dq = {pbxd(ti.x - q.x, B), pbyd(ti.y - q.y, B)};
    Dual gm;
    if (gm_dual<POT>(P, dsqrt(dot(dq, dq)), radSum, gm)) {
      const D2 jt = {pbxd(tj.x - ti.x, B), pbyd(tj.y - ti.y, B)};
      const Dual cross = jt.x * sg.y - sg.x * jt.y;
      const double sign = cross.v > 0 ? 1.0 : (cross.v < 0 ? -1.0 : 0.0);
      const Dual il = 1.0 / dsqrt(len2);
      const Dual c = (sign * cross) * il * il;
      const Dual gl = gm * il;
      fi = fi + D2{gl * (sign * -sg.y), gl * (sign * sg.x)};
      fj = fj + D2{gl * (sign * ip.y - c * sg.x), gl * (sign * -ip.x - c * sg.y)};
      fp = fp + D2{gl * (sign * jt.y + c * sg.x), gl * (sign * -jt.x + c * sg.y)};
    }
    if (previousProj >= 0 && previousProj < 1) pair(ip, -1.0);  // concave: the pair (p, i) subtracted
  } else if (proj.v < 0 && previousProj >= 1) {
    pair(ip, 1.0);  // convex corner: the pair (i, p) added
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
template <int INTER, int POT>
__global__ __launch_bounds__(WAVE* MAXSPLIT) void k_adj_hvp(View V, Params P, CellGrid G, AdjArgs A) {
  // partials of gather waves 1..nw-1: red[((w - 1) * NR + q) * WAVE + lane], sized by the launch
  extern __shared__ double red[];
  constexpr int NR = INTER == INT_SMOOTH ? 4 : 2;
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
    if (rebuild && A.rowsLdsW == V.W && !G.brute[L.s])  // the tile's waves share the cell search
      build_rows_tile(V, P, G, A.count, G.slots, L.lane, L.v, L.r, L.s, L.act, w, nw, B, x, y, (int*)red, A.err,
                      A.maxcnt);
    else if (rebuild && w == 0 && L.act)
      build_row(V, P, G, A.count, G.slots, L.v, L.s, B, A.err, A.maxcnt);
    if (rebuild) __syncthreads();
  }

  // pair tangent along w: of sum_j g(d) (r_v - r_j) / d (kernelCalcVertexInteraction,
  // cudaKernel.cuh:690-722), or of the vertex-segment forces on v gathered as the step kernel
  // gathers them (kernelCalcSmoothInteraction, 922-975): for every u in list(v), the force on v
  // from the segment (prev(u), u) and, from the pair (u, segment (prev(v), v)), the force on v
  // and the share (tqx, tqy) owed to prev(v), handed to prev(v)'s lane below
  double tx = 0, ty = 0, tqx = 0, tqy = 0;
  const int g0 = nw > 1 ? 1 : 0;
  if (w >= g0 && L.act) {
    const double radv = V.rad[L.v];
    const int c = V.ncnt[L.v];
    if constexpr (INTER == INT_SMOOTH) {
      const int vp = L.v0 + L.pl;
      const int vpp = L.v0 + ((L.pl > L.segS) ? L.pl - 1 : L.segE);
      const D2 tv = {mk(x, wvx), mk(y, wvy)};
      const D2 tvp = {mk(V.x[vp], A.wx[vp]), mk(V.y[vp], A.wy[vp])};
      const D2 tvpp = {mk(V.x[vpp], A.wx[vpp]), mk(V.y[vpp], A.wy[vpp])};
      for (int k = w - g0; k < c; k += nw - g0) {
        const int u = V.nbr[(long)k * V.nv + L.v];
        const int2 pp = V.nbrp[(long)k * V.nv + L.v];
        const D2 tu = {mk(V.x[u], A.wx[u]), mk(V.y[u], A.wy[u])};
        const D2 tup = {mk(V.x[pp.x], A.wx[pp.x]), mk(V.y[pp.x], A.wy[pp.x])};
        const D2 tupp = {mk(V.x[pp.y], A.wx[pp.y]), mk(V.y[pp.y], A.wy[pp.y])};
        const double radSum = radv + V.rad[u];
        D2 fv{}, fu{}, fup{};
        smooth_dual<POT>(P, B, tv, tu, tup, tupp, radSum, fv, fu, fup);  // v against (prev(u), u)
        D2 gu{}, gv{}, gvp{};
        smooth_dual<POT>(P, B, tu, tv, tvp, tvpp, radSum, gu, gv, gvp);  // u against (prev(v), v)
        tx += fv.x.t + gv.x.t;
        ty += fv.y.t + gv.y.t;
        tqx += gvp.x.t;
        tqy += gvp.y.t;
      }
    } else {
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
  }
  if (nw > 1 && w > 0) {
    double* r = red + (w - 1) * NR * WAVE + L.lane;
    r[0] = tx;
    r[WAVE] = ty;
    if constexpr (NR == 4) {
      r[2 * WAVE] = tqx;
      r[3 * WAVE] = tqy;
    }
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
      const double* r = red + (q - 1) * NR * WAVE + L.lane;
      tx += r[0];
      ty += r[WAVE];
      if constexpr (NR == 4) {
        tqx += r[2 * WAVE];
        tqy += r[3 * WAVE];
      }
    }
  }
  if constexpr (INTER == INT_SMOOTH) {  // the shares owed to this lane's vertex by its successor
    tx += shfl_d(tqx, L.nl);
    ty += shfl_d(tqy, L.nl);
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
  if (A.mode == 0 || A.mode == 2) {
    A.hx[L.v] = hwx;
    A.hy[L.v] = hwy;
    if (A.mode == 0) return;
  }
  // parameter partials (values of the unit-cost forces) contracted with w, weighted by cd
  const long nv = V.nv;
  const double cd = A.cd;
  A.g[L.v] += cd * (wvx * Fa.x.v + wvy * Fa.y.v);
  A.g[nv + L.v] += cd * (wvx * Fp.x.v + wvy * Fp.y.v);
  A.g[2 * nv + L.v] += cd * (wvx * Fb.x.v + wvy * Fb.y.v);
  // rest length l0_v of the segment (v, next): its force is +f on v and -f on next with
  // f = kl (1 / l0^2 - 1 / (l0 |n|)) n, so d/dl0 contracts with (w_v - w_next)
  const double dg = -2.0 / (l0v * l0v * l0v) + 1.0 / (l0v * l0v * nl);
  A.g[3 * nv + L.v] += cd * kl * dg * ((wvx - wnx) * nSx + (wvy - wny) * nSy);
  if (A.mode == 2) return;
  // reverse-pass update of velocity Verlet
  const double ax = A.ax[L.v] - cd * hwx, ay = A.ay[L.v] - cd * hwy;
  A.ax[L.v] = ax;
  A.ay[L.v] = ay;
  A.nwx[L.v] = A.update ? wvx + A.dt * ax : wvx;
  A.nwy[L.v] = A.update ? wvy + A.dt * ay : wvy;
}

// ------------------------------------------------------ FIRE reverse pass ---
// One FIRE iteration (FIRE.cu:338-411) with its branch held at the recorded one: the system's dt
// and alpha after the parameter update and whether it restarted (N+ = 0).  With F = F(x):
//   restart: x_a = x - dt/2 v, v_a = 0;  otherwise x_a = x, v_a = v
//   v_b = v_a + dt/2 F;  v_c = (1 - a) v_b + a F |v_b| / |F|  (|.| over the system; v_c = v_b if
//   F = 0);  x' = x_a + dt v_c;  v' = v_c + dt/2 F(x').
// Reverse, given the cotangents (ax, bv) of (x', v') and pend of F(x') from the next iteration:
//   Fc = dt/2 bv + pend;  x'bar = ax - H(x') Fc  (Hessian-vector product, k_adj_hvp mode 2);
//   v_c bar = bv + dt x'bar;  v_b bar = (1 - a) v_c bar + a (F . v_c bar) v_b / (|v_b| |F|);
//   F bar = a |v_b| (v_c bar / |F| - F (F . v_c bar) / |F|^3) + dt/2 v_b bar;
//   x bar = x'bar, v bar = restart ? -dt/2 x'bar : v_b bar; pend <- F bar.
// Vectors are 2V SoA (all x components, then all y); a tape slot holds x, y, vx, vy, fx, fy.
__global__ void k_fadj_prep(View V, const FireState* st, const double* bv, const double* pend, double* fc) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V.nv) return;
  const double h = 0.5 * st[V.vsys[v]].dt;
  fc[v] = h * bv[v] + pend[v];
  fc[V.nv + v] = h * bv[V.nv + v] + pend[V.nv + v];
}

__device__ __forceinline__ double tile_sys_sum(const double* part, int q, int t0, int t1, int lane) {
  double acc = 0;
  for (int t = t0 + lane; t < t1; t += WAVE) acc += part[t * 4 + q];
  return wave_sum(acc);
}

// per-tile partials of the system sums F . v_c bar, |v_b|^2, |F|^2
__global__ void k_fadj_partials(View V, const FireState* st, const double* slot, const double* ax, const double* bv,
                                const double* hw, double* part) {
  const int tile = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
  if (tile >= V.ntiles) return;
  Lane L;
  lane_ctx(V, tile, threadIdx.x & (WAVE - 1), L);
  const long nv = V.nv;
  double s1 = 0, n2 = 0, f2 = 0;
  if (L.act) {
    const FireState fs = st[L.s];
    const double dt = fs.dt, h = 0.5 * dt;
    const bool restart = fs.numStep == 0;
    const int v = L.v;
    const double fx = slot[4 * nv + v], fy = slot[5 * nv + v];
    const double vbx = (restart ? 0.0 : slot[2 * nv + v]) + h * fx;
    const double vby = (restart ? 0.0 : slot[3 * nv + v]) + h * fy;
    const double cx = bv[v] + dt * (ax[v] - hw[v]), cy = bv[nv + v] + dt * (ax[nv + v] - hw[nv + v]);
    s1 = fx * cx + fy * cy;
    n2 = vbx * vbx + vby * vby;
    f2 = fx * fx + fy * fy;
  }
  s1 = wave_sum(s1);
  n2 = wave_sum(n2);
  f2 = wave_sum(f2);
  if (L.lane == 0) {
    part[tile * 4 + 0] = s1;
    part[tile * 4 + 1] = n2;
    part[tile * 4 + 2] = f2;
  }
}

// the reverse step of one iteration: (ax, bv, pend) of (x', v', F(x')) -> those of (x, v, F(x))
__global__ void k_fadj_update(View V, const FireState* st, const double* slot, double* ax, double* bv, double* pend,
                              const double* hw, const double* part) {
  const int tile = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
  if (tile >= V.ntiles) return;
  Lane L;
  lane_ctx(V, tile, threadIdx.x & (WAVE - 1), L);
  const int t0 = V.stile[L.s], t1 = V.stile[L.s + 1];
  const double s1 = tile_sys_sum(part, 0, t0, t1, L.lane);
  const double n2 = tile_sys_sum(part, 1, t0, t1, L.lane);
  const double f2 = tile_sys_sum(part, 2, t0, t1, L.lane);
  if (!L.act) return;
  const long nv = V.nv;
  const int v = L.v;
  const FireState fs = st[L.s];
  const double dt = fs.dt, h = 0.5 * dt, a = fs.a;
  const bool restart = fs.numStep == 0;
  const double fx = slot[4 * nv + v], fy = slot[5 * nv + v];
  const double vbx = (restart ? 0.0 : slot[2 * nv + v]) + h * fx;
  const double vby = (restart ? 0.0 : slot[3 * nv + v]) + h * fy;
  const double xbx = ax[v] - hw[v], xby = ax[nv + v] - hw[nv + v];
  const double cx = bv[v] + dt * xbx, cy = bv[nv + v] + dt * xby;
  double bbx = cx, bby = cy, Fbx = 0, Fby = 0;
  if (f2 > 0) {
    const double nvb = sqrt(n2), nf = sqrt(f2);
    const double k = nvb > 0 ? a * s1 / (nvb * nf) : 0.0;
    bbx = (1 - a) * cx + k * vbx;
    bby = (1 - a) * cy + k * vby;
    const double q = a * nvb / nf, r = a * nvb * s1 / (f2 * nf);
    Fbx = q * cx - r * fx;
    Fby = q * cy - r * fy;
  }
  pend[v] = Fbx + h * bbx;
  pend[nv + v] = Fby + h * bby;
  ax[v] = xbx;
  ax[nv + v] = xby;
  bv[v] = restart ? -h * xbx : bbx;
  bv[nv + v] = restart ? -h * xby : bby;
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

__global__ void k_axpy(double* y, const double* x, double alpha, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += alpha * x[i];
}

__global__ void k_copy_d(double* dst, const double* src, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// several copies in one launch: thread t copies word t of every segment, grid-stride
__global__ void k_copy_segs(CopySegs c) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (int q = 0; q < c.n; q++) {
    const unsigned* src = c.src[q];
    unsigned* dst = c.dst[q];
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < c.words[q]; i += stride) dst[i] = src[i];
  }
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
  hipStream_t st = S->stream;
  // rows rebuilt by the whole workgroup need 64 + 64 W ints of LDS (launches of more than one
  // wave, W <= 64), in the region the gather partials use afterwards
  const bool tileRows = A.flag && S->split >= 2 && S->V.W <= 64;
  A.rowsLdsW = tileRows ? S->V.W : 0;
  const size_t rows = tileRows ? (size_t)(64 + 64 * S->V.W) * sizeof(int) : 0;
  if (S->P.interaction == INT_SMOOTH) {
    const size_t sh = std::max((size_t)(S->split - 1) * 4 * WAVE * sizeof(double), rows);
    switch (S->P.potential) {
      case POT_HARMONIC: k_adj_hvp<INT_SMOOTH, POT_HARMONIC><<<g, b, sh, st>>>(Vn, S->P, S->G, A); break;
      case POT_LJ: k_adj_hvp<INT_SMOOTH, POT_LJ><<<g, b, sh, st>>>(Vn, S->P, S->G, A); break;
      case POT_ADHESIVE: k_adj_hvp<INT_SMOOTH, POT_ADHESIVE><<<g, b, sh, st>>>(Vn, S->P, S->G, A); break;
      default: k_adj_hvp<INT_SMOOTH, POT_WCA><<<g, b, sh, st>>>(Vn, S->P, S->G, A); break;
    }
    return hipGetLastError();
  }
  const size_t sh = std::max((size_t)(S->split - 1) * 2 * WAVE * sizeof(double), rows);
  switch (S->P.potential) {
    case POT_HARMONIC: k_adj_hvp<INT_VV, POT_HARMONIC><<<g, b, sh, st>>>(Vn, S->P, S->G, A); break;
    case POT_LJ: k_adj_hvp<INT_VV, POT_LJ><<<g, b, sh, st>>>(Vn, S->P, S->G, A); break;
    case POT_ADHESIVE: k_adj_hvp<INT_VV, POT_ADHESIVE><<<g, b, sh, st>>>(Vn, S->P, S->G, A); break;
    default: k_adj_hvp<INT_VV, POT_WCA><<<g, b, sh, st>>>(Vn, S->P, S->G, A); break;
  }
  return hipGetLastError();
}

hipError_t launch_adj_seg_check(dpm_set_s* S, const double* seg, int slots, int* bad) {
  k_adj_seg_check<<<grid_n(S->nv), 256, 0, S->stream>>>(S->V, S->P, seg, slots, bad);
  return hipGetLastError();
}

hipError_t launch_fadj_prep(dpm_set_s* S, const FireState* st, const double* bv, const double* pend, double* fc) {
  k_fadj_prep<<<grid_n(S->nv), 256, 0, S->stream>>>(S->V, st, bv, pend, fc);
  return hipGetLastError();
}

hipError_t launch_fadj_partials(dpm_set_s* S, const FireState* st, const double* slot, const double* ax,
                                const double* bv, const double* hw) {
  k_fadj_partials<<<grid_n((long)S->V.ntiles * WAVE), 256, 0, S->stream>>>(S->V, st, slot, ax, bv, hw, S->d_part);
  return hipGetLastError();
}

hipError_t launch_fadj_update(dpm_set_s* S, const FireState* st, const double* slot, double* ax, double* bv,
                              double* pend, const double* hw) {
  k_fadj_update<<<grid_n((long)S->V.ntiles * WAVE), 256, 0, S->stream>>>(S->V, st, slot, ax, bv, pend, hw,
                                                                         S->d_part);
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

hipError_t launch_copy_segs(const CopySegs& segs, hipStream_t st) {
  long most = 0;
  for (int q = 0; q < segs.n; q++) most = segs.words[q] > most ? segs.words[q] : most;
  if (most <= 0) return hipSuccess;
  const long blocks = (most + 255) / 256;
  k_copy_segs<<<(unsigned)(blocks < 2048 ? blocks : 2048), 256, 0, st>>>(segs);
  return hipGetLastError();
}

hipError_t launch_axpy(double* y, const double* x, double alpha, long n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  k_axpy<<<grid_n(n), 256, 0, st>>>(y, x, alpha, n);
  return hipGetLastError();
}

hipError_t launch_copy_d(double* dst, const double* src, long n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  k_copy_d<<<grid_n(n), 256, 0, st>>>(dst, src, n);
  return hipGetLastError();
}

}  // namespace dpm
