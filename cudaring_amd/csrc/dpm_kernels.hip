// dpm_kernels.hip -- gfx950 kernels of the DPM2D vertex engine.
//
// Work decomposition: vertices are cut into ring-aligned wave tiles of at most
// 64 vertices (whole rings, one system per tile).  One wavefront owns a tile:
// ring perimeters, areas and centres are segmented wave reductions, ring
// neighbours (prev/next vertex, their angles and rest lengths) come from lane
// shuffles, so the reference's three per-ring/per-vertex shape kernels collapse
// into one pass with no second read of the positions.  The interparticle force
// is a deterministic gather over an ELL (column-major) int32 neighbour list.
#include "dpm_internal.h"

namespace dpm {

#define WAVE 64
#define BLOCK 256

__device__ __forceinline__ double shfl_d(double v, int src) { return __shfl(v, src, WAVE); }
__device__ __forceinline__ int shfl_i(int v, int src) { return __shfl(v, src, WAVE); }

// segmented inclusive scan over lanes [segS, lane], broadcast of the segment total
__device__ __forceinline__ double seg_sum(double v, int lane, int segS, int segE) {
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    double t = __shfl_up(v, d, WAVE);
    if (lane - d >= segS) v += t;
  }
  return __shfl(v, segE, WAVE);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, WAVE);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v = fmax(v, __shfl_xor(v, d, WAVE));
  return v;
}

// Lane context of a ring-aligned tile.
struct Lane {
  int lane, tile, v0, v, r, s, segS, segE, nl, pl, n;
  bool act;
};

__device__ __forceinline__ bool lane_ctx(const View& V, Lane& L) {
  L.lane = threadIdx.x & (WAVE - 1);
  L.tile = (blockIdx.x * BLOCK + threadIdx.x) / WAVE;
  if (L.tile >= V.ntiles) return false;
  L.v0 = V.tile_v0[L.tile];
  const int v1 = V.tile_v1[L.tile];
  L.s = V.tile_sys[L.tile];
  L.act = (L.v0 + L.lane) < v1;
  L.v = L.act ? L.v0 + L.lane : L.v0;
  L.r = V.vring[L.v];
  const int first = V.rfirst[L.r];
  L.n = V.rnv[L.r];
  L.segS = L.act ? first - L.v0 : L.lane;
  const int segN = L.act ? L.n : 1;
  L.segE = L.segS + segN - 1;
  L.nl = (L.lane < L.segE) ? L.lane + 1 : L.segS;
  L.pl = (L.lane > L.segS) ? L.lane - 1 : L.segE;
  return true;
}

// Ring geometry + shape force for one lane.  Restates kernelCalcParticleShape
// (cudaKernel.cuh:1571-1606), calcParticlePos (365-388) and
// kernelCalcShapeForceEnergy (474-529) with the area/perimeter/bending terms of
// 429-471.  Writes theta, length and the ring area/perimeter/centre.
struct ShapeOut {
  double fx, fy, e;
  double x, y, xn, yn, xp, yp;  // raw positions of this, next and previous vertex
};

__device__ __forceinline__ ShapeOut ring_shape(const View& V, const Params& P, const Lane& L, double Lx,
                                               double Ly, bool write) {
  ShapeOut o;
  const double x = V.x[L.v], y = V.y[L.v];
  const double xn = shfl_d(x, L.nl), yn = shfl_d(y, L.nl);
  const double xp = shfl_d(x, L.pl), yp = shfl_d(y, L.pl);
  const double xf = shfl_d(x, L.segS), yf = shfl_d(y, L.segS);
  o.x = x; o.y = y; o.xn = xn; o.yn = yn; o.xp = xp; o.yp = yp;
  // segments of the unwrapped walk
  const double dnx = pbc(xn - x, Lx), dny = pbc(yn - y, Ly);
  const double psx = -pbc(xp - x, Lx), psy = -pbc(yp - y, Ly);
  const double theta = atan2(dnx * psy - dny * psx, dnx * psx + dny * psy);  // calcAngle 423-427
  const double seglen = sqrt(dnx * dnx + dny * dny);
  const double ux = pbc(x - xf, Lx), uy = pbc(y - yf, Ly);
  const double shoe = L.act ? (ux * dny - dnx * uy) : 0.0;  // shoelace: cur x next = cur x delta
  const double perim = seg_sum(L.act ? seglen : 0.0, L.lane, L.segS, L.segE);
  const double shoeSum = seg_sum(shoe, L.lane, L.segS, L.segE);
  const double sux = seg_sum(ux, L.lane, L.segS, L.segE);
  const double suy = seg_sum(uy, L.lane, L.segS, L.segE);
  const double nd = (double)L.n;
  const double area = fabs(shoeSum) * 0.5;
  const double comx = xf + sux / nd, comy = yf + suy / nd;
  if (write && L.act) {
    V.th[L.v] = theta;
    V.len[L.v] = seglen;
    if (L.lane == L.segS) {
      V.area[L.r] = area;
      V.perim[L.r] = perim;
      V.cx[L.r] = comx;
      V.cy[L.r] = comy;
    }
  }
  // positions relative to the ring centre (getRelativeVertexPos 273-279)
  const double rvx = pbc(x - comx, Lx), rvy = pbc(y - comy, Ly);
  const double rnx = pbc(xn - comx, Lx), rny = pbc(yn - comy, Ly);
  const double rpx = pbc(xp - comx, Lx), rpy = pbc(yp - comy, Ly);
  const double a0 = V.a0[L.r];
  // area term (calcAreaForceEnergy 429-435)
  const double deltaA = (area / a0) - 1.;
  const double gmA = P.ea * deltaA / a0;
  double fx = 0.5 * gmA * (rpy - rny);
  double fy = 0.5 * gmA * (rnx - rpx);
  double e = (0.5 * P.ea * deltaA * deltaA) / nd;
  // perimeter term (calcPerimeterForceEnergy 437-447 / FENE 449-455)
  const double nsx = rnx - rvx, nsy = rny - rvy;
  const double qsx = rvx - rpx, qsy = rvy - rpy;
  const double pLen = sqrt(qsx * qsx + qsy * qsy);
  const double nLen = sqrt(nsx * nsx + nsy * nsy);
  const double l0v = V.l0[L.v];
  const double l0p = shfl_d(l0v, L.pl);
  if (P.monomer == MONO_HARMONIC) {
    const double pDelta = (pLen / l0p) - 1.;
    const double nDelta = (nLen / l0v) - 1.;
    fx += P.el * ((nDelta * (rnx - rvx) / (l0v * nLen)) - (pDelta * (rvx - rpx) / (l0p * pLen)));
    fy += P.el * ((nDelta * (rny - rvy) / (l0v * nLen)) - (pDelta * (rvy - rpy) / (l0p * pLen)));
    e += (0.5 * P.el * pDelta * pDelta);
  } else {
    const double c = 2 * P.el * P.stiff * P.extSq;
    fx += c * ((nLen / (P.extSq * l0v * l0v - nLen * nLen)) * (rnx - rvx) / nLen -
               (pLen / (P.extSq * l0p * l0p - pLen * pLen)) * (rvx - rpx) / pLen);
    fy += c * ((nLen / (P.extSq * l0v * l0v - nLen * nLen)) * (rny - rvy) / nLen -
               (pLen / (P.extSq * l0p * l0p - pLen * pLen)) * (rvy - rpy) / pLen);
    e += P.el * P.stiff * P.extSq * log(1 - pLen * pLen / (P.extSq * l0p * l0p));
  }
  // bending term (calcBendingForceEnergy 457-471)
  const double D = theta - V.th0[L.v];
  const double Dp = shfl_d(D, L.pl), Dn = shfl_d(D, L.nl);
  const double preVar = (D - Dp) / (qsx * qsx + qsy * qsy);
  const double nextVar = (D - Dn) / (nsx * nsx + nsy * nsy);
  fx += P.eb * (preVar * qsy + nextVar * nsy);
  fy += P.eb * (preVar * (-qsx) + nextVar * (-nsx));
  e += (0.5 * P.eb * D * D);
  o.fx = fx; o.fy = fy; o.e = e;
  return o;
}

// -------------------------------------------------------------- step ---
template <int INTER, int POT, int MODE>
__global__ __launch_bounds__(BLOCK) void k_step(View V, Params P, StepArgs A) {
  Lane L;
  if (!lane_ctx(V, L)) return;
  const double Lx = V.bx[L.s], Ly = V.by[L.s];
  ShapeOut so = ring_shape(V, P, L, Lx, Ly, true);
  double fx = so.fx, fy = so.fy, e = so.e;
  const double x = so.x, y = so.y;
  const double radv = V.rad[L.v];
  double es = 0.0;
  if constexpr (INTER == INT_VV) {
    // kernelCalcVertexInteraction, cudaKernel.cuh:690-722
    const int c = L.act ? V.ncnt[L.v] : 0;
    for (int k = 0; k < c; k++) {
      const int j = V.nbr[(long)k * V.nv + L.v];
      e += pair_force<POT>(P, pbc(x - V.x[j], Lx), pbc(y - V.y[j], Ly), radv + V.rad[j], fx, fy);
    }
  } else if constexpr (INTER == INT_SMOOTH) {
    // kernelCalcSmoothInteraction, cudaKernel.cuh:922-975, as a gather over three roles
    const int ppl = shfl_i(L.pl, L.pl);
    const double xpp = shfl_d(x, ppl), ypp = shfl_d(y, ppl);
    const double radn = shfl_d(radv, L.nl);
    const int c = L.act ? V.ncnt[L.v] : 0;
    for (int k = 0; k < c; k++) {
      const int u = V.nbr[(long)k * V.nv + L.v];
      const double xu = V.x[u], yu = V.y[u], ru = V.rad[u];
      // previous and second previous vertex of u in its ring (getPreviousId 233-243)
      const int ru_ring = V.vring[u];
      const int uf = V.rfirst[ru_ring], un = V.rnv[ru_ring];
      const int p1 = (u == uf) ? uf + un - 1 : u - 1;
      const int p2 = (p1 == uf) ? uf + un - 1 : p1 - 1;
      const double radSum = radv + ru;
      es += smooth_pair<POT, ROLE_THIS>(P, Lx, Ly, x, y, xu, yu, V.x[p1], V.y[p1], V.x[p2], V.y[p2], radSum, fx, fy);
      es += smooth_pair<POT, ROLE_OTHER>(P, Lx, Ly, xu, yu, x, y, so.xp, so.yp, xpp, ypp, radSum, fx, fy);
    }
    const int vn = L.v0 + L.nl;
    const int cn = L.act ? V.ncnt[vn] : 0;
    for (int k = 0; k < cn; k++) {
      const int u = V.nbr[(long)k * V.nv + vn];
      const double radSum = V.rad[u] + radn;
      smooth_pair<POT, ROLE_PREV>(P, Lx, Ly, V.x[u], V.y[u], so.xn, so.yn, x, y, so.xp, so.yp, radSum, fx, fy);
    }
    const double pe = seg_sum(L.act ? es : 0.0, L.lane, L.segS, L.segE);
    if (L.act) {
      V.es[L.v] = es;
      if (L.lane == L.segS) V.pe[L.r] = pe;
    }
  } else {
    // kernelCalcAllToAllVertexInteraction, cudaKernel.cuh:725-757 (test mode)
    if (L.act) {
      const int b0 = V.svoff[L.s], b1 = V.svoff[L.s + 1];
      for (int j = b0; j < b1; j++) {
        if (V.vring[j] == L.r) continue;
        e += pair_force<POT>(P, pbc(x - V.x[j], Lx), pbc(y - V.y[j], Ly), radv + V.rad[j], fx, fy);
      }
    }
  }
  // active self propulsion, addSelfPropulsion DPM2D.cu:1995-2035
  if (P.particle == PART_ACTIVE) {
    const double amplitude = (P.taup != 0) ? sqrt(2.0 * P.dt / P.taup) : 0.0;
    double z0, z1;
    gauss_pair(P.seed, (uint32_t)L.s, TAG_ACTIVE, A.ctr[1] - 1, (uint32_t)(L.r - V.sroff[L.s]), z0, z1);
    double a = V.ang[L.r];
    a += amplitude * z0;
    a = a + kPI;
    a = a - 2.0 * kPI * floor(a / (2.0 * kPI));
    a = a - kPI;
    fx += P.driving * ((1 - 0) * cos(a) + 0 * sin(a));
    fy += P.driving * ((1 - 1) * cos(a) + 1 * sin(a));
    // every lane of the ring has read the old angle before the head lane writes
    __builtin_amdgcn_wave_barrier();
    if (L.act && L.lane == L.segS) V.ang[L.r] = a;
  }
  if (!L.act) {
    if constexpr ((MODE & MODE_FIRE) != 0) {
      // inactive lanes still take part in the tile reductions below
    } else {
      return;
    }
  }
  double vx = 0, vy = 0;
  if constexpr (MODE != MODE_FORCE) {
    vx = V.vx[L.v];
    vy = V.vy[L.v];
  }
  if constexpr ((MODE & MODE_LANGEVIN) != 0) {
    // Langevin::updateThermalVel, Simulator.cu:54-77: one Gaussian pair per ring
    double z0, z1;
    gauss_pair(P.seed, (uint32_t)L.s, TAG_THERMAL, A.ctr[0] - 1, (uint32_t)(L.r - V.sroff[L.s]), z0, z1);
    fx += P.noise * z0 - P.gamma * vx;
    fy += P.noise * z1 - P.gamma * vy;
  }
  if constexpr ((MODE & MODE_FIRE) != 0) {
    const FireState fs = A.fire[L.s];
    const double h = 0.5 * fs.dt;
    if (L.act && !fs.converged) {
      vx += h * fx;
      vy += h * fy;
    }
    const double pvf = L.act ? (vx * fx + vy * fy) : 0.0;
    const double f2 = L.act ? fmax(fx * fx, fy * fy) : -1.0;
    const double tP = wave_sum(pvf);
    const double tM = wave_max(f2);
    if (L.lane == 0) {
      A.part[L.tile * 4 + 0] = tP;
      A.part[L.tile * 4 + 3] = tM;
    }
    if (!L.act) return;
    V.fx[L.v] = fx; V.fy[L.v] = fy; V.ev[L.v] = e;
    V.vx[L.v] = vx; V.vy[L.v] = vy;
    return;
  }
  V.fx[L.v] = fx;
  V.fy[L.v] = fy;
  V.ev[L.v] = e;
  if constexpr (MODE == MODE_FORCE) return;
  if constexpr ((MODE & MODE_KICK) != 0) {
    vx += A.h * fx;
    vy += A.h * fy;
  }
  if constexpr ((MODE & MODE_START) != 0) {
    vx += A.h * fx;
    vy += A.h * fy;
    const double xn = x + A.dt * vx, yn = y + A.dt * vy;
    A.xout[L.v] = xn;
    A.yout[L.v] = yn;
    const double ddx = pbc(xn - V.lx[L.v], Lx), ddy = pbc(yn - V.ly[L.v], Ly);
    if (3 * sqrt(ddx * ddx + ddy * ddy) > P.cutoff) atomicOr(&A.flagNext[L.s], 1);
  }
  V.vx[L.v] = vx;
  V.vy[L.v] = vy;
}

// geometry only (calcParticleShape + calcParticlePositions)
__global__ __launch_bounds__(BLOCK) void k_shape(View V, Params P) {
  Lane L;
  if (!lane_ctx(V, L)) return;
  ring_shape(V, P, L, V.bx[L.s], V.by[L.s], true);
}

// first half step of a velocity-Verlet call: v += h F; x' = x + dt v; displacement test
__global__ void k_start(View V, Params P, double h, double dt, double* xo, double* yo, int* flagNext) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V.nv) return;
  const int s = V.vsys[v];
  double vx = V.vx[v] + h * V.fx[v];
  double vy = V.vy[v] + h * V.fy[v];
  const double xn = V.x[v] + dt * vx, yn = V.y[v] + dt * vy;
  V.vx[v] = vx; V.vy[v] = vy;
  xo[v] = xn; yo[v] = yn;
  const double ddx = pbc(xn - V.lx[v], V.bx[s]), ddy = pbc(yn - V.ly[v], V.by[s]);
  if (3 * sqrt(ddx * ddx + ddy * ddy) > P.cutoff) atomicOr(&flagNext[s], 1);
}

// --------------------------------------------------------- neighbours ---
// Prologue of a step: bump RNG call counters, latch rebuild requests, clear the
// cells of systems that rebuild.
__global__ void k_prologue(int nsys, int* flagCur, const int* flagNext, unsigned long long* ctr, int bumpT,
                           int bumpA, CellGrid G) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) {
    ctr[0] += bumpT;
    ctr[1] += bumpA;
  }
  if (t < nsys) flagCur[t] = flagNext[t];
  if (t < G.ncells && flagNext[G.cellSys[t]]) G.count[t] = 0;
}

__device__ __forceinline__ void cell_of(double x, double y, double Lx, double Ly, int ncx, int ncy, int& cx,
                                        int& cy) {
  double wx = x - Lx * floor(x / Lx), wy = y - Ly * floor(y / Ly);
  cx = (int)(wx * ncx / Lx);
  cy = (int)(wy * ncy / Ly);
  cx = cx < 0 ? 0 : (cx >= ncx ? ncx - 1 : cx);
  cy = cy < 0 ? 0 : (cy >= ncy ? ncy - 1 : cy);
}

__global__ void k_bin(View V, const int* flagCur, CellGrid G, int* err) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V.nv) return;
  const int s = V.vsys[v];
  if (!flagCur[s] || G.brute[s]) return;
  int cx, cy;
  cell_of(V.x[v], V.y[v], V.bx[s], V.by[s], G.ncx[s], G.ncy[s], cx, cy);
  const int c = G.coff[s] + cy * G.ncx[s] + cx;
  const int slot = atomicAdd(&G.count[c], 1);
  if (slot < G.cap) G.slots[(long)c * G.cap + slot] = v;
  else atomicOr(err, ERR_CELL_OVERFLOW);
}

// insert j into the ascending ELL row of v (count c entries present)
__device__ __forceinline__ void row_insert(int* nbr, long nv, int v, int c, int j) {
  int k = c;
  while (k > 0) {
    const int prev = nbr[(long)(k - 1) * nv + v];
    if (prev <= j) break;
    nbr[(long)k * nv + v] = prev;
    k--;
  }
  nbr[(long)k * nv + v] = j;
}

// kernelCalcNeighborList predicate (cudaKernel.cuh:1735-1763): j in list(i) iff rings differ
// and -(1 - d/radSum) < cutDistance.  Rows are kept ascending as in the reference.
__global__ void k_build(View V, Params P, const int* flagCur, int* flagNext, int* updateCount, CellGrid G,
                        int* err, int* maxcnt) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < V.nsys) {
    flagNext[v] = 0;
    if (flagCur[v]) updateCount[v] += 1;
  }
  if (v >= V.nv) return;
  const int s = V.vsys[v];
  if (!flagCur[s]) return;
  const double Lx = V.bx[s], Ly = V.by[s];
  const double x = V.x[v], y = V.y[v], rv = V.rad[v];
  const int r = V.vring[v];
  const int W = V.W;
  int c = 0;
  if (G.brute[s]) {
    const int b0 = V.svoff[s], b1 = V.svoff[s + 1];
    for (int j = b0; j < b1; j++) {
      if (V.vring[j] == r) continue;
      const double dx = pbc(x - V.x[j], Lx), dy = pbc(y - V.y[j], Ly);
      const double d = sqrt(dx * dx + dy * dy);
      if (-(1 - d / (rv + V.rad[j])) < P.cutDistance) {
        if (c < W) V.nbr[(long)c * V.nv + v] = j;
        c++;
      }
    }
  } else {
    const int ncx = G.ncx[s], ncy = G.ncy[s];
    int cx, cy;
    cell_of(x, y, Lx, Ly, ncx, ncy, cx, cy);
    for (int oy = -1; oy <= 1; oy++) {
      const int yy = (cy + oy + ncy) % ncy;
      for (int ox = -1; ox <= 1; ox++) {
        const int xx = (cx + ox + ncx) % ncx;
        const int cell = G.coff[s] + yy * ncx + xx;
        const int m = min(G.count[cell], G.cap);
        for (int q = 0; q < m; q++) {
          const int j = G.slots[(long)cell * G.cap + q];
          if (V.vring[j] == r) continue;
          const double dx = pbc(x - V.x[j], Lx), dy = pbc(y - V.y[j], Ly);
          const double d = sqrt(dx * dx + dy * dy);
          if (-(1 - d / (rv + V.rad[j])) < P.cutDistance) {
            if (c < W) row_insert(V.nbr, V.nv, v, c, j);
            c++;
          }
        }
      }
    }
  }
  if (c > W) atomicOr(err, ERR_LIST_OVERFLOW);
  atomicMax(maxcnt, c);
  V.ncnt[v] = c < W ? c : W;
  V.lx[v] = x;
  V.ly[v] = y;
}

__global__ void k_fill_int(int* p, int value, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = value;
}

__global__ void k_copy_lastpos(View V) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V.nv) return;
  V.lx[v] = V.x[v];
  V.ly[v] = V.y[v];
}

// ---------------------------------------------------------- reductions ---
// One block per system, fixed striding and tree order (deterministic).
// out[s*8 + q]: 0 sum e_v, 1 sum pe, 2 sum v^2, 3 max F_c^2, 4 sum F^2,
//               5 sum area, 6 sum a0, 7 sum vertex-disk area r^2 (n/2 - 1)
__global__ __launch_bounds__(BLOCK) void k_reduce(View V, double* out) {
  const int s = blockIdx.x;
  __shared__ double sh[8][BLOCK / WAVE];
  double acc[8] = {0, 0, 0, -1, 0, 0, 0, 0};
  const int b0 = V.svoff[s], b1 = V.svoff[s + 1];
  for (int v = b0 + threadIdx.x; v < b1; v += BLOCK) {
    acc[0] += V.ev[v];
    acc[2] += V.vx[v] * V.vx[v] + V.vy[v] * V.vy[v];
    acc[3] = fmax(acc[3], fmax(V.fx[v] * V.fx[v], V.fy[v] * V.fy[v]));
    acc[4] += V.fx[v] * V.fx[v] + V.fy[v] * V.fy[v];
  }
  const int r0 = V.sroff[s], r1 = V.sroff[s + 1];
  for (int r = r0 + threadIdx.x; r < r1; r += BLOCK) {
    acc[1] += V.pe[r];
    acc[5] += V.area[r];
    acc[6] += V.a0[r];
    const double rr = V.rad[V.rfirst[r]];
    acc[7] += rr * rr * (0.5 * V.rnv[r] - 1);
  }
  const int w = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
#pragma unroll
  for (int q = 0; q < 8; q++) {
    double t = (q == 3) ? wave_max(acc[q]) : wave_sum(acc[q]);
    if (lane == 0) sh[q][w] = t;
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    const int q = threadIdx.x;
    double t = sh[q][0];
    for (int i = 1; i < BLOCK / WAVE; i++) t = (q == 3) ? fmax(t, sh[q][i]) : t + sh[q][i];
    out[s * 8 + q] = t;
  }
}

// Langevin::injectKineticEnergy, Simulator.cu:32-52: per-ring Gaussian velocity
__global__ void k_inject(View V, Params P, double amplitude, unsigned long long call) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V.nv) return;
  const int r = V.vring[v], s = V.vsys[v];
  double z0, z1;
  gauss_pair(P.seed, (uint32_t)s, TAG_INJECT, call, (uint32_t)(r - V.sroff[s]), z0, z1);
  V.vx[v] = amplitude * z0;
  V.vy[v] = amplitude * z1;
}

// subtract the per-system mean velocity; sums come from a dedicated reduction
__global__ __launch_bounds__(BLOCK) void k_sub_mean(View V) {
  const int s = blockIdx.x;
  __shared__ double sh[2][BLOCK / WAVE];
  __shared__ double mean[2];
  const int b0 = V.svoff[s], b1 = V.svoff[s + 1];
  double ax = 0, ay = 0;
  for (int v = b0 + threadIdx.x; v < b1; v += BLOCK) { ax += V.vx[v]; ay += V.vy[v]; }
  const int w = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
  ax = wave_sum(ax); ay = wave_sum(ay);
  if (lane == 0) { sh[0][w] = ax; sh[1][w] = ay; }
  __syncthreads();
  if (threadIdx.x < 2) {
    double t = 0;
    for (int i = 0; i < BLOCK / WAVE; i++) t += sh[threadIdx.x][i];
    mean[threadIdx.x] = t / (b1 - b0);
  }
  __syncthreads();
  for (int v = b0 + threadIdx.x; v < b1; v += BLOCK) { V.vx[v] -= mean[0]; V.vy[v] -= mean[1]; }
}

// v *= scale[s]
__global__ void k_scale_vel(View V, const double* scale) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V.nv) return;
  const double sc = scale[V.vsys[v]];
  V.vx[v] *= sc;
  V.vy[v] *= sc;
}

// scaleVertices, DPM2D.cu:1802-1811 + kernelScaleVertexPositions, cudaKernel.cuh:1644-1658
__global__ void k_scale_vertices(View V, const double* scale) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < V.nv) {
    const int r = V.vring[v];
    const double sc = scale[V.vsys[v]];
    V.x[v] += (sc - 1) * (V.x[v] - V.cx[r]);
    V.y[v] += (sc - 1) * (V.y[v] - V.cy[r]);
    V.l0[v] *= sc;
    V.rad[v] *= sc;
  }
  if (v < V.np) {
    const double sc = scale[V.rsys[v]];
    V.a0[v] *= sc * sc;
    V.area[v] *= sc * sc;
  }
}

// --------------------------------------------------------------- FIRE ---
// FIRE::minimizerVertexLoop, FIRE.cu:397-411, split at its two global reductions.
// fire_begin: checkFIREParams (338-371) + half kick + partial |v|^2, |F|^2.
__global__ __launch_bounds__(BLOCK) void k_fire_begin(View V, FireParams fp, const FireState* fin, FireState* fout,
                                                      const double* part, double* partOut, const int* stile) {
  Lane L;
  if (!lane_ctx(V, L)) return;
  FireState fs = fin[L.s];
  // P = v.F over the system, summed tile by tile in fixed order
  double Psum = 0;
  for (int t = stile[L.s]; t < stile[L.s + 1]; t++) Psum += part[t * 4 + 0];
  bool reset = false;
  if (!fs.converged) {
    fs.P = Psum;
    if (Psum > 0) {
      fs.numStep++;
      if (fs.numStep > fp.minStep) {
        fs.dt = fmin(fs.dt * fp.f_inc, fp.dt_max);
        fs.a *= fp.f_a;
      }
    } else {
      fs.numStep = 0;
      fs.dt = fmax(fs.dt * fp.f_dec, fp.dt_max / 10);
      fs.a = fp.a_start;
      reset = true;
    }
    fs.iters++;
  }
  if (L.tile == stile[L.s] && L.lane == 0) fout[L.s] = fs;
  double vx = 0, vy = 0, fx = 0, fy = 0;
  if (L.act) {
    vx = V.vx[L.v]; vy = V.vy[L.v]; fx = V.fx[L.v]; fy = V.fy[L.v];
    if (!fs.converged) {
      if (reset) {
        const double hb = 0.5 * fs.dt;
        V.x[L.v] -= hb * vx;
        V.y[L.v] -= hb * vy;
        vx = 0; vy = 0;
      }
      const double h = 0.5 * fs.dt;
      vx += h * fx;
      vy += h * fy;
      V.vx[L.v] = vx;
      V.vy[L.v] = vy;
    }
  }
  const double tv = wave_sum(vx * vx + vy * vy);
  const double tf = wave_sum(fx * fx + fy * fy);
  if (L.lane == 0) {
    partOut[L.tile * 4 + 1] = tv;
    partOut[L.tile * 4 + 2] = tf;
  }
}

// fire_mid: bendVertexVelocityTowardsForce (374-394) + drift + displacement test
__global__ __launch_bounds__(BLOCK) void k_fire_mid(View V, Params P, const FireState* fcur, const double* part,
                                                    const int* stile, double* xo, double* yo, int* flagNext) {
  Lane L;
  if (!lane_ctx(V, L)) return;
  if (!L.act) return;
  const FireState fs = fcur[L.s];
  double vx = V.vx[L.v], vy = V.vy[L.v];
  const double x = V.x[L.v], y = V.y[L.v];
  if (fs.converged) {
    xo[L.v] = x;
    yo[L.v] = y;
    return;
  }
  double vv = 0, ff = 0;
  for (int t = stile[L.s]; t < stile[L.s + 1]; t++) { vv += part[t * 4 + 1]; ff += part[t * 4 + 2]; }
  if (ff > 0) {
    const double ratio = sqrt(vv / ff);
    const double fa = fs.a;
    vx = (1 - fa) * vx + fa * V.fx[L.v] * ratio;
    vy = (1 - fa) * vy + fa * V.fy[L.v] * ratio;
  }
  const double xn = x + fs.dt * vx, yn = y + fs.dt * vy;
  V.vx[L.v] = vx;
  V.vy[L.v] = vy;
  xo[L.v] = xn;
  yo[L.v] = yn;
  const double Lx = V.bx[L.s], Ly = V.by[L.s];
  const double ddx = pbc(xn - V.lx[L.v], Lx), ddy = pbc(yn - V.ly[L.v], Ly);
  if (3 * sqrt(ddx * ddx + ddy * ddy) > P.cutoff) atomicOr(&flagNext[L.s], 1);
}

// per-tile partials of P = v.F and max F_c^2 for the current state (start of a FIRE call)
__global__ __launch_bounds__(BLOCK) void k_fire_partials(View V, double* part) {
  Lane L;
  if (!lane_ctx(V, L)) return;
  double pvf = 0, f2 = -1;
  if (L.act) {
    const double vx = V.vx[L.v], vy = V.vy[L.v], fx = V.fx[L.v], fy = V.fy[L.v];
    pvf = vx * fx + vy * fy;
    f2 = fmax(fx * fx, fy * fy);
  }
  const double tP = wave_sum(pvf);
  const double tM = wave_max(f2);
  if (L.lane == 0) {
    part[L.tile * 4 + 0] = tP;
    part[L.tile * 4 + 3] = tM;
  }
}

// --------------------------------------------------------- launchers ---
static inline int grid_v(long n) { return (int)((n + 255) / 256); }
static inline int grid_t(int ntiles) { return (ntiles + 3) / 4; }

template <int INTER, int POT>
static hipError_t step_mode(dpm_set_s* S, int mode, const StepArgs& A) {
  const int g = grid_t(S->V.ntiles);
  switch (mode) {
    case MODE_FORCE: k_step<INTER, POT, MODE_FORCE><<<g, BLOCK, 0, S->stream>>>(S->V, S->P, A); break;
    case MODE_KICK: k_step<INTER, POT, MODE_KICK><<<g, BLOCK, 0, S->stream>>>(S->V, S->P, A); break;
    case MODE_KICK | MODE_LANGEVIN:
      k_step<INTER, POT, MODE_KICK | MODE_LANGEVIN><<<g, BLOCK, 0, S->stream>>>(S->V, S->P, A);
      break;
    case MODE_KICK | MODE_START:
      k_step<INTER, POT, MODE_KICK | MODE_START><<<g, BLOCK, 0, S->stream>>>(S->V, S->P, A);
      break;
    case MODE_KICK | MODE_LANGEVIN | MODE_START:
      k_step<INTER, POT, MODE_KICK | MODE_LANGEVIN | MODE_START><<<g, BLOCK, 0, S->stream>>>(S->V, S->P, A);
      break;
    case MODE_FIRE: k_step<INTER, POT, MODE_FIRE><<<g, BLOCK, 0, S->stream>>>(S->V, S->P, A); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int INTER>
static hipError_t step_pot(dpm_set_s* S, int mode, const StepArgs& A) {
  switch (S->P.potential) {
    case POT_HARMONIC: return step_mode<INTER, POT_HARMONIC>(S, mode, A);
    case POT_LJ: return step_mode<INTER, POT_LJ>(S, mode, A);
    case POT_ADHESIVE: return step_mode<INTER, POT_ADHESIVE>(S, mode, A);
    case POT_WCA: return step_mode<INTER, POT_WCA>(S, mode, A);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_step(dpm_set_s* S, int mode, const StepArgs& A) {
  switch (S->P.interaction) {
    case INT_VV: return step_pot<INT_VV>(S, mode, A);
    case INT_SMOOTH: return step_pot<INT_SMOOTH>(S, mode, A);
    case INT_ALL: {
      if (mode != MODE_FORCE) return hipErrorInvalidValue;
      const int g = grid_t(S->V.ntiles);
      switch (S->P.potential) {
        case POT_HARMONIC: k_step<INT_ALL, POT_HARMONIC, MODE_FORCE><<<g, BLOCK, 0, S->stream>>>(S->V, S->P, A); break;
        case POT_LJ: k_step<INT_ALL, POT_LJ, MODE_FORCE><<<g, BLOCK, 0, S->stream>>>(S->V, S->P, A); break;
        case POT_ADHESIVE: k_step<INT_ALL, POT_ADHESIVE, MODE_FORCE><<<g, BLOCK, 0, S->stream>>>(S->V, S->P, A); break;
        case POT_WCA: k_step<INT_ALL, POT_WCA, MODE_FORCE><<<g, BLOCK, 0, S->stream>>>(S->V, S->P, A); break;
      }
      return hipGetLastError();
    }
  }
  return hipErrorInvalidValue;
}

hipError_t launch_shape_only(dpm_set_s* S) {
  k_shape<<<grid_t(S->V.ntiles), BLOCK, 0, S->stream>>>(S->V, S->P);
  return hipGetLastError();
}

hipError_t launch_start(dpm_set_s* S, double h, double dt) {
  const int nb = 1 - S->cur;
  k_start<<<grid_v(S->nv), 256, 0, S->stream>>>(S->V, S->P, h, dt, S->pos[nb][0], S->pos[nb][1], S->d_flagNext);
  return hipGetLastError();
}

hipError_t launch_prologue(dpm_set_s* S, int bumpThermal, int bumpActive) {
  long n = S->nsys > S->G.ncells ? S->nsys : S->G.ncells;
  if (n < 1) n = 1;
  k_prologue<<<grid_v(n), 256, 0, S->stream>>>((int)S->nsys, S->d_flagCur, S->d_flagNext, S->d_ctr, bumpThermal,
                                                bumpActive, S->G);
  return hipGetLastError();
}

hipError_t launch_bin(dpm_set_s* S) {
  k_bin<<<grid_v(S->nv), 256, 0, S->stream>>>(S->V, S->d_flagCur, S->G, S->d_err);
  return hipGetLastError();
}

hipError_t launch_build(dpm_set_s* S) {
  long n = S->nv > S->nsys ? S->nv : S->nsys;
  k_build<<<grid_v(n), 256, 0, S->stream>>>(S->V, S->P, S->d_flagCur, S->d_flagNext, S->d_updateCount, S->G,
                                             S->d_err, S->d_maxcnt);
  return hipGetLastError();
}

hipError_t launch_fill_int(int* p, int value, long n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  k_fill_int<<<grid_v(n), 256, 0, st>>>(p, value, n);
  return hipGetLastError();
}

hipError_t launch_reduce(dpm_set_s* S) {
  k_reduce<<<(int)S->nsys, BLOCK, 0, S->stream>>>(S->V, S->d_red);
  return hipGetLastError();
}

hipError_t launch_inject(dpm_set_s* S, double amplitude, unsigned long long call) {
  k_inject<<<grid_v(S->nv), 256, 0, S->stream>>>(S->V, S->P, amplitude, call);
  return hipGetLastError();
}

hipError_t launch_sub_mean(dpm_set_s* S) {
  k_sub_mean<<<(int)S->nsys, BLOCK, 0, S->stream>>>(S->V);
  return hipGetLastError();
}

hipError_t launch_scale_vel(dpm_set_s* S) {
  k_scale_vel<<<grid_v(S->nv), 256, 0, S->stream>>>(S->V, S->d_scale);
  return hipGetLastError();
}

hipError_t launch_scale_vertices(dpm_set_s* S) {
  long n = S->nv > S->np ? S->nv : S->np;
  k_scale_vertices<<<grid_v(n), 256, 0, S->stream>>>(S->V, S->d_scale);
  return hipGetLastError();
}

hipError_t launch_copy_lastpos(dpm_set_s* S) {
  k_copy_lastpos<<<grid_v(S->nv), 256, 0, S->stream>>>(S->V);
  return hipGetLastError();
}

hipError_t launch_fire_begin(dpm_set_s* S) {
  k_fire_begin<<<grid_t(S->V.ntiles), BLOCK, 0, S->stream>>>(S->V, S->fp, S->d_fire[S->fireCur],
                                                              S->d_fire[1 - S->fireCur], S->d_part, S->d_part,
                                                              S->d_stile);
  return hipGetLastError();
}

hipError_t launch_fire_mid(dpm_set_s* S) {
  const int nb = 1 - S->cur;
  k_fire_mid<<<grid_t(S->V.ntiles), BLOCK, 0, S->stream>>>(S->V, S->P, S->d_fire[1 - S->fireCur], S->d_part,
                                                            S->d_stile, S->pos[nb][0], S->pos[nb][1],
                                                            S->d_flagNext);
  return hipGetLastError();
}

hipError_t launch_fire_partials(dpm_set_s* S) {
  k_fire_partials<<<grid_t(S->V.ntiles), BLOCK, 0, S->stream>>>(S->V, S->d_part);
  return hipGetLastError();
}

hipError_t launch_fire_step(dpm_set_s* S) {
  StepArgs A{};
  A.ctr = S->d_ctr;
  A.fire = S->d_fire[1 - S->fireCur];
  A.part = S->d_part;
  A.err = S->d_err;
  return launch_step(S, MODE_FIRE, A);
}

}  // namespace dpm
