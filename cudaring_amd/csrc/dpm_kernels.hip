// dpm_kernels.hip -- gfx950 kernels of the DPM2D vertex engine.
//
// Work decomposition.  Vertices are cut into ring-aligned tiles of at most 64
// vertices (whole rings, one system per tile).  A step launches one workgroup per
// tile with nw wavefronts (the autotuned "waves per tile"):
//   * wave 0, one lane per vertex, does the ring geometry and the shape force:
//     ring perimeters, areas and centres are segmented wave scans, ring
//     neighbours (prev/next vertex, their angles and rest lengths) come from lane
//     shuffles, so the reference's three shape kernels collapse into one pass;
//     it also draws the ring's Philox numbers;
//   * meanwhile waves 1..nw-1 share the interparticle gather over the ELL
//     neighbour list and reduce their partial forces through LDS in a fixed
//     order (with nw = 1 wave 0 does the gather itself);
//   * wave 0 applies the integrator epilogue (thermostat, kicks, drift,
//     displacement test).
// Neighbour rows are rebuilt inside the step kernel when the system's rebuild
// flag is set, and a tile's rows are all built by its own workgroup before its
// gather starts.  In the step loops the cells searched are those of the list
// positions (the positions of the system's last rebuild, binned by that rebuild:
// md_cells), so a step is one launch with no binning and no host
// synchronisation; the FIRE graph path bins the current positions in a k_bin
// launch before each step kernel.
#include "dpm_internal.h"
#include "dpm_rows.h"
#include "dpm_tile.h"

namespace dpm {

// wave-per-tile kernels (4 tiles per 256-thread block)
__device__ __forceinline__ bool wave_tile_ctx(const View& V, Lane& L) {
  const int tile = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
  if (tile >= V.ntiles) return false;
  lane_ctx(V, tile, threadIdx.x & (WAVE - 1), L);
  return true;
}

// Ring geometry + shape force for one lane (wave-uniform call).  Restates
// kernelCalcParticleShape (cudaKernel.cuh:1571-1606), calcParticlePos (365-388)
// and kernelCalcShapeForceEnergy (474-529) with the area/perimeter/bending
// terms of 429-471.  Writes theta, length and the ring area/perimeter/centre.
struct ShapeOut {
  double fx, fy, e;
  double cx, cy;  // ring centre (calcParticlePos, cudaKernel.cuh:365-388)
};

// turning angle at a vertex from its own, next and previous positions (calcAngle 423-427)
__device__ __forceinline__ double vertex_theta(const Box& B, double x, double y, double xn, double yn, double xp,
                                               double yp) {
  const double dnx = pbx(xn - x, B), dny = pby(yn - y, B);
  const double psx = -pbx(xp - x, B), psy = -pby(yp - y, B);
  return atan2(dnx * psy - dny * psx, dnx * psx + dny * psy);
}

__device__ __forceinline__ ShapeOut ring_shape(const View& V, const Params& P, const Lane& L, const Box& B,
                                               double x, double y, bool write) {
  ShapeOut o;
  double xn, yn, xp, yp, theta, seglen, area, perim, comx, comy;
  int vn = 0, vp = 0;
  if (!L.wide) {
    xn = shfl_d(x, L.nl); yn = shfl_d(y, L.nl);
    xp = shfl_d(x, L.pl); yp = shfl_d(y, L.pl);
    const double xf = shfl_d(x, L.segS), yf = shfl_d(y, L.segS);
    // segments of the reference's unwrapped walk
    const double dnx = pbx(xn - x, B), dny = pby(yn - y, B);
    const double psx = -pbx(xp - x, B), psy = -pby(yp - y, B);
    theta = atan2(dnx * psy - dny * psx, dnx * psx + dny * psy);  // calcAngle 423-427
    seglen = sqrt(dnx * dnx + dny * dny);
    // Unwrapped position relative to the first vertex: the reference walks the ring one
    // minimum-image segment at a time; here that walk is an exclusive segmented scan of the
    // segment vectors, valid for rings wider than half the box.
    const double Ix = seg_scan(L.act ? dnx : 0.0, L.lane, L.segS);
    const double Iy = seg_scan(L.act ? dny : 0.0, L.lane, L.segS);
    const double ux = L.act ? Ix - dnx : 0.0;
    const double uy = L.act ? Iy - dny : 0.0;
    const double shoe = L.act ? (ux * dny - dnx * uy) : 0.0;  // shoelace: cur x next = cur x delta
    perim = seg_sum(L.act ? seglen : 0.0, L.lane, L.segS, L.segE);
    const double shoeSum = seg_sum(shoe, L.lane, L.segS, L.segE);
    const double sux = seg_sum(ux, L.lane, L.segS, L.segE);
    const double suy = seg_sum(uy, L.lane, L.segS, L.segE);
    const double ind = 1.0 / (double)L.n;
    area = fabs(shoeSum) * 0.5;
    comx = xf + sux * ind;
    comy = yf + suy * ind;
  } else {
    // a slice of a wide ring: ring neighbours from memory (this launch reads the current positions
    // only), the ring's area, perimeter and centre from k_wide_rings, launched just before
    vn = ring_next(L.v, L.f, L.n);
    vp = ring_prev(L.v, L.f, L.n);
    xn = V.x[vn]; yn = V.y[vn];
    xp = V.x[vp]; yp = V.y[vp];
    theta = vertex_theta(B, x, y, xn, yn, xp, yp);
    const double dnx = pbx(xn - x, B), dny = pby(yn - y, B);
    seglen = sqrt(dnx * dnx + dny * dny);
    area = V.area[L.r];
    perim = V.perim[L.r];
    comx = V.cx[L.r];
    comy = V.cy[L.r];
  }
  o.cx = comx;
  o.cy = comy;
  if (write && L.act) {
    V.th[L.v] = theta;
    V.len[L.v] = seglen;
    if (L.lane == L.segS && !L.wide) {
      V.area[L.r] = area;
      V.perim[L.r] = perim;
      V.cx[L.r] = comx;
      V.cy[L.r] = comy;
    }
  }
  const double ind = 1.0 / (double)L.n;
  // positions relative to the ring centre (getRelativeVertexPos 273-279)
  const double rvx = pbx(x - comx, B), rvy = pby(y - comy, B);
  const double rnx = pbx(xn - comx, B), rny = pby(yn - comy, B);
  const double rpx = pbx(xp - comx, B), rpy = pby(yp - comy, B);
  const double a0 = V.a0[L.r];
  // shape costs: the global ea, el, eb, or per-ring values when the caller set them
  const double ea = V.kr ? V.kr[L.r] : P.ea;
  const double el = V.kr ? V.kr[V.np + L.r] : P.el;
  const double eb = V.kr ? V.kr[2 * V.np + L.r] : P.eb;
  // area term (calcAreaForceEnergy 429-435)
  const double deltaA = (area / a0) - 1.;
  const double gmA = ea * deltaA / a0;
  double fx = 0.5 * gmA * (rpy - rny);
  double fy = 0.5 * gmA * (rnx - rpx);
  double e = (0.5 * ea * deltaA * deltaA) * ind;
  // perimeter term (calcPerimeterForceEnergy 437-447 / FENE 449-455)
  const double nsx = rnx - rvx, nsy = rny - rvy;
  const double qsx = rvx - rpx, qsy = rvy - rpy;
  const double pLen = sqrt(qsx * qsx + qsy * qsy);
  const double nLen = sqrt(nsx * nsx + nsy * nsy);
  const double l0v = V.l0[L.v];
  const double l0s = shfl_d(l0v, L.pl);
  const double l0p = L.wide ? V.l0[vp] : l0s;
  if (P.monomer == MONO_HARMONIC) {
    const double pDelta = (pLen / l0p) - 1.;
    const double nDelta = (nLen / l0v) - 1.;
    const double cn = nDelta / (l0v * nLen), cp = pDelta / (l0p * pLen);
    fx += el * ((cn * (rnx - rvx)) - (cp * (rvx - rpx)));
    fy += el * ((cn * (rny - rvy)) - (cp * (rvy - rpy)));
    e += (0.5 * el * pDelta * pDelta);
  } else {
    const double c = 2 * el * P.stiff * P.extSq;
    fx += c * ((nLen / (P.extSq * l0v * l0v - nLen * nLen)) * (rnx - rvx) / nLen -
               (pLen / (P.extSq * l0p * l0p - pLen * pLen)) * (rvx - rpx) / pLen);
    fy += c * ((nLen / (P.extSq * l0v * l0v - nLen * nLen)) * (rny - rvy) / nLen -
               (pLen / (P.extSq * l0p * l0p - pLen * pLen)) * (rvy - rpy) / pLen);
    e += el * P.stiff * P.extSq * log(1 - pLen * pLen / (P.extSq * l0p * l0p));
  }
  // bending term (calcBendingForceEnergy 457-471)
  const double D = theta - V.th0[L.v];
  double Dp = shfl_d(D, L.pl), Dn = shfl_d(D, L.nl);
  if (L.wide) {  // the neighbours' angles from their own neighbours, as their lanes form them
    const int vpp = ring_prev(vp, L.f, L.n), vnn = ring_next(vn, L.f, L.n);
    Dp = vertex_theta(B, xp, yp, x, y, V.x[vpp], V.y[vpp]) - V.th0[vp];
    Dn = vertex_theta(B, xn, yn, V.x[vnn], V.y[vnn], x, y) - V.th0[vn];
  }
  const double preVar = (D - Dp) / (qsx * qsx + qsy * qsy);
  const double nextVar = (D - Dn) / (nsx * nsx + nsy * nsy);
  fx += eb * (preVar * qsy + nextVar * nsy);
  fy += eb * (preVar * (-qsx) + nextVar * (-nsx));
  e += (0.5 * eb * D * D);
  o.fx = fx; o.fy = fy; o.e = e;
  return o;
}

// The ring centre alone (calcParticlePos, cudaKernel.cuh:365-388), as ring_shape computes it,
// stored by the ring's first lane: the rigid rings' shape pass
__device__ __forceinline__ ShapeOut ring_centre(const View& V, const Lane& L, const Box& B, double x, double y) {
  ShapeOut o{0, 0, 0, 0, 0};
  const double xn = shfl_d(x, L.nl), yn = shfl_d(y, L.nl);
  const double xf = shfl_d(x, L.segS), yf = shfl_d(y, L.segS);
  const double dnx = pbx(xn - x, B), dny = pby(yn - y, B);
  const double Ix = seg_scan(L.act ? dnx : 0.0, L.lane, L.segS);
  const double Iy = seg_scan(L.act ? dny : 0.0, L.lane, L.segS);
  const double ux = L.act ? Ix - dnx : 0.0;
  const double uy = L.act ? Iy - dny : 0.0;
  const double sux = seg_sum(ux, L.lane, L.segS, L.segE);
  const double suy = seg_sum(uy, L.lane, L.segS, L.segE);
  const double ind = 1.0 / (double)L.n;
  o.cx = xf + sux * ind;
  o.cy = yf + suy * ind;
  if (L.act && L.lane == L.segS) {
    V.cx[L.r] = o.cx;
    V.cy[L.r] = o.cy;
  }
  return o;
}

// --------------------------------------------------------- neighbours ---
// cell_of, prev_pair, row_insert and build_row live in dpm_rows.h (shared with the adjoint).

// Bin the positions of flagged systems into the cells of this parity.  Thread 0 first bumps
// the RNG call counters of the step and clears the other parity's flags (they were consumed
// by the previous step).
__global__ void k_bin(View V, CellGrid G, int* count, const int* flagCur, const int* anyCur, int* flagOther,
                      int* anyOther, unsigned long long* ctr, int bumpT, int bumpA, int* err) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) {
    ctr[0] += bumpT;
    ctr[1] += bumpA;
    *anyOther = 0;
  }
  if (t < V.nsys) flagOther[t] = 0;
  if (*anyCur == 0) return;
  if (t >= V.nv) return;
  const int s = V.vsys[t];
  if (!flagCur[s] || G.brute[s]) return;
  const Box B = box_of(V, s);
  int cx, cy;
  cell_of(V.x[t], V.y[t], B, G.ncx[s], G.ncy[s], cx, cy);
  const int c = G.coff[s] + cy * G.ncx[s] + cx;
  const int slot = atomicAdd(&count[c], 1);
  if (slot < G.cap) G.slots[(long)c * G.cap + slot] = t;
  else atomicOr(err, ERR_CELL_OVERFLOW);
}

// Host-driven rebuild of the flagged systems' rows (calcNeighborList / checkNeighbors).
__global__ void k_build_rows(View V, Params P, CellGrid G, const int* count, const int* flagCur, int* updateCount,
                             int countUpdates, int* err, int* maxcnt) {
  Lane L;
  if (!wave_tile_ctx(V, L)) return;
  if (!flagCur[L.s]) return;
  if (countUpdates && L.tile == V.stile[L.s] && L.lane == 0) updateCount[L.s] += 1;
  if (!L.act) return;
  build_row(V, P, G, count, G.slots, L.v, L.s, box_of(V, L.s), err, maxcnt);
}

// Interparticle gather of lane L's vertex over entries k0, k0 + kstep, ... of its row:
// accumulates the force on v, the vertex-vertex energy e, the smooth energy es and the share
// (qx, qy) that v's `other` evaluation owes prev(v) (handed to prev(v)'s lane by the caller).
// On the last lane of a wide ring's slice, whose next vertex lies in another tile, (qnx, qny)
// accumulates the share that next(v)'s evaluations owe v, from next(v)'s row (complete before
// the launch: wide sets rebuild rows in their own launch).
template <int INTER, int POT>
__device__ __forceinline__ void gather_pairs(const View& V, const Params& P, const Lane& L, const Box& B, double x,
                                             double y, int k0, int kstep, double& fx, double& fy, double& e,
                                             double& es, double& qx, double& qy, double& qnx, double& qny) {
  const double radv = V.rad[L.v];
  // Row entries are fetched one iteration ahead, the first one before the row length is known:
  // rows are W entries wide (unused entries hold -1), so k < W keeps every fetch in bounds, and
  // an entry is only used once k < ncnt is known.
  if constexpr (INTER == INT_VV) {
    // kernelCalcVertexInteraction, cudaKernel.cuh:690-722
    int j = k0 < V.W ? V.nbr[(long)k0 * V.nv + L.v] : 0;
    const int c = L.act ? V.ncnt[L.v] : 0;
    for (int k = k0; k < c; k += kstep) {
      const int kn = k + kstep;
      const int jn = kn < V.W ? V.nbr[(long)kn * V.nv + L.v] : 0;
      e += pair_force<POT>(P, pbx(x - V.x[j], B), pby(y - V.y[j], B), radv + V.rad[j], fx, fy);
      j = jn;
    }
  } else if constexpr (INTER == INT_SMOOTH) {
    // kernelCalcSmoothInteraction, cudaKernel.cuh:922-975, as a gather: for every u in list(v),
    // (v, segment of u) gives v's `this` share, and (u, segment (prev(v), v)) gives v's
    // `other` share and prev(v)'s `previous` share (handed to prev(v)'s lane by the caller).
    // Ring neighbours of v; inactive lanes past the tile end use v itself (in bounds).
    int vp = L.act ? L.v0 + L.pl : L.v;
    int vpp = L.act ? L.v0 + ((L.pl > L.segS) ? L.pl - 1 : L.segE) : L.v;
    if (L.wide && L.act) {
      vp = ring_prev(L.v, L.f, L.n);
      vpp = ring_prev(vp, L.f, L.n);
    }
    const double xp = V.x[vp], yp = V.y[vp];
    const double xpp = V.x[vpp], ypp = V.y[vpp];
    int u = 0;
    int2 pp = make_int2(0, 0);
    if (k0 < V.W) {
      u = V.nbr[(long)k0 * V.nv + L.v];
      pp = V.nbrp[(long)k0 * V.nv + L.v];
    }
    const int c = L.act ? V.ncnt[L.v] : 0;
    for (int k = k0; k < c; k += kstep) {
      const int kn = k + kstep;
      int un = 0;
      int2 ppn = make_int2(0, 0);
      if (kn < V.W) {
        un = V.nbr[(long)kn * V.nv + L.v];
        ppn = V.nbrp[(long)kn * V.nv + L.v];
      }
      const double xu = V.x[u], yu = V.y[u];
      const double radSum = radv + V.rad[u];
      es += smooth_pair<POT, ROLE_THIS>(P, B, x, y, xu, yu, V.x[pp.x], V.y[pp.x], V.x[pp.y], V.y[pp.y], radSum,
                                        fx, fy);
      es += smooth_pair_op<POT>(P, B, xu, yu, x, y, xp, yp, xpp, ypp, radSum, fx, fy, qx, qy);
      u = un;
      pp = ppn;
    }
    if (L.wide && L.act && L.lane == L.segE) {
      // (u, segment (v, next(v))) for u in list(next(v)): next(v)'s `other` share is its own
      // lane's; the `previous` share is v's (the energy is counted by next(v)'s lane)
      const int j = ring_next(L.v, L.f, L.n);
      const double xj = V.x[j], yj = V.y[j], radj = V.rad[j];
      const int cj = V.ncnt[j];
      for (int k = k0; k < cj; k += kstep) {
        const int w = V.nbr[(long)k * V.nv + j];
        double ox = 0, oy = 0;
        (void)smooth_pair_op<POT>(P, B, V.x[w], V.y[w], xj, yj, x, y, xp, yp, radj + V.rad[w], ox, oy, qnx, qny);
      }
    }
  } else {
    // kernelCalcAllToAllVertexInteraction, cudaKernel.cuh:725-757 (test mode)
    if (L.act) {
      const int b0 = V.svoff[L.s], b1 = V.svoff[L.s + 1];
      for (int j = b0 + k0; j < b1; j += kstep) {
        if (V.vring[j] == L.r) continue;
        e += pair_force<POT>(P, pbx(x - V.x[j], B), pby(y - V.y[j], B), radv + V.rad[j], fx, fy);
      }
    }
  }
}

// -------------------------------------------------------------- step ---
// Waves of a tile: with nw > 1, wave 0 does the ring geometry, shape force and random draws
// while waves 1..nw-1 share the neighbour gather; with nw == 1 wave 0 does both in turn.
// A drift that moved a vertex more than cutoff / 3 marks its system FLAG_FAR: the next rebuild
// cannot rely on the cells of its list positions (md_cells)
__device__ __forceinline__ void far_test(const Params& P, const Box& B, double dx, double dy, int* flag) {
  dx = pbx(dx, B);
  dy = pby(dy, B);
  if (3 * sqrt(dx * dx + dy * dy) > P.cutoff) atomicOr(flag, FLAG_FAR);
}

// The step loop's cell lists of a system whose list positions (the positions of its last
// rebuild) have generation g: its rows are rebuilt from list g % 3, the rebuild bins its own
// positions (the next list positions) into list (g + 1) % 3, cleared by the rebuild before, and
// clears list (g + 2) % 3, which that rebuild read.  The cells are wide enough for a vertex 2
// cutoff / 3 from its list position: at most cutoff / 3 before the drift that triggers the
// rebuild (the displacement test) plus cutoff / 3 in that drift (else FLAG_FAR, all pairs).
struct MdCells {
  const int *cnt, *slots;  // read
  int *binCnt, *binSlots;  // binned into
  int* clear;              // cleared
};
__device__ __forceinline__ MdCells md_cells(const StepArgs& A, const CellGrid& G, int g) {
  const long C = G.ncells, CS = (long)G.ncells * G.cap;
  const int k0 = g % 3, k1 = (g + 1) % 3, k2 = (g + 2) % 3;
  return MdCells{A.cellCur + k0 * C, A.slotsCur + k0 * CS, A.cellCur + k1 * C, A.slotsCur + k1 * CS,
                 A.cellCur + k2 * C};
}

// a rebuilding vertex at (x, y) farther than 2 cutoff / 3 from its list position: its cells may
// miss a neighbour (cannot happen while the displacement and far tests hold)
__device__ __forceinline__ void md_stale_test(const View& V, const Params& P, const Box& B, int v, double x, double y,
                                              int* err) {
  const double ddx = pbx(x - V.lx[v], B), ddy = pby(y - V.ly[v], B);
  if (3 * sqrt(ddx * ddx + ddy * ddy) > 2 * P.cutoff) atomicOr(err, ERR_STALE_BINS);
}

// ---------------------------------------------------------- rigid rings ---
// Half kick of a rigid ring: RigidNVE::updateVelocity (Simulator.cu:348-366), or
// RigidLangevin::updateVelocity (263-286) with its unit Gaussian pair (z0, z1) of the step and
// d_thermalVel zero; the angular velocity takes the torque / moment of inertia.
__device__ __forceinline__ void rigid_kick(bool langevin, double h, double gamma, double Fx, double Fy, double T,
                                           double z0, double z1, double& vx, double& vy, double& w) {
  if (langevin) {
    vx += h * (Fx - vx * gamma);
    vx -= 0.5 * h * h * (Fx - vx * gamma) / gamma;
    vx += z0;
    vy += h * (Fy - vy * gamma);
    vy -= 0.5 * h * h * (Fy - vy * gamma) / gamma;
    vy += z1;
  } else {
    vx += h * Fx;
    vy += h * Fy;
  }
  w += h * T;
}

// A ring's vertex after the ring moved from centre (cx, cy) to (ncx, ncy) and turned by dth:
// translateVertices then rotateVertices (kernelTranslateVertices / kernelRotateVertices,
// cudaKernel.cuh:1678-1713), each difference a box distance.
__device__ __forceinline__ void rigid_move(const Box& B, double x, double y, double cx, double cy, double ncx,
                                           double ncy, double c, double sn, double& xn, double& yn) {
  const double x1 = x + pbx(ncx - cx, B), y1 = y + pby(ncy - cy, B);
  const double dx = pbx(x1 - ncx, B), dy = pby(y1 - ncy, B);
  xn = x1 + ((dx * c - dy * sn) - dx);
  yn = y1 + ((dx * sn + dy * c) - dy);
}

// Per-tile partials of the rigid FIRE sums over the rings whose first vertex is in the tile:
// [0] v.F, [1] |v|^2, [2] |F|^2 (ring velocities and forces)
__device__ __forceinline__ void rigid_fire_partials(const Lane& L, bool head, double vx, double vy, double Fx,
                                                    double Fy, double* part) {
  const double p0 = wave_sum(head ? vx * Fx + vy * Fy : 0.0);
  const double p1 = wave_sum(head ? vx * vx + vy * vy : 0.0);
  const double p2 = wave_sum(head ? Fx * Fx + Fy * Fy : 0.0);
  if (L.lane == 0) {
    part[L.tile * 4 + 0] = p0;
    part[L.tile * 4 + 1] = p1;
    part[L.tile * 4 + 2] = p2;
  }
}

// Step-kernel epilogue of rigid rings (wave 0, one ring per lane segment): the interaction
// forces' transfer to the ring (calcRigidForceEnergy, DPM2D.cu:2128-2142: torque r x f about
// the centre, moment of inertia sum |r|^2, ring force sum f, ring torque sum (r x f) / I, ring
// energy sum e for vertexVertex), then the second half kick of this step and, with MODE_START,
// the next step's first half: last centre and angle stored, half kick with the next step's
// draw, centre and angle advanced, vertices moved (RigidNVE / RigidLangevin::integrate,
// Simulator.cu:241-252, 316-326).  With MODE_FIRE the half kick takes the system's FIRE step.
template <int MODE, bool VV>
__device__ __forceinline__ void rigid_epilogue(const View& V, const Params& P, const StepArgs& A, const Lane& L,
                                               const Box& B, double x, double y, double fx, double fy, double e,
                                               double cx, double cy, double lx, double ly, double z0, double z1) {
  const Rigid R = *V.rg;
  const double rx = pbx(x - cx, B), ry = pby(y - cy, B);
  const double tq = rx * fy - ry * fx;
  const double I = seg_sum(L.act ? rx * rx + ry * ry : 0.0, L.lane, L.segS, L.segE);
  const double Fx = seg_sum(L.act ? fx : 0.0, L.lane, L.segS, L.segE);
  const double Fy = seg_sum(L.act ? fy : 0.0, L.lane, L.segS, L.segE);
  const double T = seg_sum(L.act ? tq : 0.0, L.lane, L.segS, L.segE) / I;
  double E = 0;
  if constexpr (VV) E = seg_sum(L.act ? e : 0.0, L.lane, L.segS, L.segE);
  const bool head = L.act && L.lane == L.segS;
  double vx = 0, vy = 0, w = 0;
  if constexpr (MODE != MODE_RIGID) {
    if (L.act) {
      vx = R.pvx[L.r];
      vy = R.pvy[L.r];
      w = R.pw[L.r];
    }
  }
  if constexpr ((MODE & MODE_FIRE) != 0) {
    // the second half kick with the system's FIRE step, then the tile partials of the next
    // iteration's sums (every lane takes part in the wave sums)
    if (L.act) rigid_kick(false, 0.5 * A.fire[L.s].dt, 0.0, Fx, Fy, T, 0.0, 0.0, vx, vy, w);
    rigid_fire_partials(L, head, vx, vy, Fx, Fy, A.part);
  }
  if (!L.act) return;
  V.fx[L.v] = fx;
  V.fy[L.v] = fy;
  V.ev[L.v] = e;
  R.vtq[L.v] = tq;
  if (head) {
    R.pfx[L.r] = Fx;
    R.pfy[L.r] = Fy;
    R.ptq[L.r] = T;
    R.moi[L.r] = I;
    if constexpr (VV) V.pe[L.r] = E;
  }
  if constexpr (MODE == MODE_RIGID) return;
  if constexpr ((MODE & MODE_FIRE) == 0) {
    rigid_kick(A.rigidLangevin, A.h, P.gamma, Fx, Fy, T, z0, z1, vx, vy, w);
    if constexpr ((MODE & MODE_START) != 0) {
      double n0 = 0, n1 = 0;
      if (A.rigidLangevin)
        gauss_pair(P.seed, (uint32_t)L.s, TAG_THERMAL, A.ctr[0] + (unsigned long long)A.ctrOffT + 1ull,
                   (uint32_t)(L.r - V.sroff[L.s]), n0, n1);
      rigid_kick(A.rigidLangevin, A.h, P.gamma, Fx, Fy, T, n0, n1, vx, vy, w);
      const double ang = V.ang[L.r];
      const double ncx = cx + A.dt * vx, ncy = cy + A.dt * vy, nang = ang + A.dt * w;
      const double dth = nang - ang;
      double xn, yn;
      rigid_move(B, x, y, cx, cy, ncx, ncy, cos(dth), sin(dth), xn, yn);
      A.xout[L.v] = xn;
      A.yout[L.v] = yn;
      const double ddx = pbx(xn - lx, B), ddy = pby(yn - ly, B);
      if (3 * sqrt(ddx * ddx + ddy * ddy) > P.cutoff) {
        atomicOr(&A.flagOut[L.s], 1);
        atomicOr(A.anyOut, 1);
      }
      far_test(P, B, xn - x, yn - y, A.flagOut + L.s);
      if (head) {
        V.ang[L.r] = nang;
        V.cx[L.r] = ncx;
        V.cy[L.r] = ncy;
      }
    }
  }
  if (head) {
    R.pvx[L.r] = vx;
    R.pvy[L.r] = vy;
    R.pw[L.r] = w;
  }
}

template <int INTER, int POT, int MODE>
__global__ __launch_bounds__(WAVE* MAXSPLIT) void k_step(View V, Params P, StepArgs A, CellGrid G) {
  // partials of gather waves 1..nw-1: red[((w - 1) * 6 + q) * WAVE + lane], sized by the launch
  extern __shared__ double red[];
  const int w = threadIdx.x / WAVE;
  const int nw = blockDim.x / WAVE;
  Lane L;
  constexpr bool WIDE = (MODE & MODE_WIDE) != 0;
  constexpr int MB = MODE & ~MODE_WIDE;  // the integrator mode
  // A step of a velocity-Verlet, Langevin or BAOAB loop other than its last consumes the force it
  // computes in this launch (the closing and next opening half kicks); its force, energies, ring
  // geometry and wall forces are stored by the loop's last step only.
  constexpr bool STORE = (MODE & MODE_START) == 0 || (MODE & (MODE_BROWN | MODE_RIGID)) != 0;
  lane_ctx<WIDE>(V, A.xcd ? xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x, threadIdx.x & (WAVE - 1), L);
  const Box B = box_of(V, L.s);
  // own position first: the row rebuild below does not change it, so its load overlaps the
  // rebuild-flag load instead of waiting behind it
  const double x = V.x[L.v], y = V.y[L.v];

  // loop bookkeeping: clear flags (and cells) for later steps, count this step's rebuild, rebuild rows
  if (A.manage) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (A.cellOther)
      for (int c = gid; c < G.ncells; c += gridDim.x * blockDim.x) A.cellOther[c] = 0;
    if (A.flagClear)
      for (int t = gid; t < V.nsys; t += gridDim.x * blockDim.x) A.flagClear[t] = 0;
    const int flag = A.flagCur[L.s];
    const bool rebuild = (flag & 1) != 0;
    // a system flagged FLAG_FAR searches all its vertices instead of cells that may miss a
    // neighbour (the reference's all-pairs search)
    const bool allPairs = A.staleCheck && (flag & FLAG_FAR) != 0;
    const int* cells = A.cellCur;
    const int* slots = A.slotsCur;
    if (A.genCur) {
      // step loop: rows from the cells of the list positions (md_cells)
      const int g = A.genCur[L.s];
      if (L.tile == V.stile[L.s] && threadIdx.x == 0) A.genOut[L.s] = rebuild ? g + 1 : g;
      if (rebuild) {
        const MdCells M = md_cells(A, G, g);
        cells = M.cnt;
        slots = M.slots;
        if (!A.rowsBuilt) {
          if (w == 0 && L.act && !allPairs) md_stale_test(V, P, B, L.v, x, y, A.err);
          // these positions are the system's next list positions (its rows are built from them)
          if (w == nw - 1 && L.act) bin_vertex(G, B, M.binCnt, M.binSlots, L.v, L.s, x, y, A.err);
          if (!G.brute[L.s]) {
            const int t0 = V.stile[L.s], nt = V.stile[L.s + 1] - t0;
            const int n = G.ncx[L.s] * G.ncy[L.s];
            int* clr = M.clear + G.coff[L.s];
            for (int c = (L.tile - t0) * (int)blockDim.x + (int)threadIdx.x; c < n; c += nt * (int)blockDim.x)
              clr[c] = 0;
          }
        }
      }
    }
    if (rebuild && L.tile == V.stile[L.s] && threadIdx.x == 0) A.updateCount[L.s] += 1;
    // the cell search shared by the tile's waves (rows in LDS) where the launch sized LDS for it
    const bool tileBuild = rebuild && !A.rowsBuilt && !allPairs && A.rowsLdsW == V.W && !G.brute[L.s];
    if (tileBuild) {
      build_rows_tile(V, P, G, cells, slots, L.lane, L.v, L.r, L.s, L.act, w, nw, B, x, y, (int*)red, A.err,
                      A.maxcnt);
    } else if (rebuild && w == 0 && L.act && !A.rowsBuilt) {
      build_row(V, P, G, cells, slots, L.v, L.s, B, A.err, A.maxcnt, allPairs);
    }
    if (rebuild) __syncthreads();
  }
  // wave 0's epilogue operands, fetched before the force work (lastPos after the rebuild above)
  double vx = 0, vy = 0, lx = 0, ly = 0;
  if (w == 0) {
    if constexpr (MB != MODE_FORCE && (MODE & MODE_RIGID) == 0) {
      vx = V.vx[L.v];
      vy = V.vy[L.v];
    }
    if constexpr ((MODE & MODE_START) != 0) {
      lx = V.lx[L.v];
      ly = V.ly[L.v];
    }
  }

  // gather partials: force on v, vertex-vertex energy, smooth energy, share for prev(v)
  double fx = 0, fy = 0, e = 0, es = 0, qx = 0, qy = 0, qnx = 0, qny = 0;
  const int g0 = nw > 1 ? 1 : 0;
  const int ng = nw - g0;
  // partials per gather wave: six, eight when the tile is a slice of a wide ring
  constexpr int nq = (INTER == INT_SMOOTH && WIDE) ? 8 : 6;
  if (w >= g0) gather_pairs<INTER, POT>(V, P, L, B, x, y, w - g0, ng, fx, fy, e, es, qx, qy, qnx, qny);
  if (nw > 1 && w > 0) {
    double* r = red + (w - 1) * nq * WAVE + L.lane;
    r[0 * WAVE] = fx;
    r[1 * WAVE] = fy;
    if constexpr (STORE) {  // energies are outputs only: a step that stores none leaves them out
      r[2 * WAVE] = e;
      r[3 * WAVE] = es;
    }
    r[4 * WAVE] = qx;
    r[5 * WAVE] = qy;
    if (nq == 8) {
      r[6 * WAVE] = qnx;
      r[7 * WAVE] = qny;
    }
  }

  // ---- wave 0: work that does not need the gather, overlapped with it
  ShapeOut so{0, 0, 0};
  double zt0 = 0, zt1 = 0, ang = 0;
  if (w == 0) {
    if constexpr ((MODE & MODE_RIGID) != 0) so = ring_centre(V, L, B, x, y);
    else so = ring_shape(V, P, L, B, x, y, STORE);
    // one thermal draw per integrate call and ring: Langevin::updateThermalVel (Simulator.cu:54-77)
    // and Brownian::updateThermalVel (156-179) after this call's force, counter = call index;
    // BAOAB::updateThermalVel (122-144) in the next call's first half, whose index is one more
    const unsigned long long callT = A.ctr[0] + (unsigned long long)A.ctrOffT;
    const unsigned long long callA = A.ctr[1] + (unsigned long long)A.ctrOffA;
    if constexpr ((MODE & MODE_LANGEVIN) != 0 || ((MODE & MODE_BAOAB) != 0 && (MODE & MODE_START) != 0))
      gauss_pair(P.seed, (uint32_t)L.s, TAG_THERMAL, callT, (uint32_t)(L.r - V.sroff[L.s]), zt0, zt1);
    // rigid Langevin: this step's pair (RigidLangevin::integrate draws it before its first half)
    if constexpr ((MODE & MODE_RIGID) != 0 && (MODE & MODE_KICK) != 0)
      if (A.rigidLangevin)
        gauss_pair(P.seed, (uint32_t)L.s, TAG_THERMAL, callT, (uint32_t)(L.r - V.sroff[L.s]), zt0, zt1);
    if (P.geometry != GEOM_NORMAL) {  // walls, after the interactions (DPM2D.cu:2090-2103)
      double wfx, wfy;
      so.e += wall_forces(P, B, x, y, V.rad[L.v], so.fx, so.fy, wfx, wfy);
      if (STORE && L.act) {
        V.wfx[L.v] = wfx;
        V.wfy[L.v] = wfy;
      }
    }
    if (P.particle == PART_ACTIVE_SHAPE && L.act) {  // updateRestLength, DPM2D.cu:2037-2056
      double z0, z1;
      gauss_pair(P.seed, (uint32_t)L.s, TAG_SHAPE, callA, (uint32_t)(L.r - V.sroff[L.s]), z0, z1);
      const double l0v = V.l0[L.v];
      V.l0[L.v] = l0v + ((l0v - V.l0r[L.v]) / P.l0Tau + sqrt(2. * P.l0Diff / P.dt) * z0) * P.dt;
    }
    if (P.particle == PART_ACTIVE) {  // addSelfPropulsion, DPM2D.cu:1995-2035
      const double amplitude = (P.taup != 0) ? sqrt(2.0 * P.dt / P.taup) : 0.0;
      double z0, z1;
      gauss_pair(P.seed, (uint32_t)L.s, TAG_ACTIVE, callA, (uint32_t)(L.r - V.sroff[L.s]), z0, z1);
      ang = V.ang[L.r];
      ang += amplitude * z0;
      ang = ang + kPI;
      ang = ang - 2.0 * kPI * floor(ang / (2.0 * kPI));
      ang = ang - kPI;
      // the store depends on every lane's load of the old angle, which has returned
      if (L.act && L.lane == L.segS) V.ang[L.r] = ang;
    }
  }
  if (nw > 1) {
    __syncthreads();
    if (w > 0) return;
    for (int q = 1; q < nw; q++) {
      const double* r = red + (q - 1) * nq * WAVE + L.lane;
      fx += r[0 * WAVE];
      fy += r[1 * WAVE];
      if constexpr (STORE) {
        e += r[2 * WAVE];
        es += r[3 * WAVE];
      }
      qx += r[4 * WAVE];
      qy += r[5 * WAVE];
      if (nq == 8) {
        qnx += r[6 * WAVE];
        qny += r[7 * WAVE];
      }
    }
  }

  // ---- wave 0: assemble F = shape + interparticle (+ shares from the next lane)
  if constexpr (INTER == INT_SMOOTH) {
    const double sx = shfl_d(qx, L.nl), sy = shfl_d(qy, L.nl);
    const bool fromMemory = L.wide && L.lane == L.segE;  // next(v) in another tile
    fx += fromMemory ? qnx : sx;
    fy += fromMemory ? qny : sy;
    if constexpr (STORE) {
      const double pe = seg_sum(L.act ? es : 0.0, L.lane, L.segS, L.segE);
      if (L.act) {
        V.es[L.v] = es;
        if (L.lane == L.segS && !L.wide) V.pe[L.r] = pe;  // wide rings: k_wide_pe sums es
      }
    }
  }
  if constexpr ((MODE & MODE_RIGID) != 0) {
    rigid_epilogue<MB, INTER == INT_VV>(V, P, A, L, B, x, y, fx, fy, e, so.cx, so.cy, lx, ly, zt0, zt1);
    return;
  }
  fx = so.fx + fx;
  fy = so.fy + fy;
  e = so.e + e;
  if (P.particle == PART_ACTIVE) {
    fx += P.driving * ((1 - 0) * cos(ang) + 0 * sin(ang));
    fy += P.driving * ((1 - 1) * cos(ang) + 1 * sin(ang));
  }

  if constexpr ((MODE & MODE_LANGEVIN) != 0 && (MODE & MODE_BROWN) == 0) {
    fx += P.noise * zt0 - P.gamma * vx;
    fy += P.noise * zt1 - P.gamma * vy;
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
  if (!L.act) return;
  if constexpr (STORE) {
    V.fx[L.v] = fx;
    V.fy[L.v] = fy;
    V.ev[L.v] = e;
  }
  if constexpr (MB == MODE_FORCE) return;
  if constexpr ((MODE & MODE_BROWN) != 0) {
    // overdamped velocity (Brownian: (F + noise xi) / gamma; driven: F / gamma), then the drift
    if constexpr ((MODE & MODE_LANGEVIN) != 0) {
      vx = (fx + P.noise * zt0) / P.gamma;
      vy = (fy + P.noise * zt1) / P.gamma;
    } else {
      vx = fx / P.gamma;
      vy = fy / P.gamma;
    }
  }
  if constexpr ((MODE & MODE_KICK) != 0) {
    vx += A.h * fx;
    vy += A.h * fy;
  }
  if constexpr ((MODE & MODE_START) != 0) {
    double xn, yn;
    if constexpr ((MODE & MODE_BROWN) != 0) {
      xn = x + vx * A.dt;
      yn = y + vy * A.dt;
    } else if constexpr ((MODE & MODE_BAOAB) != 0) {
      // B, A (half), O, A (half)
      vx += A.h * fx;
      vy += A.h * fy;
      const double xa = x + vx * A.h, ya = y + vy * A.h;
      vx = P.gamma * vx + P.noise * zt0;
      vy = P.gamma * vy + P.noise * zt1;
      xn = xa + vx * A.h;
      yn = ya + vy * A.h;
    } else {
      vx += A.h * fx;
      vy += A.h * fy;
      xn = x + A.dt * vx;
      yn = y + A.dt * vy;
    }
    A.xout[L.v] = xn;
    A.yout[L.v] = yn;
    if (A.recx) {
      A.recx[L.v] = xn;
      A.recy[L.v] = yn;
    }
    const double ddx = pbx(xn - lx, B), ddy = pby(yn - ly, B);
    if (3 * sqrt(ddx * ddx + ddy * ddy) > P.cutoff) {
      atomicOr(&A.flagOut[L.s], 1);
      atomicOr(A.anyOut, 1);
    }
    far_test(P, B, xn - x, yn - y, A.flagOut + L.s);
  }
  V.vx[L.v] = vx;
  V.vy[L.v] = vy;
}

// ------------------------------------------------------------ wide rings ---
// Ring sums of the rings of more than 64 vertices, whose slices k_step and k_shape read: area,
// perimeter and centre by ring_shape's unwrapped walk (exclusive prefix sums of the minimum-image
// segments from the ring's first vertex), here a block-wide scan over chunks of 256 vertices.
// One 256-thread block per wide ring.
__global__ __launch_bounds__(256) void k_wide_rings(View V, const int* wide) {
  __shared__ double wsum[2][4];
  __shared__ double red[4][4];
  const int r = wide[blockIdx.x];
  const int f = V.rfirst[r], n = V.rnv[r];
  const Box B = box_of(V, V.rsys[r]);
  const int t = threadIdx.x, lane = t & (WAVE - 1), wv = t / WAVE;
  double carryX = 0, carryY = 0;  // unwrapped offset of the chunk's first vertex
  double aP = 0, aS = 0, aX = 0, aY = 0;
  for (int b = 0; b < n; b += 256) {
    const int i = b + t;
    const bool act = i < n;
    double dnx = 0, dny = 0;
    if (act) {
      const int v = f + i, vn = (i + 1 == n) ? f : v + 1;
      dnx = pbx(V.x[vn] - V.x[v], B);
      dny = pby(V.y[vn] - V.y[v], B);
    }
    double ix = dnx, iy = dny;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
      const double tx = __shfl_up(ix, d, WAVE), ty = __shfl_up(iy, d, WAVE);
      if (lane >= d) {
        ix += tx;
        iy += ty;
      }
    }
    if (lane == WAVE - 1) {
      wsum[0][wv] = ix;
      wsum[1][wv] = iy;
    }
    __syncthreads();
    double ox = carryX, oy = carryY;
    for (int q = 0; q < wv; q++) {
      ox += wsum[0][q];
      oy += wsum[1][q];
    }
    const double ux = ox + (ix - dnx), uy = oy + (iy - dny);
    for (int q = 0; q < 4; q++) {
      carryX += wsum[0][q];
      carryY += wsum[1][q];
    }
    __syncthreads();
    if (act) {
      aP += sqrt(dnx * dnx + dny * dny);
      aS += ux * dny - dnx * uy;
      aX += ux;
      aY += uy;
    }
  }
  aP = wave_sum(aP);
  aS = wave_sum(aS);
  aX = wave_sum(aX);
  aY = wave_sum(aY);
  if (lane == 0) {
    red[0][wv] = aP;
    red[1][wv] = aS;
    red[2][wv] = aX;
    red[3][wv] = aY;
  }
  __syncthreads();
  if (t == 0) {
    const double perim = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    const double shoe = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    const double sux = red[2][0] + red[2][1] + red[2][2] + red[2][3];
    const double suy = red[3][0] + red[3][1] + red[3][2] + red[3][3];
    const double ind = 1.0 / (double)n;
    V.area[r] = fabs(shoe) * 0.5;
    V.perim[r] = perim;
    V.cx[r] = V.x[f] + sux * ind;
    V.cy[r] = V.y[f] + suy * ind;
  }
}

// smooth interaction energy of each wide ring: the sum of its vertices' shares, which the
// slices of k_step store per vertex
__global__ __launch_bounds__(256) void k_wide_pe(View V, const int* wide) {
  __shared__ double red[4];
  const int r = wide[blockIdx.x];
  const int f = V.rfirst[r], n = V.rnv[r];
  double acc = 0;
  for (int i = threadIdx.x; i < n; i += 256) acc += V.es[f + i];
  acc = wave_sum(acc);
  if ((threadIdx.x & (WAVE - 1)) == 0) red[threadIdx.x / WAVE] = acc;
  __syncthreads();
  if (threadIdx.x == 0) V.pe[r] = red[0] + red[1] + red[2] + red[3];
}

// The step loop's row rebuild of k_step as a launch of its own, for sets with wide rings under
// the vertex-segment interaction: the last lane of a slice reads the row of a vertex in another
// tile, so every flagged row is complete before the step kernel starts (which then skips its
// rebuild, A.rowsBuilt).  One thread per vertex.
__global__ void k_rows_loop(View V, Params P, StepArgs A, CellGrid G) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V.nv) return;
  const int s = V.vsys[v];
  const int flag = A.flagCur[s];
  if ((flag & 1) == 0) return;
  const Box B = box_of(V, s);
  const bool allPairs = A.staleCheck && (flag & FLAG_FAR) != 0;
  const int* cnt = A.cellCur;  // FIRE: cells of these positions
  const int* slots = A.slotsCur;
  if (A.genCur) {
    // step loop: the cells of the list positions (md_cells); the step kernel after this launch
    // advances the system's generation
    const MdCells M = md_cells(A, G, A.genCur[s]);
    const double x = V.x[v], y = V.y[v];
    if (!allPairs) md_stale_test(V, P, B, v, x, y, A.err);
    bin_vertex(G, B, M.binCnt, M.binSlots, v, s, x, y, A.err);
    if (!G.brute[s]) {
      const int n = G.ncx[s] * G.ncy[s];
      for (int c = v - V.svoff[s]; c < n; c += V.svoff[s + 1] - V.svoff[s]) M.clear[G.coff[s] + c] = 0;
    }
    cnt = M.cnt;
    slots = M.slots;
  }
  build_row(V, P, G, cnt, slots, v, s, B, A.err, A.maxcnt, allPairs);
}

// geometry only (calcParticleShape + calcParticlePositions)
template <bool WIDE>
__global__ void k_shape(View V, Params P) {
  Lane L;
  const int tile = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
  if (tile >= V.ntiles) return;
  lane_ctx<WIDE>(V, tile, threadIdx.x & (WAVE - 1), L);
  ring_shape(V, P, L, box_of(V, L.s), V.x[L.v], V.y[L.v], true);
}

// Prologue of a step loop's cell protocol for vertex v of system s at positions (x, y): its list
// position binned into list 0 (generation 0), and the rebuild flag of (x, y) (the displacement
// test, with FLAG_FAR where (x, y) lies more than 2 cutoff / 3 from the list position, beyond what
// the cells allow for)
__device__ __forceinline__ void md_prologue(const View& V, const Params& P, const CellGrid& G, const StepArgs& A,
                                            const Box& B, int v, int s, double x, double y, int* flag, int* any) {
  const double lx = V.lx[v], ly = V.ly[v];
  bin_vertex(G, B, A.cellCur, A.slotsCur, v, s, lx, ly, A.err);
  const double ddx = pbx(x - lx, B), ddy = pby(y - ly, B);
  const double d3 = 3 * sqrt(ddx * ddx + ddy * ddy);
  if (d3 > P.cutoff) {
    atomicOr(flag, d3 > 2 * P.cutoff ? 1 | FLAG_FAR : 1);
    if (any) atomicOr(any, 1);
  }
}

// The step loop's first half step (velocity Verlet, or BAOAB's B A O A with the draw of call
// ctr[0] + ctrOffT) from the stored force, with the loop's cell protocol (md_prologue on the
// destination, whose flags go to A.flagOut; generation 0 into A.genOut; A.flagClear cleared)
// (BAOAB::integrate, Simulator.cu:112-120, up to its force evaluation).
template <bool BAOAB>
__global__ void k_start_loop(View V, Params P, CellGrid G, double h, double dt, StepArgs A) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  for (int t = v; t < V.nsys; t += gridDim.x * blockDim.x) {
    A.flagClear[t] = 0;
    A.genOut[t] = 0;
  }
  if (v >= V.nv) return;
  const int s = V.vsys[v];
  const Box B = box_of(V, s);
  double vx = V.vx[v] + h * V.fx[v];
  double vy = V.vy[v] + h * V.fy[v];
  double xn, yn;
  if constexpr (BAOAB) {
    double z0, z1;
    gauss_pair(P.seed, (uint32_t)s, TAG_THERMAL, A.ctr[0] + (unsigned long long)A.ctrOffT,
               (uint32_t)(V.vring[v] - V.sroff[s]), z0, z1);
    const double xa = V.x[v] + vx * h, ya = V.y[v] + vy * h;
    vx = P.gamma * vx + P.noise * z0;
    vy = P.gamma * vy + P.noise * z1;
    xn = xa + vx * h;
    yn = ya + vy * h;
  } else {
    xn = V.x[v] + dt * vx;
    yn = V.y[v] + dt * vy;
  }
  V.vx[v] = vx;
  V.vy[v] = vy;
  A.xout[v] = xn;
  A.yout[v] = yn;
  if (A.recx) {
    A.recx[v] = xn;
    A.recy[v] = yn;
  }
  md_prologue(V, P, G, A, B, v, s, xn, yn, A.flagOut + s, A.anyOut);
  far_test(P, B, xn - V.x[v], yn - V.y[v], A.flagOut + s);
}

// The rigid loop's first half step from the stored ring force and torque, one thread per ring
// (RigidNVE / RigidLangevin::integrate, Simulator.cu:241-252, 316-326, up to the vertex move):
// last centre and angle stored, half kick with the draw of call ctr[0] + ctrOffT, centre and
// angle advanced, and each vertex moved into A.xout and tested into A.flagOut; in the step loop
// (A.genCur) with its cell protocol (md_prologue, generation 0 into A.genOut), else A.cellOther
// cleared; A.flagClear cleared when given.
__global__ void k_rigid_start(View V, Params P, CellGrid G, StepArgs A) {
  const Rigid R = *V.rg;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (A.cellOther)
    for (int c = t; c < G.ncells; c += gridDim.x * blockDim.x) A.cellOther[c] = 0;
  if (A.flagClear)
    for (int q = t; q < V.nsys; q += gridDim.x * blockDim.x) A.flagClear[q] = 0;
  if (A.genCur)
    for (int q = t; q < V.nsys; q += gridDim.x * blockDim.x) A.genOut[q] = 0;
  if (t >= V.np) return;
  const int r = t, s = V.rsys[r];
  const Box B = box_of(V, s);
  double vx = R.pvx[r], vy = R.pvy[r], w = R.pw[r];
  double z0 = 0, z1 = 0;
  if (A.rigidLangevin)
    gauss_pair(P.seed, (uint32_t)s, TAG_THERMAL, A.ctr[0] + (unsigned long long)A.ctrOffT, (uint32_t)(r - V.sroff[s]),
               z0, z1);
  rigid_kick(A.rigidLangevin, A.h, P.gamma, R.pfx[r], R.pfy[r], R.ptq[r], z0, z1, vx, vy, w);
  const double cx = V.cx[r], cy = V.cy[r], ang = V.ang[r];
  const double ncx = cx + A.dt * vx, ncy = cy + A.dt * vy, nang = ang + A.dt * w;
  const double dth = nang - ang, c = cos(dth), sn = sin(dth);
  bool flag = false;
  for (int v = V.rfirst[r]; v < V.rfirst[r] + V.rnv[r]; v++) {
    const double x = V.x[v], y = V.y[v];
    double xn, yn;
    rigid_move(B, x, y, cx, cy, ncx, ncy, c, sn, xn, yn);
    A.xout[v] = xn;
    A.yout[v] = yn;
    if (A.genCur) {
      md_prologue(V, P, G, A, B, v, s, xn, yn, A.flagOut + s, nullptr);
    } else {
      const double ddx = pbx(xn - V.lx[v], B), ddy = pby(yn - V.ly[v], B);
      flag = flag || 3 * sqrt(ddx * ddx + ddy * ddy) > P.cutoff;
    }
    far_test(P, B, xn - x, yn - y, A.flagOut + s);
  }
  if (flag) {
    atomicOr(&A.flagOut[s], 1);
    atomicOr(A.anyOut, 1);
  }
  R.plx[r] = cx;
  R.ply[r] = cy;
  R.pla[r] = ang;
  R.pvx[r] = vx;
  R.pvy[r] = vy;
  R.pw[r] = w;
  V.ang[r] = nang;
  V.cx[r] = ncx;
  V.cy[r] = ncy;
}

// RigidLangevin::injectKineticEnergy (Simulator.cu:254-261), one workgroup per system: a
// Gaussian pair of the given amplitude per ring (counter `call`), minus the system's mean ring
// velocity (kernelConserveParticleMomentum, cudaKernel.cuh:1998-2027, over every ring)
__global__ __launch_bounds__(256) void k_rigid_inject(View V, Params P, double amplitude, unsigned long long call) {
  const Rigid R = *V.rg;
  __shared__ double red[2][256 / WAVE];
  const int s = blockIdx.x, tid = threadIdx.x;
  const int r0 = V.sroff[s], r1 = V.sroff[s + 1];
  double mx = 0, my = 0;
  for (int r = r0 + tid; r < r1; r += blockDim.x) {
    double z0, z1;
    gauss_pair(P.seed, (uint32_t)s, TAG_INJECT, call, (uint32_t)(r - r0), z0, z1);
    R.pvx[r] = amplitude * z0;
    R.pvy[r] = amplitude * z1;
    mx += R.pvx[r];
    my += R.pvy[r];
  }
  mx = wave_sum(mx);
  my = wave_sum(my);
  if ((tid & (WAVE - 1)) == 0) {
    red[0][tid / WAVE] = mx;
    red[1][tid / WAVE] = my;
  }
  __syncthreads();
  mx = 0;
  my = 0;
  for (int q = 0; q < (int)(blockDim.x / WAVE); q++) {
    mx += red[0][q];
    my += red[1][q];
  }
  mx /= (r1 - r0);
  my /= (r1 - r0);
  for (int r = r0 + tid; r < r1; r += blockDim.x) {
    R.pvx[r] -= mx;
    R.pvy[r] -= my;
  }
}

// The overdamped loops' prologue: md_prologue on the current positions (their first step
// rebuilds when flagged), flags into `flag`, generation 0 into `genOut`.
__global__ void k_bin_flag(View V, Params P, CellGrid G, StepArgs A, int* flag) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < V.nsys) A.genOut[v] = 0;
  if (v >= V.nv) return;
  const int s = V.vsys[v];
  md_prologue(V, P, G, A, box_of(V, s), v, s, V.x[v], V.y[v], flag + s, nullptr);
}

// RNG call counters advanced by the draws of a finished loop
__global__ void k_ctr_add(unsigned long long* ctr, long long addT, long long addA) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctr[0] += (unsigned long long)addT;
    ctr[1] += (unsigned long long)addA;
  }
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
__global__ __launch_bounds__(256) void k_reduce(View V, double* out) {
  const int s = blockIdx.x;
  __shared__ double sh[8][4];
  double acc[8] = {0, 0, 0, -1, 0, 0, 0, 0};
  const int b0 = V.svoff[s], b1 = V.svoff[s + 1];
  for (int v = b0 + threadIdx.x; v < b1; v += 256) {
    acc[0] += V.ev[v];
    acc[2] += V.vx[v] * V.vx[v] + V.vy[v] * V.vy[v];
    acc[3] = fmax(acc[3], fmax(V.fx[v] * V.fx[v], V.fy[v] * V.fy[v]));
    acc[4] += V.fx[v] * V.fx[v] + V.fy[v] * V.fy[v];
  }
  const int r0 = V.sroff[s], r1 = V.sroff[s + 1];
  for (int r = r0 + threadIdx.x; r < r1; r += 256) {
    acc[1] += V.pe[r];
    acc[5] += V.area[r];
    acc[6] += V.a0[r];
    const double rr = V.rad[V.rfirst[r]];
    acc[7] += rr * rr * (0.5 * V.rnv[r] - 1);
  }
  const int w = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const double t = (q == 3) ? wave_max(acc[q]) : wave_sum(acc[q]);
    if (lane == 0) sh[q][w] = t;
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    const int q = threadIdx.x;
    double t = sh[q][0];
    for (int i = 1; i < 4; i++) t = (q == 3) ? fmax(t, sh[q][i]) : t + sh[q][i];
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

// Stages of the reference's Langevin integrator for callers that drive them one by one
// (Simulator.cu:85-95 kernelUpdateVertexPos / kernelUpdateVertexVel, 54-77 updateThermalVel):
// x += dt v; v += dt F; F += noise xi_ring - gamma v with the ring's draw of call `call`
__global__ void k_update_pos(View V, double dt) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V.nv) return;
  V.x[v] += dt * V.vx[v];
  V.y[v] += dt * V.vy[v];
}

__global__ void k_update_vel(View V, double dt) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V.nv) return;
  V.vx[v] += dt * V.fx[v];
  V.vy[v] += dt * V.fy[v];
}

__global__ void k_thermal_force(View V, Params P, const unsigned long long* ctr) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V.nv) return;
  const int r = V.vring[v], s = V.vsys[v];
  double z0, z1;
  gauss_pair(P.seed, (uint32_t)s, TAG_THERMAL, ctr[0], (uint32_t)(r - V.sroff[s]), z0, z1);
  V.fx[v] += P.noise * z0 - P.gamma * V.vx[v];
  V.fy[v] += P.noise * z1 - P.gamma * V.vy[v];
}

// subtract the per-system mean velocity
__global__ __launch_bounds__(256) void k_sub_mean(View V) {
  const int s = blockIdx.x;
  __shared__ double sh[2][4];
  __shared__ double mean[2];
  const int b0 = V.svoff[s], b1 = V.svoff[s + 1];
  double ax = 0, ay = 0;
  for (int v = b0 + threadIdx.x; v < b1; v += 256) { ax += V.vx[v]; ay += V.vy[v]; }
  const int w = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
  ax = wave_sum(ax); ay = wave_sum(ay);
  if (lane == 0) { sh[0][w] = ax; sh[1][w] = ay; }
  __syncthreads();
  if (threadIdx.x < 2) {
    double t = 0;
    for (int i = 0; i < 4; i++) t += sh[threadIdx.x][i];
    mean[threadIdx.x] = t / (b1 - b0);
  }
  __syncthreads();
  for (int v = b0 + threadIdx.x; v < b1; v += 256) { V.vx[v] -= mean[0]; V.vy[v] -= mean[1]; }
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

// scaleBox, DPM2D.cu:1693-1719, with translateVertices (kernelTranslateVertices,
// cudaKernel.cuh:1678-1693): every vertex moves by the shift of its ring's centre to
// centre / scale, a box distance in the new box (uploaded before this launch); the centres
// themselves are divided in k_scale_centres, launched after
__global__ void k_scale_box(View V, const double* scale) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V.nv) return;
  const int r = V.vring[v], s = V.vsys[v];
  const Box B = box_of(V, s);
  const double cx = V.cx[r], cy = V.cy[r], sc = scale[s];
  V.x[v] += pbx(cx / sc - cx, B);
  V.y[v] += pby(cy / sc - cy, B);
}

__global__ void k_scale_centres(View V, const double* scale) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= V.np) return;
  const double sc = scale[V.rsys[r]];
  V.cx[r] /= sc;
  V.cy[r] /= sc;
}

// --------------------------------------------------------------- FIRE ---
// sum of the per-tile partials q of system s, all 64 lanes striding over the tiles; the
// xor-butterfly leaves the same value in every lane (addition is commutative)
__device__ __forceinline__ double sys_sum(const double* part, int q, int t0, int t1, int lane) {
  double acc = 0;
  for (int t = t0 + lane; t < t1; t += WAVE) acc += part[t * 4 + q];
  return wave_sum(acc);
}

__device__ __forceinline__ double sys_max(const double* part, int q, int t0, int t1, int lane) {
  double acc = -1;
  for (int t = t0 + lane; t < t1; t += WAVE) acc = fmax(acc, part[t * 4 + q]);
  return wave_max(acc);
}

// --------------------------------------------------------- rigid FIRE ---
__global__ void k_rigid_fire_partials(View V, double* part) {
  const Rigid R = *V.rg;
  Lane L;
  if (!wave_tile_ctx(V, L)) return;
  const bool head = L.act && L.lane == L.segS;
  double vx = 0, vy = 0, Fx = 0, Fy = 0;
  if (head) {
    vx = R.pvx[L.r];
    vy = R.pvy[L.r];
    Fx = R.pfx[L.r];
    Fy = R.pfy[L.r];
  }
  rigid_fire_partials(L, head, vx, vy, Fx, Fy, part);
}

// One rigid FIRE iteration up to the vertex move, one wave per tile (FIRE::minimizerRigidLoop,
// FIRE.cu:279-291, with checkRigidFIREParams 217-254 and bendRigidVelocityTowardsForce 256-277).
// Every wave derives its system's FIRE update from the tile partials of the state after the
// previous iteration (power P = sum v.F decides; on a restart the centres and angles step back by
// half the already-decreased step and both velocities clear), and the bend ratio from
// |v'|^2 = h^2 |F|^2 after a restart, else |v|^2 + 2 h P + h^2 |F|^2, the half-kicked velocities'
// norm.  Each lane moves its ring redundantly and its vertex with it; the ring's first lane
// stores the ring, the system's first tile the new FIRE state (fout; fin stays as it was).
// The last centres and angles are the ones at the start of the iteration.
__global__ void k_rigid_fire_move(View V, Params P, FireParams fp, const FireState* fin, FireState* fout,
                                  const double* part, double* xo, double* yo, int* flagOut, int* anyOut) {
  const Rigid R = *V.rg;
  Lane L;
  if (!wave_tile_ctx(V, L)) return;
  const int t0 = V.stile[L.s], t1 = V.stile[L.s + 1];
  const double P0 = sys_sum(part, 0, t0, t1, L.lane);
  const double VV = sys_sum(part, 1, t0, t1, L.lane);
  const double FF = sys_sum(part, 2, t0, t1, L.lane);
  FireState f = fin[L.s];
  const bool restart = !(P0 > 0);
  if (!restart) {
    f.numStep++;
    if (f.numStep > fp.minStep) {
      f.dt = fmin(f.dt * fp.f_inc, fp.dt_max);
      f.a *= fp.f_a;
    }
  } else {
    f.numStep = 0;
    f.dt = fmax(f.dt * fp.f_dec, fp.dt_max / 10);
    f.a = fp.a_start;
  }
  const double h = 0.5 * f.dt;
  if (L.tile == t0 && L.lane == 0) {
    FireState o = f;
    o.P = P0;
    o.iters = f.iters + 1;
    fout[L.s] = o;
  }
  if (!L.act) return;
  const double VV1 = restart ? h * h * FF : VV + 2 * h * P0 + h * h * FF;
  double cx = V.cx[L.r], cy = V.cy[L.r], ang = V.ang[L.r];
  double vx = R.pvx[L.r], vy = R.pvy[L.r], w = R.pw[L.r];
  const double Fx = R.pfx[L.r], Fy = R.pfy[L.r];
  const double lcx = cx, lcy = cy, lang = ang;
  if (restart) {
    cx -= h * vx;
    cy -= h * vy;
    ang -= h * w;
    vx = vy = w = 0;
  }
  rigid_kick(false, h, 0.0, Fx, Fy, R.ptq[L.r], 0.0, 0.0, vx, vy, w);
  if (FF > 0) {
    const double ratio = sqrt(VV1 / FF);
    vx = (1 - f.a) * vx + f.a * Fx * ratio;
    vy = (1 - f.a) * vy + f.a * Fy * ratio;
  }
  cx += f.dt * vx;
  cy += f.dt * vy;
  ang += f.dt * w;
  const Box B = box_of(V, L.s);
  const double dth = ang - lang;
  double xn, yn;
  rigid_move(B, V.x[L.v], V.y[L.v], lcx, lcy, cx, cy, cos(dth), sin(dth), xn, yn);
  xo[L.v] = xn;
  yo[L.v] = yn;
  const double ddx = pbx(xn - V.lx[L.v], B), ddy = pby(yn - V.ly[L.v], B);
  if (3 * sqrt(ddx * ddx + ddy * ddy) > P.cutoff) {
    atomicOr(&flagOut[L.s], 1);
    atomicOr(anyOut, 1);
  }
  far_test(P, B, xn - V.x[L.v], yn - V.y[L.v], flagOut + L.s);
  if (L.lane == L.segS) {
    R.plx[L.r] = lcx;
    R.ply[L.r] = lcy;
    R.pla[L.r] = lang;
    V.cx[L.r] = cx;
    V.cy[L.r] = cy;
    V.ang[L.r] = ang;
    R.pvx[L.r] = vx;
    R.pvy[L.r] = vy;
    R.pw[L.r] = w;
  }
}

// FIRE::minimizerVertexLoop, FIRE.cu:397-411, split at its two global reductions.
// fire_begin: checkFIREParams (338-371) + half kick + partial |v|^2, |F|^2.
// With tol >= 0 a system whose max |F_c| <= tol (or that reached maxIter) is frozen.
__global__ void k_fire_begin(View V, FireParams fp, const FireState* fin, FireState* fout, double* part) {
  Lane L;
  if (!wave_tile_ctx(V, L)) return;
  FireState fs = fin[L.s];
  const int t0 = V.stile[L.s], t1 = V.stile[L.s + 1];
  const double Psum = sys_sum(part, 0, t0, t1, L.lane);
  if (!fs.converged && fp.tol >= 0) {
    const double fmax2 = sys_max(part, 3, t0, t1, L.lane);
    if (sqrt(fmax2) <= fp.tol || fs.iters >= fp.maxIter) fs.converged = 1;
  }
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
  if (L.tile == t0 && L.lane == 0) fout[L.s] = fs;
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
    part[L.tile * 4 + 1] = tv;
    part[L.tile * 4 + 2] = tf;
  }
}

// fire_mid: bendVertexVelocityTowardsForce (374-394) + drift + displacement test
__global__ void k_fire_mid(View V, Params P, const FireState* fcur, const double* part, double* xo, double* yo,
                           int* flagOut, int* anyOut) {
  Lane L;
  if (!wave_tile_ctx(V, L)) return;
  const FireState fs = fcur[L.s];
  const int t0 = V.stile[L.s], t1 = V.stile[L.s + 1];
  const double vv = sys_sum(part, 1, t0, t1, L.lane);
  const double ff = sys_sum(part, 2, t0, t1, L.lane);
  if (!L.act) return;
  double vx = V.vx[L.v], vy = V.vy[L.v];
  const double x = V.x[L.v], y = V.y[L.v];
  if (fs.converged) {
    xo[L.v] = x;
    yo[L.v] = y;
    return;
  }
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
  const Box B = box_of(V, L.s);
  const double ddx = pbx(xn - V.lx[L.v], B), ddy = pby(yn - V.ly[L.v], B);
  if (3 * sqrt(ddx * ddx + ddy * ddy) > P.cutoff) {
    atomicOr(&flagOut[L.s], 1);
    atomicOr(anyOut, 1);
  }
}

// per-tile partials of P = v.F and max F_c^2 for the current state (start of a FIRE call)
__global__ void k_fire_partials(View V, double* part) {
  Lane L;
  if (!wave_tile_ctx(V, L)) return;
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

// per-system FIRE summary after a minimisation: all-converged flag
__global__ void k_fire_done(const FireState* fs, int nsys, int* notDone) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nsys && !fs[s].converged) atomicOr(notDone, 1);
}

// a new minimisation call tests every system's stopping rule again (the driver's loop condition
// is evaluated per call, compressFIRE.cpp:121-131)
__global__ void k_fire_resume(FireState* fs, int nsys) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nsys) fs[s].converged = 0;
}

// ------------------------------------------------------- persistent FIRE ---
// One workgroup per system runs whole FIRE iterations (FIRE::minimizerVertexLoop,
// FIRE.cu:397-411) back to back, separated by workgroup barriers only: systems share nothing,
// so no grid-wide synchronisation is needed and a converged system's workgroup simply exits.
// Each wave walks the system's ring-aligned tiles; the force phase is the step kernel's
// (ring_shape + gather_pairs).  At most `chunk` iterations per launch; every thread decides the
// loop exit from the same FIRE state, so all waves leave together.
// Two builds: up to 8 waves per system (about 190 VGPRs, no spills) and up to 16 (the 1024-lane
// workgroup caps a wave at 128 VGPRs, so 40-75 of them spill to scratch; halving the tiles each
// wave walks still makes an iteration about 12% shorter on systems of 76 tiles).
#define FS_WAVES 8
#define FS_WAVES_WIDE 16

// Workgroup reductions of the persistent kernel, with one barrier each: a caller alternates two
// exchange buffers `red`, so that between two exchanges through the same buffer every thread
// passes another barrier (no exchange overwrites one that a wave is still reading).
// workgroup sum of a and maximum of b (fixed order: lanes by butterfly, waves by index)
__device__ __forceinline__ void block_sum_max(double a, double b, double (*red)[2], double& sa, double& mb) {
  const int w = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
  a = wave_sum(a);
  b = wave_max(b);
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    red[w][0] = a;
    red[w][1] = b;
  }
  __syncthreads();
  double s = 0, m = -1;
  for (int q = 0; q < nw; q++) {
    s += red[q][0];
    m = fmax(m, red[q][1]);
  }
  sa = s;
  mb = m;
}

// workgroup sums of a and of b in one exchange (the order of block_sum_max's sum for each)
__device__ __forceinline__ void block_sum_sum(double a, double b, double (*red)[2], double& sa, double& sb) {
  const int w = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
  a = wave_sum(a);
  b = wave_sum(b);
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    red[w][0] = a;
    red[w][1] = b;
  }
  __syncthreads();
  double s = 0, t = 0;
  for (int q = 0; q < nw; q++) {
    s += red[q][0];
    t += red[q][1];
  }
  sa = s;
  sb = t;
}

// P = v.F and max F_c^2 over the system's vertices
__device__ __forceinline__ void fire_power(const View& V, int t0, int t1, double (*red)[2], double& P, double& Fm) {
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
  double p = 0, m = -1;
  for (int t = t0 + w; t < t1; t += nw) {
    const int v = V.tile_v0[t] + lane;
    if (v < V.tile_v1[t]) {
      const double fx = V.fx[v], fy = V.fy[v];
      p += V.vx[v] * fx + V.vy[v] * fy;
      m = fmax(m, fmax(fx * fx, fy * fy));
    }
  }
  block_sum_max(p, m, red, P, Fm);
}

template <int INTER, int POT, int MAXW>
__global__ __launch_bounds__(WAVE* MAXW) void k_fire_sys(View V, Params P, FireParams fp, CellGrid G,
                                                         FireState* fst, int chunk, int* updateCount, int* err,
                                                         int* maxcnt, const int* order) {
  // exchange buffers: [0] for |v|^2 and |F|^2 after the half kick, [1] for P and max F_c^2 after
  // the force (between two exchanges through one buffer every thread passes the other's barrier
  // and the displacement test's)
  __shared__ double red[2][MAXW][2];
  const int s = order[blockIdx.x];
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
  const int t0 = V.stile[s], t1 = V.stile[s + 1];
  const Box B = box_of(V, s);
  // every thread carries the system's FIRE state and updates it from the same reduced values in
  // the same order, so all hold the same state and take the same exit without a barrier
  FireState fs = fst[s];
  double Pv, Fm;
  fire_power(V, t0, t1, red[1], Pv, Fm);
  for (int it = 0; it < chunk; it++) {
    // ---- checkFIREParams (FIRE.cu:338-371) and the stopping rule
    if (!fs.converged && (sqrt(Fm) <= fp.tol || fs.iters >= fp.maxIter)) fs.converged = 1;
    if (fs.converged) break;
    int reset = 0;
    fs.P = Pv;
    if (Pv > 0) {
      fs.numStep++;
      if (fs.numStep > fp.minStep) {
        fs.dt = fmin(fs.dt * fp.f_inc, fp.dt_max);
        fs.a *= fp.f_a;
      }
    } else {
      fs.numStep = 0;
      fs.dt = fmax(fs.dt * fp.f_dec, fp.dt_max / 10);
      fs.a = fp.a_start;
      reset = 1;
    }
    fs.iters++;
    const double dt = fs.dt, h = 0.5 * dt, fa = fs.a;
    // ---- back-step on reset (with the decreased dt), half kick, |v|^2 and |F|^2
    double vv = 0, ff = 0;
    for (int t = t0 + w; t < t1; t += nw) {
      const int v = V.tile_v0[t] + lane;
      if (v < V.tile_v1[t]) {
        double vx = V.vx[v], vy = V.vy[v];
        const double fx = V.fx[v], fy = V.fy[v];
        if (reset) {
          V.x[v] -= h * vx;
          V.y[v] -= h * vy;
          vx = 0;
          vy = 0;
        }
        vx += h * fx;
        vy += h * fy;
        V.vx[v] = vx;
        V.vy[v] = vy;
        vv += vx * vx + vy * vy;
        ff += fx * fx + fy * fy;
      }
    }
    double VV, FF;
    block_sum_sum(vv, ff, red[0], VV, FF);
    // ---- bendVertexVelocityTowardsForce (FIRE.cu:374-394), drift, displacement test
    int moved = 0;
    for (int t = t0 + w; t < t1; t += nw) {
      const int v = V.tile_v0[t] + lane;
      if (v < V.tile_v1[t]) {
        double vx = V.vx[v], vy = V.vy[v];
        if (FF > 0) {
          const double ratio = sqrt(VV / FF);
          vx = (1 - fa) * vx + fa * V.fx[v] * ratio;
          vy = (1 - fa) * vy + fa * V.fy[v] * ratio;
        }
        const double xn = V.x[v] + dt * vx, yn = V.y[v] + dt * vy;
        V.vx[v] = vx;
        V.vy[v] = vy;
        V.x[v] = xn;
        V.y[v] = yn;
        const double ddx = pbx(xn - V.lx[v], B), ddy = pby(yn - V.ly[v], B);
        if (3 * sqrt(ddx * ddx + ddy * ddy) > P.cutoff) moved = 1;
      }
    }
    // ---- checkNeighbors: rebuild the system's rows from cells binned here (DPM2D.cu:1408-1423)
    if (__syncthreads_or(moved)) {
      const int ncx = G.ncx[s], ncy = G.ncy[s];
      int* count = G.count[0];
      if (!G.brute[s]) {
        for (int t = t0 + w; t < t1; t += nw) {
          const int v = V.tile_v0[t] + lane;
          if (v < V.tile_v1[t]) {
            int cx, cy;
            cell_of(V.x[v], V.y[v], B, ncx, ncy, cx, cy);
            const int c = G.coff[s] + cy * ncx + cx;
            const int slot = atomicAdd(&count[c], 1);
            if (slot < G.cap) G.slots[(long)c * G.cap + slot] = v;
            else atomicOr(err, ERR_CELL_OVERFLOW);
          }
        }
      }
      __syncthreads();
      for (int t = t0 + w; t < t1; t += nw) {
        const int v = V.tile_v0[t] + lane;
        if (v < V.tile_v1[t]) build_row(V, P, G, count, G.slots, v, s, B, err, maxcnt);
      }
      __syncthreads();
      if (!G.brute[s])
        for (int c = threadIdx.x; c < ncx * ncy; c += blockDim.x) count[G.coff[s] + c] = 0;
      if (threadIdx.x == 0) updateCount[s] += 1;
    }
    // ---- calcForceEnergy (shape + interaction) and the second half kick
    double p = 0, m = -1;
    for (int t = t0 + w; t < t1; t += nw) {
      Lane L;
      lane_ctx(V, t, lane, L);
      const double x = V.x[L.v], y = V.y[L.v];
      double fx = 0, fy = 0, e = 0, es = 0, qx = 0, qy = 0;
      double qnx = 0, qny = 0;  // sets with wide rings do not run this kernel
      gather_pairs<INTER, POT>(V, P, L, B, x, y, 0, 1, fx, fy, e, es, qx, qy, qnx, qny);
      const ShapeOut so = ring_shape(V, P, L, B, x, y, true);
      if constexpr (INTER == INT_SMOOTH) {
        fx += shfl_d(qx, L.nl);
        fy += shfl_d(qy, L.nl);
        const double pe = seg_sum(L.act ? es : 0.0, L.lane, L.segS, L.segE);
        if (L.act) {
          V.es[L.v] = es;
          if (L.lane == L.segS) V.pe[L.r] = pe;
        }
      }
      fx = so.fx + fx;
      fy = so.fy + fy;
      if (L.act) {
        const double vx = V.vx[L.v] + h * fx, vy = V.vy[L.v] + h * fy;
        V.fx[L.v] = fx;
        V.fy[L.v] = fy;
        V.ev[L.v] = so.e + e;
        V.vx[L.v] = vx;
        V.vy[L.v] = vy;
        p += vx * fx + vy * fy;
        m = fmax(m, fmax(fx * fx, fy * fy));
      }
    }
    block_sum_max(p, m, red[1], Pv, Fm);
  }
  if (threadIdx.x == 0) fst[s] = fs;
}

// --------------------------------------------------------- launchers ---
static inline int grid_v(long n) { return (int)((n + 255) / 256); }
static inline int grid_w(int ntiles) { return (ntiles + 3) / 4; }

// LDS of the step kernel: six partials per lane of each gather wave (eight with wide rings)
static inline size_t step_lds(const dpm_set_s* S) {
  const size_t red = (size_t)(S->split - 1) * (S->nwide > 0 ? 8 : 6) * WAVE * sizeof(double);
  const size_t rows = (S->split >= 2 && S->V.W <= 64) ? (size_t)(64 + 64 * S->V.W) * sizeof(int) : 0;
  return red > rows ? red : rows;
}

// rigid-ring modes (sets of rings of at most 64 vertices only)
template <int INTER, int POT>
static hipError_t step_rigid(dpm_set_s* S, int mode, const StepArgs& A) {
  const int g = S->V.ntiles;
  const int b = WAVE * S->split;
  const size_t sh = step_lds(S);
  hipStream_t st = S->stream;
  switch (mode) {
    case MODE_RIGID: k_step<INTER, POT, MODE_RIGID><<<g, b, sh, st>>>(S->V, S->P, A, S->G); break;
    case MODE_RIGID | MODE_KICK:
      k_step<INTER, POT, MODE_RIGID | MODE_KICK><<<g, b, sh, st>>>(S->V, S->P, A, S->G);
      break;
    case MODE_RIGID | MODE_KICK | MODE_START:
      k_step<INTER, POT, MODE_RIGID | MODE_KICK | MODE_START><<<g, b, sh, st>>>(S->V, S->P, A, S->G);
      break;
    case MODE_RIGID | MODE_FIRE:
      k_step<INTER, POT, MODE_RIGID | MODE_FIRE><<<g, b, sh, st>>>(S->V, S->P, A, S->G);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int INTER, int POT, int WB>
static hipError_t step_mode(dpm_set_s* S, int mode, const StepArgs& A) {
  if constexpr (WB != 0) {
    if (mode & MODE_RIGID) return hipErrorInvalidValue;  // refused for wide sets before this point
  } else {
    if (mode & MODE_RIGID) return step_rigid<INTER, POT>(S, mode, A);
  }
  const int g = S->V.ntiles;
  const int b = WAVE * S->split;
  const size_t sh = step_lds(S);
  hipStream_t st = S->stream;
  switch (mode) {
    case MODE_FORCE: k_step<INTER, POT, MODE_FORCE | WB><<<g, b, sh, st>>>(S->V, S->P, A, S->G); break;
    case MODE_KICK: k_step<INTER, POT, MODE_KICK | WB><<<g, b, sh, st>>>(S->V, S->P, A, S->G); break;
    case MODE_KICK | MODE_LANGEVIN:
      k_step<INTER, POT, MODE_KICK | MODE_LANGEVIN | WB><<<g, b, sh, st>>>(S->V, S->P, A, S->G);
      break;
    case MODE_KICK | MODE_START:
      k_step<INTER, POT, MODE_KICK | MODE_START | WB><<<g, b, sh, st>>>(S->V, S->P, A, S->G);
      break;
    case MODE_KICK | MODE_LANGEVIN | MODE_START:
      k_step<INTER, POT, MODE_KICK | MODE_LANGEVIN | MODE_START | WB><<<g, b, sh, st>>>(S->V, S->P, A, S->G);
      break;
    case MODE_FIRE: k_step<INTER, POT, MODE_FIRE | WB><<<g, b, sh, st>>>(S->V, S->P, A, S->G); break;
    case MODE_KICK | MODE_BAOAB:
      k_step<INTER, POT, MODE_KICK | MODE_BAOAB | WB><<<g, b, sh, st>>>(S->V, S->P, A, S->G);
      break;
    case MODE_KICK | MODE_BAOAB | MODE_START:
      k_step<INTER, POT, MODE_KICK | MODE_BAOAB | MODE_START | WB><<<g, b, sh, st>>>(S->V, S->P, A, S->G);
      break;
    case MODE_BROWN | MODE_START:
      k_step<INTER, POT, MODE_BROWN | MODE_START | WB><<<g, b, sh, st>>>(S->V, S->P, A, S->G);
      break;
    case MODE_BROWN | MODE_LANGEVIN | MODE_START:
      k_step<INTER, POT, MODE_BROWN | MODE_LANGEVIN | MODE_START | WB><<<g, b, sh, st>>>(S->V, S->P, A, S->G);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int INTER, int WB>
static hipError_t step_pot(dpm_set_s* S, int mode, const StepArgs& A) {
  switch (S->P.potential) {
    case POT_HARMONIC: return step_mode<INTER, POT_HARMONIC, WB>(S, mode, A);
    case POT_LJ: return step_mode<INTER, POT_LJ, WB>(S, mode, A);
    case POT_ADHESIVE: return step_mode<INTER, POT_ADHESIVE, WB>(S, mode, A);
    case POT_WCA: return step_mode<INTER, POT_WCA, WB>(S, mode, A);
  }
  return hipErrorInvalidValue;
}

template <int WB>
static hipError_t step_all(dpm_set_s* S, const StepArgs& A) {
  const int g = S->V.ntiles, b = WAVE * S->split;
  const size_t sh = step_lds(S);
  hipStream_t st = S->stream;
  switch (S->P.potential) {
    case POT_HARMONIC: k_step<INT_ALL, POT_HARMONIC, MODE_FORCE | WB><<<g, b, sh, st>>>(S->V, S->P, A, S->G); break;
    case POT_LJ: k_step<INT_ALL, POT_LJ, MODE_FORCE | WB><<<g, b, sh, st>>>(S->V, S->P, A, S->G); break;
    case POT_ADHESIVE: k_step<INT_ALL, POT_ADHESIVE, MODE_FORCE | WB><<<g, b, sh, st>>>(S->V, S->P, A, S->G); break;
    case POT_WCA: k_step<INT_ALL, POT_WCA, MODE_FORCE | WB><<<g, b, sh, st>>>(S->V, S->P, A, S->G); break;
  }
  return hipGetLastError();
}

static hipError_t launch_step_kernel(dpm_set_s* S, int mode, const StepArgs& A);

// Sets with wide rings (more than 64 vertices) run, around each step kernel, the ring sums
// of their slices before it (and, under the vertex-segment interaction inside loops, the row
// rebuild), and the wide rings' interaction energies after it.
hipError_t launch_step(dpm_set_s* S, int mode, const StepArgs& A0) {
  if (S->nwide == 0) return launch_step_kernel(S, mode, A0);
  if ((mode & MODE_RIGID) || S->P.particle == PART_ACTIVE || S->P.particle == PART_ACTIVE_SHAPE)
    return hipErrorInvalidValue;  // refused by the entry points before this point
  StepArgs A = A0;
  const bool smooth = S->P.interaction == INT_SMOOTH;
  if (smooth && A.manage) {
    if (!A.flagCur || !A.cellCur || !A.slotsCur || !A.err || !A.maxcnt) return hipErrorInvalidValue;
    k_rows_loop<<<grid_v(S->nv), 256, 0, S->stream>>>(S->V, S->P, A, S->G);
    A.rowsBuilt = 1;
  }
  k_wide_rings<<<(int)S->nwide, 256, 0, S->stream>>>(S->V, S->d_wide);
  hipError_t e = launch_step_kernel(S, mode, A);
  if (e != hipSuccess) return e;
  if (smooth) k_wide_pe<<<(int)S->nwide, 256, 0, S->stream>>>(S->V, S->d_wide);
  return hipGetLastError();
}

static hipError_t launch_step_kernel(dpm_set_s* S, int mode, const StepArgs& A0) {
  // refuse argument sets the chosen epilogue would dereference as null
  if (!A0.ctr) return hipErrorInvalidValue;
  // in-loop rebuilds by the whole workgroup need LDS for the tile's rows: 64 + 64 W ints, in the
  // region the gather partials use afterwards (launches of more than one wave, W <= 64)
  StepArgs A = A0;
  A.rowsLdsW = (A.manage && S->split >= 2 && S->V.W <= 64) ? S->V.W : 0;
  if (S->P.particle == PART_ACTIVE_SHAPE && !S->V.l0r) return hipErrorInvalidValue;
  if ((mode & MODE_FIRE) && (!A.fire || !A.part)) return hipErrorInvalidValue;
  if ((mode & MODE_RIGID) && S->P.interaction == INT_ALL) return hipErrorInvalidValue;
  if ((mode & MODE_START) && (!A.xout || !A.yout || !A.flagOut || !A.anyOut)) return hipErrorInvalidValue;
  if (A.manage && (!A.flagCur || !A.cellCur || !A.slotsCur || (A.genCur && !A.genOut) || !A.err || !A.maxcnt ||
                   !A.updateCount))
    return hipErrorInvalidValue;
  const bool wide = S->nwide > 0;
  switch (S->P.interaction) {
    case INT_VV: return wide ? step_pot<INT_VV, MODE_WIDE>(S, mode, A) : step_pot<INT_VV, 0>(S, mode, A);
    case INT_SMOOTH:
      return wide ? step_pot<INT_SMOOTH, MODE_WIDE>(S, mode, A) : step_pot<INT_SMOOTH, 0>(S, mode, A);
    case INT_ALL:
      if (mode != MODE_FORCE) return hipErrorInvalidValue;
      return wide ? step_all<MODE_WIDE>(S, A) : step_all<0>(S, A);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_shape_only(dpm_set_s* S) {
  if (S->nwide > 0) {
    k_wide_rings<<<(int)S->nwide, 256, 0, S->stream>>>(S->V, S->d_wide);
    k_shape<true><<<grid_w(S->V.ntiles), 256, 0, S->stream>>>(S->V, S->P);
  } else {
    k_shape<false><<<grid_w(S->V.ntiles), 256, 0, S->stream>>>(S->V, S->P);
  }
  return hipGetLastError();
}

hipError_t launch_start_baoab(dpm_set_s* S, double h, const StepArgs& A) {
  if (!A.genOut || !A.flagClear || !A.flagOut || !A.anyOut || !A.cellCur || !A.slotsCur || !A.ctr)
    return hipErrorInvalidValue;
  k_start_loop<true><<<grid_v(S->nv > S->G.ncells ? S->nv : S->G.ncells), 256, 0, S->stream>>>(S->V, S->P, S->G, h,
                                                                                           S->P.dt, A);
  return hipGetLastError();
}

hipError_t launch_start_md(dpm_set_s* S, double h, double dt, const StepArgs& A) {
  if (!A.genOut || !A.flagClear || !A.flagOut || !A.anyOut || !A.cellCur || !A.slotsCur) return hipErrorInvalidValue;
  k_start_loop<false><<<grid_v(S->nv > S->G.ncells ? S->nv : S->G.ncells), 256, 0, S->stream>>>(S->V, S->P, S->G, h,
                                                                                            dt, A);
  return hipGetLastError();
}

hipError_t launch_bin_flag(dpm_set_s* S, int* count, int* slots, int* flag, int* genOut) {
  StepArgs A{};
  A.cellCur = count;
  A.slotsCur = slots;
  A.genOut = genOut;
  A.err = S->d_err;
  k_bin_flag<<<grid_v(S->nv > S->nsys ? S->nv : S->nsys), 256, 0, S->stream>>>(S->V, S->P, S->G, A, flag);
  return hipGetLastError();
}

hipError_t launch_ctr_add(dpm_set_s* S, long long addT, long long addA) {
  k_ctr_add<<<1, 64, 0, S->stream>>>(S->d_ctr, addT, addA);
  return hipGetLastError();
}

hipError_t launch_bin(dpm_set_s* S, int bumpThermal, int bumpActive) {
  const int b = S->cur;
  long n = S->nv > S->nsys ? S->nv : S->nsys;
  k_bin<<<grid_v(n), 256, 0, S->stream>>>(S->V, S->G, S->G.count[b], S->d_flag[b], S->d_any + b,
                                          S->d_flag[1 - b], S->d_any + 1 - b, S->d_ctr, bumpThermal, bumpActive,
                                          S->d_err);
  return hipGetLastError();
}

// k_bin on positions other than the set's current buffer (the adjoint's recorded positions),
// with the flags and cells of `parity`; no RNG counter bump.
hipError_t launch_bin_view(dpm_set_s* S, const View& Vn, int parity) {
  long n = S->nv > S->nsys ? S->nv : S->nsys;
  k_bin<<<grid_v(n), 256, 0, S->stream>>>(Vn, S->G, S->G.count[parity], S->d_flag[parity], S->d_any + parity,
                                          S->d_flag[1 - parity], S->d_any + 1 - parity, S->d_ctr, 0, 0, S->d_err);
  return hipGetLastError();
}

hipError_t launch_build_rows(dpm_set_s* S, int countUpdates) {
  const int b = S->cur;
  k_build_rows<<<grid_w(S->V.ntiles), 256, 0, S->stream>>>(S->V, S->P, S->G, S->G.count[b], S->d_flag[b],
                                                             S->d_updateCount, countUpdates, S->d_err,
                                                             S->d_maxcnt);
  return hipGetLastError();
}

// k_build_rows on positions other than the set's current buffer (the list positions), with the
// flags and cells of `parity`
hipError_t launch_build_rows_view(dpm_set_s* S, const View& Vn, int parity, int countUpdates) {
  k_build_rows<<<grid_w(S->V.ntiles), 256, 0, S->stream>>>(Vn, S->P, S->G, S->G.count[parity], S->d_flag[parity],
                                                             S->d_updateCount, countUpdates, S->d_err,
                                                             S->d_maxcnt);
  return hipGetLastError();
}

hipError_t launch_fill_int(int* p, int value, long n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  k_fill_int<<<grid_v(n), 256, 0, st>>>(p, value, n);
  return hipGetLastError();
}

hipError_t launch_reduce(dpm_set_s* S) {
  k_reduce<<<(int)S->nsys, 256, 0, S->stream>>>(S->V, S->d_red);
  return hipGetLastError();
}

hipError_t launch_inject(dpm_set_s* S, double amplitude, unsigned long long call) {
  k_inject<<<grid_v(S->nv), 256, 0, S->stream>>>(S->V, S->P, amplitude, call);
  return hipGetLastError();
}

hipError_t launch_update_pos(dpm_set_s* S, double dt) {
  k_update_pos<<<grid_v(S->nv), 256, 0, S->stream>>>(S->V, dt);
  return hipGetLastError();
}

hipError_t launch_update_vel(dpm_set_s* S, double dt) {
  k_update_vel<<<grid_v(S->nv), 256, 0, S->stream>>>(S->V, dt);
  return hipGetLastError();
}

hipError_t launch_thermal_force(dpm_set_s* S) {
  k_thermal_force<<<grid_v(S->nv), 256, 0, S->stream>>>(S->V, S->P, S->d_ctr);
  return hipGetLastError();
}

hipError_t launch_sub_mean(dpm_set_s* S) {
  k_sub_mean<<<(int)S->nsys, 256, 0, S->stream>>>(S->V);
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

hipError_t launch_scale_box(dpm_set_s* S) {
  k_scale_box<<<grid_v(S->nv), 256, 0, S->stream>>>(S->V, S->d_scale);
  k_scale_centres<<<grid_v(S->np), 256, 0, S->stream>>>(S->V, S->d_scale);
  return hipGetLastError();
}

hipError_t launch_rigid_start(dpm_set_s* S, bool langevin, const StepArgs& A) {
  if (!A.ctr || (A.genCur && (!A.genOut || !A.cellCur || !A.slotsCur)) || (!A.genCur && !A.cellOther) ||
      !A.flagOut || !A.anyOut || !A.xout || !A.yout)
    return hipErrorInvalidValue;
  StepArgs B = A;
  B.rigidLangevin = langevin ? 1 : 0;
  long n = S->np > S->G.ncells ? S->np : S->G.ncells;
  if (n < S->nsys) n = S->nsys;
  k_rigid_start<<<grid_v(n), 256, 0, S->stream>>>(S->V, S->P, S->G, B);
  return hipGetLastError();
}

hipError_t launch_rigid_inject(dpm_set_s* S, double amplitude, unsigned long long call) {
  k_rigid_inject<<<(int)S->nsys, 256, 0, S->stream>>>(S->V, S->P, amplitude, call);
  return hipGetLastError();
}

hipError_t launch_rigid_fire_partials(dpm_set_s* S) {
  k_rigid_fire_partials<<<grid_w(S->V.ntiles), 256, 0, S->stream>>>(S->V, S->d_part);
  return hipGetLastError();
}

hipError_t launch_rigid_fire_move(dpm_set_s* S, int* flagOut, int* anyOut) {
  if (!flagOut || !anyOut) return hipErrorInvalidValue;
  const int nb = 1 - S->cur;
  k_rigid_fire_move<<<grid_w(S->V.ntiles), 256, 0, S->stream>>>(
      S->V, S->P, S->fp, S->d_rfire[S->rfireCur], S->d_rfire[1 - S->rfireCur], S->d_part, S->pos[nb][0],
      S->pos[nb][1], flagOut, anyOut);
  return hipGetLastError();
}

hipError_t launch_copy_lastpos(dpm_set_s* S) {
  k_copy_lastpos<<<grid_v(S->nv), 256, 0, S->stream>>>(S->V);
  return hipGetLastError();
}

hipError_t launch_fire_begin(dpm_set_s* S) {
  k_fire_begin<<<grid_w(S->V.ntiles), 256, 0, S->stream>>>(S->V, S->fp, S->d_fire[S->fireCur],
                                                            S->d_fire[1 - S->fireCur], S->d_part);
  return hipGetLastError();
}

hipError_t launch_fire_mid(dpm_set_s* S) {
  const int nb = 1 - S->cur;
  k_fire_mid<<<grid_w(S->V.ntiles), 256, 0, S->stream>>>(S->V, S->P, S->d_fire[S->fireCur], S->d_part,
                                                          S->pos[nb][0], S->pos[nb][1], S->d_flag[nb],
                                                          S->d_any + nb);
  return hipGetLastError();
}

hipError_t launch_fire_partials(dpm_set_s* S) {
  k_fire_partials<<<grid_w(S->V.ntiles), 256, 0, S->stream>>>(S->V, S->d_part);
  return hipGetLastError();
}

template <int INTER, int POT>
static hipError_t fire_sys_t(dpm_set_s* S, int chunk, int threads) {
  if (threads > WAVE * FS_WAVES)
    k_fire_sys<INTER, POT, FS_WAVES_WIDE><<<(unsigned)S->nsys, threads, 0, S->stream>>>(
        S->V, S->P, S->fp, S->G, S->d_fire[S->fireCur], chunk, S->d_updateCount, S->d_err, S->d_maxcnt,
        S->d_sysOrder);
  else
    k_fire_sys<INTER, POT, FS_WAVES><<<(unsigned)S->nsys, threads, 0, S->stream>>>(
        S->V, S->P, S->fp, S->G, S->d_fire[S->fireCur], chunk, S->d_updateCount, S->d_err, S->d_maxcnt,
        S->d_sysOrder);
  return hipGetLastError();
}

template <int INTER>
static hipError_t fire_sys_pot(dpm_set_s* S, int chunk, int threads) {
  switch (S->P.potential) {
    case POT_HARMONIC: return fire_sys_t<INTER, POT_HARMONIC>(S, chunk, threads);
    case POT_LJ: return fire_sys_t<INTER, POT_LJ>(S, chunk, threads);
    case POT_ADHESIVE: return fire_sys_t<INTER, POT_ADHESIVE>(S, chunk, threads);
    default: return fire_sys_t<INTER, POT_WCA>(S, chunk, threads);
  }
}

hipError_t launch_fire_sys(dpm_set_s* S, int chunk, int waves) {
  const int threads = WAVE * (waves < 1 ? 1 : (waves > FS_WAVES_WIDE ? FS_WAVES_WIDE : waves));
  switch (S->P.interaction) {
    case INT_VV: return fire_sys_pot<INT_VV>(S, chunk, threads);
    case INT_SMOOTH: return fire_sys_pot<INT_SMOOTH>(S, chunk, threads);
    default: return fire_sys_pot<INT_ALL>(S, chunk, threads);
  }
}

hipError_t launch_fire_resume(dpm_set_s* S) {
  k_fire_resume<<<grid_v(S->nsys), 256, 0, S->stream>>>(S->d_fire[S->fireCur], (int)S->nsys);
  return hipGetLastError();
}

hipError_t launch_fire_done(dpm_set_s* S, int* notDone) {
  k_fire_done<<<grid_v(S->nsys), 256, 0, S->stream>>>(S->d_fire[S->fireCur], (int)S->nsys, notDone);
  return hipGetLastError();
}

StepArgs loop_args(dpm_set_s* S) {
  const int b = S->cur;
  StepArgs A{};
  A.h = 0.5 * S->P.dt;
  A.dt = S->P.dt;
  A.xout = S->pos[1 - b][0];
  A.yout = S->pos[1 - b][1];
  A.flagCur = S->d_flag[b];
  A.flagOut = S->d_flag[1 - b];
  A.anyOut = S->d_any + 1 - b;
  A.cellCur = S->G.count[b];
  A.slotsCur = S->G.slots;
  A.cellOther = S->G.count[1 - b];
  A.ctr = S->d_ctr;
  A.ctrOffT = -1;  // k_bin has bumped the counters of this step
  A.ctrOffA = -1;
  A.err = S->d_err;
  A.maxcnt = S->d_maxcnt;
  A.updateCount = S->d_updateCount;
  A.manage = 1;
  A.xcd = S->xcdMap;
  return A;
}

hipError_t launch_fire_step(dpm_set_s* S) {
  StepArgs A = loop_args(S);
  A.fire = S->d_fire[S->fireCur];
  A.part = S->d_part;
  return launch_step(S, MODE_FIRE, A);
}

}  // namespace dpm
