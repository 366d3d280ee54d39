// dpm_rows.h -- neighbour-row construction shared by the step kernels and the adjoint kernels.
#pragma once

#include "dpm_internal.h"

namespace dpm {

__device__ __forceinline__ Box box_of(const View& V, int s) {
  return Box{V.bx[s], V.by[s], V.ibx[s], V.iby[s], V.mbx[s], V.mby[s]};
}

__device__ __forceinline__ void cell_of(double x, double y, const Box& B, int ncx, int ncy, int& cx, int& cy) {
  const double wx = x - B.Lx * floor(x * B.iLx), wy = y - B.Ly * floor(y * B.iLy);
  // clamp in floating point first (fmax maps NaN to 0), so the conversion is always defined
  const double fx = fmin(fmax(wx * B.iLx * ncx, 0.0), ncx - 1.0);
  const double fy = fmin(fmax(wy * B.iLy * ncy, 0.0), ncy - 1.0);
  cx = (int)fx;
  cy = (int)fy;
}

// previous and second previous vertex of u in its ring (getPreviousId 233-243)
__device__ __forceinline__ int2 prev_pair(const View& V, int u) {
  const int ring = V.vring[u];
  const int f = V.rfirst[ring], n = V.rnv[ring];
  const int p1 = (u == f) ? f + n - 1 : u - 1;
  const int p2 = (p1 == f) ? f + n - 1 : p1 - 1;
  return make_int2(p1, p2);
}

// insert j (with its prev pair) into the ascending ELL row of v that holds c entries
__device__ __forceinline__ void row_insert(const View& V, int v, int c, int j, int2 pj) {
  const long nv = V.nv;
  int k = c;
  while (k > 0) {
    const int prev = V.nbr[(long)(k - 1) * nv + v];
    if (prev <= j) break;
    V.nbr[(long)k * nv + v] = prev;
    V.nbrp[(long)k * nv + v] = V.nbrp[(long)(k - 1) * nv + v];
    k--;
  }
  V.nbr[(long)k * nv + v] = j;
  V.nbrp[(long)k * nv + v] = pj;
}

// Rebuild the row of vertex v.  kernelCalcNeighborList predicate (cudaKernel.cuh:1735-1763):
// j in list(v) iff rings differ and -(1 - d/radSum) < cutDistance; rows stay ascending as in
// the reference.  Candidates come from the 3x3 cell block, or the whole system when the box
// is narrower than three cells or `allPairs` is set (cells that may miss a vertex).  Also
// resets lastPos (resetLastPositions).
__device__ void build_row(const View& V, const Params& P, const CellGrid& G, const int* count, const int* slots, int v,
                          int s, const Box& B, int* err, int* maxcnt, bool allPairs = false) {
  const double x = V.x[v], y = V.y[v], rv = V.rad[v];
  const int r = V.vring[v];
  const int W = V.W;
  int c = 0;
  if (G.brute[s] || allPairs) {
    const int b0 = V.svoff[s], b1 = V.svoff[s + 1];
    for (int j = b0; j < b1; j++) {
      if (V.vring[j] == r) continue;
      const double dx = pbx(x - V.x[j], B), dy = pby(y - V.y[j], B);
      const double d = sqrt(dx * dx + dy * dy);
      if (-(1 - d / (rv + V.rad[j])) < P.cutDistance) {
        if (c < W) {
          V.nbr[(long)c * V.nv + v] = j;
          V.nbrp[(long)c * V.nv + v] = prev_pair(V, j);
        }
        c++;
      }
    }
  } else {
    const int ncx = G.ncx[s], ncy = G.ncy[s];
    int cx, cy;
    cell_of(x, y, B, ncx, ncy, cx, cy);
    for (int oy = -1; oy <= 1; oy++) {
      const int yy = (cy + oy + ncy) % ncy;
      for (int ox = -1; ox <= 1; ox++) {
        const int xx = (cx + ox + ncx) % ncx;
        const int cell = G.coff[s] + yy * ncx + xx;
        const int m = min(count[cell], G.cap);
        for (int q = 0; q < m; q++) {
          const int j = slots[(long)cell * G.cap + q];
          if (V.vring[j] == r) continue;
          const double dx = pbx(x - V.x[j], B), dy = pby(y - V.y[j], B);
          const double d = sqrt(dx * dx + dy * dy);
          if (-(1 - d / (rv + V.rad[j])) < P.cutDistance) {
            if (c < W) row_insert(V, v, c, j, prev_pair(V, j));
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

// The rows of a tile's vertices (lane = vertex, L.act lanes), built by all nw waves of its
// workgroup: thread (w, lane) scans the cells w, w + nw, ... of the 3 x 3 block around its lane's
// vertex and appends the hits to the lane's list in LDS (an LDS counter per lane, so the order
// of arrival varies); after a barrier each lane of wave 0 sorts its list ascending, and all waves
// write the sorted entries with their prev pairs into the ELL rows.  The same rows as build_row
// from the same cells (its predicate, ascending order), with the dependent steps of an insertion
// in LDS instead of memory.  Every thread of the workgroup calls it (it holds barriers); lds holds
// 64 + 64 W ints; the caller synchronises before reading the rows or reusing lds.
__device__ void build_rows_tile(const View& V, const Params& P, const CellGrid& G, const int* count,
                                const int* slots, int lane, int v, int r, int s, bool act, int w, int nw,
                                const Box& B, double x, double y, int* lds, int* err, int* maxcnt) {
  int* lcnt = lds;
  int* lrow = lds + 64;
  const int W = V.W;
  if (w == 0) lcnt[lane] = 0;
  __syncthreads();
  if (act) {
    const double rv = V.rad[v];
    const int ncx = G.ncx[s], ncy = G.ncy[s];
    int cx, cy;
    cell_of(x, y, B, ncx, ncy, cx, cy);
    for (int k = w; k < 9; k += nw) {
      const int yy = (cy + k / 3 - 1 + ncy) % ncy;
      const int xx = (cx + k % 3 - 1 + ncx) % ncx;
      const int cell = G.coff[s] + yy * ncx + xx;
      const int m = min(count[cell], G.cap);
      for (int q = 0; q < m; q++) {
        const int j = slots[(long)cell * G.cap + q];
        if (V.vring[j] == r) continue;
        const double dx = pbx(x - V.x[j], B), dy = pby(y - V.y[j], B);
        const double d = sqrt(dx * dx + dy * dy);
        if (-(1 - d / (rv + V.rad[j])) < P.cutDistance) {
          const int idx = atomicAdd(&lcnt[lane], 1);
          if (idx < W) lrow[idx * 64 + lane] = j;
        }
      }
    }
  }
  __syncthreads();
  const int c = act ? lcnt[lane] : 0;
  const int n = c < W ? c : W;
  if (w == 0 && act) {
    for (int a = 1; a < n; a++) {
      const int key = lrow[a * 64 + lane];
      int b = a - 1;
      while (b >= 0 && lrow[b * 64 + lane] > key) {
        lrow[(b + 1) * 64 + lane] = lrow[b * 64 + lane];
        b--;
      }
      lrow[(b + 1) * 64 + lane] = key;
    }
    if (c > W) atomicOr(err, ERR_LIST_OVERFLOW);
    atomicMax(maxcnt, c);
    V.ncnt[v] = n;
    V.lx[v] = x;
    V.ly[v] = y;
  }
  __syncthreads();
  if (act)
    for (int k = w; k < n; k += nw) {
      const int j = lrow[k * 64 + lane];
      V.nbr[(long)k * V.nv + v] = j;
      V.nbrp[(long)k * V.nv + v] = prev_pair(V, j);
    }
}

// Bin vertex v at (x, y) of system s into a cell list (counts + slots of capacity G.cap);
// systems searched by brute force have no cells.
__device__ __forceinline__ void bin_vertex(const CellGrid& G, const Box& B, int* count, int* slots, int v, int s,
                                           double x, double y, int* err) {
  if (G.brute[s]) return;
  int cx, cy;
  cell_of(x, y, B, G.ncx[s], G.ncy[s], cx, cy);
  const int c = G.coff[s] + cy * G.ncx[s] + cx;
  const int slot = atomicAdd(&count[c], 1);
  if (slot < G.cap) slots[(long)c * G.cap + slot] = v;
  else atomicOr(err, ERR_CELL_OVERFLOW);
}

}  // namespace dpm
