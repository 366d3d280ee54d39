// dpm_tile.h -- wavefront helpers and the lane context of a ring-aligned tile, shared by the
// step kernels and the adjoint kernels.
#pragma once

#include "dpm_internal.h"

namespace dpm {

#define WAVE 64
#define MAXSPLIT 8

__device__ __forceinline__ double shfl_d(double v, int src) { return __shfl(v, src, WAVE); }
__device__ __forceinline__ int shfl_i(int v, int src) { return __shfl(v, src, WAVE); }

// segmented inclusive scan over lanes [segS, lane]
__device__ __forceinline__ double seg_scan(double v, int lane, int segS) {
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    double t = __shfl_up(v, d, WAVE);
    if (lane - d >= segS) v += t;
  }
  return v;
}

// segment total, broadcast to every lane of the segment
__device__ __forceinline__ double seg_sum(double v, int lane, int segS, int segE) {
  return __shfl(seg_scan(v, lane, segS), segE, WAVE);
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

// Lane context of a ring-aligned tile: the lane's vertex v (the tile's first vertex on lanes
// past the tile end, act = false), its ring r and system s, the lanes [segS, segE] holding the
// ring, and the lanes of the next and previous vertex of the ring.
// A ring of more than 64 vertices (wide) is cut into tiles of 64 consecutive vertices, one
// slice each: then [segS, segE] is the slice, the slice's first lane has pl = itself and its
// last lane nl = itself (their ring neighbours lie in other tiles and are read from memory).
struct Lane {
  int lane, tile, v0, v, r, s, segS, segE, nl, pl, n, f;  // f: the ring's first vertex
  bool act, wide;
};

// WIDE: the launch may hold slices of wide rings (sets with rings of more than 64 vertices); the
// other kernels either never run on such sets or use only lane, tile, v, r, s and act.
template <bool WIDE = false>
__device__ __forceinline__ void lane_ctx(const View& V, int tile, int lane, Lane& L) {
  L.lane = lane;
  L.tile = tile;
  L.v0 = V.tile_v0[tile];
  const int v1 = V.tile_v1[tile];
  L.s = V.tile_sys[tile];
  L.act = (L.v0 + lane) < v1;
  L.v = L.act ? L.v0 + lane : L.v0;
  const int4 vi = V.vinfo[L.v];
  L.r = vi.x;
  const int first = vi.y;
  L.f = first;
  L.n = vi.z;
  L.wide = WIDE && L.n > WAVE;
  if (L.wide) {
    L.segS = L.act ? 0 : lane;
    L.segE = L.act ? v1 - L.v0 - 1 : lane;
    L.nl = (lane < L.segE) ? lane + 1 : lane;
    L.pl = (lane > L.segS) ? lane - 1 : lane;
    return;
  }
  L.segS = L.act ? first - L.v0 : lane;
  const int segN = L.act ? L.n : 1;
  L.segE = L.segS + segN - 1;
  L.nl = (lane < L.segE) ? lane + 1 : L.segS;
  L.pl = (lane > L.segS) ? lane - 1 : L.segE;
}

// next and previous vertex of v in its ring (first vertex f, n vertices)
__device__ __forceinline__ int ring_next(int v, int f, int n) { return (v + 1 == f + n) ? f : v + 1; }
__device__ __forceinline__ int ring_prev(int v, int f, int n) { return (v == f) ? f + n - 1 : v - 1; }

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on XCD b mod 8), each
// with its own L2.  This maps the workgroups of XCD k onto the k-th contiguous eighth of the tiles,
// so that rings stored near each other (spatially sorted sets) share one L2.  A bijection of
// [0, nb).
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  const int q = nb >> 3, r = nb & 7, k = b & 7;
  return k * q + (k < r ? k : r) + (b >> 3);
}

}  // namespace dpm
