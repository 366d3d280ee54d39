// dpm_api.hip -- the C ABI (include/dpmring.h) over the gfx950 kernels.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <initializer_list>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/dpmring.h"
#include "dpm_internal.h"

using namespace dpm;

// Scratch of dpm_energy_param_grad_noalloc: the four per-vertex partials (shares of dU/dka,
// dU/dkl, dU/dkb and dU/dl0) before the per-ring sums, kScratchDoubles doubles = 8 MiB, enough
// for sets of DPM_NOALLOC_MAX_VERTICES vertices.  Every GPU the library runs on holds its own
// copy until the process ends; every call on one GPU shares it, so calls that overlap on one GPU
// race.  In host code `scratch` names a host-side stand-in, not a device address: kernels and
// copies take the per-device address that scratch_address() looks up.
constexpr long kScratchDoubles = 4L * DPM_NOALLOC_MAX_VERTICES;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

constexpr double kWCA = 1.122462048;  // defs.h:17

int fail(dpm_set_s* S, int code, const std::string& msg) {
  if (S) S->err = msg;
  return code;
}

std::mutex g_scratchMu;
std::vector<double*> g_scratchAddr;  // device address of `scratch`, per device (null: not looked up)

// Look up (once per device, with `device` current) the device address of `scratch`.
hipError_t scratch_address(int device, double** out) {
  std::lock_guard<std::mutex> lock(g_scratchMu);
  if (device < 0) return hipErrorInvalidDevice;
  if ((int)g_scratchAddr.size() <= device) g_scratchAddr.resize(device + 1, nullptr);
  if (!g_scratchAddr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratchAddr[device] = static_cast<double*>(p);
  }
  *out = g_scratchAddr[device];
  return hipSuccess;
}

#define HIPCHK(expr)                                                                       \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return fail(S, DPM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));     \
  } while (0)

#define CHECK_HANDLE(h)                 \
  if (!(h)) return DPM_ERR_ARG;         \
  dpm_set_s* S = (h);                   \
  if (hipSetDevice(S->device) != hipSuccess) return fail(S, DPM_ERR_HIP, "hipSetDevice");

template <class T>
hipError_t dalloc(T** p, long n) {
  if (n <= 0) n = 1;
  return hipMalloc((void**)p, sizeof(T) * n);
}

template <class T>
hipError_t up(T* d, const std::vector<T>& h, hipStream_t st) {
  if (h.empty()) return hipSuccess;
  return hipMemcpyAsync(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, st);
}

long pow2ceil(long x) {
  long p = 1;
  while (p < x) p <<= 1;
  return p;
}

void bump(dpm_set_s* S) { S->version++; }

void set_pos_view(dpm_set_s* S) {
  S->V.x = S->pos[S->cur][0];
  S->V.y = S->pos[S->cur][1];
  S->V.xn = S->pos[1 - S->cur][0];
  S->V.yn = S->pos[1 - S->cur][1];
}

void swap_pos(dpm_set_s* S) {
  S->cur = 1 - S->cur;
  set_pos_view(S);
}

// scalar/vector field -> device component pointers
bool field_ptrs(dpm_set_s* S, int f, double** c0, double** c1, long* n, int* ncomp) {
  const View& V = S->V;
  *c1 = nullptr;
  *ncomp = 1;
  switch (f) {
    case DPM_FIELD_POS: *c0 = S->pos[S->cur][0]; *c1 = S->pos[S->cur][1]; *n = S->nv; *ncomp = 2; return true;
    case DPM_FIELD_VEL: *c0 = V.vx; *c1 = V.vy; *n = S->nv; *ncomp = 2; return true;
    case DPM_FIELD_FORCE: *c0 = V.fx; *c1 = V.fy; *n = S->nv; *ncomp = 2; return true;
    case DPM_FIELD_ENERGY: *c0 = V.ev; *n = S->nv; return true;
    case DPM_FIELD_RAD: *c0 = V.rad; *n = S->nv; return true;
    case DPM_FIELD_L0: *c0 = V.l0; *n = S->nv; return true;
    case DPM_FIELD_THETA0: *c0 = V.th0; *n = S->nv; return true;
    case DPM_FIELD_THETA: *c0 = V.th; *n = S->nv; return true;
    case DPM_FIELD_LENGTH: *c0 = V.len; *n = S->nv; return true;
    case DPM_FIELD_A0: *c0 = V.a0; *n = S->np; return true;
    case DPM_FIELD_AREA: *c0 = V.area; *n = S->np; return true;
    case DPM_FIELD_PERIMETER: *c0 = V.perim; *n = S->np; return true;
    case DPM_FIELD_PPOS: *c0 = V.cx; *c1 = V.cy; *n = S->np; *ncomp = 2; return true;
    case DPM_FIELD_PENERGY: *c0 = V.pe; *n = S->np; return true;
    case DPM_FIELD_PANGLE: *c0 = V.ang; *n = S->np; return true;
    case DPM_FIELD_LASTPOS: *c0 = V.lx; *c1 = V.ly; *n = S->nv; *ncomp = 2; return true;
    case DPM_FIELD_BOX: *c0 = V.bx; *c1 = V.by; *n = S->nsys; *ncomp = 2; return true;
    case DPM_FIELD_WALLFORCE: *c0 = V.wfx; *c1 = V.wfy; *n = S->nv; *ncomp = 2; return true;
    case DPM_FIELD_PVEL: *c0 = S->R.pvx; *c1 = S->R.pvy; *n = S->np; *ncomp = 2; return true;
    case DPM_FIELD_PANGVEL: *c0 = S->R.pw; *n = S->np; return true;
    case DPM_FIELD_PFORCE: *c0 = S->R.pfx; *c1 = S->R.pfy; *n = S->np; *ncomp = 2; return true;
    case DPM_FIELD_PTORQUE: *c0 = S->R.ptq; *n = S->np; return true;
    case DPM_FIELD_VTORQUE: *c0 = S->R.vtq; *n = S->nv; return true;
    case DPM_FIELD_PRAD: *c0 = S->R.prad; *n = S->np; return true;
  }
  return false;
}

// box reciprocals, and the image multipliers of the geometry (periodic x for normal and
// fixedSides, periodic y for normal only)
int upload_box(dpm_set_s* S) {
  std::vector<double> ix(S->nsys), iy(S->nsys), mx(S->nsys), my(S->nsys);
  const bool px = S->P.geometry == GEOM_NORMAL || S->P.geometry == GEOM_FIXED_SIDES;
  const bool py = S->P.geometry == GEOM_NORMAL;
  for (long s = 0; s < S->nsys; s++) {
    ix[s] = 1.0 / S->h_box[2 * s];
    iy[s] = 1.0 / S->h_box[2 * s + 1];
    mx[s] = px ? ix[s] : 0.0;
    my[s] = py ? iy[s] : 0.0;
  }
  HIPCHK(up(S->V.ibx, ix, S->stream));
  HIPCHK(up(S->V.iby, iy, S->stream));
  HIPCHK(up(S->V.mbx, mx, S->stream));
  HIPCHK(up(S->V.mby, my, S->stream));
  HIPCHK(hipStreamSynchronize(S->stream));
  return DPM_OK;
}

// the step loop's three cell lists (counts + slots at the current capacity, one allocation
// each), four flag rows and three generation rows
int alloc_md_cells(dpm_set_s* S) {
  HIPCHK(hipStreamSynchronize(S->stream));
  const long C = S->G.ncells > 0 ? S->G.ncells : 1;
  hipFree(S->mdCount[0]);
  hipFree(S->mdSlots[0]);
  for (int i = 0; i < 3; i++) S->mdCount[i] = S->mdSlots[i] = nullptr;
  HIPCHK(dalloc(&S->mdCount[0], 3 * C));
  HIPCHK(dalloc(&S->mdSlots[0], 3 * C * S->G.cap));
  for (int i = 1; i < 3; i++) {
    S->mdCount[i] = S->mdCount[0] + i * C;
    S->mdSlots[i] = S->mdSlots[0] + i * C * S->G.cap;
  }
  HIPCHK(launch_fill_int(S->mdCount[0], 0, 3 * C, S->stream));
  for (int i = 0; i < 4; i++)
    if (!S->mdFlag[i]) {
      HIPCHK(dalloc(&S->mdFlag[i], S->nsys));
      HIPCHK(launch_fill_int(S->mdFlag[i], 0, S->nsys, S->stream));
    }
  for (int i = 0; i < 3; i++)
    if (!S->mdGen[i]) {
      HIPCHK(dalloc(&S->mdGen[i], S->nsys));
      HIPCHK(launch_fill_int(S->mdGen[i], 0, S->nsys, S->stream));
    }
  HIPCHK(hipStreamSynchronize(S->stream));
  return DPM_OK;
}

// (Re)configure the per-system cell grids for the current cutDistance, boxes and radii.
int configure_cells(dpm_set_s* S) {
  const long B = S->nsys;
  std::vector<int> ncx(B), ncy(B), coff(B), brute(B);
  long C = 0;
  double maxMean = 1;
  for (long s = 0; s < B; s++) {
    // list radius, plus the 2 cutoff / 3 a vertex may lie from its list position (where the step
    // loop binned it) when its system rebuilds: at most cutoff / 3 before the drift that
    // triggers the rebuild (the displacement test) and cutoff / 3 in that drift (FLAG_FAR)
    const double rc = ((1 + S->P.cutDistance) * 2 * S->h_maxrad[s] + 2 * S->P.cutoff / 3) * S->cellScale;
    const double Lx = S->h_box[2 * s], Ly = S->h_box[2 * s + 1];
    int nx = rc > 0 ? (int)floor(Lx / rc) : 0;
    int ny = rc > 0 ? (int)floor(Ly / rc) : 0;
    if (nx > 4096) nx = 4096;
    if (ny > 4096) ny = 4096;
    if (nx < 3 || ny < 3) {
      brute[s] = 1;
      nx = ny = 0;
    }
    ncx[s] = nx;
    ncy[s] = ny;
    coff[s] = (int)C;
    C += (long)nx * ny;
    if (nx > 0) {
      const double mean = (double)(S->h_svoff[s + 1] - S->h_svoff[s]) / ((double)nx * ny);
      maxMean = std::max(maxMean, mean);
    }
  }
  int cap = (int)pow2ceil((long)(4 * maxMean + 16));
  if (S->G.cap > cap && S->G.ncells == C) cap = S->G.cap;  // keep a grown capacity
  std::vector<int> cellSys(C);
  for (long s = 0; s < B; s++)
    for (long c = 0; c < (long)ncx[s] * ncy[s]; c++) cellSys[coff[s] + c] = (int)s;
  HIPCHK(hipStreamSynchronize(S->stream));
  hipFree(S->G.count[0]); hipFree(S->G.count[1]); hipFree(S->G.slots); hipFree(S->G.cellSys);
  S->G.count[0] = S->G.count[1] = S->G.slots = S->G.cellSys = nullptr;
  HIPCHK(dalloc(&S->G.count[0], C));
  HIPCHK(dalloc(&S->G.count[1], C));
  HIPCHK(dalloc(&S->G.slots, C * cap));
  HIPCHK(dalloc(&S->G.cellSys, C));
  HIPCHK(up(S->G.cellSys, cellSys, S->stream));
  HIPCHK(up(S->G.ncx, ncx, S->stream));
  HIPCHK(up(S->G.ncy, ncy, S->stream));
  HIPCHK(up(S->G.coff, coff, S->stream));
  HIPCHK(up(S->G.brute, brute, S->stream));
  HIPCHK(launch_fill_int(S->G.count[0], 0, C, S->stream));
  HIPCHK(launch_fill_int(S->G.count[1], 0, C, S->stream));
  HIPCHK(hipStreamSynchronize(S->stream));
  S->G.ncells = (int)C;
  S->G.cap = cap;
  int rc = alloc_md_cells(S);
  if (rc) return rc;
  bump(S);
  return DPM_OK;
}

int alloc_list(dpm_set_s* S, int W) {
  HIPCHK(hipStreamSynchronize(S->stream));
  hipFree(S->V.nbr);
  hipFree(S->V.nbrp);
  S->V.nbr = nullptr;
  S->V.nbrp = nullptr;
  HIPCHK(dalloc(&S->V.nbr, (long)W * S->nv));
  HIPCHK(dalloc(&S->V.nbrp, (long)W * S->nv));
  HIPCHK(launch_fill_int(S->V.nbr, -1, (long)W * S->nv, S->stream));
  HIPCHK(launch_fill_int((int*)S->V.nbrp, 0, 2L * W * S->nv, S->stream));
  HIPCHK(launch_fill_int(S->V.ncnt, 0, S->nv, S->stream));
  S->V.W = W;
  S->listBuilt = false;
  bump(S);
  return DPM_OK;
}

// Rebuild the rows of the systems flagged in `flags` (host array) from the current positions,
// or with `atListPositions` from the positions of the last rebuild (lastPos: the rows as they
// were built then).  calcNeighborList (countUpdates = 0) and checkNeighbors (countUpdates = 1)
// semantics.  A wider row needs new storage: then every system is rebuilt into it before
// returning.
int rebuild(dpm_set_s* S, const std::vector<int>& flags, int countUpdates, bool atListPositions = false) {
  const int b = S->cur;
  View Vl = S->V;
  if (atListPositions) {
    Vl.x = S->V.lx;
    Vl.y = S->V.ly;
  }
  for (int attempt = 0; attempt < 8; attempt++) {
    int any = 0;
    for (int f : flags) any |= f;
    HIPCHK(hipMemcpyAsync(S->d_flag[b], flags.data(), sizeof(int) * S->nsys, hipMemcpyHostToDevice, S->stream));
    HIPCHK(hipMemcpyAsync(S->d_any + b, &any, sizeof(int), hipMemcpyHostToDevice, S->stream));
    HIPCHK(launch_fill_int(S->d_err, 0, 1, S->stream));
    HIPCHK(launch_fill_int(S->d_maxcnt, 0, 1, S->stream));
    HIPCHK(launch_fill_int(S->G.count[b], 0, S->G.ncells, S->stream));
    if (atListPositions) {
      HIPCHK(launch_bin_view(S, Vl, b));
      HIPCHK(launch_build_rows_view(S, Vl, b, countUpdates));
    } else {
      HIPCHK(launch_bin(S, 0, 0));
      HIPCHK(launch_build_rows(S, countUpdates));
    }
    int err = 0, maxcnt = 0;
    HIPCHK(hipMemcpyAsync(&err, S->d_err, sizeof(int), hipMemcpyDeviceToHost, S->stream));
    HIPCHK(hipMemcpyAsync(&maxcnt, S->d_maxcnt, sizeof(int), hipMemcpyDeviceToHost, S->stream));
    HIPCHK(hipStreamSynchronize(S->stream));
    // leave both parities' flags and the cells clean
    HIPCHK(launch_fill_int(S->d_flag[0], 0, S->nsys, S->stream));
    HIPCHK(launch_fill_int(S->d_flag[1], 0, S->nsys, S->stream));
    HIPCHK(launch_fill_int(S->d_any, 0, 2, S->stream));
    HIPCHK(launch_fill_int(S->G.count[b], 0, S->G.ncells, S->stream));
    HIPCHK(hipStreamSynchronize(S->stream));
    if (err & ERR_CELL_OVERFLOW) {
      S->G.cap *= 2;
      hipFree(S->G.slots);
      S->G.slots = nullptr;
      HIPCHK(dalloc(&S->G.slots, (long)S->G.ncells * S->G.cap));
      int rc = alloc_md_cells(S);
      if (rc) return rc;
      bump(S);
      continue;
    }
    // keep head room so device-side rebuilds during a loop do not overflow
    const int want = (int)pow2ceil(std::max(8, 2 * maxcnt));
    if (want > S->V.W || (err & ERR_LIST_OVERFLOW)) {
      int rc = alloc_list(S, std::max(want, S->V.W));
      if (rc) return rc;
      std::vector<int> all(S->nsys, 1);
      return rebuild(S, all, 0, atListPositions);
    }
    S->listBuilt = true;
    return DPM_OK;
  }
  return fail(S, DPM_ERR_OVERFLOW, "neighbour list rebuild did not converge");
}

// status of a loop whose kernels left the error bits `err`
int loop_status(dpm_set_s* S, int err) {
  if (err & ERR_STALE_BINS)
    return fail(S, DPM_ERR_STATE, "a rebuilding vertex lay more than 2 cutoff / 3 from the position its cell "
                                  "was binned at: the neighbour-list skin does not cover the time step");
  if (err) return fail(S, DPM_ERR_OVERFLOW, err & ERR_LIST_OVERFLOW ? "neighbour list overflow during loop"
                                                                    : "cell overflow during loop");
  return DPM_OK;
}

int check_err_after_loop(dpm_set_s* S) {
  int err = 0;
  HIPCHK(hipMemcpyAsync(&err, S->d_err, sizeof(int), hipMemcpyDeviceToHost, S->stream));
  HIPCHK(hipStreamSynchronize(S->stream));
  return loop_status(S, err);
}

// area of system s's domain: the box, or pi R^2 for the round box (getPhi, DPM2D.cu:1208-1246)
double box_area(const dpm_set_s* S, long s) {
  if (S->P.geometry == GEOM_ROUND_BOX) return kPI * S->P.boxRadius * S->P.boxRadius;
  return S->h_box[2 * s] * S->h_box[2 * s + 1];
}

int reduce_host(dpm_set_s* S) {
  HIPCHK(launch_reduce(S));
  S->h_red.resize(S->nsys * 8);
  HIPCHK(hipMemcpyAsync(S->h_red.data(), S->d_red, sizeof(double) * S->nsys * 8, hipMemcpyDeviceToHost,
                        S->stream));
  HIPCHK(hipStreamSynchronize(S->stream));
  return DPM_OK;
}

int require_list(dpm_set_s* S) {
  if (S->P.particle == PART_ACTIVE_SHAPE && !S->V.l0r)
    return fail(S, DPM_ERR_STATE, "activeShape: dpm_set_active_tension not called");
  if (S->P.interaction == INT_ALL) return DPM_OK;
  if (!S->listBuilt) return fail(S, DPM_ERR_STATE, "neighbour list not built: call dpm_calc_neighbor_list first");
  return DPM_OK;
}

// ------------------------------------------------------- replay after a list overflow ---
// A loop's in-kernel rebuild that finds a row longer than the list width W keeps W entries and
// flags ERR_LIST_OVERFLOW; the loop's later forces would miss pairs.  Each loop therefore runs
// from a copy of what it changes: on that error the set is put back to the copy, the list is
// widened with its rows rebuilt from the lastPos positions (the rows the loop started with) and
// the loop runs again, so no force ever uses a truncated row.  The reference grows its width in
// calcNeighborList before any force reads the rows (DPM2D.cu:3479-3507).
enum : int { SNAP_FIRE = 1, SNAP_RIGID = 2, SNAP_RIGID_FIRE = 4 };

// the (device pointer, 4-byte words) pairs a loop of kind `what` changes
std::vector<std::pair<unsigned*, long>> snap_fields(dpm_set_s* S, int what) {
  const long nv = S->nv, np = S->np, ns = S->nsys;
  auto w = [](void* p, long bytes) { return std::make_pair((unsigned*)p, bytes / 4); };
  std::vector<std::pair<unsigned*, long>> f = {
      w(S->V.x, 8 * nv),  w(S->V.y, 8 * nv),  w(S->V.vx, 8 * nv), w(S->V.vy, 8 * nv),
      w(S->V.fx, 8 * nv), w(S->V.fy, 8 * nv), w(S->V.lx, 8 * nv), w(S->V.ly, 8 * nv),
      w(S->V.l0, 8 * nv), w(S->V.ang, 8 * np), w(S->V.cx, 8 * np), w(S->V.cy, 8 * np),
      w(S->d_ctr, 4 * sizeof(unsigned long long)), w(S->d_updateCount, 4 * ns),
      w(S->d_flag[0], 4 * ns), w(S->d_flag[1], 4 * ns), w(S->d_any, 4 * 3)};
  if (what & SNAP_FIRE) f.push_back(w(S->d_fire[S->fireCur], sizeof(FireState) * ns));
  if (what & (SNAP_RIGID | SNAP_RIGID_FIRE)) {
    const Rigid& R = S->R;
    for (double* p : {R.pvx, R.pvy, R.pw, R.pfx, R.pfy, R.ptq, R.moi, R.plx, R.ply, R.pla}) f.push_back(w(p, 8 * np));
    f.push_back(w(R.vtq, 8 * nv));
  }
  if (what & SNAP_RIGID_FIRE) f.push_back(w(S->d_rfire[S->rfireCur], sizeof(FireState) * ns));
  return f;
}

int snap_copy(dpm_set_s* S, int what, bool save) {
  const auto f = snap_fields(S, what);
  long words = 0;
  for (auto& q : f) words += q.second;
  if (save && words > S->snapWords) {
    HIPCHK(hipStreamSynchronize(S->stream));
    hipFree(S->d_snap);
    S->d_snap = nullptr;
    HIPCHK(dalloc(&S->d_snap, words));
    S->snapWords = words;
  }
  CopySegs c{};
  long off = 0;
  for (auto& q : f) {
    if (c.n == kMaxCopySegs) return fail(S, DPM_ERR_STATE, "loop snapshot: too many fields");
    c.src[c.n] = save ? q.first : S->d_snap + off;
    c.dst[c.n] = save ? S->d_snap + off : q.first;
    c.words[c.n] = q.second;
    c.n++;
    off += q.second;
  }
  HIPCHK(launch_copy_segs(c, S->stream));
  return DPM_OK;
}

int snap_save(dpm_set_s* S, int what) {
  S->snapCur = S->cur;
  S->snapFireCur = S->fireCur;
  S->snapRfireCur = S->rfireCur;
  return snap_copy(S, what, true);
}

int snap_restore(dpm_set_s* S, int what) {
  S->cur = S->snapCur;
  S->fireCur = S->snapFireCur;
  S->rfireCur = S->snapRfireCur;
  set_pos_view(S);
  return snap_copy(S, what, false);
}

// Run `loop` (which leaves its errors in d_err) from a copy of the state it changes, and again on
// a wider list for as long as it reports ERR_LIST_OVERFLOW; returns the loop's status (one host
// synchronisation per run).
template <class F>
int with_replay(dpm_set_s* S, int what, F loop) {
  int rc = snap_save(S, what);
  if (rc) return rc;
  for (int attempt = 0; attempt < 6; attempt++) {
    HIPCHK(launch_fill_int(S->d_err, 0, 1, S->stream));
    rc = loop();
    if (rc) return rc;
    int err = 0;
    HIPCHK(hipMemcpyAsync(&err, S->d_err, sizeof(int), hipMemcpyDeviceToHost, S->stream));
    HIPCHK(hipStreamSynchronize(S->stream));
    if (!(err & ERR_LIST_OVERFLOW)) return loop_status(S, err);
    rc = snap_restore(S, what);
    if (rc) return rc;
    rc = alloc_list(S, 2 * S->V.W);
    if (rc) return rc;
    std::vector<int> all(S->nsys, 1);
    rc = rebuild(S, all, 0, true);
    if (rc) return rc;
    rc = snap_restore(S, what);  // the rebuild cleared the rebuild flags the loop started with
    if (rc) return rc;
  }
  return fail(S, DPM_ERR_OVERFLOW, "neighbour list still overflowing after widening it six times");
}

// One loop step's launches at position parity cur: bin (normally idle) + step.  The FIRE step
// also needs the per-system FIRE state and the tile partial buffer.
int loop_step(dpm_set_s* S, int mode, int bumpT, double* recx = nullptr, double* recy = nullptr) {
  HIPCHK(launch_bin(S, bumpT, S->P.particle != PART_PASSIVE ? 1 : 0));
  if (mode == MODE_FIRE) {
    HIPCHK(launch_fire_step(S));
    return DPM_OK;
  }
  StepArgs A = loop_args(S);
  A.recx = recx;
  A.recy = recy;
  HIPCHK(launch_step(S, mode, A));
  return DPM_OK;
}

// Arguments of the step-loop kernel of rotation index r: it reads the rebuild flags F[r] of its
// positions and the generations G[r] of its systems' list-position cells, writes the flags
// F[r + 1] of its drift destination and the generations G[r + 1], and clears F[r + 2] for the
// kernel after it.  A flagged system rebuilds its rows from the cells of its list positions
// (widened by the 2 cutoff / 3 a vertex may have moved from them), bins its new list positions
// into the next cell list and clears the one after; no other step touches the cells, so the
// loop needs no binning, in a launch of its own or in the step kernel.  Its draws use the call
// indices ctr[k] + offset.
StepArgs md_args(dpm_set_s* S, long r, long long offT, long long offA) {
  StepArgs A = loop_args(S);
  const int a = (int)(r % 3), b = (int)((r + 1) % 3), c = (int)((r + 2) % 3);
  A.flagCur = S->mdFlag[a];
  A.cellCur = S->mdCount[0];
  A.slotsCur = S->mdSlots[0];
  A.cellOther = nullptr;
  A.genCur = S->mdGen[a];
  A.genOut = S->mdGen[b];
  A.flagOut = S->mdFlag[b];
  A.flagClear = S->mdFlag[c];
  A.staleCheck = 1;
  A.anyOut = S->d_any + 2;  // not read by the loop
  A.ctrOffT = offT;
  A.ctrOffA = offA;
  return A;
}

// A step loop's cell protocol at its start: cell lists 0 and 1 cleared (the start kernel bins
// the list positions into list 0, the first rebuild bins into list 1) and the flags of the first
// step's drift destination F[1] cleared.
int md_start(dpm_set_s* S) {
  HIPCHK(launch_fill_int(S->mdCount[0], 0, 2L * (S->G.ncells > 0 ? S->G.ncells : 1), S->stream));
  HIPCHK(launch_fill_int(S->mdFlag[1], 0, S->nsys, S->stream));
  return DPM_OK;
}

long long active_draws(const dpm_set_s* S, long steps) { return S->P.particle != PART_PASSIVE ? steps : 0; }

// Launch sequence of a velocity-Verlet loop (NVE / Langevin):
//   start: list positions binned; v += dt/2 F, x' = x + dt v, displacement test of x'
//          (buffer b -> 1-b)
//   per step: step [rebuild rows if flagged; F(x'); thermostat; v += dt/2 F;
//             unless last: v += dt/2 F, x'' = x' + dt v, displacement test of x'']
//   end: RNG call counters advanced by the loop's draws.
// The fused form performs the reference's operations in the reference's order.
// With `rec` (slots of 2 nv doubles, x then y), the position after step i is also written to
// slot i + 1 (the adjoint's recomputed segment).
int verlet_sequence(dpm_set_s* S, long steps, bool langevin, double* rec = nullptr) {
  const double h = 0.5 * S->P.dt;
  const long nv = S->nv;
  int rc = md_start(S);
  if (rc) return rc;
  if (rec) {
    // a recomputed segment starts from a restored checkpoint without a valid list: its list
    // positions become the start positions (whose cells the start kernel bins), and every
    // system rebuilds its rows in the first step (below)
    HIPCHK(launch_copy_d(S->V.lx, S->V.x, nv, S->stream));
    HIPCHK(launch_copy_d(S->V.ly, S->V.y, nv, S->stream));
  }
  StepArgs A0 = md_args(S, 0, 0, 0);
  A0.recx = rec ? rec + 2 * nv : nullptr;
  A0.recy = rec ? rec + 3 * nv : nullptr;
  HIPCHK(launch_start_md(S, h, S->P.dt, A0));
  if (rec) HIPCHK(launch_fill_int(S->mdFlag[1], 1, S->nsys, S->stream));
  swap_pos(S);
  for (long i = 1; i <= steps; i++) {
    const bool more = i < steps;
    const int mode = MODE_KICK | (langevin ? MODE_LANGEVIN : 0) | (more ? MODE_START : 0);
    StepArgs A = md_args(S, i, i - 1, i - 1);
    if (rec && more) {
      A.recx = rec + (i + 1) * 2 * nv;
      A.recy = A.recx + nv;
    }
    HIPCHK(launch_step(S, mode, A));
    if (more) swap_pos(S);
  }
  HIPCHK(launch_ctr_add(S, langevin ? steps : 0, active_draws(S, steps)));
  return DPM_OK;
}

// BAOAB loop (BAOAB::integrate, Simulator.cu:112-120): the first call's B A O A from the stored
// force, then per step [F(x'); v += dt/2 F; unless last: B A O A].  The O step of call k of the
// loop draws at counter ctr[0] + k.
int baoab_sequence(dpm_set_s* S, long steps) {
  const double h = 0.5 * S->P.dt;
  int rc = md_start(S);
  if (rc) return rc;
  HIPCHK(launch_start_baoab(S, h, md_args(S, 0, 0, 0)));
  swap_pos(S);
  for (long i = 1; i <= steps; i++) {
    const bool more = i < steps;
    HIPCHK(launch_step(S, MODE_KICK | MODE_BAOAB | (more ? MODE_START : 0), md_args(S, i, i, i - 1)));
    if (more) swap_pos(S);
  }
  HIPCHK(launch_ctr_add(S, steps, active_draws(S, steps)));
  return DPM_OK;
}

// Overdamped loop (Brownian::integrate 147-152, DrivenBrownian::integrate 182-187): the cells of
// the list positions and the list validity of the current positions first, then per step
// [rebuild if flagged; F(x); v = (F + noise xi) / gamma or F / gamma; x' = x + dt v; displacement
// test of x'].
int brown_sequence(dpm_set_s* S, long steps, bool noisy) {
  // step i has rotation index i - 1: the prologue writes the flags F[0] and generations G[0]
  int rc = md_start(S);
  if (rc) return rc;
  HIPCHK(launch_fill_int(S->mdFlag[0], 0, S->nsys, S->stream));
  HIPCHK(launch_bin_flag(S, S->mdCount[0], S->mdSlots[0], S->mdFlag[0], S->mdGen[0]));
  for (long i = 1; i <= steps; i++) {
    const int mode = MODE_BROWN | MODE_START | (noisy ? MODE_LANGEVIN : 0);
    StepArgs A = md_args(S, i - 1, i - 1, i - 1);
    HIPCHK(launch_step(S, mode, A));
    swap_pos(S);
  }
  HIPCHK(launch_ctr_add(S, noisy ? steps : 0, active_draws(S, steps)));
  return DPM_OK;
}

void drop_graphs(dpm_set_s* S) {
  for (auto& kv : S->graphs) hipGraphExecDestroy(kv.second);
  S->graphs.clear();
  S->graphVersion = S->version;
}

// Capture `seq` into a graph cached under `key`, or replay the cached one.  `seq` must leave
// the host-side parity state exactly as a direct run would; the caller advances it.
template <class F>
int run_graph(dpm_set_s* S, std::tuple<int, long, int> key, F seq) {
  if (S->graphVersion != S->version) drop_graphs(S);
  auto it = S->graphs.find(key);
  if (it == S->graphs.end()) {
    const int cur0 = S->cur, f0 = S->fireCur, rf0 = S->rfireCur;
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(S->stream, hipStreamCaptureModeThreadLocal));
    int r2 = seq();
    hipError_t e = hipStreamEndCapture(S->stream, &g);
    S->cur = cur0;
    S->fireCur = f0;
    S->rfireCur = rf0;
    set_pos_view(S);
    if (r2) return r2;
    HIPCHK(e);
    hipGraphExec_t ge;
    HIPCHK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    hipGraphDestroy(g);
    S->graphs[key] = ge;
    it = S->graphs.find(key);
  }
  HIPCHK(hipGraphLaunch(it->second, S->stream));
  return DPM_OK;
}

bool graphs_enabled() { return getenv("DPM_NO_GRAPH") == nullptr; }

// Run `steps` Verlet steps (a captured graph per (kind, steps, parity) when steps >= 4).
int run_verlet_once(dpm_set_s* S, long steps, bool langevin) {
  int rc;
  const int noise = langevin ? S->noiseType : -1;
  auto seq = [&]() -> int {
    switch (noise) {
      case NOISE_BAOAB: return baoab_sequence(S, steps);
      case NOISE_BROWNIAN: return brown_sequence(S, steps, true);
      case NOISE_DRIVEN_BROWNIAN: return brown_sequence(S, steps, false);
      default: return verlet_sequence(S, steps, langevin);
    }
  };
  // every sequence swaps the position buffers once per step
  if (steps < 4 || !graphs_enabled()) return seq();
  const int kind = noise < 0 ? 2 : (noise == NOISE_LANGEVIN ? 1 : 10 + noise);
  rc = run_graph(S, std::make_tuple(kind, steps, S->cur), seq);
  if (rc) return rc;
  if ((steps % 2) == 1) swap_pos(S);
  return DPM_OK;
}

int run_verlet(dpm_set_s* S, long steps, bool langevin) {
  if (steps <= 0) return DPM_OK;
  int rc = require_list(S);
  if (rc) return rc;
  return with_replay(S, 0, [&]() { return run_verlet_once(S, steps, langevin); });
}

int fire_iteration(dpm_set_s* S) {
  HIPCHK(launch_fire_begin(S));
  S->fireCur = 1 - S->fireCur;  // fire_mid and the step read the state fire_begin wrote
  HIPCHK(launch_fire_mid(S));
  swap_pos(S);
  return loop_step(S, MODE_FIRE, 0);
}

int fire_sequence(dpm_set_s* S, long iterations) {
  HIPCHK(launch_fill_int(S->d_flag[1 - S->cur], 0, S->nsys, S->stream));
  HIPCHK(launch_fill_int(S->d_any + 1 - S->cur, 0, 1, S->stream));
  HIPCHK(launch_fire_partials(S));
  for (long i = 0; i < iterations; i++) {
    int rc = fire_iteration(S);
    if (rc) return rc;
  }
  return DPM_OK;
}

// FIRE iterations; the sequence flips the FIRE state and position buffers once per iteration.
int run_fire(dpm_set_s* S, long iterations) {
  if (iterations <= 0) return DPM_OK;
  if (iterations < 4 || !graphs_enabled()) return fire_sequence(S, iterations);
  int rc = run_graph(S, std::make_tuple(8 + S->fireCur, iterations, S->cur),
                     [&]() { return fire_sequence(S, iterations); });
  if (rc) return rc;
  if ((iterations % 2) == 1) {
    swap_pos(S);
    S->fireCur = 1 - S->fireCur;
  }
  return DPM_OK;
}

int temperature_scale(dpm_set_s* S, double T) {
  int rc = reduce_host(S);
  if (rc) return rc;
  std::vector<double> sc(S->nsys);
  for (long s = 0; s < S->nsys; s++) {
    const double n = S->h_svoff[s + 1] - S->h_svoff[s];
    const double temp = 2. * (0.5 * S->h_red[s * 8 + 2]) / (n * 2);
    sc[s] = sqrt(T / temp);
  }
  HIPCHK(hipMemcpyAsync(S->d_scale, sc.data(), sizeof(double) * S->nsys, hipMemcpyHostToDevice, S->stream));
  HIPCHK(launch_scale_vel(S));
  HIPCHK(hipStreamSynchronize(S->stream));
  return DPM_OK;
}

// ------------------------------------------------------------------ adjoint ---
// Flags, any-flags and both parities' cell counts cleared: the state the forward loops and the
// adjoint's own rebuild protocol both start from.
int clean_protocol(dpm_set_s* S) {
  HIPCHK(launch_fill_int(S->d_flag[0], 0, S->nsys, S->stream));
  HIPCHK(launch_fill_int(S->d_flag[1], 0, S->nsys, S->stream));
  HIPCHK(launch_fill_int(S->d_any, 0, 2, S->stream));
  HIPCHK(launch_fill_int(S->G.count[0], 0, S->G.ncells, S->stream));
  HIPCHK(launch_fill_int(S->G.count[1], 0, S->G.ncells, S->stream));
  return DPM_OK;
}

// Self-propelled rings (`active`: the Hessian-vector product, the energy's parameter gradient and
// the NVE reverse pass) add a force that does not depend on positions or parameters; their angles
// advance by per-step draws, which the NVE tape's checkpoints carry (save_state).  FIRE's reverse
// pass takes passive rings only.
int adjoint_supported(dpm_set_s* S, bool active = false) {
  if (S->nwide > 0) return fail(S, DPM_ERR_UNSUPPORTED, "adjoint: rings of at most 64 vertices only");
  if (S->P.interaction != INT_VV && S->P.interaction != INT_SMOOTH)
    return fail(S, DPM_ERR_UNSUPPORTED, "adjoint: vertexVertex or vertexSmooth interaction only");
  if (S->P.monomer != MONO_HARMONIC) return fail(S, DPM_ERR_UNSUPPORTED, "adjoint: harmonic monomers only");
  if (S->P.particle != PART_PASSIVE && !(active && S->P.particle == PART_ACTIVE))
    return fail(S, DPM_ERR_UNSUPPORTED, active ? "adjoint: passive or self-propelled (active) particles only"
                                               : "adjoint: passive particles only");
  if (S->P.geometry != GEOM_NORMAL) return fail(S, DPM_ERR_UNSUPPORTED, "adjoint: periodic geometry only");
  return require_list(S);
}

// work buffers: w0, w1, ax (2nv each), partials g (4nv), ring sums (3np)
int adj_alloc(dpm_set_s* S) {
  if (S->d_adj && S->adjNv == S->nv) return DPM_OK;
  HIPCHK(hipStreamSynchronize(S->stream));
  hipFree(S->d_adj);
  S->d_adj = nullptr;
  HIPCHK(dalloc(&S->d_adj, 10 * S->nv + 3 * S->np));
  S->adjNv = S->nv;
  drop_graphs(S);  // captured sequences hold the old addresses
  return DPM_OK;
}
double* adj_w(dpm_set_s* S, int i) { return S->d_adj + (long)i * 2 * S->nv; }
double* adj_ax(dpm_set_s* S) { return S->d_adj + 4 * S->nv; }
double* adj_g(dpm_set_s* S) { return S->d_adj + 6 * S->nv; }
double* adj_ring(dpm_set_s* S) { return S->d_adj + 10 * S->nv; }

View view_at(const dpm_set_s* S, double* x, double* y) {
  View V = S->V;
  V.x = x;
  V.y = y;
  return V;
}

// one H w evaluation at the positions of Vn with the rebuild protocol of parity p
int hvp_launches(dpm_set_s* S, const View& Vn, int p, AdjArgs A) {
  HIPCHK(launch_adj_flag(S, Vn, p));
  HIPCHK(launch_bin_view(S, Vn, p));
  A.flag = S->d_flag[p];
  A.count = S->G.count[p];
  HIPCHK(launch_adj_hvp(S, Vn, A));
  return DPM_OK;
}

// state copies: x, y (current buffer), vx, vy, fx, fy <-> 6 nv doubles at `dst` / `src`; with
// `rings`, then the ring angles (np doubles) and the RNG call counters (4 words of 8 bytes, copied
// bit for bit), which a recomputed segment of self-propelled rings replays its draws from
long state_words(const dpm_set_s* S, bool rings) { return 6 * S->nv + (rings ? S->np + 4 : 0); }
int save_state(dpm_set_s* S, double* dst, bool rings = false) {
  const long nv = S->nv;
  const double* src[6] = {S->V.x, S->V.y, S->V.vx, S->V.vy, S->V.fx, S->V.fy};
  for (int q = 0; q < 6; q++) HIPCHK(launch_copy_d(dst + q * nv, src[q], nv, S->stream));
  if (rings) {
    HIPCHK(launch_copy_d(dst + 6 * nv, S->V.ang, S->np, S->stream));
    HIPCHK(launch_copy_d(dst + 6 * nv + S->np, (const double*)S->d_ctr, 4, S->stream));
  }
  return DPM_OK;
}
int load_state(dpm_set_s* S, const double* src, bool rings = false) {
  const long nv = S->nv;
  double* dst[6] = {S->V.x, S->V.y, S->V.vx, S->V.vy, S->V.fx, S->V.fy};
  for (int q = 0; q < 6; q++) HIPCHK(launch_copy_d(dst[q], src + q * nv, nv, S->stream));
  if (rings) {
    HIPCHK(launch_copy_d(S->V.ang, src + 6 * nv, S->np, S->stream));
    HIPCHK(launch_copy_d((double*)S->d_ctr, src + 6 * nv + S->np, 4, S->stream));
  }
  return DPM_OK;
}

// positions x_a .. x_{a+len} of a recomputed segment into d_seg slots 0..len
int record_segment(dpm_set_s* S, long len) {
  const long nv = S->nv;
  HIPCHK(launch_copy_d(S->d_seg, S->V.x, nv, S->stream));
  HIPCHK(launch_copy_d(S->d_seg + nv, S->V.y, nv, S->stream));
  int rc;
  if (len < 4 || !graphs_enabled()) {
    rc = verlet_sequence(S, len, false, S->d_seg);
    return rc;
  }
  rc = run_graph(S, std::make_tuple(4, len, S->cur), [&]() { return verlet_sequence(S, len, false, S->d_seg); });
  if (rc) return rc;
  if ((len % 2) == 1) swap_pos(S);
  return DPM_OK;
}

// reverse pass over positions n = top .. a of the recorded segment starting at step a
int backward_segment(dpm_set_s* S, long a, long top, long N) {
  const long nv = S->nv;
  const long count = top - a + 1;
  const bool hasTop = (top == N), hasZero = (a == 0);
  // One list for the whole segment when it can serve every recorded position: rows built at
  // the middle slot, checked against all slots.  Otherwise each position runs the rebuild
  // protocol (displacement flags, binning, in-kernel row rebuild).
  const long mid = (count - 1) / 2;
  HIPCHK(launch_copy_d(S->V.x, S->d_seg + mid * 2 * nv, nv, S->stream));
  HIPCHK(launch_copy_d(S->V.y, S->d_seg + mid * 2 * nv + nv, nv, S->stream));
  std::vector<int> all(S->nsys, 1);
  int rc = rebuild(S, all, 0);
  if (rc) return rc;
  int bad = 0;
  HIPCHK(launch_fill_int(S->d_notDone, 0, 1, S->stream));
  HIPCHK(launch_adj_seg_check(S, S->d_seg, (int)count, S->d_notDone));
  HIPCHK(hipMemcpyAsync(&bad, S->d_notDone, sizeof(int), hipMemcpyDeviceToHost, S->stream));
  HIPCHK(hipStreamSynchronize(S->stream));
  const bool single = !bad;
  auto seq = [&]() -> int {
    int p = 0;
    if (!single) {  // list validity at the first position processed, x_top
      double* st = S->d_seg + (top - a) * 2 * nv;
      HIPCHK(launch_adj_flag(S, view_at(S, st, st + nv), 0));
    }
    for (long n = top; n >= a; n--) {
      double* sx = S->d_seg + (n - a) * 2 * nv;
      const View Vn = view_at(S, sx, sx + nv);
      AdjArgs A{};
      A.wx = adj_w(S, S->wCur);
      A.wy = A.wx + nv;
      A.ax = adj_ax(S);
      A.ay = A.ax + nv;
      A.nwx = adj_w(S, 1 - S->wCur);
      A.nwy = A.nwx + nv;
      A.g = adj_g(S);
      A.cd = ((n == N || n == 0) ? 0.5 : 1.0) * S->tapeDt;
      A.dt = S->tapeDt;
      A.mode = 1;
      A.update = n != 0;
      if (!single) {
        // rebuild protocol of the step loop: bin the flagged systems, rebuild their rows in the
        // kernel, which also tests the next position and clears the other parity's cells
        HIPCHK(launch_bin_view(S, Vn, p));
        A.flag = S->d_flag[p];
        A.count = S->G.count[p];
        A.cellOther = S->G.count[1 - p];
        A.ncells = S->G.ncells;
        if (n > a) {
          A.px = sx - 2 * nv;
          A.py = sx - nv;
          A.flagOut = S->d_flag[1 - p];
          A.anyOut = S->d_any + 1 - p;
        }
      }
      HIPCHK(launch_adj_hvp(S, Vn, A));
      p ^= 1;
      S->wCur ^= 1;
    }
    return single ? DPM_OK : clean_protocol(S);
  };
  const int w0 = S->wCur;
  if (!graphs_enabled()) return seq();
  const int kind = 256 | (hasTop ? 1 : 0) | (hasZero ? 2 : 0) | (single ? 4 : 0);
  rc = run_graph(S, std::make_tuple(kind, count, w0), [&]() {
    int r = seq();
    S->wCur = w0;
    return r;
  });
  if (rc) return rc;
  if (count % 2) S->wCur ^= 1;
  return DPM_OK;
}

// scratch of at least n doubles for the observables
int obs_scratch(dpm_set_s* S, long n) {
  if (S->obsLen >= n) return DPM_OK;
  HIPCHK(hipStreamSynchronize(S->stream));
  hipFree(S->d_obs);
  S->d_obs = nullptr;
  HIPCHK(dalloc(&S->d_obs, n));
  S->obsLen = n;
  return DPM_OK;
}

// setInitialPositions + setInitialParticlePositions, DPM2D.cu:921-930
int store_initial_positions(dpm_set_s* S) {
  if (!S->d_x0) HIPCHK(dalloc(&S->d_x0, 2 * S->nv + 2 * S->np));
  HIPCHK(launch_shape_only(S));
  HIPCHK(launch_copy_d(S->d_x0, S->V.x, S->nv, S->stream));
  HIPCHK(launch_copy_d(S->d_x0 + S->nv, S->V.y, S->nv, S->stream));
  HIPCHK(launch_copy_d(S->d_x0 + 2 * S->nv, S->V.cx, S->np, S->stream));
  HIPCHK(launch_copy_d(S->d_x0 + 2 * S->nv + S->np, S->V.cy, S->np, S->stream));
  return DPM_OK;
}

// per-system sums of the displacement kernel's two outputs: (|dr|^2, sin(k|dr|)/(k|dr|))
int displacement_sums(dpm_set_s* S, bool rings, double waveNum, std::vector<double>& d2, std::vector<double>& sf) {
  if (!S->d_x0) return fail(S, DPM_ERR_STATE, "initial positions not set (init a simulation or call "
                                               "dpm_set_initial_positions)");
  const long n = rings ? S->np : S->nv;
  int rc = obs_scratch(S, 2 * n);
  if (rc) return rc;
  if (rings) HIPCHK(launch_shape_only(S));
  const double* x = rings ? S->V.cx : S->V.x;
  const double* y = rings ? S->V.cy : S->V.y;
  const double* x0 = S->d_x0 + (rings ? 2 * S->nv : 0);
  const double* y0 = x0 + (rings ? S->np : S->nv);
  HIPCHK(launch_displacement(S, x, y, x0, y0, n, rings ? 1 : 0, waveNum, S->d_obs, S->d_obs + n));
  std::vector<double> h(2 * n);
  HIPCHK(hipMemcpyAsync(h.data(), S->d_obs, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, S->stream));
  HIPCHK(hipStreamSynchronize(S->stream));
  d2.assign(S->nsys, 0.0);
  sf.assign(S->nsys, 0.0);
  for (long s = 0; s < S->nsys; s++) {
    const long a = rings ? S->h_sroff[s] : S->h_svoff[s], b = rings ? S->h_sroff[s + 1] : S->h_svoff[s + 1];
    for (long i = a; i < b; i++) {
      d2[s] += h[i];
      sf[s] += h[n + i];
    }
  }
  return DPM_OK;
}

// ------------------------------------------------------------- rigid rings ---
int rigid_supported(dpm_set_s* S) {
  if (S->nwide > 0) return fail(S, DPM_ERR_UNSUPPORTED, "rigid rings: rings of at most 64 vertices only");
  if (S->P.interaction != INT_VV && S->P.interaction != INT_SMOOTH)
    return fail(S, DPM_ERR_UNSUPPORTED, "rigid rings: vertexVertex or vertexSmooth interaction only");
  if (S->P.geometry != GEOM_NORMAL) return fail(S, DPM_ERR_UNSUPPORTED, "rigid rings: periodic geometry only");
  if (S->P.particle != PART_PASSIVE) return fail(S, DPM_ERR_UNSUPPORTED, "rigid rings: passive particles only");
  return require_list(S);
}

// rotational state cleared (initRotationalVariables / initDeltaVariables, DPM2D.cu:197-221)
int rigid_clear(dpm_set_s* S) {
  double* rp[] = {S->V.ang, S->R.pw, S->R.ptq, S->R.moi, S->R.pla};
  for (double* p : rp) HIPCHK(launch_fill_d(p, 0.0, S->np, S->stream));
  HIPCHK(launch_fill_d(S->R.vtq, 0.0, S->nv, S->stream));
  return DPM_OK;
}

// Launch sequence of a rigid loop (RigidNVE / RigidLangevin):
//   start: half kick of the rings, centres and angles advanced, vertices moved (buffer b -> 1-b)
//   per step: step [rebuild flagged rows; interaction F(x'); ring force, torque, moment of
//             inertia; half kick; unless last: the next step's half kick and move]
//   end: thermal call counter advanced by the loop's draws.
// Both run the velocity-Verlet loop's protocol (a rebuilding step builds its rows from the cells
// of the list positions and bins the new ones).  Langevin's kick adds unit Gaussian draws to the
// ring velocities whatever the time step, so a ring can move far in one step: a drift longer than
// cutoff / 3 flags its system FLAG_FAR, whose next rebuild searches all the system's vertices.
int rigid_sequence(dpm_set_s* S, long steps, bool langevin) {
  int rc = md_start(S);
  if (rc) return rc;
  HIPCHK(launch_rigid_start(S, langevin, md_args(S, 0, 0, 0)));
  swap_pos(S);
  for (long i = 1; i <= steps; i++) {
    const bool more = i < steps;
    StepArgs A = md_args(S, i, i - 1, 0);
    A.rigidLangevin = langevin ? 1 : 0;
    HIPCHK(launch_step(S, MODE_RIGID | MODE_KICK | (more ? MODE_START : 0), A));
    if (more) swap_pos(S);
  }
  if (langevin) HIPCHK(launch_ctr_add(S, steps, 0));
  return DPM_OK;
}

int run_rigid(dpm_set_s* S, long steps, bool langevin) {
  if (steps <= 0) return DPM_OK;
  int rc = rigid_supported(S);
  if (rc) return rc;
  auto seq = [&]() -> int { return rigid_sequence(S, steps, langevin); };
  return with_replay(S, SNAP_RIGID, [&]() -> int {
    if (steps < 4 || !graphs_enabled()) return seq();
    int r = run_graph(S, std::make_tuple(langevin ? 20 : 21, steps, S->cur), seq);
    if (!r && (steps % 2) == 1) swap_pos(S);
    return r;
  });
}

// rigid FIRE iterations: the tile partials of the current state, then per iteration the ring
// update and vertex move (into the other position buffer, with its flags; the FIRE state flips
// buffers), the bin launch and the step kernel (rebuild of flagged rows, forces, ring sums, the
// second half kick, the next iteration's partials)
int rigid_fire_sequence(dpm_set_s* S, long iterations) {
  // the step loop's cell protocol: the list positions binned into cell list 0 and the current
  // positions' list validity in F[0]; iteration i's move ORs its moved positions' validity into
  // F[i] (a drift longer than cutoff / 3 flags FLAG_FAR), and its step (rotation index i) reads
  // F[i], rebuilds flagged systems from the list-position cells (all pairs on FLAG_FAR) and clears
  // F[i + 2] for the move after next
  int rc = md_start(S);
  if (rc) return rc;
  HIPCHK(launch_fill_int(S->mdFlag[0], 0, S->nsys, S->stream));
  HIPCHK(launch_bin_flag(S, S->mdCount[0], S->mdSlots[0], S->mdFlag[0], S->mdGen[0]));
  HIPCHK(launch_rigid_fire_partials(S));
  for (long it = 0; it < iterations; it++) {
    HIPCHK(launch_rigid_fire_move(S, S->mdFlag[it % 3], S->d_any + 2));
    S->rfireCur = 1 - S->rfireCur;
    swap_pos(S);
    StepArgs A = md_args(S, it, 0, 0);
    A.fire = S->d_rfire[S->rfireCur];
    A.part = S->d_part;
    HIPCHK(launch_step(S, MODE_RIGID | MODE_FIRE, A));
  }
  return DPM_OK;
}

}  // namespace

// ======================================================================= ABI
extern "C" {

int dpm_create(dpm_set_t* out, int device, long numSystems, const long* ringsPerSystem,
               const long* numVertexInParticleList) {
  if (!out || numSystems <= 0 || !ringsPerSystem || !numVertexInParticleList) return DPM_ERR_ARG;
  *out = nullptr;
  if (hipSetDevice(device) != hipSuccess) return DPM_ERR_HIP;
  // the allocation-free entry point's buffer on this device, looked up once per device (a
  // failure is reported by that entry point, which looks it up again)
  double* scratchAddr = nullptr;
  (void)scratch_address(device, &scratchAddr);
  dpm_set_s* S = new dpm_set_s();
  S->device = device;
  S->nsys = numSystems;
  long np = 0, nv = 0;
  S->h_svoff.push_back(0);
  S->h_sroff.push_back(0);
  for (long s = 0; s < numSystems; s++) {
    if (ringsPerSystem[s] <= 0) { delete S; return DPM_ERR_ARG; }
    for (long k = 0; k < ringsPerSystem[s]; k++) {
      const long n = numVertexInParticleList[np];
      if (n < 3 || n > (1L << 24)) { delete S; return DPM_ERR_UNSUPPORTED; }
      if (n > 64) S->h_wide.push_back((int)np);
      S->h_rnv.push_back((int)n);
      S->h_rfirst.push_back((int)nv);
      S->h_rsys.push_back((int)s);
      for (long q = 0; q < n; q++) {
        S->h_vring.push_back((int)np);
        S->h_vsys.push_back((int)s);
      }
      nv += n;
      np++;
    }
    S->h_svoff.push_back((int)nv);
    S->h_sroff.push_back((int)np);
  }
  if (nv > (1L << 30)) { delete S; return DPM_ERR_UNSUPPORTED; }
  S->np = np;
  S->nv = nv;
  // ring-aligned tiles of at most 64 vertices, greedy, never crossing a system
  const int cap = 64;  // smaller caps (32-48 vertices) measured slower at 1024 rings
  S->h_stile.push_back(0);
  for (long s = 0; s < numSystems; s++) {
    long r = S->h_sroff[s];
    while (r < S->h_sroff[s + 1]) {
      const int v0 = S->h_rfirst[r];
      int v1 = v0;
      while (r < S->h_sroff[s + 1] && (v1 + S->h_rnv[r] - v0) <= cap) {
        v1 += S->h_rnv[r];
        r++;
      }
      if (v1 == v0) {  // a wide ring: tiles of 64 consecutive vertices, the last one shorter
        const int end = v0 + S->h_rnv[r];
        for (int a = v0; a < end; a += cap) {
          S->h_tile_v0.push_back(a);
          S->h_tile_v1.push_back(std::min(a + cap, end));
          S->h_tile_sys.push_back((int)s);
        }
        r++;
        continue;
      }
      S->h_tile_v0.push_back(v0);
      S->h_tile_v1.push_back(v1);
      S->h_tile_sys.push_back((int)s);
    }
    S->h_stile.push_back((int)S->h_tile_v0.size());
  }
  View& V = S->V;
  V.nv = (int)nv;
  V.np = (int)np;
  V.nsys = (int)numSystems;
  V.ntiles = (int)S->h_tile_v0.size();
  S->h_box.assign(2 * numSystems, 1.0);
  S->h_maxrad.assign(numSystems, 0.0);
  if (hipStreamCreateWithFlags(&S->ownStream, hipStreamNonBlocking) != hipSuccess) {
    delete S;
    return DPM_ERR_HIP;
  }
  S->stream = S->ownStream;
#define A(p, n) if (dalloc(&(p), (n)) != hipSuccess) { dpm_destroy(S); return DPM_ERR_OOM; }
  for (int b = 0; b < 2; b++) for (int c = 0; c < 2; c++) A(S->pos[b][c], nv);
  A(V.vx, nv); A(V.vy, nv); A(V.fx, nv); A(V.fy, nv); A(V.rad, nv); A(V.l0, nv); A(V.th0, nv);
  A(V.th, nv); A(V.len, nv); A(V.ev, nv); A(V.es, nv); A(V.lx, nv); A(V.ly, nv);
  A(V.vring, nv); A(V.vsys, nv); A(V.vinfo, nv);
  A(V.rfirst, np); A(V.rnv, np); A(V.rsys, np);
  A(V.a0, np); A(V.area, np); A(V.perim, np); A(V.cx, np); A(V.cy, np); A(V.pe, np); A(V.ang, np);
  A(V.bx, numSystems); A(V.by, numSystems); A(V.ibx, numSystems); A(V.iby, numSystems);
  A(V.mbx, numSystems); A(V.mby, numSystems); A(V.wfx, nv); A(V.wfy, nv);
  A(S->R.pvx, np); A(S->R.pvy, np); A(S->R.pw, np); A(S->R.pfx, np); A(S->R.pfy, np); A(S->R.ptq, np);
  A(S->R.moi, np); A(S->R.prad, np); A(S->R.plx, np); A(S->R.ply, np); A(S->R.pla, np); A(S->R.vtq, nv);
  if (hipMalloc(&S->d_rg, sizeof(dpm::Rigid)) != hipSuccess) { dpm_destroy(S); return DPM_ERR_OOM; }
  V.rg = S->d_rg;
  A(S->d_rfire[0], numSystems); A(S->d_rfire[1], numSystems);
  A(V.svoff, numSystems + 1); A(V.sroff, numSystems + 1);
  A(V.ncnt, nv);
  A(V.tile_v0, V.ntiles); A(V.tile_v1, V.ntiles); A(V.tile_sys, V.ntiles);
  A(V.stile, numSystems + 1);
  A(S->G.ncx, numSystems); A(S->G.ncy, numSystems); A(S->G.coff, numSystems); A(S->G.brute, numSystems);
  A(S->d_flag[0], numSystems); A(S->d_flag[1], numSystems); A(S->d_any, 3); A(S->d_updateCount, numSystems);
  A(S->d_err, 1); A(S->d_maxcnt, 1); A(S->d_ctr, 4); A(S->d_notDone, 1); A(S->d_sysOrder, numSystems);
  A(S->d_fire[0], numSystems); A(S->d_fire[1], numSystems);
  A(S->d_part, (long)V.ntiles * 4);
  S->nwide = (long)S->h_wide.size();
  if (S->nwide > 0) A(S->d_wide, S->nwide);
  A(S->d_red, numSystems * 8); A(S->d_scale, numSystems);
#undef A
  set_pos_view(S);
  hipStream_t st = S->stream;
  std::vector<int4> vinfo(nv);
  for (long v = 0; v < nv; v++) {
    const int r = S->h_vring[v];
    vinfo[v] = make_int4(r, S->h_rfirst[r], S->h_rnv[r], S->h_vsys[v]);
  }
  bool ok = up(V.vinfo, vinfo, st) == hipSuccess &&
            up(V.vring, S->h_vring, st) == hipSuccess && up(V.vsys, S->h_vsys, st) == hipSuccess &&
            up(V.rfirst, S->h_rfirst, st) == hipSuccess && up(V.rnv, S->h_rnv, st) == hipSuccess &&
            up(V.rsys, S->h_rsys, st) == hipSuccess && up(V.svoff, S->h_svoff, st) == hipSuccess &&
            up(V.sroff, S->h_sroff, st) == hipSuccess && up(V.tile_v0, S->h_tile_v0, st) == hipSuccess &&
            up(V.tile_v1, S->h_tile_v1, st) == hipSuccess && up(V.tile_sys, S->h_tile_sys, st) == hipSuccess &&
            up(V.stile, S->h_stile, st) == hipSuccess &&
            (S->nwide == 0 || up(S->d_wide, S->h_wide, st) == hipSuccess);
  // the persistent FIRE kernel's workgroup b runs system order[b]: the systems with the most tiles
  // first, so that the longest iterations start in the first wave of workgroups
  std::vector<int> order(numSystems);
  for (long s = 0; s < numSystems; s++) order[s] = (int)s;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return S->h_stile[a + 1] - S->h_stile[a] > S->h_stile[b + 1] - S->h_stile[b];
  });
  ok = ok && up(S->d_sysOrder, order, st) == hipSuccess;
  std::vector<double> ones(numSystems, 1.0);
  ok = ok && up(V.bx, ones, st) == hipSuccess && up(V.by, ones, st) == hipSuccess &&
       up(V.ibx, ones, st) == hipSuccess && up(V.iby, ones, st) == hipSuccess &&
       up(V.mbx, ones, st) == hipSuccess && up(V.mby, ones, st) == hipSuccess;
  double* zd[] = {S->pos[0][0], S->pos[0][1], S->pos[1][0], S->pos[1][1], V.vx, V.vy, V.fx, V.fy, V.rad, V.l0,
                  V.th0, V.th, V.len, V.ev, V.es, V.lx, V.ly, V.wfx, V.wfy, S->R.vtq};
  for (double* p : zd) ok = ok && hipMemsetAsync(p, 0, sizeof(double) * nv, st) == hipSuccess;
  double* zp[] = {V.a0, V.area, V.perim, V.cx, V.cy, V.pe, V.ang, S->R.pvx, S->R.pvy, S->R.pw, S->R.pfx,
                  S->R.pfy, S->R.ptq, S->R.moi, S->R.prad, S->R.plx, S->R.ply, S->R.pla};
  for (double* p : zp) ok = ok && hipMemsetAsync(p, 0, sizeof(double) * np, st) == hipSuccess;
  ok = ok && hipMemcpyAsync(S->d_rg, &S->R, sizeof(dpm::Rigid), hipMemcpyHostToDevice, st) == hipSuccess;
  ok = ok && hipMemsetAsync(V.ncnt, 0, sizeof(int) * nv, st) == hipSuccess;
  ok = ok && hipMemsetAsync(S->d_flag[0], 0, sizeof(int) * numSystems, st) == hipSuccess;
  ok = ok && hipMemsetAsync(S->d_flag[1], 0, sizeof(int) * numSystems, st) == hipSuccess;
  ok = ok && hipMemsetAsync(S->d_any, 0, sizeof(int) * 2, st) == hipSuccess;
  ok = ok && hipMemsetAsync(S->d_updateCount, 0, sizeof(int) * numSystems, st) == hipSuccess;
  ok = ok && hipMemsetAsync(S->d_err, 0, sizeof(int), st) == hipSuccess;
  ok = ok && hipMemsetAsync(S->d_ctr, 0, sizeof(unsigned long long) * 4, st) == hipSuccess;
  ok = ok && hipMemsetAsync(S->d_part, 0, sizeof(double) * V.ntiles * 4, st) == hipSuccess;
  ok = ok && hipStreamSynchronize(st) == hipSuccess;
  if (!ok) {
    dpm_destroy(S);
    return DPM_ERR_HIP;
  }
  // defaults of the DPM2D constructor, DPM2D.cu:41-94
  S->P.potential = POT_HARMONIC;
  S->P.interaction = INT_VV;
  S->P.monomer = MONO_HARMONIC;
  S->P.particle = PART_PASSIVE;
  S->P.dt = 1e-3;
  S->P.ea = 1e5; S->P.el = 20; S->P.eb = 10; S->P.ec = 1;
  S->P.gamma = 1; S->P.noise = 0;
  S->P.cutDistance = 1;
  S->P.ew = 1e3;  // DPM2D.cu:70
  S->P.cutoff = 0;
  S->P.seed = 0;
  S->fp.tol = -1;
  if (const char* e = getenv("DPM_SPLIT")) S->split = std::max(1, std::min(8, atoi(e)));
  if (const char* e = getenv("DPM_XCD_MAP")) S->xcdMap = atoi(e) != 0;
  if (alloc_list(S, 8) != DPM_OK) { dpm_destroy(S); return DPM_ERR_OOM; }
  *out = S;
  return DPM_OK;
}

int dpm_destroy(dpm_set_t h) {
  if (!h) return DPM_ERR_ARG;
  dpm_set_s* S = h;
  hipSetDevice(S->device);
  if (S->stream) hipStreamSynchronize(S->stream);
  for (auto& kv : S->graphs) hipGraphExecDestroy(kv.second);
  View& V = S->V;
  void* ptrs[] = {S->pos[0][0], S->pos[0][1], S->pos[1][0], S->pos[1][1], V.vx, V.vy, V.fx, V.fy, V.rad,
                  V.l0, V.th0, V.th, V.len, V.ev, V.es, V.lx, V.ly, V.vring, V.vsys, V.rfirst, V.rnv, V.rsys,
                  V.a0, V.area, V.perim, V.cx, V.cy, V.pe, V.ang, V.bx, V.by, V.ibx, V.iby, V.svoff, V.sroff,
                  V.ncnt, V.nbr, V.nbrp, V.tile_v0, V.tile_v1, V.tile_sys, V.stile, S->G.ncx, S->G.ncy,
                  S->G.coff, S->G.brute, S->G.count[0], S->G.count[1], S->G.slots, S->G.cellSys, S->d_flag[0],
                  S->d_flag[1], S->d_any, S->d_updateCount, S->d_err, S->d_maxcnt, S->d_ctr, S->d_notDone,
                  S->d_sysOrder, S->d_fire[0], S->d_fire[1], S->d_part, S->d_red, S->d_scale, V.kr,
                  S->d_ck, S->d_fin, S->d_seg, S->d_adj, V.vinfo, V.l0r, V.mbx, V.mby, V.wfx, V.wfy,
                  S->d_contacts, S->d_ccount, S->d_x0, S->d_obs, S->mdCount[0], S->mdSlots[0], S->mdFlag[0],
                  S->mdFlag[1], S->mdFlag[2], S->mdFlag[3], S->mdGen[0], S->mdGen[1], S->mdGen[2], S->R.pvx, S->R.pvy, S->R.pw, S->R.pfx, S->R.pfy, S->R.ptq, S->R.moi,
                  S->R.prad, S->R.plx, S->R.ply, S->R.pla, S->R.vtq, S->d_rfire[0], S->d_rfire[1], S->d_rg,
                  S->d_snap, S->d_ftape, S->d_ftapeState, S->d_fadj, S->d_wide};
  for (void* p : ptrs)
    if (p) hipFree(p);
  if (S->ownStream) hipStreamDestroy(S->ownStream);
  delete S;
  return DPM_OK;
}

int dpm_set_stream(dpm_set_t h, void* stream) {
  CHECK_HANDLE(h);
  HIPCHK(hipStreamSynchronize(S->stream));
  S->stream = stream ? (hipStream_t)stream : S->ownStream;
  bump(S);
  return DPM_OK;
}

const char* dpm_last_error(dpm_set_t h) { return h ? h->err.c_str() : "null handle"; }

int dpm_get_sizes(dpm_set_t h, long* numSystems, long* numParticles, long* numVertices) {
  if (!h) return DPM_ERR_ARG;
  if (numSystems) *numSystems = h->nsys;
  if (numParticles) *numParticles = h->np;
  if (numVertices) *numVertices = h->nv;
  return DPM_OK;
}

int dpm_synchronize(dpm_set_t h) {
  CHECK_HANDLE(h);
  HIPCHK(hipStreamSynchronize(S->stream));
  return DPM_OK;
}

int dpm_set_launch_config(dpm_set_t h, int wavesPerTile) {
  CHECK_HANDLE(h);
  if (wavesPerTile < 1 || wavesPerTile > 8) return fail(S, DPM_ERR_ARG, "wavesPerTile must be in 1..8");
  S->split = wavesPerTile;
  bump(S);
  return DPM_OK;
}

int dpm_set_field(dpm_set_t h, int field, const double* host, long n) {
  CHECK_HANDLE(h);
  double *c0, *c1;
  long cnt;
  int nc;
  if (!host || !field_ptrs(S, field, &c0, &c1, &cnt, &nc)) return fail(S, DPM_ERR_ARG, "bad field");
  if (n != cnt * nc) return fail(S, DPM_ERR_ARG, "field size mismatch");
  if (field == DPM_FIELD_RAD || field == DPM_FIELD_L0 || field == DPM_FIELD_THETA0 || field == DPM_FIELD_A0 ||
      field == DPM_FIELD_BOX)
    S->paramSerial++;
  if (nc == 1) {
    HIPCHK(hipMemcpyAsync(c0, host, sizeof(double) * cnt, hipMemcpyHostToDevice, S->stream));
  } else {
    std::vector<double> a(cnt), b(cnt);
    for (long i = 0; i < cnt; i++) { a[i] = host[2 * i]; b[i] = host[2 * i + 1]; }
    HIPCHK(hipMemcpyAsync(c0, a.data(), sizeof(double) * cnt, hipMemcpyHostToDevice, S->stream));
    HIPCHK(hipMemcpyAsync(c1, b.data(), sizeof(double) * cnt, hipMemcpyHostToDevice, S->stream));
    HIPCHK(hipStreamSynchronize(S->stream));
  }
  HIPCHK(hipStreamSynchronize(S->stream));
  if (field == DPM_FIELD_BOX) {
    for (long i = 0; i < 2 * cnt; i++) S->h_box[i] = host[i];
    int rc = upload_box(S);
    if (rc) return rc;
    if (S->cutSet) { rc = configure_cells(S); if (rc) return rc; }
  }
  if (field == DPM_FIELD_RAD) {
    for (long s = 0; s < S->nsys; s++) {
      double m = 0;
      for (int v = S->h_svoff[s]; v < S->h_svoff[s + 1]; v++) m = std::max(m, host[v]);
      S->h_maxrad[s] = m;
    }
    if (S->cutSet) { int rc = configure_cells(S); if (rc) return rc; }
  }
  return DPM_OK;
}

int dpm_get_field(dpm_set_t h, int field, double* host, long n) {
  CHECK_HANDLE(h);
  double *c0, *c1;
  long cnt;
  int nc;
  if (!host || !field_ptrs(S, field, &c0, &c1, &cnt, &nc)) return fail(S, DPM_ERR_ARG, "bad field");
  if (n != cnt * nc) return fail(S, DPM_ERR_ARG, "field size mismatch");
  if (nc == 1) {
    HIPCHK(hipMemcpyAsync(host, c0, sizeof(double) * cnt, hipMemcpyDeviceToHost, S->stream));
    HIPCHK(hipStreamSynchronize(S->stream));
  } else {
    std::vector<double> a(cnt), b(cnt);
    HIPCHK(hipMemcpyAsync(a.data(), c0, sizeof(double) * cnt, hipMemcpyDeviceToHost, S->stream));
    HIPCHK(hipMemcpyAsync(b.data(), c1, sizeof(double) * cnt, hipMemcpyDeviceToHost, S->stream));
    HIPCHK(hipStreamSynchronize(S->stream));
    for (long i = 0; i < cnt; i++) { host[2 * i] = a[i]; host[2 * i + 1] = b[i]; }
  }
  return DPM_OK;
}

int dpm_device_ptr(dpm_set_t h, int field, int component, void** ptr) {
  CHECK_HANDLE(h);
  double *c0, *c1;
  long cnt;
  int nc;
  if (!ptr || !field_ptrs(S, field, &c0, &c1, &cnt, &nc) || component < 0 || component >= nc)
    return fail(S, DPM_ERR_ARG, "bad field/component");
  *ptr = component == 0 ? (void*)c0 : (void*)c1;
  return DPM_OK;
}

int dpm_set_types(dpm_set_t h, int potential, int interaction, int monomer, int particle) {
  CHECK_HANDLE(h);
  if (potential < 0 || potential > 3 ||
      (interaction != INT_VV && interaction != INT_SMOOTH && interaction != INT_ALL) || monomer < 0 ||
      monomer > 1 || particle < 0 || particle > 2)
    return fail(S, DPM_ERR_ARG, "bad type");
  if (S->nwide > 0 && particle != PART_PASSIVE)
    return fail(S, DPM_ERR_UNSUPPORTED, "active rings: rings of at most 64 vertices only");
  S->P.potential = potential;
  S->P.interaction = interaction;
  S->P.monomer = monomer;
  S->P.particle = particle;
  bump(S);
  return DPM_OK;
}

// setGeometryType, DPM2D.cu:338-358
int dpm_set_geometry(dpm_set_t h, int geometry) {
  CHECK_HANDLE(h);
  if (geometry < GEOM_NORMAL || geometry > GEOM_ROUND_BOX) return fail(S, DPM_ERR_ARG, "bad geometry");
  S->P.geometry = geometry;
  HIPCHK(launch_fill_d(S->V.wfx, 0.0, S->nv, S->stream));
  HIPCHK(launch_fill_d(S->V.wfy, 0.0, S->nv, S->stream));
  int rc = upload_box(S);
  if (rc) return rc;
  bump(S);
  return DPM_OK;
}

// setBoxEnergyCost, DPM2D.cu:1911-1914
int dpm_set_box_energy_cost(dpm_set_t h, double ew) {
  CHECK_HANDLE(h);
  S->P.ew = ew;
  bump(S);
  return DPM_OK;
}

// setBoxRadius / getBoxRadius, DPM2D.cu:627-664
int dpm_set_box_radius(dpm_set_t h, double boxRadius) {
  CHECK_HANDLE(h);
  if (!(boxRadius > 0)) return fail(S, DPM_ERR_ARG, "boxRadius must be positive");
  S->P.boxRadius = boxRadius;
  bump(S);
  return DPM_OK;
}

int dpm_get_box_radius(dpm_set_t h, double* boxRadius) {
  CHECK_HANDLE(h);
  if (!boxRadius) return fail(S, DPM_ERR_ARG, "null output");
  *boxRadius = S->P.boxRadius;
  return DPM_OK;
}

int dpm_set_energy_costs(dpm_set_t h, double ea, double el, double eb, double ec) {
  CHECK_HANDLE(h);
  S->P.ea = ea; S->P.el = el; S->P.eb = eb; S->P.ec = ec;
  bump(S);
  return DPM_OK;
}

int dpm_set_ring_costs(dpm_set_t h, const double* ka, const double* kl, const double* kb) {
  CHECK_HANDLE(h);
  if (!ka && !kl && !kb) {
    if (S->V.kr) {
      HIPCHK(hipStreamSynchronize(S->stream));
      hipFree(S->V.kr);
      S->V.kr = nullptr;
      S->paramSerial++;
      bump(S);
    }
    return DPM_OK;
  }
  if (!ka || !kl || !kb) return fail(S, DPM_ERR_ARG, "ring costs: pass all three arrays or none");
  const long np = S->np;
  if (!S->V.kr) HIPCHK(dalloc(&S->V.kr, 3 * np));
  std::vector<double> k(3 * np);
  for (long r = 0; r < np; r++) {
    k[r] = ka[r];
    k[np + r] = kl[r];
    k[2 * np + r] = kb[r];
  }
  HIPCHK(hipMemcpyAsync(S->V.kr, k.data(), sizeof(double) * 3 * np, hipMemcpyHostToDevice, S->stream));
  HIPCHK(hipStreamSynchronize(S->stream));
  S->paramSerial++;
  bump(S);
  return DPM_OK;
}

int dpm_set_attraction_constants(dpm_set_t h, double l1, double l2) {
  CHECK_HANDLE(h);
  S->P.l1 = l1; S->P.l2 = l2;
  bump(S);
  return DPM_OK;
}

// setLJcutoff, DPM2D.cu:1922-1932
int dpm_set_lj_cutoff(dpm_set_t h, double LJcutoff) {
  CHECK_HANDLE(h);
  S->P.LJcutoff = LJcutoff;
  const double ratio6 = 1 / pow(LJcutoff, 6);
  S->P.LJecut = 4 * S->P.ec * (ratio6 * ratio6 - ratio6);
  S->P.LJfshift = 24 * S->P.ec * (2 * ratio6 - 1) * ratio6 / LJcutoff;
  bump(S);
  return DPM_OK;
}

int dpm_set_fene_constants(dpm_set_t h, double stiff, double ext) {
  CHECK_HANDLE(h);
  S->P.stiff = stiff;
  S->P.extSq = ext * ext;
  bump(S);
  return DPM_OK;
}

int dpm_set_time_step(dpm_set_t h, double dt) {
  CHECK_HANDLE(h);
  S->P.dt = dt;
  if (S->simKind == 1) S->P.noise = sqrt(2. * S->Tinject * S->P.gamma / dt);
  bump(S);
  return DPM_OK;
}

int dpm_set_self_propulsion(dpm_set_t h, double driving, double taup) {
  CHECK_HANDLE(h);
  S->P.driving = driving;
  S->P.taup = taup;
  bump(S);
  return DPM_OK;
}

int dpm_set_seed(dpm_set_t h, uint64_t seed) {
  CHECK_HANDLE(h);
  S->P.seed = seed;
  S->injectCalls = 0;
  HIPCHK(hipMemsetAsync(S->d_ctr, 0, sizeof(unsigned long long) * 4, S->stream));
  HIPCHK(hipStreamSynchronize(S->stream));
  bump(S);
  return DPM_OK;
}

// setDisplacementCutoff, DPM2D.cu:1300-1329
int dpm_set_displacement_cutoff(dpm_set_t h, double cutoff, double size, double* cutDistance) {
  CHECK_HANDLE(h);
  double cd = 1;
  switch (S->P.potential) {
    case POT_HARMONIC: cd = 1; break;
    case POT_LJ: cd = S->P.LJcutoff; break;
    case POT_WCA: cd = kWCA; break;
    case POT_ADHESIVE: cd = 1 + S->P.l2; break;
  }
  cd += cutoff;
  S->P.cutDistance = cd;
  S->P.cutoff = cutoff * size;
  S->cutSet = true;
  if (cutDistance) *cutDistance = cd;
  bump(S);
  return configure_cells(S);
}

int dpm_set_cell_scale(dpm_set_t h, double scale) {
  CHECK_HANDLE(h);
  if (!(scale >= 1.0) || scale > 64.0) return fail(S, DPM_ERR_ARG, "dpm_set_cell_scale: scale in [1, 64]");
  if (scale == S->cellScale) return DPM_OK;
  S->cellScale = scale;
  return S->cutSet ? configure_cells(S) : DPM_OK;
}

int dpm_calc_neighbor_list(dpm_set_t h, double cutDistance) {
  CHECK_HANDLE(h);
  if (cutDistance != S->P.cutDistance || !S->cutSet) {
    S->P.cutDistance = cutDistance;
    S->cutSet = true;
    int rc = configure_cells(S);
    if (rc) return rc;
  }
  std::vector<int> all(S->nsys, 1);
  return rebuild(S, all, 0);
}

// checkNeighbors: displacement test against lastPos (kernelCheckVertexDisplacement) per system
int dpm_check_neighbors(dpm_set_t h) {
  CHECK_HANDLE(h);
  int rc = require_list(S);
  if (rc) return rc;
  std::vector<double> x(S->nv), y(S->nv), lx(S->nv), ly(S->nv);
  HIPCHK(hipMemcpyAsync(x.data(), S->V.x, sizeof(double) * S->nv, hipMemcpyDeviceToHost, S->stream));
  HIPCHK(hipMemcpyAsync(y.data(), S->V.y, sizeof(double) * S->nv, hipMemcpyDeviceToHost, S->stream));
  HIPCHK(hipMemcpyAsync(lx.data(), S->V.lx, sizeof(double) * S->nv, hipMemcpyDeviceToHost, S->stream));
  HIPCHK(hipMemcpyAsync(ly.data(), S->V.ly, sizeof(double) * S->nv, hipMemcpyDeviceToHost, S->stream));
  HIPCHK(hipStreamSynchronize(S->stream));
  std::vector<int> flags(S->nsys, 0);
  for (long v = 0; v < S->nv; v++) {
    const int s = S->h_vsys[v];
    const double Lx = S->h_box[2 * s], Ly = S->h_box[2 * s + 1];
    double dx = x[v] - lx[v], dy = y[v] - ly[v];
    dx -= Lx * round(dx / Lx);
    dy -= Ly * round(dy / Ly);
    if (3 * sqrt(dx * dx + dy * dy) > S->P.cutoff) flags[s] = 1;
  }
  return rebuild(S, flags, 1);
}

int dpm_get_neighbor_list(dpm_set_t h, long* list, long* counts, long cap, long* width) {
  CHECK_HANDLE(h);
  const long nv = S->nv, W = S->V.W;
  std::vector<int> nbr((size_t)W * nv), cnt(nv);
  HIPCHK(hipMemcpyAsync(nbr.data(), S->V.nbr, sizeof(int) * W * nv, hipMemcpyDeviceToHost, S->stream));
  HIPCHK(hipMemcpyAsync(cnt.data(), S->V.ncnt, sizeof(int) * nv, hipMemcpyDeviceToHost, S->stream));
  HIPCHK(hipStreamSynchronize(S->stream));
  for (long v = 0; v < nv; v++) {
    if (counts) counts[v] = cnt[v];
    if (list)
      for (long k = 0; k < cap; k++) list[v * cap + k] = (k < cnt[v] && k < W) ? nbr[k * nv + v] : -1;
  }
  if (width) *width = W;
  return DPM_OK;
}

int dpm_reset_last_positions(dpm_set_t h) {
  CHECK_HANDLE(h);
  HIPCHK(launch_copy_lastpos(S));
  HIPCHK(hipStreamSynchronize(S->stream));
  return DPM_OK;
}

int dpm_get_update_count(dpm_set_t h, long* perSystem) {
  CHECK_HANDLE(h);
  std::vector<int> c(S->nsys);
  HIPCHK(hipMemcpyAsync(c.data(), S->d_updateCount, sizeof(int) * S->nsys, hipMemcpyDeviceToHost, S->stream));
  HIPCHK(hipStreamSynchronize(S->stream));
  for (long s = 0; s < S->nsys; s++) perSystem[s] = c[s];
  return DPM_OK;
}

int dpm_reset_update_count(dpm_set_t h) {
  CHECK_HANDLE(h);
  HIPCHK(launch_fill_int(S->d_updateCount, 0, S->nsys, S->stream));
  HIPCHK(hipStreamSynchronize(S->stream));
  return DPM_OK;
}

int dpm_calc_particle_shape(dpm_set_t h) {
  CHECK_HANDLE(h);
  HIPCHK(launch_shape_only(S));
  HIPCHK(hipStreamSynchronize(S->stream));
  return DPM_OK;
}

int dpm_calc_force_energy(dpm_set_t h) {
  CHECK_HANDLE(h);
  int rc = require_list(S);
  if (rc) return rc;
  // a standalone force call counts as one active-noise draw (addSelfPropulsion runs inside
  // calcForceEnergy in the reference); with the any-flag cleared the bin launch only bumps it
  HIPCHK(launch_fill_int(S->d_any + S->cur, 0, 1, S->stream));
  HIPCHK(launch_bin(S, 0, S->P.particle != PART_PASSIVE ? 1 : 0));
  StepArgs A{};
  A.ctr = S->d_ctr;
  A.ctrOffT = A.ctrOffA = -1;  // the bin launch has bumped the counters
  A.err = S->d_err;
  A.manage = 0;
  HIPCHK(launch_step(S, MODE_FORCE, A));
  HIPCHK(hipStreamSynchronize(S->stream));
  return DPM_OK;
}

int dpm_get_interaction_forces(dpm_set_t h, double* host, long n) {
  CHECK_HANDLE(h);
  if (!host || n != 2L * S->nv) return fail(S, DPM_ERR_ARG, "interaction forces: n must be 2 * numVertices");
  int rc = require_list(S);
  if (rc) return rc;
  const size_t nv = S->nv, np = S->np;
  // keep the stored forces and energies: this call must not change the state
  std::vector<double> fx(nv), fy(nv), ev(nv), pe(np);
  HIPCHK(hipMemcpyAsync(fx.data(), S->V.fx, nv * sizeof(double), hipMemcpyDeviceToHost, S->stream));
  HIPCHK(hipMemcpyAsync(fy.data(), S->V.fy, nv * sizeof(double), hipMemcpyDeviceToHost, S->stream));
  HIPCHK(hipMemcpyAsync(ev.data(), S->V.ev, nv * sizeof(double), hipMemcpyDeviceToHost, S->stream));
  HIPCHK(hipMemcpyAsync(pe.data(), S->V.pe, np * sizeof(double), hipMemcpyDeviceToHost, S->stream));
  HIPCHK(hipStreamSynchronize(S->stream));
  // the same step kernel with the shape costs and self-propulsion switched off
  const Params keep = S->P;
  double* keepKr = S->V.kr;
  S->P.ea = S->P.el = S->P.eb = 0;
  S->P.particle = PART_PASSIVE;
  S->V.kr = nullptr;
  HIPCHK(launch_fill_int(S->d_any + S->cur, 0, 1, S->stream));
  HIPCHK(launch_bin(S, 0, 0));
  StepArgs A{};
  A.ctr = S->d_ctr;
  A.ctrOffT = A.ctrOffA = -1;
  A.err = S->d_err;
  A.manage = 0;
  hipError_t e = launch_step(S, MODE_FORCE, A);
  S->P = keep;
  S->V.kr = keepKr;
  HIPCHK(e);
  std::vector<double> ix(nv), iy(nv);
  HIPCHK(hipMemcpyAsync(ix.data(), S->V.fx, nv * sizeof(double), hipMemcpyDeviceToHost, S->stream));
  HIPCHK(hipMemcpyAsync(iy.data(), S->V.fy, nv * sizeof(double), hipMemcpyDeviceToHost, S->stream));
  HIPCHK(hipStreamSynchronize(S->stream));
  for (size_t v = 0; v < nv; v++) {
    host[2 * v] = ix[v];
    host[2 * v + 1] = iy[v];
  }
  HIPCHK(hipMemcpyAsync(S->V.fx, fx.data(), nv * sizeof(double), hipMemcpyHostToDevice, S->stream));
  HIPCHK(hipMemcpyAsync(S->V.fy, fy.data(), nv * sizeof(double), hipMemcpyHostToDevice, S->stream));
  HIPCHK(hipMemcpyAsync(S->V.ev, ev.data(), nv * sizeof(double), hipMemcpyHostToDevice, S->stream));
  HIPCHK(hipMemcpyAsync(S->V.pe, pe.data(), np * sizeof(double), hipMemcpyHostToDevice, S->stream));
  HIPCHK(hipStreamSynchronize(S->stream));
  return DPM_OK;
}

// setNoiseType, DPM2D.cu:410-423 (simControlStruct::noiseEnum order)
int dpm_set_noise_type(dpm_set_t h, int noiseType) {
  CHECK_HANDLE(h);
  if (noiseType < NOISE_LANGEVIN || noiseType > NOISE_DRIVEN_BROWNIAN) return fail(S, DPM_ERR_ARG, "bad noise type");
  S->noiseType = noiseType;
  bump(S);
  return DPM_OK;
}

// setActiveTensionParams, DPM2D.cu:1898-1903: the rest lengths become the current l0
int dpm_set_active_tension(dpm_set_t h, double l0Tau, double l0Diff) {
  CHECK_HANDLE(h);
  if (!(l0Tau != 0) || l0Diff < 0) return fail(S, DPM_ERR_ARG, "l0Tau must be non-zero and l0Diff >= 0");
  if (!S->V.l0r) HIPCHK(dalloc(&S->V.l0r, S->nv));
  HIPCHK(launch_copy_d(S->V.l0r, S->V.l0, S->nv, S->stream));
  HIPCHK(hipStreamSynchronize(S->stream));
  S->P.l0Tau = l0Tau;
  S->P.l0Diff = l0Diff;
  bump(S);  // captured sequences hold the old view
  return DPM_OK;
}

int dpm_init_langevin(dpm_set_t h, double T, double gamma, int readState, int zeroMomentum) {
  CHECK_HANDLE(h);
  S->simKind = 1;
  S->Tinject = T;
  // initLangevin, DPM2D.cu:3713-3753: BAOAB keeps exp(-gamma dt) and its O-step amplitude
  if (S->noiseType == NOISE_BAOAB) {
    S->P.gamma = exp(-gamma * S->P.dt);
    S->P.noise = sqrt(1 - exp(-2 * gamma * S->P.dt)) * sqrt(T);
  } else {
    S->P.gamma = gamma;
    S->P.noise = sqrt(2. * T * gamma / S->P.dt);
  }
  bump(S);
  HIPCHK(launch_copy_lastpos(S));
  { int r0 = store_initial_positions(S); if (r0) return r0; }
  if (!readState) {
    HIPCHK(launch_inject(S, sqrt(T), S->injectCalls++));
    if (zeroMomentum) HIPCHK(launch_sub_mean(S));
  }
  HIPCHK(hipStreamSynchronize(S->stream));
  return DPM_OK;
}

int dpm_inject_kinetic_energy(dpm_set_t h, double T) {
  CHECK_HANDLE(h);
  HIPCHK(launch_inject(S, sqrt(T), S->injectCalls++));
  HIPCHK(hipStreamSynchronize(S->stream));
  return DPM_OK;
}

int dpm_update_position(dpm_set_t h, double timeStep) {
  CHECK_HANDLE(h);
  HIPCHK(launch_update_pos(S, timeStep));
  HIPCHK(hipStreamSynchronize(S->stream));
  return DPM_OK;
}

int dpm_update_velocity(dpm_set_t h, double timeStep) {
  CHECK_HANDLE(h);
  HIPCHK(launch_update_vel(S, timeStep));
  HIPCHK(hipStreamSynchronize(S->stream));
  return DPM_OK;
}

int dpm_update_thermal_vel(dpm_set_t h) {
  CHECK_HANDLE(h);
  if (S->simKind != 1 || S->noiseType != NOISE_LANGEVIN)
    return fail(S, DPM_ERR_STATE, "dpm_update_thermal_vel: dpm_init_langevin (Langevin noise) not called");
  HIPCHK(launch_thermal_force(S));
  HIPCHK(launch_ctr_add(S, 1, 0));
  HIPCHK(hipStreamSynchronize(S->stream));
  return DPM_OK;
}

int dpm_conserve_momentum(dpm_set_t h) {
  CHECK_HANDLE(h);
  HIPCHK(launch_sub_mean(S));
  HIPCHK(hipStreamSynchronize(S->stream));
  return DPM_OK;
}

int dpm_init_nve(dpm_set_t h, double T, int readState, int zeroMomentum) {
  CHECK_HANDLE(h);
  S->simKind = 2;
  S->Tinject = T;
  bump(S);
  HIPCHK(launch_copy_lastpos(S));
  { int r0 = store_initial_positions(S); if (r0) return r0; }
  if (!readState) {
    HIPCHK(launch_inject(S, sqrt(T), S->injectCalls++));
    if (zeroMomentum) HIPCHK(launch_sub_mean(S));
  }
  HIPCHK(hipStreamSynchronize(S->stream));
  return DPM_OK;
}

int dpm_langevin_loop(dpm_set_t h, long steps) {
  CHECK_HANDLE(h);
  if (S->simKind != 1) return fail(S, DPM_ERR_STATE, "dpm_init_langevin not called");
  return run_verlet(S, steps, true);
}

int dpm_nve_loop(dpm_set_t h, long steps) {
  CHECK_HANDLE(h);
  if (S->simKind != 2 && S->simKind != 3) return fail(S, DPM_ERR_STATE, "dpm_init_nve not called");
  return run_verlet(S, steps, false);
}

int dpm_init_nve_rescale(dpm_set_t h, double T) {
  CHECK_HANDLE(h);
  S->simKind = 3;
  S->Tinject = T;
  bump(S);
  HIPCHK(launch_copy_lastpos(S));
  { int r0 = store_initial_positions(S); if (r0) return r0; }
  HIPCHK(launch_inject(S, sqrt(T), S->injectCalls++));
  return temperature_scale(S, T);
}

int dpm_nve_rescale_loop(dpm_set_t h, long steps) {
  CHECK_HANDLE(h);
  if (S->simKind != 3) return fail(S, DPM_ERR_STATE, "dpm_init_nve_rescale not called");
  int rc = require_list(S);
  if (rc) return rc;
  HIPCHK(launch_fill_int(S->d_err, 0, 1, S->stream));
  for (long i = 0; i < steps; i++) {
    rc = temperature_scale(S, S->Tinject);  // NVERescale::injectKineticEnergy, Simulator.cu:222-238
    if (rc) return rc;
    rc = with_replay(S, 0, [&]() { return verlet_sequence(S, 1, false); });
    if (rc) return rc;
  }
  return check_err_after_loop(S);
}

int dpm_adjust_temperature(dpm_set_t h, double T) {
  CHECK_HANDLE(h);
  return temperature_scale(S, T);
}

int dpm_adjust_kinetic_energy(dpm_set_t h, const double* prevEtot, double* scaleOut) {
  CHECK_HANDLE(h);
  if (!prevEtot) return fail(S, DPM_ERR_ARG, "dpm_adjust_kinetic_energy: null prevEtot");
  int rc = reduce_host(S);
  if (rc) return rc;
  std::vector<double> sc(S->nsys);
  for (long s = 0; s < S->nsys; s++) {
    const double ekin = 0.5 * S->h_red[s * 8 + 2];
    const double epot = S->h_red[s * 8 + 0] + (S->P.interaction == INT_SMOOTH ? S->h_red[s * 8 + 1] : 0.0);
    const double dE = (epot + ekin) - prevEtot[s];
    const bool adjust = ekin > dE;
    sc[s] = adjust ? sqrt((ekin - dE) / ekin) : 1.0;
    if (scaleOut) scaleOut[s] = adjust ? sc[s] : 0.0;
  }
  HIPCHK(hipMemcpyAsync(S->d_scale, sc.data(), sizeof(double) * S->nsys, hipMemcpyHostToDevice, S->stream));
  HIPCHK(launch_scale_vel(S));
  HIPCHK(hipStreamSynchronize(S->stream));
  return DPM_OK;
}

int dpm_init_fire(dpm_set_t h, const double* p, long minStep, long numStep) {
  CHECK_HANDLE(h);
  if (!p) return fail(S, DPM_ERR_ARG, "null FIRE params");
  S->fp.a_start = p[0]; S->fp.f_dec = p[1]; S->fp.f_inc = p[2]; S->fp.f_a = p[3];
  S->fp.dt_max = p[5];
  S->fp.minStep = minStep;
  S->fp.tol = -1;
  S->fp.maxIter = 0;
  std::vector<FireState> fs(S->nsys);
  for (auto& f : fs) {
    f.dt = p[4]; f.a = p[6]; f.P = 0; f.numStep = numStep; f.iters = 0; f.converged = 0; f.pad = 0;
  }
  for (int b = 0; b < 2; b++)
    HIPCHK(hipMemcpyAsync(S->d_fire[b], fs.data(), sizeof(FireState) * S->nsys, hipMemcpyHostToDevice, S->stream));
  S->fireCur = 0;
  S->fireInit = true;
  bump(S);
  HIPCHK(hipStreamSynchronize(S->stream));
  return DPM_OK;
}

int dpm_fire_loop(dpm_set_t h, long iterations) {
  CHECK_HANDLE(h);
  if (!S->fireInit) return fail(S, DPM_ERR_STATE, "dpm_init_fire not called");
  int rc = require_list(S);
  if (rc) return rc;
  if (S->fp.tol >= 0) {
    S->fp.tol = -1;
    bump(S);
  }
  return with_replay(S, SNAP_FIRE, [&]() { return run_fire(S, iterations); });
}

// The compressFIRE.cpp:121-131 loop for every system at once.  A system whose max |F_c| has
// dropped to tol (or that has done maxIter iterations) freezes while the others continue; the
// host only polls an all-converged flag between chunks of iterations.
int dpm_fire_minimize(dpm_set_t h, double tol, long maxIter, long* iterations) {
  CHECK_HANDLE(h);
  if (!S->fireInit) return fail(S, DPM_ERR_STATE, "dpm_init_fire not called");
  int rc = require_list(S);
  if (rc) return rc;
  if (tol < 0 || maxIter < 0) return fail(S, DPM_ERR_ARG, "tol and maxIter must be >= 0");
  if (S->fp.tol != tol || S->fp.maxIter != maxIter) {
    S->fp.tol = tol;
    S->fp.maxIter = maxIter;
    bump(S);
  }
  HIPCHK(launch_fill_int(S->d_err, 0, 1, S->stream));
  HIPCHK(launch_fire_resume(S));
  long done = 0;
  if (S->fireKernel == 1 && S->P.particle == PART_PASSIVE && S->P.geometry == GEOM_NORMAL && S->nwide == 0) {
    // one workgroup per system, at most `pchunk` iterations per launch
    rc = clean_protocol(S);
    if (rc) return rc;
    int waves = S->fireWaves;
    if (waves <= 0) {
      int maxTiles = 1;
      for (long s = 0; s < S->nsys; s++) maxTiles = std::max(maxTiles, S->h_stile[s + 1] - S->h_stile[s]);
      waves = std::min(16, maxTiles);
    }
    const long pchunk = 2048;
    while (done <= maxIter) {
      rc = with_replay(S, SNAP_FIRE, [&]() -> int {
        HIPCHK(launch_fire_sys(S, (int)pchunk, waves));
        return DPM_OK;
      });
      if (rc) return rc;
      done += pchunk;
      int notDone = 0;
      HIPCHK(launch_fill_int(S->d_notDone, 0, 1, S->stream));
      HIPCHK(launch_fire_done(S, S->d_notDone));
      HIPCHK(hipMemcpyAsync(&notDone, S->d_notDone, sizeof(int), hipMemcpyDeviceToHost, S->stream));
      HIPCHK(hipStreamSynchronize(S->stream));
      if (!notDone) break;
    }
    done = maxIter + 1;  // the graph path below is skipped
  }
  const long chunk = 256;
  while (done <= maxIter) {
    rc = with_replay(S, SNAP_FIRE, [&]() { return run_fire(S, chunk); });
    if (rc) return rc;
    done += chunk;
    int notDone = 0;
    HIPCHK(launch_fill_int(S->d_notDone, 0, 1, S->stream));
    HIPCHK(launch_fire_done(S, S->d_notDone));
    HIPCHK(hipMemcpyAsync(&notDone, S->d_notDone, sizeof(int), hipMemcpyDeviceToHost, S->stream));
    HIPCHK(hipStreamSynchronize(S->stream));
    if (!notDone) break;
  }
  if (iterations) {
    std::vector<FireState> fs(S->nsys);
    HIPCHK(hipMemcpyAsync(fs.data(), S->d_fire[S->fireCur], sizeof(FireState) * S->nsys, hipMemcpyDeviceToHost,
                          S->stream));
    HIPCHK(hipStreamSynchronize(S->stream));
    for (long s = 0; s < S->nsys; s++) iterations[s] = fs[s].iters;
  }
  return check_err_after_loop(S);
}

int dpm_set_fire_kernel(dpm_set_t h, int persistent, int wavesPerSystem) {
  CHECK_HANDLE(h);
  if (persistent < 0 || persistent > 1 || wavesPerSystem < 0 || wavesPerSystem > 16)
    return fail(S, DPM_ERR_ARG, "dpm_set_fire_kernel: persistent in {0, 1}, waves in 0..16");
  if (persistent && S->nwide > 0)
    return fail(S, DPM_ERR_UNSUPPORTED, "one-workgroup-per-system FIRE: rings of at most 64 vertices only");
  S->fireKernel = persistent;
  S->fireWaves = wavesPerSystem;
  return DPM_OK;
}

int dpm_get_fire_state(dpm_set_t h, double* out) {
  CHECK_HANDLE(h);
  std::vector<FireState> fs(S->nsys);
  HIPCHK(hipMemcpyAsync(fs.data(), S->d_fire[S->fireCur], sizeof(FireState) * S->nsys, hipMemcpyDeviceToHost,
                        S->stream));
  HIPCHK(hipStreamSynchronize(S->stream));
  for (long s = 0; s < S->nsys; s++) {
    out[5 * s + 0] = fs[s].dt;
    out[5 * s + 1] = fs[s].a;
    out[5 * s + 2] = (double)fs[s].numStep;
    out[5 * s + 3] = fs[s].P;
    out[5 * s + 4] = (double)fs[s].iters;
  }
  return DPM_OK;
}

int dpm_get_potential_energy(dpm_set_t h, double* out) {
  CHECK_HANDLE(h);
  int rc = reduce_host(S);
  if (rc) return rc;
  for (long s = 0; s < S->nsys; s++)
    out[s] = S->h_red[s * 8 + 0] + (S->P.interaction == INT_SMOOTH ? S->h_red[s * 8 + 1] : 0.0);
  return DPM_OK;
}

int dpm_get_kinetic_energy(dpm_set_t h, double* out) {
  CHECK_HANDLE(h);
  int rc = reduce_host(S);
  if (rc) return rc;
  for (long s = 0; s < S->nsys; s++) out[s] = 0.5 * S->h_red[s * 8 + 2];
  return DPM_OK;
}

int dpm_get_temperature(dpm_set_t h, double* out) {
  CHECK_HANDLE(h);
  int rc = reduce_host(S);
  if (rc) return rc;
  for (long s = 0; s < S->nsys; s++) {
    const double n = S->h_svoff[s + 1] - S->h_svoff[s];
    out[s] = 2. * (0.5 * S->h_red[s * 8 + 2]) / (n * 2);
  }
  return DPM_OK;
}

int dpm_get_max_unbalanced_force(dpm_set_t h, double* out) {
  CHECK_HANDLE(h);
  int rc = reduce_host(S);
  if (rc) return rc;
  for (long s = 0; s < S->nsys; s++) out[s] = sqrt(S->h_red[s * 8 + 3]);
  return DPM_OK;
}

int dpm_get_total_force_magnitude(dpm_set_t h, double* out) {
  CHECK_HANDLE(h);
  int rc = reduce_host(S);
  if (rc) return rc;
  for (long s = 0; s < S->nsys; s++) {
    const double n = S->h_svoff[s + 1] - S->h_svoff[s];
    out[s] = sqrt(S->h_red[s * 8 + 4] / (n * 2));
  }
  return DPM_OK;
}

int dpm_get_phi(dpm_set_t h, double* out) {
  CHECK_HANDLE(h);
  int rc = reduce_host(S);
  if (rc) return rc;
  for (long s = 0; s < S->nsys; s++)
    out[s] = (S->h_red[s * 8 + 5] + kPI * S->h_red[s * 8 + 7]) / box_area(S, s);
  return DPM_OK;
}

int dpm_get_preferred_phi(dpm_set_t h, double* out) {
  CHECK_HANDLE(h);
  int rc = reduce_host(S);
  if (rc) return rc;
  for (long s = 0; s < S->nsys; s++)
    out[s] = (S->h_red[s * 8 + 6] + kPI * S->h_red[s * 8 + 7]) / box_area(S, s);
  return DPM_OK;
}

// setInitialPositions + setInitialParticlePositions, DPM2D.cu:921-930 (also done by the inits)
int dpm_set_initial_positions(dpm_set_t h) {
  CHECK_HANDLE(h);
  int rc = store_initial_positions(S);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(S->stream));
  return DPM_OK;
}

// calcContacts, DPM2D.cu:3408-3428: rows widen (doubling from 16) until every ring's contacts fit
int dpm_calc_contacts(dpm_set_t h) {
  CHECK_HANDLE(h);
  int rc = require_list(S);
  if (rc) return rc;
  if (!S->d_ccount) HIPCHK(dalloc(&S->d_ccount, S->np));
  int width = S->contactWidth > 0 ? S->contactWidth : 16;
  for (;;) {
    if (width != S->contactWidth || !S->d_contacts) {
      HIPCHK(hipStreamSynchronize(S->stream));
      hipFree(S->d_contacts);
      S->d_contacts = nullptr;
      HIPCHK(dalloc(&S->d_contacts, (long)S->np * width));
      S->contactWidth = width;
    }
    HIPCHK(launch_fill_int(S->d_notDone, 0, 1, S->stream));
    HIPCHK(launch_contacts(S, S->d_contacts, S->d_ccount, width, S->d_notDone));
    int over = 0;
    HIPCHK(hipMemcpyAsync(&over, S->d_notDone, sizeof(int), hipMemcpyDeviceToHost, S->stream));
    HIPCHK(hipStreamSynchronize(S->stream));
    if (!over) return DPM_OK;
    if (width >= (1 << 20)) return fail(S, DPM_ERR_OVERFLOW, "contact rows do not fit");
    width *= 2;
  }
}

// calcGeometricContacts, DPM2D.cu:3430-3450: gap 1, rows of 4 * nDim = 8, extra contacts dropped
int dpm_calc_geometric_contacts(dpm_set_t h) {
  CHECK_HANDLE(h);
  int rc = require_list(S);
  if (rc) return rc;
  if (!S->d_ccount) HIPCHK(dalloc(&S->d_ccount, S->np));
  const int width = 8;
  if (width != S->contactWidth || !S->d_contacts) {
    HIPCHK(hipStreamSynchronize(S->stream));
    hipFree(S->d_contacts);
    S->d_contacts = nullptr;
    HIPCHK(dalloc(&S->d_contacts, (long)S->np * width));
    S->contactWidth = width;
  }
  HIPCHK(launch_contacts(S, S->d_contacts, S->d_ccount, width, S->d_notDone, 1.0));
  HIPCHK(hipStreamSynchronize(S->stream));
  return DPM_OK;
}

// getContacts, DPM2D.cu:1019-1024: rows of `cap` ring ids (-1 padded) and counts; the device
// row width in *width.  Before any dpm_calc_contacts every ring has no contact.
int dpm_get_contacts(dpm_set_t h, long* list, long* counts, long cap, long* width) {
  CHECK_HANDLE(h);
  const long np = S->np, W = S->contactWidth;
  std::vector<int> rows((size_t)np * (W > 0 ? W : 1), -1), cnt(np, 0);
  if (S->d_contacts) {
    HIPCHK(hipMemcpyAsync(rows.data(), S->d_contacts, sizeof(int) * np * W, hipMemcpyDeviceToHost, S->stream));
    HIPCHK(hipMemcpyAsync(cnt.data(), S->d_ccount, sizeof(int) * np, hipMemcpyDeviceToHost, S->stream));
    HIPCHK(hipStreamSynchronize(S->stream));
  }
  for (long p = 0; p < np; p++) {
    if (counts) counts[p] = cnt[p];
    if (list)
      for (long k = 0; k < cap; k++) list[p * cap + k] = (k < cnt[p] && k < W) ? rows[p * W + k] : -1;
  }
  if (width) *width = W;
  return DPM_OK;
}

// getParticleHexaticOrder, DPM2D.cu:1505-1510: mean over each system's rings of psi6 from the
// contacts of the last dpm_calc_contacts (0 for rings without contacts)
int dpm_get_hexatic_order(dpm_set_t h, double* perSystem) {
  CHECK_HANDLE(h);
  if (!perSystem) return fail(S, DPM_ERR_ARG, "null output");
  int rc = obs_scratch(S, S->np);
  if (rc) return rc;
  std::vector<double> psi(S->np, 0.0);
  if (S->d_contacts) {
    HIPCHK(launch_shape_only(S));
    HIPCHK(launch_hexatic(S, S->d_contacts, S->d_ccount, S->contactWidth, S->d_obs));
    HIPCHK(hipMemcpyAsync(psi.data(), S->d_obs, sizeof(double) * S->np, hipMemcpyDeviceToHost, S->stream));
    HIPCHK(hipStreamSynchronize(S->stream));
  }
  for (long s = 0; s < S->nsys; s++) {
    double t = 0;
    for (long p = S->h_sroff[s]; p < S->h_sroff[s + 1]; p++) t += psi[p];
    perSystem[s] = t / (S->h_sroff[s + 1] - S->h_sroff[s]);
  }
  return DPM_OK;
}

// getVertexMSD / getParticleMSD, DPM2D.cu:1284-1298: sum of squared displacements over the
// system's vertices (ring centres) / (count * Lx * Ly)
int dpm_get_vertex_msd(dpm_set_t h, double* perSystem) {
  CHECK_HANDLE(h);
  std::vector<double> d2, sf;
  int rc = displacement_sums(S, false, 1.0, d2, sf);
  if (rc) return rc;
  for (long s = 0; s < S->nsys; s++)
    perSystem[s] = d2[s] / ((S->h_svoff[s + 1] - S->h_svoff[s]) * S->h_box[2 * s] * S->h_box[2 * s + 1]);
  return DPM_OK;
}

int dpm_get_particle_msd(dpm_set_t h, double* perSystem) {
  CHECK_HANDLE(h);
  std::vector<double> d2, sf;
  int rc = displacement_sums(S, true, 1.0, d2, sf);
  if (rc) return rc;
  for (long s = 0; s < S->nsys; s++)
    perSystem[s] = d2[s] / ((S->h_sroff[s + 1] - S->h_sroff[s]) * S->h_box[2 * s] * S->h_box[2 * s + 1]);
  return DPM_OK;
}

// getVertexISF (wave number pi / (2 r) of the system's first vertex) / getParticleISF(k),
// DPM2D.cu:1486-1503: mean of sin(k d) / (k d)
int dpm_get_vertex_isf(dpm_set_t h, double* perSystem) {
  CHECK_HANDLE(h);
  std::vector<double> d2, sf;
  int rc = displacement_sums(S, false, 0.0, d2, sf);
  if (rc) return rc;
  for (long s = 0; s < S->nsys; s++) perSystem[s] = sf[s] / (S->h_svoff[s + 1] - S->h_svoff[s]);
  return DPM_OK;
}

int dpm_get_particle_isf(dpm_set_t h, double waveNumber, double* perSystem) {
  CHECK_HANDLE(h);
  if (!(waveNumber > 0)) return fail(S, DPM_ERR_ARG, "waveNumber must be positive");
  std::vector<double> d2, sf;
  int rc = displacement_sums(S, true, waveNumber, d2, sf);
  if (rc) return rc;
  for (long s = 0; s < S->nsys; s++) perSystem[s] = sf[s] / (S->h_sroff[s + 1] - S->h_sroff[s]);
  return DPM_OK;
}

// getPerParticleStressTensor, DPM2D.cu:962-967 / 2189-2198: 4 per ring (xx, xy, yx, yy)
int dpm_get_per_particle_stress(dpm_set_t h, double* out) {
  CHECK_HANDLE(h);
  int rc = require_list(S);
  if (rc) return rc;
  rc = obs_scratch(S, 4 * S->np);
  if (rc) return rc;
  HIPCHK(launch_shape_only(S));
  HIPCHK(launch_particle_stress(S, S->d_obs));
  HIPCHK(hipMemcpyAsync(out, S->d_obs, sizeof(double) * 4 * S->np, hipMemcpyDeviceToHost, S->stream));
  HIPCHK(hipStreamSynchronize(S->stream));
  return DPM_OK;
}

// getStressTensor, DPM2D.cu:969-974 / 2181-2187: per system the sum of its rings' tensors (the
// reference's kernelCalcStressTensor adds ring 0's entries once per ring without atomics;
// the sum is what it is meant to form); getPressure, DPM2D.cu:976-983: trace / (2 Nv)
int dpm_get_stress_tensor(dpm_set_t h, double* out4) {
  CHECK_HANDLE(h);
  std::vector<double> per(4 * S->np);
  int rc = dpm_get_per_particle_stress(h, per.data());
  if (rc) return rc;
  for (long s = 0; s < S->nsys; s++) {
    for (int c = 0; c < 4; c++) out4[4 * s + c] = 0;
    for (long p = S->h_sroff[s]; p < S->h_sroff[s + 1]; p++)
      for (int c = 0; c < 4; c++) out4[4 * s + c] += per[4 * p + c];
  }
  return DPM_OK;
}

int dpm_get_pressure(dpm_set_t h, double* perSystem) {
  CHECK_HANDLE(h);
  std::vector<double> st(4 * S->nsys);
  int rc = dpm_get_stress_tensor(h, st.data());
  if (rc) return rc;
  for (long s = 0; s < S->nsys; s++)
    perSystem[s] = (st[4 * s] + st[4 * s + 3]) / (2.0 * (S->h_svoff[s + 1] - S->h_svoff[s]));
  return DPM_OK;
}

int dpm_get_particle_shapes(dpm_set_t h, double* shape3) {
  CHECK_HANDLE(h);
  std::vector<double> a(S->np), p(S->np);
  HIPCHK(hipMemcpyAsync(a.data(), S->V.area, sizeof(double) * S->np, hipMemcpyDeviceToHost, S->stream));
  HIPCHK(hipMemcpyAsync(p.data(), S->V.perim, sizeof(double) * S->np, hipMemcpyDeviceToHost, S->stream));
  HIPCHK(hipStreamSynchronize(S->stream));
  for (long r = 0; r < S->np; r++) {
    shape3[3 * r] = a[r];
    shape3[3 * r + 1] = p[r];
    shape3[3 * r + 2] = p[r] * p[r] / (4 * kPI * a[r]);
  }
  return DPM_OK;
}

int dpm_scale_vertices(dpm_set_t h, const double* scale) {
  CHECK_HANDLE(h);
  HIPCHK(hipMemcpyAsync(S->d_scale, scale, sizeof(double) * S->nsys, hipMemcpyHostToDevice, S->stream));
  HIPCHK(launch_scale_vertices(S));
  HIPCHK(hipStreamSynchronize(S->stream));
  for (long s = 0; s < S->nsys; s++) S->h_maxrad[s] *= scale[s];
  if (S->cutSet) return configure_cells(S);
  return DPM_OK;
}

// scaleBox, DPM2D.cu:1693-1719.  Ring centres are those of the current positions (recomputed
// here; the reference holds those of its last force evaluation, the same after any loop or
// force call).  The round box takes one radius for all systems, so its scales must agree.
int dpm_scale_box(dpm_set_t h, const double* scale) {
  CHECK_HANDLE(h);
  if (!scale) return fail(S, DPM_ERR_ARG, "null scale");
  for (long s = 0; s < S->nsys; s++)
    if (!(scale[s] > 0)) return fail(S, DPM_ERR_ARG, "scale must be positive");
  const bool round = S->P.geometry == GEOM_ROUND_BOX;
  if (round)
    for (long s = 1; s < S->nsys; s++)
      if (scale[s] != scale[0]) return fail(S, DPM_ERR_ARG, "roundBox: one scale for all systems");
  HIPCHK(launch_shape_only(S));
  if (round) {
    S->P.boxRadius /= scale[0];
  } else {
    for (long s = 0; s < S->nsys; s++) {
      S->h_box[2 * s] /= scale[s];
      S->h_box[2 * s + 1] /= scale[s];
    }
    std::vector<double> bx(S->nsys), by(S->nsys);
    for (long s = 0; s < S->nsys; s++) { bx[s] = S->h_box[2 * s]; by[s] = S->h_box[2 * s + 1]; }
    HIPCHK(up(S->V.bx, bx, S->stream));
    HIPCHK(up(S->V.by, by, S->stream));
    int rc = upload_box(S);
    if (rc) return rc;
  }
  HIPCHK(hipMemcpyAsync(S->d_scale, scale, sizeof(double) * S->nsys, hipMemcpyHostToDevice, S->stream));
  HIPCHK(launch_scale_box(S));
  HIPCHK(hipStreamSynchronize(S->stream));
  S->paramSerial++;
  bump(S);
  if (S->cutSet && !round) return configure_cells(S);
  return DPM_OK;
}

int dpm_time_step_kernel(dpm_set_t h, long reps, double* avgUs, double* meanNeighbors) {
  CHECK_HANDLE(h);
  int rc = require_list(S);
  if (rc) return rc;
  if (reps <= 0) return fail(S, DPM_ERR_ARG, "reps must be positive");
  double *svx = nullptr, *svy = nullptr;
  HIPCHK(dalloc(&svx, S->nv));
  HIPCHK(dalloc(&svy, S->nv));
  HIPCHK(hipMemcpyAsync(svx, S->V.vx, sizeof(double) * S->nv, hipMemcpyDeviceToDevice, S->stream));
  HIPCHK(hipMemcpyAsync(svy, S->V.vy, sizeof(double) * S->nv, hipMemcpyDeviceToDevice, S->stream));
  // the Langevin step kernel as launched inside loops, with rebuild flags that stay clear: no
  // rebuild, and the drift writes into the other position buffer only
  HIPCHK(launch_fill_int(S->mdFlag[3], 0, S->nsys, S->stream));
  for (int r = 0; r < 3; r++) HIPCHK(launch_fill_int(S->mdGen[r], 0, S->nsys, S->stream));
  auto args = [&](long i) {
    StepArgs A = md_args(S, i, 0, 0);
    A.flagCur = S->mdFlag[3];
    A.flagOut = S->d_flag[1 - S->cur];
    A.flagClear = nullptr;
    return A;
  };
  const int mode = MODE_KICK | MODE_LANGEVIN | MODE_START;
  HIPCHK(launch_step(S, mode, args(0)));  // warm caches and code
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, S->stream));
  for (long i = 0; i < reps; i++) HIPCHK(launch_step(S, mode, args(i + 1)));
  HIPCHK(hipEventRecord(e1, S->stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  *avgUs = 1000.0 * ms / reps;
  HIPCHK(hipMemcpyAsync(S->V.vx, svx, sizeof(double) * S->nv, hipMemcpyDeviceToDevice, S->stream));
  HIPCHK(hipMemcpyAsync(S->V.vy, svy, sizeof(double) * S->nv, hipMemcpyDeviceToDevice, S->stream));
  HIPCHK(launch_fill_int(S->d_flag[1 - S->cur], 0, S->nsys, S->stream));
  HIPCHK(launch_fill_int(S->d_any, 0, 2, S->stream));
  std::vector<int> cnt(S->nv);
  HIPCHK(hipMemcpyAsync(cnt.data(), S->V.ncnt, sizeof(int) * S->nv, hipMemcpyDeviceToHost, S->stream));
  HIPCHK(hipStreamSynchronize(S->stream));
  hipFree(svx);
  hipFree(svy);
  double t = 0;
  for (int c : cnt) t += c;
  if (meanNeighbors) *meanNeighbors = t / S->nv;
  return DPM_OK;
}

int dpm_force_hvp(dpm_set_t h, const double* ux, const double* uy, double* outx, double* outy) {
  CHECK_HANDLE(h);
  if (!ux || !uy || !outx || !outy) return fail(S, DPM_ERR_ARG, "dpm_force_hvp: null vector");
  int rc = adjoint_supported(S, true);
  if (rc) return rc;
  rc = adj_alloc(S);
  if (rc) return rc;
  for (int attempt = 0; attempt < 2; attempt++) {
    HIPCHK(launch_fill_int(S->d_err, 0, 1, S->stream));
    rc = clean_protocol(S);
    if (rc) return rc;
    AdjArgs A{};
    A.wx = ux;
    A.wy = uy;
    A.hx = outx;
    A.hy = outy;
    A.mode = 0;
    rc = hvp_launches(S, S->V, 0, A);
    if (rc) return rc;
    rc = clean_protocol(S);
    if (rc) return rc;
    int err = 0;
    HIPCHK(hipMemcpyAsync(&err, S->d_err, sizeof(int), hipMemcpyDeviceToHost, S->stream));
    HIPCHK(hipStreamSynchronize(S->stream));
    if (!err) return DPM_OK;
    std::vector<int> all(S->nsys, 1);  // rows outgrew the list: widen, then evaluate again
    rc = rebuild(S, all, 0);
    if (rc) return rc;
  }
  return fail(S, DPM_ERR_OVERFLOW, "dpm_force_hvp: neighbour list overflow");
}

int dpm_energy_param_grad(dpm_set_t h, double* gring, double* gl0) {
  CHECK_HANDLE(h);
  int rc = adjoint_supported(S, true);
  if (rc) return rc;
  rc = adj_alloc(S);
  if (rc) return rc;
  double* tmp = adj_g(S);
  HIPCHK(launch_energy_param_partials(S, tmp));
  if (gring) HIPCHK(launch_ring_sum3(S, tmp, gring));
  if (gl0) HIPCHK(launch_copy_d(gl0, tmp + 3 * S->nv, S->nv, S->stream));
  HIPCHK(launch_fill_d(tmp, 0.0, 4 * S->nv, S->stream));
  HIPCHK(hipStreamSynchronize(S->stream));
  return DPM_OK;
}

int dpm_energy_param_grad_noalloc(dpm_set_t h, double* gring, double* gl0) {
  CHECK_HANDLE(h);
  if (S->nv > DPM_NOALLOC_MAX_VERTICES)
    return fail(S, DPM_ERR_ARG, "dpm_energy_param_grad_noalloc: the set has " + std::to_string(S->nv) +
                                    " vertices, the scratch buffer holds " +
                                    std::to_string((long)DPM_NOALLOC_MAX_VERTICES));
  int rc = adjoint_supported(S, true);
  if (rc) return rc;
  double* tmp = nullptr;
  HIPCHK(scratch_address(S->device, &tmp));
  // every partial of the set is overwritten before it is read: nothing needs clearing
  HIPCHK(launch_energy_param_partials(S, tmp));
  if (gring) HIPCHK(launch_ring_sum3(S, tmp, gring));
  if (gl0) HIPCHK(launch_copy_d(gl0, tmp + 3 * S->nv, S->nv, S->stream));
  HIPCHK(hipStreamSynchronize(S->stream));
  return DPM_OK;
}

int dpm_nve_forward(dpm_set_t h, long steps, long every) {
  CHECK_HANDLE(h);
  if (steps < 1 || every < 1) return fail(S, DPM_ERR_ARG, "dpm_nve_forward: steps and every must be >= 1");
  int rc = adjoint_supported(S, true);
  if (rc) return rc;
  const long nv = S->nv;
  const long nck = (steps + every - 1) / every;
  const bool rings = S->P.particle == PART_ACTIVE;
  const long words = state_words(S, rings);
  if (nck > S->tapeNck || every > S->tapeEvery || !S->d_fin || words != S->ckWords) {
    HIPCHK(hipStreamSynchronize(S->stream));
    hipFree(S->d_ck);
    hipFree(S->d_seg);
    hipFree(S->d_fin);
    S->d_ck = S->d_seg = S->d_fin = nullptr;
    HIPCHK(dalloc(&S->d_ck, nck * words));
    HIPCHK(dalloc(&S->d_seg, (every + 1) * 2 * nv));
    HIPCHK(dalloc(&S->d_fin, words));
    S->ckWords = words;
    drop_graphs(S);
  }
  S->tapeSteps = 0;
  for (long k = 0; k < nck; k++) {
    const long len = std::min(every, steps - k * every);
    rc = save_state(S, S->d_ck + k * words, rings);
    if (rc) return rc;
    rc = run_verlet(S, len, false);
    if (rc) return rc;
  }
  rc = save_state(S, S->d_fin, rings);
  if (rc) return rc;
  rc = check_err_after_loop(S);
  if (rc) return rc;
  S->tapeSteps = steps;
  S->tapeEvery = every;
  S->tapeNck = nck;
  S->tapeDt = S->P.dt;
  S->tapeP = S->P;
  S->tapeSerial = S->paramSerial;
  return DPM_OK;
}

int dpm_fire_forward(dpm_set_t h, long iterations) {
  CHECK_HANDLE(h);
  if (iterations < 1) return fail(S, DPM_ERR_ARG, "dpm_fire_forward: iterations must be >= 1");
  if (!S->fireInit) return fail(S, DPM_ERR_STATE, "dpm_init_fire not called");
  int rc = adjoint_supported(S);
  if (rc) return rc;
  const long nv = S->nv, ns = S->nsys;
  if (iterations > S->fireTapeCap || !S->d_fadj) {
    HIPCHK(hipStreamSynchronize(S->stream));
    hipFree(S->d_ftape);
    hipFree(S->d_ftapeState);
    hipFree(S->d_fadj);
    S->d_ftape = nullptr;
    S->d_ftapeState = nullptr;
    S->d_fadj = nullptr;
    HIPCHK(dalloc(&S->d_ftape, (iterations + 1) * 6 * nv));
    HIPCHK(dalloc(&S->d_ftapeState, iterations * ns));
    HIPCHK(dalloc(&S->d_fadj, 10 * nv));
    S->fireTapeCap = iterations;
  }
  if (S->fp.tol >= 0) {  // a fixed number of iterations: no system stops
    S->fp.tol = -1;
    bump(S);
  }
  // every iteration one at a time, its starting state recorded before it and its FIRE scalars
  // after its parameter update; replayed on a wider list as the loops are
  rc = with_replay(S, SNAP_FIRE, [&]() -> int {
    for (long k = 0; k < iterations; k++) {
      int r = save_state(S, S->d_ftape + k * 6 * nv);
      if (r) return r;
      r = fire_sequence(S, 1);
      if (r) return r;
      HIPCHK(hipMemcpyAsync(S->d_ftapeState + k * ns, S->d_fire[S->fireCur], sizeof(FireState) * ns,
                            hipMemcpyDeviceToDevice, S->stream));
    }
    return save_state(S, S->d_ftape + iterations * 6 * nv);
  });
  if (rc) return rc;
  S->fireTapeN = iterations;
  S->fireTapeP = S->P;
  S->fireTapeSerial = S->paramSerial;
  return DPM_OK;
}

int dpm_fire_backward(dpm_set_t h, const double* ax, const double* av, double* gx0, double* gv0, double* gring,
                      double* gl0) {
  CHECK_HANDLE(h);
  if (!ax || !av) return fail(S, DPM_ERR_ARG, "dpm_fire_backward: null adjoint");
  if (S->fireTapeN < 1) return fail(S, DPM_ERR_STATE, "dpm_fire_backward: no FIRE trajectory recorded");
  if (memcmp(&S->P, &S->fireTapeP, sizeof(Params)) != 0 || S->paramSerial != S->fireTapeSerial)
    return fail(S, DPM_ERR_STATE, "dpm_fire_backward: parameters changed since dpm_fire_forward");
  int rc = adjoint_supported(S);
  if (rc) return rc;
  rc = adj_alloc(S);
  if (rc) return rc;
  const long nv = S->nv, ns = S->nsys, N = S->fireTapeN;
  double *xa = S->d_fadj, *bv = xa + 2 * nv, *pend = xa + 4 * nv, *fc = xa + 6 * nv, *hw = xa + 8 * nv;
  HIPCHK(launch_copy_d(xa, ax, 2 * nv, S->stream));
  HIPCHK(launch_copy_d(bv, av, 2 * nv, S->stream));
  HIPCHK(launch_fill_d(pend, 0.0, 2 * nv, S->stream));
  HIPCHK(launch_fill_d(adj_g(S), 0.0, 4 * nv, S->stream));
  HIPCHK(launch_fill_int(S->d_err, 0, 1, S->stream));
  rc = clean_protocol(S);
  if (rc) return rc;
  // positions x_N, x_{N-1}, ..., x_0, each with the step loop's row protocol: the displacement
  // test of the next position is made by the kernel of this one
  auto slot = [&](long k) { return S->d_ftape + k * 6 * nv; };
  auto hvp = [&](long n, int p, const double* w) -> int {
    const View Vn = view_at(S, slot(n), slot(n) + nv);
    HIPCHK(launch_bin_view(S, Vn, p));
    AdjArgs A{};
    A.wx = w;
    A.wy = w + nv;
    A.hx = hw;
    A.hy = hw + nv;
    A.g = adj_g(S);
    A.cd = 1.0;
    A.mode = 2;
    A.flag = S->d_flag[p];
    A.count = S->G.count[p];
    A.cellOther = S->G.count[1 - p];
    A.ncells = S->G.ncells;
    if (n > 0) {
      A.px = slot(n - 1);
      A.py = slot(n - 1) + nv;
      A.flagOut = S->d_flag[1 - p];
      A.anyOut = S->d_any + 1 - p;
    }
    HIPCHK(launch_adj_hvp(S, Vn, A));
    return DPM_OK;
  };
  HIPCHK(launch_adj_flag(S, view_at(S, slot(N), slot(N) + nv), 0));
  int p = 0;
  for (long k = N - 1; k >= 0; k--) {
    const FireState* st = S->d_ftapeState + k * ns;
    HIPCHK(launch_fadj_prep(S, st, bv, pend, fc));
    rc = hvp(k + 1, p, fc);
    if (rc) return rc;
    p ^= 1;
    HIPCHK(launch_fadj_partials(S, st, slot(k), xa, bv, hw));
    HIPCHK(launch_fadj_update(S, st, slot(k), xa, bv, pend, hw));
  }
  rc = hvp(0, p, pend);  // F(x_0) enters the first iteration only: x bar -= H(x_0) pend
  if (rc) return rc;
  HIPCHK(launch_axpy(xa, hw, -1.0, 2 * nv, S->stream));
  {
    int err = 0;
    HIPCHK(hipMemcpyAsync(&err, S->d_err, sizeof(int), hipMemcpyDeviceToHost, S->stream));
    HIPCHK(hipStreamSynchronize(S->stream));
    if (err) return fail(S, DPM_ERR_OVERFLOW, "dpm_fire_backward: neighbour list overflow");
  }
  if (gx0) HIPCHK(launch_copy_d(gx0, xa, 2 * nv, S->stream));
  if (gv0) HIPCHK(launch_copy_d(gv0, bv, 2 * nv, S->stream));
  if (gring) HIPCHK(launch_ring_sum3(S, adj_g(S), gring));
  if (gl0) HIPCHK(launch_copy_d(gl0, adj_g(S) + 3 * nv, nv, S->stream));
  // the rows again for the set's own positions (the reverse pass left them at x_0)
  rc = clean_protocol(S);
  if (rc) return rc;
  std::vector<int> all(S->nsys, 1);
  rc = rebuild(S, all, 0);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(S->stream));
  return DPM_OK;
}

int dpm_nve_backward(dpm_set_t h, const double* ax, const double* av, double* gx0, double* gv0, double* gring,
                     double* gl0) {
  CHECK_HANDLE(h);
  if (!ax || !av) return fail(S, DPM_ERR_ARG, "dpm_nve_backward: null adjoint");
  if (S->tapeSteps < 1) return fail(S, DPM_ERR_STATE, "dpm_nve_backward: no trajectory recorded");
  if (memcmp(&S->P, &S->tapeP, sizeof(Params)) != 0 || S->paramSerial != S->tapeSerial)
    return fail(S, DPM_ERR_STATE, "dpm_nve_backward: parameters changed since dpm_nve_forward");
  int rc = adjoint_supported(S, true);
  if (rc) return rc;
  rc = adj_alloc(S);
  if (rc) return rc;
  const long nv = S->nv, N = S->tapeSteps, every = S->tapeEvery;
  const bool rings = S->P.particle == PART_ACTIVE;
  const long words = S->ckWords;
  S->wCur = 0;
  HIPCHK(launch_copy_d(adj_w(S, 0), av, 2 * nv, S->stream));
  HIPCHK(launch_copy_d(adj_ax(S), ax, 2 * nv, S->stream));
  HIPCHK(launch_fill_d(adj_g(S), 0.0, 4 * nv, S->stream));
  HIPCHK(launch_fill_int(S->d_err, 0, 1, S->stream));
  std::vector<int> all(S->nsys, 1);
  for (long k = S->tapeNck - 1; k >= 0; k--) {
    const long a = k * every, b = std::min(a + every, N);
    // restore the checkpoint, recompute the segment (its first step rebuilds every row), then
    // the reverse pass over it; no host synchronisation until the end
    rc = load_state(S, S->d_ck + k * words, rings);
    if (rc) return rc;
    rc = clean_protocol(S);
    if (rc) return rc;
    rc = record_segment(S, b - a);
    if (rc) return rc;
    rc = clean_protocol(S);
    if (rc) return rc;
    rc = backward_segment(S, a, (k == S->tapeNck - 1) ? b : b - 1, N);
    if (rc) return rc;
  }
  {
    int err = 0;
    HIPCHK(hipMemcpyAsync(&err, S->d_err, sizeof(int), hipMemcpyDeviceToHost, S->stream));
    HIPCHK(hipStreamSynchronize(S->stream));
    if (err) return fail(S, DPM_ERR_OVERFLOW, "dpm_nve_backward: neighbour list overflow");
  }
  if (gx0) HIPCHK(launch_copy_d(gx0, adj_ax(S), 2 * nv, S->stream));
  if (gv0) HIPCHK(launch_copy_d(gv0, adj_w(S, S->wCur), 2 * nv, S->stream));
  if (gring) HIPCHK(launch_ring_sum3(S, adj_g(S), gring));
  if (gl0) HIPCHK(launch_copy_d(gl0, adj_g(S) + 3 * nv, nv, S->stream));
  // leave the set in the state the forward pass ended in
  rc = load_state(S, S->d_fin, rings);
  if (rc) return rc;
  rc = rebuild(S, all, 0);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(S->stream));
  return DPM_OK;
}

// ------------------------------------------------------------- rigid rings ---
// calcRigidForceEnergy, DPM2D.cu:2128-2142
int dpm_calc_rigid_force_energy(dpm_set_t h) {
  CHECK_HANDLE(h);
  int rc = rigid_supported(S);
  if (rc) return rc;
  StepArgs A{};
  A.ctr = S->d_ctr;
  A.err = S->d_err;
  A.manage = 0;
  HIPCHK(launch_step(S, MODE_RIGID, A));
  HIPCHK(hipStreamSynchronize(S->stream));
  return DPM_OK;
}

// initRigidLangevin / initRigidNVE, DPM2D.cu:3823-3868
static int rigid_init(dpm_set_s* S, double T, double gamma, bool langevin, int readState) {
  if (S->nwide > 0) return fail(S, DPM_ERR_UNSUPPORTED, "rigid rings: rings of at most 64 vertices only");
  S->rigidKind = langevin ? 1 : 2;
  S->Tinject = T;
  if (langevin) S->P.gamma = gamma;  // the rigid kick adds unit draws: no noise amplitude
  bump(S);
  int rc = rigid_clear(S);
  if (rc) return rc;
  HIPCHK(launch_copy_lastpos(S));
  rc = store_initial_positions(S);
  if (rc) return rc;
  if (!readState) HIPCHK(launch_rigid_inject(S, sqrt(T), S->injectCalls++));
  HIPCHK(hipStreamSynchronize(S->stream));
  return DPM_OK;
}

int dpm_init_rigid_langevin(dpm_set_t h, double T, double gamma, int readState) {
  CHECK_HANDLE(h);
  if (!(gamma > 0)) return fail(S, DPM_ERR_ARG, "gamma must be positive");
  return rigid_init(S, T, gamma, true, readState);
}

int dpm_init_rigid_nve(dpm_set_t h, double T, int readState) {
  CHECK_HANDLE(h);
  return rigid_init(S, T, 0.0, false, readState);
}

int dpm_rigid_langevin_loop(dpm_set_t h, long steps) {
  CHECK_HANDLE(h);
  if (S->rigidKind != 1) return fail(S, DPM_ERR_STATE, "dpm_init_rigid_langevin not called");
  return run_rigid(S, steps, true);
}

int dpm_rigid_nve_loop(dpm_set_t h, long steps) {
  CHECK_HANDLE(h);
  if (S->rigidKind != 2) return fail(S, DPM_ERR_STATE, "dpm_init_rigid_nve not called");
  return run_rigid(S, steps, false);
}

// initRigidFIRE, DPM2D.cu:3700-3706 (FIRE parameters as dpm_init_fire's)
int dpm_init_rigid_fire(dpm_set_t h, const double* p, long minStep, long numStep) {
  CHECK_HANDLE(h);
  if (S->nwide > 0) return fail(S, DPM_ERR_UNSUPPORTED, "rigid rings: rings of at most 64 vertices only");
  if (!p) return fail(S, DPM_ERR_ARG, "null FIRE params");
  S->fp.a_start = p[0]; S->fp.f_dec = p[1]; S->fp.f_inc = p[2]; S->fp.f_a = p[3];
  S->fp.dt_max = p[5];
  S->fp.minStep = minStep;
  S->fp.tol = -1;
  S->fp.maxIter = 0;
  std::vector<FireState> fs(S->nsys);
  for (auto& f : fs) {
    f.dt = p[4]; f.a = p[6]; f.P = 0; f.numStep = numStep; f.iters = 0; f.converged = 0; f.pad = 0;
  }
  for (int b = 0; b < 2; b++)
    HIPCHK(hipMemcpyAsync(S->d_rfire[b], fs.data(), sizeof(FireState) * S->nsys, hipMemcpyHostToDevice, S->stream));
  S->rfireCur = 0;
  int rc = rigid_clear(S);
  if (rc) return rc;
  S->rigidFireInit = true;
  bump(S);
  HIPCHK(hipStreamSynchronize(S->stream));
  return DPM_OK;
}

// rigidFIRELoop, DPM2D.cu:3708-3710 (FIRE::minimizerRigidLoop, FIRE.cu:279-298), `iterations` times
int dpm_rigid_fire_loop(dpm_set_t h, long iterations) {
  CHECK_HANDLE(h);
  if (!S->rigidFireInit) return fail(S, DPM_ERR_STATE, "dpm_init_rigid_fire not called");
  int rc = rigid_supported(S);
  if (rc) return rc;
  if (iterations <= 0) return DPM_OK;
  return with_replay(S, SNAP_RIGID_FIRE, [&]() -> int {
    if (iterations < 4 || !graphs_enabled()) return rigid_fire_sequence(S, iterations);
    // the sequence flips the position and FIRE-state buffers once per iteration
    int r = run_graph(S, std::make_tuple(22 + S->rfireCur, iterations, S->cur),
                      [&]() { return rigid_fire_sequence(S, iterations); });
    if (!r && (iterations % 2) == 1) {
      swap_pos(S);
      S->rfireCur = 1 - S->rfireCur;
    }
    return r;
  });
}

// FIRE scalars of the rigid minimiser per system: dt, a, numStep, last power, iterations
int dpm_get_rigid_fire_state(dpm_set_t h, double* out5) {
  CHECK_HANDLE(h);
  if (!out5) return fail(S, DPM_ERR_ARG, "null output");
  std::vector<FireState> fs(S->nsys);
  HIPCHK(hipMemcpyAsync(fs.data(), S->d_rfire[S->rfireCur], sizeof(FireState) * S->nsys, hipMemcpyDeviceToHost,
                        S->stream));
  HIPCHK(hipStreamSynchronize(S->stream));
  for (long s = 0; s < S->nsys; s++) {
    out5[5 * s] = fs[s].dt;
    out5[5 * s + 1] = fs[s].a;
    out5[5 * s + 2] = (double)fs[s].numStep;
    out5[5 * s + 3] = fs[s].P;
    out5[5 * s + 4] = (double)fs[s].iters;
  }
  return DPM_OK;
}

static int ring_fields(dpm_set_s* S, std::initializer_list<std::pair<const double*, std::vector<double>*>> f) {
  for (auto& kv : f) {
    kv.second->resize(S->np);
    HIPCHK(hipMemcpyAsync(kv.second->data(), kv.first, sizeof(double) * S->np, hipMemcpyDeviceToHost, S->stream));
  }
  HIPCHK(hipStreamSynchronize(S->stream));
  return DPM_OK;
}

// getRigidMaxUnbalancedForce, DPM2D.cu:3643-3652, per system
int dpm_get_rigid_max_unbalanced_force(dpm_set_t h, double* perSystem) {
  CHECK_HANDLE(h);
  if (!perSystem) return fail(S, DPM_ERR_ARG, "null output");
  std::vector<double> fx, fy, tq;
  int rc = ring_fields(S, {{S->R.pfx, &fx}, {S->R.pfy, &fy}, {S->R.ptq, &tq}});
  if (rc) return rc;
  for (long s = 0; s < S->nsys; s++) {
    double mf = -1, mt = -1;
    for (int r = S->h_sroff[s]; r < S->h_sroff[s + 1]; r++) {
      mf = std::max(mf, std::max(fx[r] * fx[r], fy[r] * fy[r]));
      mt = std::max(mt, tq[r] * tq[r]);
    }
    perSystem[s] = std::max(sqrt(mf), sqrt(mt));
  }
  return DPM_OK;
}

// getRigidKineticEnergy, DPM2D.cu:1149-1159, per system
int dpm_get_rigid_kinetic_energy(dpm_set_t h, double* perSystem) {
  CHECK_HANDLE(h);
  if (!perSystem) return fail(S, DPM_ERR_ARG, "null output");
  std::vector<double> vx, vy, w, rad;
  int rc = ring_fields(S, {{S->R.pvx, &vx}, {S->R.pvy, &vy}, {S->R.pw, &w}, {S->R.prad, &rad}});
  if (rc) return rc;
  for (long s = 0; s < S->nsys; s++) {
    double k = 0, r2 = 0;
    for (int r = S->h_sroff[s]; r < S->h_sroff[s + 1]; r++) k += vx[r] * vx[r] + vy[r] * vy[r];
    for (int r = S->h_sroff[s]; r < S->h_sroff[s + 1]; r++) r2 += (w[r] * rad[r]) * (w[r] * rad[r]);
    perSystem[s] = 0.5 * k + 0.5 * r2;
  }
  return DPM_OK;
}

}  // extern "C"
