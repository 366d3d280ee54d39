// dpm_internal.h -- host-side handle of a system set and the launcher interface
// between the C ABI (dpm_api.hip) and the kernels (dpm_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "dpm_device.h"

namespace dpm {

// step-kernel epilogue bits
// (MODE_BAOAB: the next step's first half is B A O A; MODE_BROWN: overdamped update, with
// thermal noise when MODE_LANGEVIN is also set; MODE_RIGID: rigid rings, whose ring force and
// torque drive ring velocities, with the rigid Langevin kick when MODE_LANGEVIN is also set and
// the per-system FIRE step when MODE_FIRE is)
enum : int { MODE_FORCE = 0, MODE_KICK = 1, MODE_LANGEVIN = 2, MODE_START = 4, MODE_FIRE = 8, MODE_BAOAB = 16,
             MODE_BROWN = 32, MODE_RIGID = 64,
             MODE_WIDE = 128 };  // the launch serves a set holding rings of more than 64 vertices

// noise types, simControlStruct::noiseEnum order (include/DPM2D.h:27)
enum : int { NOISE_LANGEVIN = 0, NOISE_BAOAB = 1, NOISE_BROWNIAN = 2, NOISE_DRIVEN_BROWNIAN = 3 };

// error bits written by kernels into Set::d_err
enum : int { ERR_LIST_OVERFLOW = 1, ERR_CELL_OVERFLOW = 2, ERR_STALE_BINS = 4 };

// rebuild-flag bits (per system): 1 = rebuild, FLAG_FAR = a vertex moved more than cutoff / 3 in
// the drift that produced these positions, or (first positions of a step loop) lies more than
// 2 cutoff / 3 from its list position: the cells of the list positions may miss its neighbours
enum : int { FLAG_FAR = 2 };

// per-system FIRE scalars (double buffered by iteration parity)
struct FireState {
  double dt, a, P;
  long long numStep, iters;
  int converged, pad;
};

struct FireParams {
  double a_start, f_dec, f_inc, f_a, dt_max;
  long long minStep;
  double tol;          // < 0: no convergence test (every system advances)
  long long maxIter;   // per-system iteration cap when tol >= 0
};

// Per-system cell lists.  Counts are double buffered by position-buffer parity: the
// step that consumes count[b] clears count[1 - b] for the next step's binning.
struct CellGrid {
  int* count[2];   // [ncells] each
  int* slots;      // [ncells * cap]
  int* cellSys;    // [ncells]
  int* ncx;        // per system
  int* ncy;
  int* coff;       // per-system cell offset
  int* brute;      // per system: 1 = all-pairs within the system (box < 3 cells)
  int ncells, cap;
};

struct StepArgs {
  double h;                       // kick time step (0.5 dt)
  double dt;                      // drift time step
  double* xout;                   // drift destination (MODE_START)
  double* yout;
  double* recx;                   // optional second copy of the drift destination (adjoint tape)
  double* recy;
  int* flagCur;                   // rebuild flags of this step's positions
  int* flagOut;                   // rebuild flags of the drift destination
  int* anyOut;                    // any-flag of the drift destination
  int* cellCur;                   // cells the rows are rebuilt from (this step's positions) and
  int* slotsCur;                  //   their slots; in the step loop the base of its three lists
  int* cellOther;                 // cell counts cleared here (the next binning's target), or null
  int* flagClear;                 // rebuild flags cleared here, or null
  int* genCur;                    // step loop: per-system index of the list-position cells (the
  int* genOut;                    //   cells of the positions of the last rebuild), and its next value
  int staleCheck;                 // rows from cells of earlier positions: check the displacement
  const unsigned long long* ctr;  // [0] thermal calls, [1] active calls
  long long ctrOffT, ctrOffA;     // call index of this step's draws = ctr[k] + offset
  FireState* fire;                // MODE_FIRE: per-system state of this iteration
  double* part;                   // MODE_FIRE: per-tile partial sums [ntiles * 4]
  int* err;
  int* maxcnt;
  int* updateCount;
  int manage;                     // inside loops: rebuild flagged rows, clear cells, count updates
  int rigidLangevin;              // MODE_RIGID: the RigidLangevin kick (else RigidNVE's)
  int xcd;                        // XCD-aware workgroup -> tile mapping
  int rowsBuilt;                  // inside loops: flagged rows already rebuilt (k_rows_loop)
  int rowsLdsW;                   // set by the launcher: list width the launch's LDS holds rows of
};

// Up to kMaxCopySegs device-to-device copies in one launch (k_copy_segs): 4-byte words.
constexpr int kMaxCopySegs = 32;
struct CopySegs {
  const unsigned* src[kMaxCopySegs];
  unsigned* dst[kMaxCopySegs];
  long words[kMaxCopySegs];
  int n;
};

// Arguments of the H w kernel (dpm_adjoint.hip).
struct AdjArgs {
  const int* flag;      // rebuild flags of this parity (rows rebuilt first when set), or null
  const int* count;     // cell counts of this parity
  int* err;
  int* maxcnt;
  const double* wx;     // the vector w
  const double* wy;
  double* hx;           // mode 0: H w
  double* hy;
  double* ax;           // mode 1: adjoint of positions, updated in place
  double* ay;
  double* nwx;          // mode 1: next vector (w + dt ax when update, else w)
  double* nwy;
  double* g;            // mode 1: per-vertex parameter partials [ka | kl | kb | l0] x nv, accumulated
  double cd;            // c dt
  double dt;
  const double* px;     // positions processed next: their Verlet-list validity goes to flagOut/anyOut
  const double* py;     //   (null: no test)
  int* flagOut;
  int* anyOut;
  int* cellOther;       // cell counts of the other parity, cleared here (null: left alone)
  int ncells;
  int mode;
  int update;
  int xcd;              // XCD-aware workgroup -> tile mapping
  int rowsLdsW;         // set by the launcher: list width the launch's LDS holds rows of
};

}  // namespace dpm

struct dpm_set_s {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t ownStream = nullptr;  // created with the handle; graph capture needs a non-NULL stream
  long nsys = 0, np = 0, nv = 0;
  std::string err;
  int split = 4;  // wavefronts per tile in the step kernel (1..8)
  int xcdMap = 1; // XCD-aware workgroup -> tile mapping in the step kernel
  double cellScale = 1.0;  // cell side / (list radius + 2 cutoff / 3), >= 1 (dpm_set_cell_scale)

  // host copies of topology
  std::vector<int> h_rnv, h_rfirst, h_rsys, h_vring, h_vsys, h_svoff, h_sroff;
  std::vector<int> h_tile_v0, h_tile_v1, h_tile_sys, h_stile;  // h_stile: per-system tile offsets
  // rings of more than 64 vertices, each cut into tiles of 64-vertex slices
  std::vector<int> h_wide;
  long nwide = 0;
  int* d_wide = nullptr;
  std::vector<double> h_box;    // 2 per system
  std::vector<double> h_maxrad; // per system

  dpm::View V{};
  dpm::Rigid R{};                 // rigid-ring arrays (host copy of the pointers)
  dpm::Rigid* d_rg = nullptr;     // their device copy, V.rg
  dpm::Params P{};
  double* pos[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [buffer][component]
  int cur = 0;

  // neighbour machinery
  dpm::CellGrid G{};
  int* d_flag[2] = {nullptr, nullptr};  // per-system rebuild flags, by position parity
  // Step-loop cells: three cell lists in one allocation (mdCount[k] = mdCount[0] + k ncells,
  // mdSlots[k] likewise), holding per system the cells of its list positions (the positions
  // of its last rebuild) in list gen % 3; a rebuild reads list gen % 3, bins its positions into
  // list gen + 1 and clears list gen + 2.  Flags and generations rotate over three buffers by
  // step; mdFlag[3] is a sink for launches whose flags are not wanted.
  int* mdCount[3] = {nullptr, nullptr, nullptr};
  int* mdSlots[3] = {nullptr, nullptr, nullptr};
  int* mdFlag[4] = {nullptr, nullptr, nullptr, nullptr};
  int* mdGen[3] = {nullptr, nullptr, nullptr};
  int* d_any = nullptr;                 // [2] any-flag by parity
  int* d_updateCount = nullptr;
  int* d_err = nullptr;
  int* d_maxcnt = nullptr;
  unsigned long long* d_ctr = nullptr;  // thermal, active
  bool listBuilt = false;
  bool cutSet = false;

  // integrators
  int simKind = 0;  // 1 langevin, 2 nve, 3 nve rescale
  int rigidKind = 0;  // rigid rings: 1 langevin, 2 nve
  dpm::FireState* d_rfire[2] = {nullptr, nullptr};  // rigid FIRE state per system, by iteration parity
  int rfireCur = 0;
  bool rigidFireInit = false;
  int noiseType = 0;  // integrator of the Langevin loop (NOISE_*)
  // observables: contact rows (np x contactWidth, -1 padded) and counts, initial positions of
  // vertices and ring centres, per-ring / per-vertex scratch
  int *d_contacts = nullptr, *d_ccount = nullptr;
  int contactWidth = 0;
  double *d_x0 = nullptr, *d_obs = nullptr;  // d_x0: x0, y0 (nv each), cx0, cy0 (np each)
  long obsLen = 0;
  double Tinject = 0;
  unsigned long long injectCalls = 0;

  // FIRE
  dpm::FireState* d_fire[2] = {nullptr, nullptr};
  int fireCur = 0;
  dpm::FireParams fp{};
  double* d_part = nullptr;  // per-tile partials [ntiles * 4]
  bool fireInit = false;
  int fireKernel = 1;   // dpm_fire_minimize: 1 = one persistent workgroup per system, 0 = graph of step kernels
  int fireWaves = 0;    // waves per system workgroup (0 = from the largest system's tile count)
  int* d_notDone = nullptr;
  int* d_sysOrder = nullptr;  // persistent FIRE: workgroup b runs system d_sysOrder[b] (most tiles first)

  // reductions
  double* d_red = nullptr;   // per-system outputs [nsys * 8]
  double* d_scale = nullptr; // per-system scalars
  std::vector<double> h_red;

  // adjoint tape (dpm_nve_forward / dpm_nve_backward)
  long tapeSteps = 0, tapeEvery = 0, tapeNck = 0;
  double tapeDt = 0;
  dpm::Params tapeP{};        // parameters the trajectory was recorded with
  long tapeSerial = -1;
  long paramSerial = 0;       // bumped by set_ring_costs and by set_field of rad, l0, theta0, a0, box
  double* d_ck = nullptr;     // checkpoints: tapeNck x ckWords (x, y, vx, vy, fx, fy; self-propelled
                              //   rings: then ring angles and RNG call counters)
  long ckWords = 0;           // doubles per checkpoint of d_ck and d_fin
  double* d_fin = nullptr;    // final state, ckWords doubles
  double* d_seg = nullptr;    // one segment's positions: (tapeEvery + 1) x 2 x nv
  double* d_adj = nullptr;    // work: w[2] (2nv each), ax (2nv), partials (4nv), ring area and dA (2np)
  long adjNv = 0;
  int adjCur = 0, wCur = 0;

  // FIRE tape (dpm_fire_forward / dpm_fire_backward): slots 0..fireTapeN of x, y, vx, vy, fx, fy
  // (6 nv doubles each) and the per-system FIRE state of every iteration
  double* d_ftape = nullptr;
  dpm::FireState* d_ftapeState = nullptr;
  double* d_fadj = nullptr;   // work: ax, bv, pend, Fc, H Fc (2 nv each)
  long fireTapeN = 0, fireTapeCap = 0, fireTapeSerial = -1;
  dpm::Params fireTapeP{};

  // copy of what a loop changes, restored to replay the loop after a list overflow
  unsigned* d_snap = nullptr;
  long snapWords = 0;
  int snapCur = 0, snapFireCur = 0, snapRfireCur = 0;

  // graph cache: key = (kind, steps, cur)
  std::map<std::tuple<int, long, int>, hipGraphExec_t> graphs;
  long version = 0;  // bumped when launch arguments change
  long graphVersion = -1;
};

namespace dpm {

// launchers (dpm_kernels.hip); all return hipError_t of the last launch
hipError_t launch_shape_only(dpm_set_s* S);
hipError_t launch_step(dpm_set_s* S, int mode, const StepArgs& A);
hipError_t launch_start_baoab(dpm_set_s* S, double h, const StepArgs& A);
hipError_t launch_start_md(dpm_set_s* S, double h, double dt, const StepArgs& A);
hipError_t launch_bin_flag(dpm_set_s* S, int* count, int* slots, int* flag, int* genOut);
hipError_t launch_ctr_add(dpm_set_s* S, long long addT, long long addA);
hipError_t launch_bin(dpm_set_s* S, int bumpThermal, int bumpActive);
hipError_t launch_build_rows(dpm_set_s* S, int countUpdates);
hipError_t launch_fill_int(int* p, int value, long n, hipStream_t st);
hipError_t launch_reduce(dpm_set_s* S);
hipError_t launch_inject(dpm_set_s* S, double amplitude, unsigned long long call);
hipError_t launch_sub_mean(dpm_set_s* S);
hipError_t launch_update_pos(dpm_set_s* S, double dt);
hipError_t launch_update_vel(dpm_set_s* S, double dt);
hipError_t launch_thermal_force(dpm_set_s* S);
hipError_t launch_scale_vel(dpm_set_s* S);
hipError_t launch_scale_vertices(dpm_set_s* S);
hipError_t launch_scale_box(dpm_set_s* S);
hipError_t launch_copy_lastpos(dpm_set_s* S);
hipError_t launch_fire_begin(dpm_set_s* S);
hipError_t launch_fire_mid(dpm_set_s* S);
hipError_t launch_fire_partials(dpm_set_s* S);
hipError_t launch_fire_step(dpm_set_s* S);
hipError_t launch_fire_done(dpm_set_s* S, int* notDone);
hipError_t launch_fire_resume(dpm_set_s* S);
StepArgs loop_args(dpm_set_s* S);
hipError_t launch_bin_view(dpm_set_s* S, const View& Vn, int parity);
hipError_t launch_fire_sys(dpm_set_s* S, int chunk, int waves);
hipError_t launch_adj_flag(dpm_set_s* S, const View& Vn, int parity);
hipError_t launch_adj_seg_check(dpm_set_s* S, const double* seg, int slots, int* bad);
hipError_t launch_adj_hvp(dpm_set_s* S, const View& Vn, AdjArgs A);
hipError_t launch_energy_param_partials(dpm_set_s* S, double* out);
hipError_t launch_fadj_prep(dpm_set_s* S, const FireState* st, const double* bv, const double* pend, double* fc);
hipError_t launch_fadj_partials(dpm_set_s* S, const FireState* st, const double* slot, const double* ax,
                                const double* bv, const double* hw);
hipError_t launch_fadj_update(dpm_set_s* S, const FireState* st, const double* slot, double* ax, double* bv,
                              double* pend, const double* hw);
hipError_t launch_ring_sum3(dpm_set_s* S, const double* in, double* out);
hipError_t launch_fill_d(double* p, double value, long n, hipStream_t st);
hipError_t launch_contacts(dpm_set_s* S, int* list, int* count, int width, int* overflow, double gap = -1.0);
hipError_t launch_hexatic(dpm_set_s* S, const int* list, const int* count, int width, double* psi6);
hipError_t launch_displacement(dpm_set_s* S, const double* x, const double* y, const double* x0, const double* y0,
                               long n, int rings, double waveNum, double* d2, double* sf);
hipError_t launch_particle_stress(dpm_set_s* S, double* out);
hipError_t launch_copy_d(double* dst, const double* src, long n, hipStream_t st);
hipError_t launch_copy_segs(const CopySegs& segs, hipStream_t st);
hipError_t launch_axpy(double* y, const double* x, double alpha, long n, hipStream_t st);
hipError_t launch_build_rows_view(dpm_set_s* S, const View& Vn, int parity, int countUpdates);
hipError_t launch_rigid_start(dpm_set_s* S, bool langevin, const StepArgs& A);
hipError_t launch_rigid_inject(dpm_set_s* S, double amplitude, unsigned long long call);
hipError_t launch_rigid_fire_partials(dpm_set_s* S);
hipError_t launch_rigid_fire_move(dpm_set_s* S, int* flagOut, int* anyOut);

}  // namespace dpm
