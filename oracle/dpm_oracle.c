/*
 * dpm_oracle.c -- CPU oracle for the DPM2D vertex engine (TEST INFRASTRUCTURE).
 *
 * A host restatement of the reference's arithmetic, written for checking the
 * HIP library.  It follows the reference's expression order so that results
 * agree to rounding; where the reference sums with atomics (smooth interaction)
 * the oracle fixes the sequential order of the reference's thread loop.
 * Citations are to files of farceri/cudaRing.
 */
#include "dpm_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static const double PI = 3.141592653589793238462643383279502884197; /* defs.h:14 */
static const double WCAcut = 1.122462048;                            /* defs.h:17 */

enum { TAG_THERMAL = 1, TAG_INJECT = 2, TAG_ACTIVE = 3, TAG_SHAPE = 4 };

struct orc_state {
  long numParticles, numVertices;
  long *numVertexInParticleList, *firstVertexInParticleId, *particleIdList;
  double boxSize[2];
  int potentialType, interactionType, monomerType, particleType;
  double dt, ea, el, eb, ec, l1, l2, LJcutoff, LJecut, LJfshift, stiff, extSq;
  double driving, taup;
  double *pos, *vel, *force, *energy, *lastPos;
  double *rad, *l0, *theta0, *theta, *length;
  double *a0, *area, *perimeter, *particlePos, *particleEnergy, *particleAngle;
  double* kr; /* per-ring [ka | kl | kb] or NULL */
  long *neighborList, neighborListSize, *maxNeighborList, maxNeighbors;
  double cutDistance, cutoff;
  long updateCount;
  int shift;
  int simKind;
  double Tinject, gamma, noise;
  int noiseType;            /* simControlStruct::noiseEnum order */
  int geometryType;         /* simControlStruct::geometryEnum order */
  double ew, boxRadius;     /* wall energy scale, round-box radius */
  double* wallForce;        /* [numVertices * 2], rewritten by each force call with walls */
  long *contactList, contactListSize, *maxContactList; /* per particle, -1 padded */
  double *initialPos, *particleInitPos;                /* MSD / ISF references */
  double l0Tau, l0Diff;     /* activeShape (setActiveTensionParams) */
  double* l0Rest;
  double a_start, a, f_dec, f_inc, f_a, fire_dt_max, fire_dt, lastP;
  long minStep, numStep;
  uint64_t seed;
  uint32_t systemId;
  uint64_t thermalCalls, injectCalls, activeCalls;
  /* rigid rings (initRotationalVariables / initDeltaVariables, DPM2D.cu:197-221) */
  double *particleVel, *particleAngvel, *particleForce, *particleTorque, *momentOfInertia;
  double *particleLastPos, *particleLastAngle, *torque, *particleRad;
  int rigidKind;            /* 1 Langevin, 2 NVE */
  int listMethod;           /* 1: cell-list row build (default), 0: the reference's all-pairs loop */
  int threads;              /* OpenMP threads of the per-vertex loops (1: serial, the test default) */
};

/* ------------------------------------------------------------------ RNG --- */
/* Philox4x32-10: 10 rounds of the 4x32 Philox S-box with 9 key bumps
 * (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3"). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static double u01_from(uint32_t hi, uint32_t lo) {
  uint64_t v = ((uint64_t)hi << 32) | lo;
  return ((double)(v >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

void orc_gauss_pair(uint64_t seed, uint32_t systemId, uint32_t tag, uint64_t counter,
                    uint32_t index, double* z0, double* z1) {
  uint32_t ctr[4] = {index, (uint32_t)counter, (uint32_t)(counter >> 32),
                     (tag << 24) | (systemId & 0xFFFFFFu)};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t out[4];
  orc_philox4x32_10(ctr, key, out);
  double u1 = u01_from(out[0], out[1]);
  double u2 = u01_from(out[2], out[3]);
  double r = sqrt(-2.0 * log(u1));
  double ang = 2.0 * PI * u2;
  *z0 = r * cos(ang);
  *z1 = r * sin(ang);
}

/* ------------------------------------------------------------ lifecycle --- */
orc_state* orc_create(long numParticles, const long* nvp) {
  orc_state* s = (orc_state*)calloc(1, sizeof(orc_state));
  s->numParticles = numParticles;
  s->numVertexInParticleList = (long*)malloc(sizeof(long) * numParticles);
  s->firstVertexInParticleId = (long*)malloc(sizeof(long) * numParticles);
  long nv = 0;
  for (long p = 0; p < numParticles; p++) {
    s->numVertexInParticleList[p] = nvp[p];
    s->firstVertexInParticleId[p] = nv; /* initParticleIdList, DPM2D.cu:282-300 */
    nv += nvp[p];
  }
  s->numVertices = nv;
  s->particleIdList = (long*)malloc(sizeof(long) * nv);
  for (long p = 0; p < numParticles; p++)
    for (long k = 0; k < nvp[p]; k++) s->particleIdList[s->firstVertexInParticleId[p] + k] = p;
  long n2 = 2 * nv, np = numParticles;
  s->pos = calloc(n2, 8); s->vel = calloc(n2, 8); s->force = calloc(n2, 8);
  s->lastPos = calloc(n2, 8); s->energy = calloc(nv, 8);
  s->rad = calloc(nv, 8); s->l0 = calloc(nv, 8); s->theta0 = calloc(nv, 8);
  s->theta = calloc(nv, 8); s->length = calloc(nv, 8);
  s->a0 = calloc(np, 8); s->area = calloc(np, 8); s->perimeter = calloc(np, 8);
  s->particlePos = calloc(2 * np, 8); s->particleEnergy = calloc(np, 8);
  s->particleAngle = calloc(np, 8);
  s->maxNeighborList = (long*)calloc(nv, sizeof(long));
  s->neighborList = NULL;
  s->neighborListSize = 0;
  /* defaults of the DPM2D constructor, DPM2D.cu:41-94 */
  s->boxSize[0] = s->boxSize[1] = 1.0;
  s->dt = 1e-3; s->ea = 1e5; s->el = 20; s->eb = 10; s->ec = 1;
  s->cutDistance = 1; s->updateCount = 0; s->shift = 0;
  s->gamma = 1; s->noise = 0;
  s->ew = 1e3; /* DPM2D.cu:70 */
  s->listMethod = 1;
  s->threads = 1;
  s->wallForce = calloc(n2, 8);
  s->contactListSize = 0;
  s->contactList = NULL;
  s->maxContactList = (long*)calloc(np, sizeof(long));
  s->initialPos = calloc(n2, 8);
  s->particleInitPos = calloc(2 * np, 8);
  s->particleVel = calloc(2 * np, 8); s->particleAngvel = calloc(np, 8);
  s->particleForce = calloc(2 * np, 8); s->particleTorque = calloc(np, 8);
  s->momentOfInertia = calloc(np, 8); s->particleLastPos = calloc(2 * np, 8);
  s->particleLastAngle = calloc(np, 8); s->torque = calloc(nv, 8); s->particleRad = calloc(np, 8);
  return s;
}

void orc_destroy(orc_state* s) {
  if (!s) return;
  free(s->contactList); free(s->maxContactList); free(s->initialPos); free(s->particleInitPos);
  free(s->wallForce);
  free(s->l0Rest);
  free(s->particleVel); free(s->particleAngvel); free(s->particleForce); free(s->particleTorque);
  free(s->momentOfInertia); free(s->particleLastPos); free(s->particleLastAngle); free(s->torque);
  free(s->particleRad);
  free(s->numVertexInParticleList); free(s->firstVertexInParticleId); free(s->particleIdList);
  free(s->pos); free(s->vel); free(s->force); free(s->lastPos); free(s->energy);
  free(s->rad); free(s->l0); free(s->theta0); free(s->theta); free(s->length);
  free(s->a0); free(s->area); free(s->perimeter); free(s->particlePos);
  free(s->particleEnergy); free(s->particleAngle); free(s->maxNeighborList);
  free(s->neighborList);
  free(s->kr);
  free(s);
}

long orc_num_vertices(const orc_state* s) { return s->numVertices; }

static double* field_ptr(const orc_state* s, int f) {
  switch (f) {
    case ORC_F_POS: return s->pos; case ORC_F_VEL: return s->vel;
    case ORC_F_FORCE: return s->force; case ORC_F_ENERGY: return s->energy;
    case ORC_F_RAD: return s->rad; case ORC_F_L0: return s->l0;
    case ORC_F_THETA0: return s->theta0; case ORC_F_THETA: return s->theta;
    case ORC_F_LENGTH: return s->length; case ORC_F_A0: return s->a0;
    case ORC_F_AREA: return s->area; case ORC_F_PERIMETER: return s->perimeter;
    case ORC_F_PPOS: return s->particlePos; case ORC_F_PENERGY: return s->particleEnergy;
    case ORC_F_PANGLE: return s->particleAngle; case ORC_F_LASTPOS: return s->lastPos;
    case ORC_F_BOX: return (double*)s->boxSize;
    case ORC_F_WALLFORCE: return s->wallForce;
    case ORC_F_PVEL: return s->particleVel; case ORC_F_PANGVEL: return s->particleAngvel;
    case ORC_F_PFORCE: return s->particleForce; case ORC_F_PTORQUE: return s->particleTorque;
    case ORC_F_VTORQUE: return s->torque; case ORC_F_PRAD: return s->particleRad;
  }
  return NULL;
}

long orc_field_size(const orc_state* s, int f) {
  switch (f) {
    case ORC_F_POS: case ORC_F_VEL: case ORC_F_FORCE: case ORC_F_LASTPOS: case ORC_F_WALLFORCE:
      return 2 * s->numVertices;
    case ORC_F_ENERGY: case ORC_F_RAD: case ORC_F_L0: case ORC_F_THETA0: case ORC_F_THETA:
    case ORC_F_LENGTH: case ORC_F_VTORQUE: return s->numVertices;
    case ORC_F_A0: case ORC_F_AREA: case ORC_F_PERIMETER: case ORC_F_PENERGY:
    case ORC_F_PANGLE: case ORC_F_PANGVEL: case ORC_F_PTORQUE: case ORC_F_PRAD: return s->numParticles;
    case ORC_F_PPOS: case ORC_F_PVEL: case ORC_F_PFORCE: return 2 * s->numParticles;
    case ORC_F_BOX: return 2;
  }
  return -1;
}

int orc_set_field(orc_state* s, int f, const double* data, long n) {
  double* p = field_ptr(s, f);
  if (!p || n != orc_field_size(s, f)) return -1;
  memcpy(p, data, sizeof(double) * n);
  return 0;
}

int orc_get_field(const orc_state* s, int f, double* data, long n) {
  double* p = field_ptr(s, f);
  if (!p || n != orc_field_size(s, f)) return -1;
  memcpy(data, p, sizeof(double) * n);
  return 0;
}

/* ----------------------------------------------------------- parameters --- */
void orc_set_types(orc_state* s, int pot, int inter, int mono, int part) {
  s->potentialType = pot; s->interactionType = inter; s->monomerType = mono; s->particleType = part;
}
/* per-ring shape costs (an addition); NULL ka clears them */
void orc_set_ring_costs(orc_state* s, const double* ka, const double* kl, const double* kb) {
  long np = s->numParticles;
  if (!ka) { free(s->kr); s->kr = NULL; return; }
  if (!s->kr) s->kr = (double*)malloc(sizeof(double) * 3 * np);
  for (long p = 0; p < np; p++) { s->kr[p] = ka[p]; s->kr[np + p] = kl[p]; s->kr[2 * np + p] = kb[p]; }
}
/* setEnergyCosts, DPM2D.cu:1875-1884 */
void orc_set_energy_costs(orc_state* s, double ea, double el, double eb, double ec) {
  s->ea = ea; s->el = el; s->eb = eb; s->ec = ec;
}
/* setAttractionConstants, DPM2D.cu:1915-1920 */
void orc_set_attraction_constants(orc_state* s, double l1, double l2) { s->l1 = l1; s->l2 = l2; }
/* setLJcutoff, DPM2D.cu:1922-1932 (uses the ec in force at call time) */
void orc_set_lj_cutoff(orc_state* s, double LJcutoff) {
  s->LJcutoff = LJcutoff;
  double ratio6 = 1 / pow(LJcutoff, 6);
  s->LJecut = 4 * s->ec * (ratio6 * ratio6 - ratio6);
  s->LJfshift = 24 * s->ec * (2 * ratio6 - 1) * ratio6 / LJcutoff;
}
/* setFENEconstants, DPM2D.cu:1934-1939 */
void orc_set_fene_constants(orc_state* s, double stiff, double ext) { s->stiff = stiff; s->extSq = ext * ext; }
void orc_set_time_step(orc_state* s, double dt) { s->dt = dt; }
void orc_set_self_propulsion(orc_state* s, double driving, double taup) { s->driving = driving; s->taup = taup; }
void orc_set_rng(orc_state* s, uint64_t seed, uint32_t systemId) {
  s->seed = seed; s->systemId = systemId;
  s->thermalCalls = s->injectCalls = s->activeCalls = 0;
}

/* ------------------------------------------------------------- geometry --- */
/* pbcDistance, cudaKernel.cuh:70-74: round() is half-away-from-zero */
double orc_pbc_distance(double x1, double x2, double size) {
  double delta = x1 - x2;
  return delta - size * round(delta / size);
}

/* getNextId, cudaKernel.cuh:220-230 */
long orc_next_id(const orc_state* s, long vertexId, long particleId) {
  if (vertexId == s->numVertices - 1) return vertexId - s->numVertexInParticleList[particleId] + 1;
  long nextId = vertexId + 1;
  if (s->particleIdList[nextId] == particleId) return nextId;
  return nextId - s->numVertexInParticleList[particleId];
}

/* getPreviousId, cudaKernel.cuh:233-243 */
long orc_previous_id(const orc_state* s, long vertexId, long particleId) {
  if (vertexId == 0) return s->numVertexInParticleList[particleId] - 1;
  long previousId = vertexId - 1;
  if (s->particleIdList[previousId] == particleId) return previousId;
  return previousId + s->numVertexInParticleList[particleId];
}

/* boxDistance, cudaKernel.cuh:76-94: periodic image only along the periodic directions
 * (both for normal, x for fixedSides, none for fixedBox and roundBox) */
static inline double pbcd(const orc_state* s, double x1, double x2, int dim) {
  if (s->geometryType == ORC_GEOM_NORMAL || (s->geometryType == ORC_GEOM_FIXED_SIDES && dim == 0))
    return orc_pbc_distance(x1, x2, s->boxSize[dim]);
  return x1 - x2;
}

void orc_set_geometry(orc_state* s, int geometryType) { s->geometryType = geometryType; }
void orc_set_wall_energy_scale(orc_state* s, double ew) { s->ew = ew; }
void orc_set_box_radius(orc_state* s, double boxRadius) { s->boxRadius = boxRadius; }
double orc_get_box_radius(orc_state* s) { return s->boxRadius; }

/* calcAngle, cudaKernel.cuh:423-427 */
static double calc_angle(const double* n, const double* p) {
  double midSine = n[0] * p[1] - n[1] * p[0];
  double midCosine = n[0] * p[0] + n[1] * p[1];
  return atan2(midSine, midCosine);
}

/* kernelCalcParticleShape, cudaKernel.cuh:1571-1606 (one ring walked serially) */
void orc_calc_particle_shape(orc_state* s) {
#pragma omp parallel for num_threads(s->threads) if (s->threads > 1)
  for (long pId = 0; pId < s->numParticles; pId++) {
    double tempArea = 0.0;
    s->perimeter[pId] = 0.0;
    long firstId = s->firstVertexInParticleId[pId];
    double delta[2], nextPos[2], currentPos[2], nextSegment[2], previousPos[2], previousSegment[2];
    currentPos[0] = s->pos[firstId * 2]; currentPos[1] = s->pos[firstId * 2 + 1];
    for (long cur = firstId; cur < firstId + s->numVertexInParticleList[pId]; cur++) {
      long nextId = orc_next_id(s, cur, pId);
      long previousId = orc_previous_id(s, cur, pId);
      double segmentLength = 0.0;
      for (int dim = 0; dim < 2; dim++) {
        delta[dim] = pbcd(s, s->pos[nextId * 2 + dim], currentPos[dim], dim);
        nextPos[dim] = currentPos[dim] + delta[dim];
        segmentLength += delta[dim] * delta[dim];
        delta[dim] = pbcd(s, s->pos[previousId * 2 + dim], currentPos[dim], dim);
        previousPos[dim] = currentPos[dim] + delta[dim];
      }
      for (int dim = 0; dim < 2; dim++) {
        nextSegment[dim] = nextPos[dim] - currentPos[dim];
        previousSegment[dim] = currentPos[dim] - previousPos[dim];
      }
      s->theta[cur] = calc_angle(nextSegment, previousSegment);
      s->length[cur] = sqrt(segmentLength);
      s->perimeter[pId] += s->length[cur];
      tempArea += currentPos[0] * nextPos[1] - nextPos[0] * currentPos[1];
      currentPos[0] = nextPos[0]; currentPos[1] = nextPos[1];
    }
    s->area[pId] = fabs(tempArea) * 0.5;
  }
}

/* calcParticlePos, cudaKernel.cuh:365-388, launched by kernelCalcParticlePositions 1608-1613 */
void orc_calc_particle_positions(orc_state* s) {
#pragma omp parallel for num_threads(s->threads) if (s->threads > 1)
  for (long pId = 0; pId < s->numParticles; pId++) {
    double currentPos[2], delta[2], partPos[2];
    long firstId = s->firstVertexInParticleId[pId];
    currentPos[0] = s->pos[firstId * 2]; currentPos[1] = s->pos[firstId * 2 + 1];
    partPos[0] = currentPos[0]; partPos[1] = currentPos[1];
    for (long cur = firstId; cur < firstId + s->numVertexInParticleList[pId] - 1; cur++) {
      long nextId = orc_next_id(s, cur, pId);
      for (int dim = 0; dim < 2; dim++) {
        delta[dim] = pbcd(s, s->pos[nextId * 2 + dim], currentPos[dim], dim);
        currentPos[dim] += delta[dim];
        partPos[dim] += currentPos[dim];
      }
    }
    for (int dim = 0; dim < 2; dim++) s->particlePos[pId * 2 + dim] = partPos[dim] / s->numVertexInParticleList[pId];
  }
}

/* ---------------------------------------------------------- shape force --- */
/* calcAreaForceEnergy, cudaKernel.cuh:429-435 */
static double area_force_energy(double ea, double pA0, double pA, const double* nPos,
                                const double* pPos, double* f) {
  double deltaA = (pA / pA0) - 1.;
  double gradMultiple = ea * deltaA / pA0;
  f[0] += 0.5 * gradMultiple * (pPos[1] - nPos[1]);
  f[1] += 0.5 * gradMultiple * (nPos[0] - pPos[0]);
  return (0.5 * ea * deltaA * deltaA);
}

/* calcPerimeterForceEnergy, cudaKernel.cuh:437-447 */
static double perimeter_force_energy(double el, double tL0, double pL0, double nLength,
                                     double pLength, const double* vPos, const double* nPos,
                                     const double* pPos, double* f) {
  double pDelta = (pLength / pL0) - 1.;
  double nDelta = (nLength / tL0) - 1.;
  for (int dim = 0; dim < 2; dim++)
    f[dim] += el * ((nDelta * (nPos[dim] - vPos[dim]) / (tL0 * nLength)) -
                    (pDelta * (vPos[dim] - pPos[dim]) / (pL0 * pLength)));
  return (0.5 * el * pDelta * pDelta);
}

/* calcFENEPerimeterForceEnergy, cudaKernel.cuh:449-455 (energy sign kept as in the reference) */
static double fene_force_energy(const orc_state* s, double el, double tL0, double pL0, double nLength,
                                double pLength, const double* vPos, const double* nPos,
                                const double* pPos, double* f) {
  for (int dim = 0; dim < 2; dim++)
    f[dim] += 2 * el * s->stiff * s->extSq *
              ((nLength / (s->extSq * tL0 * tL0 - nLength * nLength)) * (nPos[dim] - vPos[dim]) / nLength -
               (pLength / (s->extSq * pL0 * pL0 - pLength * pLength)) * (vPos[dim] - pPos[dim]) / pLength);
  return el * s->stiff * s->extSq * log(1 - pLength * pLength / (s->extSq * pL0 * pL0));
}

/* calcBendingForceEnergy, cudaKernel.cuh:457-471 */
static double bending_force_energy(double eb, const double* preSeg, const double* nextSeg,
                                   double thisD, double nextD, double preD, double* f) {
  double preN[2] = {preSeg[1], -preSeg[0]};
  double nextN[2] = {nextSeg[1], -nextSeg[0]};
  double preVar = (thisD - preD) / (preSeg[0] * preSeg[0] + preSeg[1] * preSeg[1]);
  double nextVar = (thisD - nextD) / (nextSeg[0] * nextSeg[0] + nextSeg[1] * nextSeg[1]);
  for (int dim = 0; dim < 2; dim++) f[dim] += eb * (preVar * preN[dim] + nextVar * nextN[dim]);
  return (0.5 * eb * thisD * thisD);
}

/* kernelCalcShapeForceEnergy, cudaKernel.cuh:474-529: zeroes F and overwrites energy */
void orc_calc_shape_force_energy(orc_state* s) {
#pragma omp parallel for num_threads(s->threads) if (s->threads > 1)
  for (long v = 0; v < s->numVertices; v++) {
    long p = s->particleIdList[v];
    double vertexPos[2], nextPos[2], previousPos[2], partPos[2], nextSegment[2], previousSegment[2];
    double particleArea = s->area[p];
    double shapeEnergy = 0.0;
    /* global costs, or per-ring ones (an addition: the reference's are device-global scalars) */
    const double ea = s->kr ? s->kr[p] : s->ea;
    const double el = s->kr ? s->kr[s->numParticles + p] : s->el;
    const double eb = s->kr ? s->kr[2 * s->numParticles + p] : s->eb;
    long nextId = orc_next_id(s, v, p);
    long previousId = orc_previous_id(s, v, p);
    double* f = &s->force[v * 2];
    for (int dim = 0; dim < 2; dim++) {
      f[dim] = 0.0;
      partPos[dim] = s->particlePos[p * 2 + dim];
    }
    for (int dim = 0; dim < 2; dim++) { /* getRelativeVertexPos, cudaKernel.cuh:273-279 */
      vertexPos[dim] = pbcd(s, s->pos[v * 2 + dim], partPos[dim], dim);
      nextPos[dim] = pbcd(s, s->pos[nextId * 2 + dim], partPos[dim], dim);
      previousPos[dim] = pbcd(s, s->pos[previousId * 2 + dim], partPos[dim], dim);
    }
    shapeEnergy += area_force_energy(ea, s->a0[p], particleArea, nextPos, previousPos, f) /
                   s->numVertexInParticleList[p];
    for (int dim = 0; dim < 2; dim++) {
      nextSegment[dim] = nextPos[dim] - vertexPos[dim];
      previousSegment[dim] = vertexPos[dim] - previousPos[dim];
    }
    double previousLength = sqrt(previousSegment[0] * previousSegment[0] + previousSegment[1] * previousSegment[1]);
    double nextLength = sqrt(nextSegment[0] * nextSegment[0] + nextSegment[1] * nextSegment[1]);
    if (s->monomerType == ORC_MONO_HARMONIC)
      shapeEnergy += perimeter_force_energy(el, s->l0[v], s->l0[previousId], nextLength, previousLength,
                                            vertexPos, nextPos, previousPos, f);
    else
      shapeEnergy += fene_force_energy(s, el, s->l0[v], s->l0[previousId], nextLength, previousLength,
                                       vertexPos, nextPos, previousPos, f);
    double previousAngleDelta = s->theta[previousId] - s->theta0[previousId];
    double thisAngleDelta = s->theta[v] - s->theta0[v];
    double nextAngleDelta = s->theta[nextId] - s->theta0[nextId];
    shapeEnergy += bending_force_energy(eb, previousSegment, nextSegment, thisAngleDelta, nextAngleDelta,
                                        previousAngleDelta, f);
    s->energy[v] = shapeEnergy;
  }
}

/* --------------------------------------------------------- pair forces --- */
/* calcDeltaAndDistance (normal geometry), cudaKernel.cuh:160-171 */
static double delta_and_distance(const orc_state* s, const double* a, const double* b, double* d) {
  double distanceSq = 0.0;
  for (int dim = 0; dim < 2; dim++) {
    d[dim] = pbcd(s, a[dim], b[dim], dim);
    distanceSq += d[dim] * d[dim];
  }
  return sqrt(distanceSq);
}

static double calc_distance(const orc_state* s, const double* a, const double* b) {
  double d[2];
  return delta_and_distance(s, a, b, d);
}

/* Directed pair force on `this` from `other` and the directed-visit energy:
 * calcContactInteraction 531-544, calcLJInteraction 562-583,
 * calcAdhesiveInteraction 604-626, calcWCAInteraction 653-668. */
static double pair_interaction(const orc_state* s, const double* thisPos, const double* otherPos,
                               double radSum, double* f) {
  double delta[2];
  double distance = delta_and_distance(s, thisPos, otherPos, delta);
  switch (s->potentialType) {
    case ORC_HARMONIC: {
      double overlap = 1 - distance / radSum;
      if (overlap > 0) {
        double gradMultiple = s->ec * overlap / radSum;
        for (int dim = 0; dim < 2; dim++) f[dim] += gradMultiple * delta[dim] / distance;
        return (0.5 * s->ec * overlap * overlap) * 0.5;
      }
      return 0.0;
    }
    case ORC_LJ: {
      double ratio = radSum / distance;
      double ratio6 = pow(ratio, 6);
      double ratio12 = ratio6 * ratio6;
      if (distance <= (s->LJcutoff * radSum)) {
        double forceShift = s->LJfshift / radSum;
        double gradMultiple = 24 * s->ec * (2 * ratio12 - ratio6) / distance - forceShift;
        for (int dim = 0; dim < 2; dim++) f[dim] += gradMultiple * delta[dim] / distance;
        return 0.5 * (4 * s->ec * (ratio12 - ratio6) - s->LJecut -
                      fabs(forceShift) * (distance - s->LJcutoff * radSum));
      }
      return 0.0;
    }
    case ORC_ADHESIVE: {
      double gradMultiple = 0.0, epot = 0.0;
      double overlap = 1 - distance / radSum;
      if (distance < (1 + s->l2) * radSum) {
        if (distance < (1 + s->l1) * radSum) {
          gradMultiple = s->ec * overlap / radSum;
          epot = 0.5 * s->ec * (overlap * overlap - s->l1 * s->l2) * 0.5;
        } else {
          gradMultiple = -(s->ec * s->l1 / (s->l2 - s->l1)) * (overlap + s->l2) / radSum;
          epot = -(0.5 * (s->ec * s->l1 / (s->l2 - s->l1)) * (overlap + s->l2) * (overlap + s->l2)) * 0.5;
        }
        for (int dim = 0; dim < 2; dim++) f[dim] += gradMultiple * delta[dim] / distance;
        return epot;
      }
      return 0.0;
    }
    case ORC_WCA: {
      double ratio = radSum / distance;
      double ratio6 = pow(ratio, 6);
      double ratio12 = ratio6 * ratio6;
      if (distance <= (WCAcut * radSum)) {
        double gradMultiple = 24 * s->ec * (2 * ratio12 - ratio6) / distance;
        for (int dim = 0; dim < 2; dim++) f[dim] += gradMultiple * delta[dim] / distance;
        return 0.5 * s->ec * (4 * (ratio12 - ratio6) + 1);
      }
      return 0.0;
    }
  }
  return 0.0;
}

/* calcGradMultipleAndEnergy, cudaKernel.cuh:806-858 (scalar form used by the smooth kernel) */
double orc_pair_grad_multiple_and_energy(const orc_state* s, double distance, double radSum, double* epot) {
  double overlap, ratio, ratio6, ratio12;
  switch (s->potentialType) {
    case ORC_HARMONIC:
      overlap = 1 - distance / radSum;
      if (overlap > 0) {
        *epot = 0.5 * s->ec * overlap * overlap * 0.5;
        return s->ec * overlap / radSum;
      }
      return 0;
    case ORC_LJ:
      ratio = radSum / distance;
      ratio12 = pow(ratio, 12);
      ratio6 = pow(ratio, 6);
      if (distance <= (s->LJcutoff * radSum)) {
        double forceShift = s->LJfshift / radSum;
        *epot = 0.5 * (4 * s->ec * (ratio12 - ratio6) - s->LJecut - fabs(forceShift) * (distance - s->LJcutoff * radSum));
        return 24 * s->ec * (2 * ratio12 - ratio6) / distance - forceShift;
      }
      return 0;
    case ORC_ADHESIVE:
      overlap = 1 - distance / radSum;
      if (distance < (1 + s->l1) * radSum) {
        *epot = 0.5 * s->ec * (overlap * overlap - s->l1 * s->l2) * 0.5;
        return s->ec * overlap / radSum;
      } else if ((distance >= (1 + s->l1) * radSum) && (distance < (1 + s->l2) * radSum)) {
        *epot = -(0.5 * (s->ec * s->l1 / (s->l2 - s->l1)) * (overlap + s->l2) * (overlap + s->l2)) * 0.5;
        return -(s->ec * s->l1 / (s->l2 - s->l1)) * (overlap + s->l2) / radSum;
      }
      return 0;
    case ORC_WCA:
      ratio = radSum / distance;
      ratio6 = pow(ratio, 6);
      ratio12 = ratio6 * ratio6;
      if (distance <= (WCAcut * radSum)) {
        *epot = 0.5 * s->ec * (4 * (ratio12 - ratio6) + 1);
        return 4 * s->ec * (12 * ratio12 - 6 * ratio6) / distance;
      }
      return 0;
  }
  return 0;
}

/* kernelCalcVertexInteraction, cudaKernel.cuh:690-722 (gather over the Verlet row) */
void orc_calc_vertex_interaction(orc_state* s) {
  long W = s->neighborListSize;
#pragma omp parallel for num_threads(s->threads) if (s->threads > 1)
  for (long v = 0; v < s->numVertices; v++) {
    double thisPos[2] = {s->pos[v * 2], s->pos[v * 2 + 1]};
    double thisRad = s->rad[v];
    for (long k = 0; k < s->maxNeighborList[v]; k++) {
      long otherId = s->neighborList[v * W + k];
      if ((v != otherId) && (otherId != -1)) {
        double otherPos[2] = {s->pos[otherId * 2], s->pos[otherId * 2 + 1]};
        double radSum = thisRad + s->rad[otherId];
        s->energy[v] += pair_interaction(s, thisPos, otherPos, radSum, &s->force[v * 2]);
      }
    }
  }
}

/* kernelCalcAllToAllVertexInteraction, cudaKernel.cuh:725-757 */
void orc_calc_all_to_all_interaction(orc_state* s) {
#pragma omp parallel for num_threads(s->threads) if (s->threads > 1)
  for (long v = 0; v < s->numVertices; v++) {
    double thisPos[2] = {s->pos[v * 2], s->pos[v * 2 + 1]};
    double thisRad = s->rad[v];
    for (long o = 0; o < s->numVertices; o++) {
      if ((s->particleIdList[v] != s->particleIdList[o]) && (o != -1)) {
        double otherPos[2] = {s->pos[o * 2], s->pos[o * 2 + 1]};
        double radSum = thisRad + s->rad[o];
        s->energy[v] += pair_interaction(s, thisPos, otherPos, radSum, &s->force[v * 2]);
      }
    }
  }
}

/* getProjection, cudaKernel.cuh:861-864 */
static double projection(const orc_state* s, const double* thisPos, const double* otherPos,
                         const double* previousPos, double length) {
  return (pbcd(s, thisPos[0], previousPos[0], 0) * pbcd(s, otherPos[0], previousPos[0], 0) +
          pbcd(s, thisPos[1], previousPos[1], 1) * pbcd(s, otherPos[1], previousPos[1], 1)) /
         (length * length);
}

/* calcCross, cudaKernel.cuh:873-875 */
static double cross_of(const orc_state* s, const double* thisPos, const double* otherPos,
                       const double* previousPos) {
  return pbcd(s, previousPos[0], otherPos[0], 0) * pbcd(s, otherPos[1], thisPos[1], 1) -
         pbcd(s, otherPos[0], thisPos[0], 0) * pbcd(s, previousPos[1], otherPos[1], 1);
}

/* a += of the scattered smooth interaction: atomic when the vertex loop runs on several threads
 * (the reference's atomicAdd), plain in the serial order otherwise */
static inline void acc(const orc_state* s, double* p, double v) {
  if (s->threads > 1) {
#pragma omp atomic
    *p += v;
  } else {
    *p += v;
  }
}

/* calcVertexSegmentInteraction, cudaKernel.cuh:877-897.  The reference divides by |cross|;
 * a vertex exactly on the segment line gives 0/0 there, the oracle takes sign 0. */
static double vertex_segment(orc_state* s, const double* thisPos, const double* projPos,
                             const double* otherPos, const double* previousPos, double length,
                             double radSum, double* tF, double* oF, double* pF) {
  double epot = 0.0;
  double gm = orc_pair_grad_multiple_and_energy(s, calc_distance(s, thisPos, projPos), radSum, &epot);
  if (gm != 0) {
    double cross = cross_of(s, thisPos, otherPos, previousPos);
    double absCross = fabs(cross);
    double sign = absCross > 0 ? cross / absCross : 0.0;
    double L2 = length * length;
    acc(s, &tF[0], gm * sign * pbcd(s, previousPos[1], otherPos[1], 1) / length);
    acc(s, &tF[1], gm * sign * pbcd(s, otherPos[0], previousPos[0], 0) / length);
    acc(s, &oF[0], gm * (sign * pbcd(s, thisPos[1], previousPos[1], 1) + absCross * pbcd(s, previousPos[0], otherPos[0], 0) / L2) / length);
    acc(s, &oF[1], gm * (sign * pbcd(s, previousPos[0], thisPos[0], 0) + absCross * pbcd(s, previousPos[1], otherPos[1], 1) / L2) / length);
    acc(s, &pF[0], gm * (sign * pbcd(s, otherPos[1], thisPos[1], 1) - absCross * pbcd(s, previousPos[0], otherPos[0], 0) / L2) / length);
    acc(s, &pF[1], gm * (sign * pbcd(s, thisPos[0], otherPos[0], 0) - absCross * pbcd(s, previousPos[1], otherPos[1], 1) / L2) / length);
    return epot;
  }
  return 0.;
}

/* calcVertexVertexInteraction, cudaKernel.cuh:899-912 */
static double vertex_vertex(orc_state* s, const double* thisPos, const double* previousPos, double radSum,
                            double* thisF, double* otherF) {
  double delta[2], epot = 0.0;
  double gm = orc_pair_grad_multiple_and_energy(s, calc_distance(s, thisPos, previousPos), radSum, &epot);
  if (gm != 0) {
    double distance = delta_and_distance(s, thisPos, previousPos, delta);
    for (int dim = 0; dim < 2; dim++) {
      acc(s, &thisF[dim], gm * delta[dim] / distance);
      acc(s, &otherF[dim], -gm * delta[dim] / distance);
    }
    return epot;
  }
  return 0.;
}

/* kernelCalcSmoothInteraction, cudaKernel.cuh:922-975.  The reference scatters with
 * atomicAdd; the oracle applies the same updates in vertex order, then list order. */
void orc_calc_smooth_interaction(orc_state* s) {
  long W = s->neighborListSize;
#pragma omp parallel for schedule(dynamic, 64) num_threads(s->threads) if (s->threads > 1)
  for (long v = 0; v < s->numVertices; v++) {
    double thisPos[2] = {s->pos[v * 2], s->pos[v * 2 + 1]};
    double thisRad = s->rad[v];
    long particleId = s->particleIdList[v];
    for (long k = 0; k < s->maxNeighborList[v]; k++) {
      long otherId = s->neighborList[v * W + k];
      if (!((v != otherId) && (otherId != -1))) continue;
      double otherPos[2] = {s->pos[otherId * 2], s->pos[otherId * 2 + 1]};
      double radSum = thisRad + s->rad[otherId];
      long otherParticleId = s->particleIdList[otherId];
      long previousId = orc_previous_id(s, otherId, otherParticleId);
      double previousPos[2] = {s->pos[previousId * 2], s->pos[previousId * 2 + 1]};
      double segment[2] = {pbcd(s, otherPos[0], previousPos[0], 0), pbcd(s, otherPos[1], previousPos[1], 1)};
      double length = sqrt(segment[0] * segment[0] + segment[1] * segment[1]);
      double proj = projection(s, thisPos, otherPos, previousPos, length);
      long secondPreviousId = orc_previous_id(s, previousId, otherParticleId);
      double secondPreviousPos[2] = {s->pos[secondPreviousId * 2], s->pos[secondPreviousId * 2 + 1]};
      double previousSegment[2] = {pbcd(s, secondPreviousPos[0], previousPos[0], 0),
                                   pbcd(s, secondPreviousPos[1], previousPos[1], 1)};
      double previousLength = sqrt(previousSegment[0] * previousSegment[0] + previousSegment[1] * previousSegment[1]);
      double previousProj = projection(s, thisPos, previousPos, secondPreviousPos, previousLength);
      int isPreviousVertexSegment = 0;
      if (proj >= 0 && proj < 1) {
        double reduced = fmax(0.0, fmin(1.0, proj)); /* getProjectionPos, 866-871 */
        double projPos[2] = {previousPos[0] + reduced * segment[0], previousPos[1] + reduced * segment[1]};
        double inter = vertex_segment(s, thisPos, projPos, otherPos, previousPos, length, radSum,
                                      &s->force[v * 2], &s->force[otherId * 2], &s->force[previousId * 2]);
        acc(s, &s->particleEnergy[particleId], inter);
        acc(s, &s->particleEnergy[otherParticleId], inter);
        if (previousProj >= 0 && previousProj < 1) {
          isPreviousVertexSegment = 1;
          inter = vertex_vertex(s, previousPos, thisPos, radSum, &s->force[v * 2], &s->force[previousId * 2]);
          acc(s, &s->particleEnergy[particleId], -inter);
          acc(s, &s->particleEnergy[otherParticleId], -inter);
        }
      } else if (proj < 0) {
        if (previousProj >= 1 && isPreviousVertexSegment == 0) {
          double inter = vertex_vertex(s, thisPos, previousPos, radSum, &s->force[v * 2], &s->force[previousId * 2]);
          acc(s, &s->particleEnergy[particleId], inter);
          acc(s, &s->particleEnergy[otherParticleId], inter);
        }
      }
    }
  }
}

/* addSelfPropulsion, DPM2D.cu:1995-2035 (one Gaussian per ring per call) */
void orc_add_self_propulsion(orc_state* s) {
  double amplitude = 0.;
  if (s->taup != 0) amplitude = sqrt(2.0 * s->dt / s->taup);
  uint64_t call = s->activeCalls++;
  for (long p = 0; p < s->numParticles; p++) {
    double z0, z1;
    orc_gauss_pair(s->seed, s->systemId, TAG_ACTIVE, call, (uint32_t)p, &z0, &z1);
    double a = s->particleAngle[p];
    a += amplitude * z0;
    a = a + PI;
    a = a - 2.0 * PI * floor(a / (2.0 * PI));
    a = a - PI;
    s->particleAngle[p] = a;
  }
  for (long v = 0; v < s->numVertices; v++) {
    long p = s->particleIdList[v];
    for (int dim = 0; dim < 2; dim++)
      s->force[v * 2 + dim] += s->driving * ((1 - dim) * cos(s->particleAngle[p]) + dim * sin(s->particleAngle[p]));
  }
}

/* setActiveTensionParams, DPM2D.cu:1898-1903: the rest lengths become the current l0 */
void orc_set_active_tension(orc_state* s, double l0Tau, double l0Diff) {
  s->l0Tau = l0Tau;
  s->l0Diff = l0Diff;
  if (!s->l0Rest) s->l0Rest = (double*)malloc(sizeof(double) * s->numVertices);
  memcpy(s->l0Rest, s->l0, sizeof(double) * s->numVertices);
}

/* updateRestLength, DPM2D.cu:2037-2056 (one Gaussian per ring per force call):
 * l0 += ((l0 - l0Rest) / l0Tau + sqrt(2 l0Diff / dt) xi_ring) dt */
void orc_update_rest_length(orc_state* s) {
  uint64_t call = s->activeCalls++;
  double noise = sqrt(2. * s->l0Diff / s->dt);
  for (long v = 0; v < s->numVertices; v++) {
    long p = s->particleIdList[v];
    double z0, z1;
    orc_gauss_pair(s->seed, s->systemId, TAG_SHAPE, call, (uint32_t)p, &z0, &z1);
    s->l0[v] += ((s->l0[v] - s->l0Rest[v]) / s->l0Tau + noise * z0) * s->dt;
  }
}

/* calcWallWCAInteraction, cudaKernel.cuh:1221-1237: WCA of the vertex (radius = its own) with
 * its image on the wall; the wall takes the opposite force */
static double wall_wca(orc_state* s, long v, double wx, double wy) {
  double d[2] = {pbcd(s, s->pos[2 * v], wx, 0), pbcd(s, s->pos[2 * v + 1], wy, 1)};
  double radSum = s->rad[v];
  double distance = sqrt(d[0] * d[0] + d[1] * d[1]);
  double ratio = radSum / distance;
  double ratio6 = pow(ratio, 6);
  double ratio12 = ratio6 * ratio6;
  if (distance < (WCAcut * radSum)) {
    double gradMultiple = 24 * s->ew * (2 * ratio12 - ratio6) / distance;
    for (int dim = 0; dim < 2; dim++) {
      s->force[2 * v + dim] += gradMultiple * d[dim] / distance;
      s->wallForce[2 * v + dim] -= gradMultiple * d[dim] / distance;
    }
    return s->ew * (4 * (ratio12 - ratio6) + 1);
  }
  return 0.0;
}

/* kernelCalcVertexSquareWallInteraction, kernelCalcVertexSidesInteraction2D and
 * kernelCalcVertexRoundWallInteraction, cudaKernel.cuh:1324-1410 (the round box is centred on
 * the origin) */
void orc_calc_wall_forces(orc_state* s) {
  const double* L = s->boxSize;
  for (long v = 0; v < s->numVertices; v++) {
    double x = s->pos[2 * v], y = s->pos[2 * v + 1], rc = WCAcut * s->rad[v];
    s->wallForce[2 * v] = s->wallForce[2 * v + 1] = 0.;
    if (s->geometryType == ORC_GEOM_FIXED_BOX) {
      if (x < rc) s->energy[v] += wall_wca(s, v, 0, y);
      else if ((L[0] - x) < rc) s->energy[v] += wall_wca(s, v, L[0], y);
    }
    if (s->geometryType == ORC_GEOM_FIXED_BOX || s->geometryType == ORC_GEOM_FIXED_SIDES) {
      if (y < rc) s->energy[v] += wall_wca(s, v, x, 0);
      else if ((L[1] - y) < rc) s->energy[v] += wall_wca(s, v, x, L[1]);
    }
    if (s->geometryType == ORC_GEOM_ROUND_BOX) {
      double r = sqrt(x * x + y * y), theta = atan2(y, x);
      if ((s->boxRadius - r) < rc)
        s->energy[v] += wall_wca(s, v, s->boxRadius * cos(theta), s->boxRadius * sin(theta));
    }
  }
}

/* calcForceEnergyGPU, DPM2D.cu:2058-2104 with calcShapeForceEnergy 1976-1993 */
void orc_calc_force_energy(orc_state* s) {
  orc_calc_particle_shape(s);
  orc_calc_particle_positions(s);
  orc_calc_shape_force_energy(s);
  memset(s->particleEnergy, 0, sizeof(double) * s->numParticles);
  switch (s->interactionType) {
    case ORC_VERTEX_VERTEX: orc_calc_vertex_interaction(s); break;
    case ORC_VERTEX_SMOOTH: orc_calc_smooth_interaction(s); break;
    case ORC_ALL: orc_calc_all_to_all_interaction(s); break;
    default: orc_calc_smooth_interaction(s); break;
  }
  if (s->particleType == ORC_ACTIVE) orc_add_self_propulsion(s);
  if (s->particleType == ORC_ACTIVE_SHAPE) orc_update_rest_length(s);
  if (s->geometryType != ORC_GEOM_NORMAL) orc_calc_wall_forces(s);
}

/* ----------------------------------------------------------- neighbours --- */
/* setDisplacementCutoff, DPM2D.cu:1300-1329 */
double orc_set_displacement_cutoff(orc_state* s, double cutoff_, double size_) {
  switch (s->potentialType) {
    case ORC_HARMONIC: s->cutDistance = 1; break;
    case ORC_LJ: s->cutDistance = s->LJcutoff; break;
    case ORC_WCA: s->cutDistance = WCAcut; break;
    case ORC_ADHESIVE: s->cutDistance = 1 + s->l2; break;
  }
  s->cutDistance += cutoff_;
  s->cutoff = cutoff_ * size_;
  return s->cutDistance;
}

double orc_get_cut_distance(const orc_state* s) { return s->cutDistance; }

/* kernelCalcNeighborList, cudaKernel.cuh:1735-1763: ascending otherId; a non-neighbour
 * writes -1 into the next free slot (later overwritten), so the row holds the
 * neighbours followed by -1. */
static void build_list(orc_state* s, double cutDistance) {
  long W = s->neighborListSize;
  for (long v = 0; v < s->numVertices; v++) {
    long added = 0;
    double thisPos[2] = {s->pos[v * 2], s->pos[v * 2 + 1]};
    double thisRad = s->rad[v];
    long particleId = s->particleIdList[v];
    for (long o = 0; o < s->numVertices; o++) {
      if (s->particleIdList[o] == particleId) continue;
      double otherPos[2] = {s->pos[o * 2], s->pos[o * 2 + 1]};
      double radSum = thisRad + s->rad[o];
      int isNeighbor = (-(1 - calc_distance(s, thisPos, otherPos) / radSum) < cutDistance);
      if (added < W) s->neighborList[v * W + added] = isNeighbor ? o : -1;
      added += isNeighbor;
    }
    s->maxNeighborList[v] = added;
  }
}

/* The same rows from a cell list (the host cell-list concept of DPM2D.cu:3011-3026, 3558-3579):
 * vertices binned on wrapped coordinates into cells at least as wide as the largest list radius
 * (1 + cutDistance) * 2 * max radius, so every neighbour of a vertex lies in its 3 x 3 block; the
 * candidates that pass the reference's predicate are sorted, and the row keeps the smallest W ids
 * as the all-pairs loop does.  Returns 0 (nothing built) when a box side holds fewer than three
 * cells or the geometry is walled; the caller then runs the all-pairs loop. */
static int build_list_cells(orc_state* s, double cutDistance) {
  if (s->geometryType != ORC_GEOM_NORMAL) return 0;
  double maxRad = 0;
  for (long v = 0; v < s->numVertices; v++) if (s->rad[v] > maxRad) maxRad = s->rad[v];
  double rc = (1 + cutDistance) * 2 * maxRad;
  if (!(rc > 0)) return 0;
  long ncx = (long)floor(s->boxSize[0] / rc), ncy = (long)floor(s->boxSize[1] / rc);
  if (ncx < 3 || ncy < 3) return 0;
  if (ncx > 4096) ncx = 4096;
  if (ncy > 4096) ncy = 4096;
  long nv = s->numVertices, nc = ncx * ncy;
  long* cell = (long*)malloc(sizeof(long) * nv);
  long* start = (long*)calloc(nc + 1, sizeof(long));
  long* item = (long*)malloc(sizeof(long) * nv);
  for (long v = 0; v < nv; v++) {
    double wx = s->pos[2 * v] - s->boxSize[0] * floor(s->pos[2 * v] / s->boxSize[0]);
    double wy = s->pos[2 * v + 1] - s->boxSize[1] * floor(s->pos[2 * v + 1] / s->boxSize[1]);
    long cx = (long)(wx / s->boxSize[0] * ncx), cy = (long)(wy / s->boxSize[1] * ncy);
    if (cx < 0) cx = 0;
    if (cx >= ncx) cx = ncx - 1;
    if (cy < 0) cy = 0;
    if (cy >= ncy) cy = ncy - 1;
    cell[v] = cy * ncx + cx;
    start[cell[v] + 1]++;
  }
  for (long c = 0; c < nc; c++) start[c + 1] += start[c];
  long* fill = (long*)malloc(sizeof(long) * nc);
  memcpy(fill, start, sizeof(long) * nc);
  for (long v = 0; v < nv; v++) item[fill[cell[v]]++] = v; /* ascending within each cell */
  free(fill);
  long W = s->neighborListSize;
#pragma omp parallel num_threads(s->threads) if (s->threads > 1)
  {
    long cap = 64, *cand = (long*)malloc(sizeof(long) * cap);
#pragma omp for schedule(dynamic, 64)
    for (long v = 0; v < nv; v++) {
      double thisPos[2] = {s->pos[v * 2], s->pos[v * 2 + 1]};
      double thisRad = s->rad[v];
      long particleId = s->particleIdList[v];
      long cx = cell[v] % ncx, cy = cell[v] / ncx, n = 0;
      for (long oy = -1; oy <= 1; oy++)
        for (long ox = -1; ox <= 1; ox++) {
          long c = ((cy + oy + ncy) % ncy) * ncx + (cx + ox + ncx) % ncx;
          for (long q = start[c]; q < start[c + 1]; q++) {
            long o = item[q];
            if (s->particleIdList[o] == particleId) continue;
            double otherPos[2] = {s->pos[o * 2], s->pos[o * 2 + 1]};
            double radSum = thisRad + s->rad[o];
            if (-(1 - calc_distance(s, thisPos, otherPos) / radSum) < cutDistance) {
              if (n == cap) {
                cap *= 2;
                cand = (long*)realloc(cand, sizeof(long) * cap);
              }
              cand[n++] = o;
            }
          }
        }
      for (long i = 1; i < n; i++) { /* insertion sort: rows are short */
        long x = cand[i], j = i - 1;
        while (j >= 0 && cand[j] > x) { cand[j + 1] = cand[j]; j--; }
        cand[j + 1] = x;
      }
      for (long k = 0; k < n && k < W; k++) s->neighborList[v * W + k] = cand[k];
      s->maxNeighborList[v] = n;
    }
    free(cand);
  }
  free(cell); free(start); free(item);
  return 1;
}

static void build_rows(orc_state* s, double cutDistance) {
  if (!(s->listMethod == 1 && build_list_cells(s, cutDistance))) build_list(s, cutDistance);
}

void orc_set_list_method(orc_state* s, int method) { s->listMethod = method; }

void orc_set_threads(orc_state* s, int threads) { s->threads = threads > 0 ? threads : 1; }

/* calcNeighborList, DPM2D.cu:3479-3507: grow-only power-of-two row width */
void orc_calc_neighbor_list(orc_state* s, double cutDistance) {
  long nv = s->numVertices;
  memset(s->maxNeighborList, 0, sizeof(long) * nv);
  long W = s->neighborListSize;
  for (long i = 0; i < nv * W; i++) s->neighborList[i] = -1;
  build_rows(s, cutDistance);
  long maxN = -1;
  for (long v = 0; v < nv; v++) if (s->maxNeighborList[v] > maxN) maxN = s->maxNeighborList[v];
  s->maxNeighbors = maxN;
  if (maxN > s->neighborListSize) {
    s->neighborListSize = (long)pow(2, ceil(log2((double)maxN)));
    W = s->neighborListSize;
    s->neighborList = (long*)realloc(s->neighborList, sizeof(long) * nv * W);
    for (long i = 0; i < nv * W; i++) s->neighborList[i] = -1;
    build_rows(s, cutDistance);
  }
}

long orc_neighbor_list_size(const orc_state* s) { return s->neighborListSize; }

/* copies the first min(count, cap) entries of each row into list[v*cap + k] */
long orc_get_neighbors(const orc_state* s, long* list, long* counts, long cap) {
  long W = s->neighborListSize;
  for (long v = 0; v < s->numVertices; v++) {
    counts[v] = s->maxNeighborList[v];
    for (long k = 0; k < cap; k++) list[v * cap + k] = (k < counts[v] && k < W) ? s->neighborList[v * W + k] : -1;
  }
  return W;
}

void orc_reset_last_positions(orc_state* s) { memcpy(s->lastPos, s->pos, sizeof(double) * 2 * s->numVertices); }
long orc_get_update_count(const orc_state* s) { return s->updateCount; }
void orc_reset_update_count(orc_state* s) { s->updateCount = 0; }

/* checkDisplacement + kernelCheckVertexDisplacement, DPM2D.cu:1390-1406, cudaKernel.cuh:1906-1914.
 * removeCOMDrift (DPM2D.cu:1348-1374, run when `shift`) reads out of bounds in the
 * reference; it is treated as a no-op here. */
int orc_check_neighbors(orc_state* s) {
  int any = 0;
#pragma omp parallel for reduction(| : any) num_threads(s->threads) if (s->threads > 1)
  for (long v = 0; v < s->numVertices; v++) {
    double disp = calc_distance(s, &s->pos[v * 2], &s->lastPos[v * 2]);
    if (3 * disp > s->cutoff) any = 1;
  }
  if (any) {
    orc_calc_neighbor_list(s, s->cutDistance);
    orc_reset_last_positions(s);
    s->updateCount += 1;
  }
  return any;
}

/* ---------------------------------------------------------- integrators --- */
static void update_position(orc_state* s, double ts) {
#pragma omp parallel for num_threads(s->threads) if (s->threads > 1)
  for (long i = 0; i < 2 * s->numVertices; i++) s->pos[i] += ts * s->vel[i];
}
static void update_velocity(orc_state* s, double ts) {
#pragma omp parallel for num_threads(s->threads) if (s->threads > 1)
  for (long i = 0; i < 2 * s->numVertices; i++) s->vel[i] += ts * s->force[i];
}

/* Langevin::injectKineticEnergy, Simulator.cu:32-52: one Gaussian pair per ring,
 * amplitude sqrt(T).  zeroMomentum (an addition) subtracts the mean vertex velocity. */
void orc_inject_kinetic_energy(orc_state* s, int zeroMomentum) {
  double amplitude = sqrt(s->Tinject);
  uint64_t call = s->injectCalls++;
  for (long v = 0; v < s->numVertices; v++) {
    long p = s->particleIdList[v];
    double z0, z1;
    orc_gauss_pair(s->seed, s->systemId, TAG_INJECT, call, (uint32_t)p, &z0, &z1);
    s->vel[v * 2] = amplitude * z0;
    s->vel[v * 2 + 1] = amplitude * z1;
  }
  if (zeroMomentum) {
    double m[2] = {0, 0};
    for (long v = 0; v < s->numVertices; v++) { m[0] += s->vel[v * 2]; m[1] += s->vel[v * 2 + 1]; }
    m[0] /= s->numVertices; m[1] /= s->numVertices;
    for (long v = 0; v < s->numVertices; v++) { s->vel[v * 2] -= m[0]; s->vel[v * 2 + 1] -= m[1]; }
  }
}

void orc_set_noise_type(orc_state* s, int noiseType) { s->noiseType = noiseType; }

/* initLangevin, DPM2D.cu:3713-3753: the (gamma, noise) pair of each noise type */
void orc_init_langevin(orc_state* s, double T, double gamma, int readState, int zeroMomentum) {
  s->simKind = ORC_SIM_LANGEVIN;
  s->Tinject = T;
  if (s->noiseType == ORC_NOISE_BAOAB) {
    s->gamma = exp(-gamma * s->dt);                         /* the O step's velocity factor */
    s->noise = sqrt(1 - exp(-2 * gamma * s->dt)) * sqrt(T);  /* and its noise amplitude */
  } else {
    s->gamma = gamma;
    s->noise = sqrt(2. * T * gamma / s->dt);
  }
  orc_reset_last_positions(s);
  orc_set_initial_positions(s);
  if (!readState) orc_inject_kinetic_energy(s, zeroMomentum);
}

/* initNVE, DPM2D.cu:3760-3772 */
void orc_init_nve(orc_state* s, double T, int readState, int zeroMomentum) {
  s->simKind = ORC_SIM_NVE;
  s->Tinject = T;
  orc_reset_last_positions(s);
  orc_set_initial_positions(s);
  s->shift = 1;
  if (!readState) orc_inject_kinetic_energy(s, zeroMomentum);
}

double orc_get_kinetic_energy(const orc_state* s) {
  double k = 0;
  for (long i = 0; i < 2 * s->numVertices; i++) k += s->vel[i] * s->vel[i];
  return 0.5 * k;
}

double orc_get_temperature(const orc_state* s) {
  return 2. * orc_get_kinetic_energy(s) / (s->numVertices * 2);
}

/* initNVERescale, DPM2D.cu:3779-3817: per-ring Gaussian, then rescale to T */
void orc_init_nve_rescale(orc_state* s, double T) {
  s->simKind = ORC_SIM_NVE_RESCALE;
  s->Tinject = T;
  orc_reset_last_positions(s);
  orc_set_initial_positions(s);
  s->shift = 1;
  orc_inject_kinetic_energy(s, 0);
  double scale = sqrt(T / orc_get_temperature(s));
  for (long i = 0; i < 2 * s->numVertices; i++) s->vel[i] *= scale;
}

/* Langevin::updateThermalVel, Simulator.cu:54-77: F += noise*xi_ring - gamma*v */
static void thermal_forces(orc_state* s) {
  uint64_t call = s->thermalCalls++;
#pragma omp parallel for num_threads(s->threads) if (s->threads > 1)
  for (long v = 0; v < s->numVertices; v++) {
    long p = s->particleIdList[v];
    double z[2];
    orc_gauss_pair(s->seed, s->systemId, TAG_THERMAL, call, (uint32_t)p, &z[0], &z[1]);
    for (int dim = 0; dim < 2; dim++)
      s->force[v * 2 + dim] += s->noise * z[dim] - s->gamma * s->vel[v * 2 + dim];
  }
}

/* BAOAB::updateThermalVel, Simulator.cu:122-144: v = gamma v + noise xi_ring */
static void baoab_thermal_velocity(orc_state* s) {
  uint64_t call = s->thermalCalls++;
  for (long v = 0; v < s->numVertices; v++) {
    long p = s->particleIdList[v];
    double z[2];
    orc_gauss_pair(s->seed, s->systemId, TAG_THERMAL, call, (uint32_t)p, &z[0], &z[1]);
    for (int dim = 0; dim < 2; dim++) s->vel[v * 2 + dim] = s->gamma * s->vel[v * 2 + dim] + s->noise * z[dim];
  }
}

/* Brownian::updateThermalVel, Simulator.cu:156-179: v = (F + noise xi_ring) / gamma;
 * DrivenBrownian::updateThermalVel, Simulator.cu:189-203: v = F / gamma (no draw) */
static void overdamped_velocity(orc_state* s, int noisy) {
  uint64_t call = noisy ? s->thermalCalls++ : 0;
  for (long v = 0; v < s->numVertices; v++) {
    long p = s->particleIdList[v];
    double z[2] = {0, 0};
    if (noisy) orc_gauss_pair(s->seed, s->systemId, TAG_THERMAL, call, (uint32_t)p, &z[0], &z[1]);
    for (int dim = 0; dim < 2; dim++)
      s->vel[v * 2 + dim] = noisy ? (s->force[v * 2 + dim] + s->noise * z[dim]) / s->gamma
                                  : s->force[v * 2 + dim] / s->gamma;
  }
}

/* Langevin::integrate, Simulator.cu:23-30; BAOAB::integrate 112-120; Brownian::integrate
 * 147-152; DrivenBrownian::integrate 182-187 -- chosen by the noise type */
void orc_langevin_loop(orc_state* s, long steps) {
  for (long n = 0; n < steps; n++) {
    switch (s->noiseType) {
      case ORC_NOISE_BAOAB:
        update_velocity(s, 0.5 * s->dt);
        update_position(s, 0.5 * s->dt);
        baoab_thermal_velocity(s);
        update_position(s, 0.5 * s->dt);
        orc_check_neighbors(s);
        orc_calc_force_energy(s);
        update_velocity(s, 0.5 * s->dt);
        break;
      case ORC_NOISE_BROWNIAN:
      case ORC_NOISE_DRIVEN_BROWNIAN:
        orc_check_neighbors(s);
        orc_calc_force_energy(s);
        overdamped_velocity(s, s->noiseType == ORC_NOISE_BROWNIAN);
        update_position(s, s->dt);
        break;
      default:
        update_velocity(s, 0.5 * s->dt);
        update_position(s, s->dt);
        orc_check_neighbors(s);
        orc_calc_force_energy(s);
        thermal_forces(s);
        update_velocity(s, 0.5 * s->dt);
    }
  }
}

/* NVE::integrate, Simulator.cu:206-212 */
void orc_nve_loop(orc_state* s, long steps) {
  for (long n = 0; n < steps; n++) {
    update_velocity(s, 0.5 * s->dt);
    update_position(s, s->dt);
    orc_check_neighbors(s);
    orc_calc_force_energy(s);
    update_velocity(s, 0.5 * s->dt);
  }
}

/* NVERescale::integrate + injectKineticEnergy, Simulator.cu:215-238 */
void orc_nve_rescale_loop(orc_state* s, long steps) {
  for (long n = 0; n < steps; n++) {
    double scale = sqrt(s->Tinject / orc_get_temperature(s));
    for (long i = 0; i < 2 * s->numVertices; i++) s->vel[i] *= scale;
    update_velocity(s, 0.5 * s->dt);
    update_position(s, s->dt);
    orc_check_neighbors(s);
    orc_calc_force_energy(s);
    update_velocity(s, 0.5 * s->dt);
  }
}

/* ----------------------------------------------------------------- FIRE --- */
/* initFIRE + FIRE::initMinimizer, DPM2D.cu:3660-3675, FIRE.cu:29-41 */
void orc_init_fire(orc_state* s, const double* p, long minStep, long numStep) {
  s->a_start = p[0]; s->f_dec = p[1]; s->f_inc = p[2]; s->f_a = p[3];
  s->fire_dt = p[4]; s->fire_dt_max = p[5]; s->a = p[6];
  s->minStep = minStep; s->numStep = numStep; s->lastP = 0;
}

/* checkFIREParams, FIRE.cu:338-371: the back-step uses the already-decreased dt */
static void check_fire_params(orc_state* s) {
  double vDotF = 0;
  for (long i = 0; i < 2 * s->numVertices; i++) vDotF += s->vel[i] * s->force[i];
  s->lastP = vDotF;
  if (vDotF > 0) {
    s->numStep++;
    if (s->numStep > s->minStep) {
      s->fire_dt = fmin(s->fire_dt * s->f_inc, s->fire_dt_max);
      s->a *= s->f_a;
    }
  } else {
    s->numStep = 0;
    s->fire_dt = fmax(s->fire_dt * s->f_dec, s->fire_dt_max / 10);
    s->a = s->a_start;
    double h = 0.5 * s->fire_dt;
    for (long i = 0; i < 2 * s->numVertices; i++) s->pos[i] -= h * s->vel[i];
    memset(s->vel, 0, sizeof(double) * 2 * s->numVertices);
  }
}

/* bendVertexVelocityTowardsForce, FIRE.cu:374-394 */
static void bend_velocity(orc_state* s) {
  double vv = 0, ff = 0;
  for (long i = 0; i < 2 * s->numVertices; i++) vv += s->vel[i] * s->vel[i];
  for (long i = 0; i < 2 * s->numVertices; i++) ff += s->force[i] * s->force[i];
  if (ff > 0) {
    double ratio = sqrt(vv / ff);
    double fa = s->a;
    for (long i = 0; i < 2 * s->numVertices; i++)
      s->vel[i] = (1 - fa) * s->vel[i] + fa * s->force[i] * ratio;
  }
}

/* FIRE::minimizerVertexLoop, FIRE.cu:397-411 */
void orc_fire_loop(orc_state* s, long iterations) {
  for (long it = 0; it < iterations; it++) {
    check_fire_params(s);
    update_velocity(s, 0.5 * s->fire_dt);
    bend_velocity(s);
    update_position(s, s->fire_dt);
    orc_check_neighbors(s);
    orc_calc_force_energy(s);
    update_velocity(s, 0.5 * s->fire_dt);
  }
}

void orc_get_fire_state(const orc_state* s, double* out) {
  out[0] = s->fire_dt; out[1] = s->a; out[2] = (double)s->numStep; out[3] = s->lastP;
  out[4] = (double)s->updateCount;
}

/* driver loop of compressFIRE.cpp:121-131: iterate while max|F_c| > tol */
long orc_fire_minimize(orc_state* s, double tol, long maxIter) {
  long it = 0;
  double fc = orc_get_max_unbalanced_force(s);
  while (fc > tol && it != maxIter) {
    orc_fire_loop(s, 1);
    fc = orc_get_max_unbalanced_force(s);
    it++;
  }
  return it;
}

/* ---------------------------------------------------------- rigid rings --- */
/* calcRigidForceEnergy, DPM2D.cu:2128-2142: ring centres, then the interaction forces alone (no
 * shape terms, no walls) and their transfer to the rings -- per-vertex torque r x f about the
 * centre, moment of inertia sum |r|^2 (unit vertex masses), ring force sum f, ring torque
 * sum (r x f) / I (an angular acceleration) and, for vertexVertex, the ring energy sum e
 * (kernelCalcVertexForceTorque, kernelCalcTorqueAndMomentOfInertia,
 * kernelCalcParticleRigidForceEnergy: cudaKernel.cuh:1412-1492).  vertexSmooth takes the
 * deformable path's vertex-segment interaction (kernelCalcSmoothInteraction 922-975) with its
 * ring energies, where the reference calls its unfinished kernelCalcVertexSmoothInteraction
 * (1054-1141). */
void orc_calc_rigid_force_energy(orc_state* s) {
  long nv = s->numVertices, np = s->numParticles;
  orc_calc_particle_positions(s);
  memset(s->force, 0, sizeof(double) * 2 * nv);
  memset(s->energy, 0, sizeof(double) * nv);
  memset(s->particleEnergy, 0, sizeof(double) * np);
  int smooth = s->interactionType == ORC_VERTEX_SMOOTH;
  if (smooth) orc_calc_smooth_interaction(s);
  else orc_calc_vertex_interaction(s);
  for (long p = 0; p < np; p++) {
    long first = s->firstVertexInParticleId[p], end = first + s->numVertexInParticleList[p];
    double I = 0, F[2] = {0, 0}, T = 0, E = 0;
    for (long v = first; v < end; v++) {
      double r[2];
      for (int dim = 0; dim < 2; dim++) r[dim] = pbcd(s, s->pos[v * 2 + dim], s->particlePos[p * 2 + dim], dim);
      s->torque[v] = r[0] * s->force[v * 2 + 1] - r[1] * s->force[v * 2];
      I += r[0] * r[0] + r[1] * r[1];
    }
    for (long v = first; v < end; v++) {
      F[0] += s->force[v * 2];
      F[1] += s->force[v * 2 + 1];
      T += s->torque[v];
      E += s->energy[v];
    }
    s->momentOfInertia[p] = I;
    s->particleForce[p * 2] = F[0];
    s->particleForce[p * 2 + 1] = F[1];
    s->particleTorque[p] = T / I;
    if (!smooth) s->particleEnergy[p] = E;
  }
}

/* translateVertices + rotateVertices, DPM2D.cu:1834-1848 (kernelTranslateVertices
 * cudaKernel.cuh:1678-1693, kernelRotateVertices 1695-1713): each ring moves by the box distance
 * from its last centre to its centre, then turns about its centre by its angle change. */
static void rigid_move_vertices(orc_state* s) {
  long np = s->numParticles;
  for (long p = 0; p < np; p++) {
    double d[2];
    for (int dim = 0; dim < 2; dim++) d[dim] = pbcd(s, s->particlePos[p * 2 + dim], s->particleLastPos[p * 2 + dim], dim);
    long first = s->firstVertexInParticleId[p], end = first + s->numVertexInParticleList[p];
    for (long v = first; v < end; v++)
      for (int dim = 0; dim < 2; dim++) s->pos[v * 2 + dim] += d[dim];
  }
  for (long p = 0; p < np; p++) {
    double dth = s->particleAngle[p] - s->particleLastAngle[p];
    long first = s->firstVertexInParticleId[p], end = first + s->numVertexInParticleList[p];
    for (long v = first; v < end; v++) {
      double d[2], n[2];
      for (int dim = 0; dim < 2; dim++) d[dim] = pbcd(s, s->pos[v * 2 + dim], s->particlePos[p * 2 + dim], dim);
      n[0] = d[0] * cos(dth) - d[1] * sin(dth);
      n[1] = d[0] * sin(dth) + d[1] * cos(dth);
      for (int dim = 0; dim < 2; dim++) s->pos[v * 2 + dim] += (n[dim] - d[dim]);
    }
  }
}

/* initRigidLangevin / initRigidNVE, DPM2D.cu:3823-3868: angles, angular velocities, torques and
 * last angles cleared (initRotationalVariables / initDeltaVariables, 197-221), last and initial
 * positions stored, and unless the state is read, RigidLangevin::injectKineticEnergy
 * (Simulator.cu:254-261): one Gaussian pair of amplitude sqrt(T) per ring, minus the mean ring
 * velocity (kernelConserveParticleMomentum, cudaKernel.cuh:1998-2027, whose one block covers at
 * most its block size of rings; the mean here is over all of them). */
void orc_init_rigid(orc_state* s, double T, double gamma, int langevin, int readState) {
  long np = s->numParticles;
  s->rigidKind = langevin ? 1 : 2;
  s->Tinject = T;
  s->gamma = gamma;
  memset(s->particleAngle, 0, sizeof(double) * np);
  memset(s->particleAngvel, 0, sizeof(double) * np);
  memset(s->particleTorque, 0, sizeof(double) * np);
  memset(s->momentOfInertia, 0, sizeof(double) * np);
  memset(s->particleLastAngle, 0, sizeof(double) * np);
  memset(s->torque, 0, sizeof(double) * s->numVertices);
  orc_reset_last_positions(s);
  orc_set_initial_positions(s);
  if (!readState) {
    double amplitude = sqrt(T), m[2] = {0, 0};
    uint64_t call = s->injectCalls++;
    for (long p = 0; p < np; p++) {
      double z0, z1;
      orc_gauss_pair(s->seed, s->systemId, TAG_INJECT, call, (uint32_t)p, &z0, &z1);
      s->particleVel[p * 2] = amplitude * z0;
      s->particleVel[p * 2 + 1] = amplitude * z1;
      m[0] += s->particleVel[p * 2];
      m[1] += s->particleVel[p * 2 + 1];
    }
    m[0] /= np; m[1] /= np;
    for (long p = 0; p < np; p++) { s->particleVel[p * 2] -= m[0]; s->particleVel[p * 2 + 1] -= m[1]; }
  }
}

/* RigidLangevin::updateVelocity, Simulator.cu:263-286, with d_thermalVel zero; z is the step's
 * unit Gaussian pair per ring.  RigidNVE::updateVelocity 348-366 when z is NULL. */
static void rigid_kick(orc_state* s, double h, const double* z) {
  long np = s->numParticles;
  for (long p = 0; p < np; p++) {
    for (int dim = 0; dim < 2; dim++) {
      double* v = &s->particleVel[p * 2 + dim];
      double F = s->particleForce[p * 2 + dim];
      if (z) {
        *v += h * (F - *v * s->gamma);
        *v -= 0.5 * h * h * (F - *v * s->gamma) / s->gamma;
        *v += z[p * 2 + dim] - 0.0;
      } else {
        *v += h * F;
      }
    }
    s->particleAngvel[p] += h * s->particleTorque[p];
  }
}

/* RigidLangevin::integrate / RigidNVE::integrate, Simulator.cu:241-252, 316-326.  The draw of
 * Langevin::updateThermalVel (54-77) lands in the vertex forces, which the rigid force call
 * overwrites; the rings feel only its unit Gaussian pair, added unscaled in both half kicks
 * (RigidLangevin::updatePosition 288-307 adds d_rando, which stays zero). */
void orc_rigid_loop(orc_state* s, long steps) {
  long np = s->numParticles;
  double* z = s->rigidKind == 1 ? (double*)malloc(sizeof(double) * 2 * np) : NULL;
  for (long n = 0; n < steps; n++) {
    memcpy(s->particleLastPos, s->particlePos, sizeof(double) * 2 * np);
    memcpy(s->particleLastAngle, s->particleAngle, sizeof(double) * np);
    if (z) {
      uint64_t call = s->thermalCalls++;
      for (long p = 0; p < np; p++)
        orc_gauss_pair(s->seed, s->systemId, TAG_THERMAL, call, (uint32_t)p, &z[p * 2], &z[p * 2 + 1]);
    }
    rigid_kick(s, 0.5 * s->dt, z);
    for (long p = 0; p < np; p++) {
      for (int dim = 0; dim < 2; dim++) s->particlePos[p * 2 + dim] += s->dt * s->particleVel[p * 2 + dim];
      s->particleAngle[p] += s->dt * s->particleAngvel[p];
    }
    rigid_move_vertices(s);
    orc_check_neighbors(s);
    orc_calc_rigid_force_energy(s);
    rigid_kick(s, 0.5 * s->dt, z);
  }
  free(z);
}

/* initRigidFIRE, DPM2D.cu:3700-3706: initFIRE, rotational variables cleared, last centres
 * stored */
void orc_init_rigid_fire(orc_state* s, const double* p, long minStep, long numStep) {
  long np = s->numParticles;
  orc_init_fire(s, p, minStep, numStep);
  memset(s->particleAngle, 0, sizeof(double) * np);
  memset(s->particleAngvel, 0, sizeof(double) * np);
  memset(s->particleTorque, 0, sizeof(double) * np);
  memset(s->momentOfInertia, 0, sizeof(double) * np);
  memset(s->particleLastAngle, 0, sizeof(double) * np);
  memset(s->torque, 0, sizeof(double) * s->numVertices);
  memcpy(s->particleLastPos, s->particlePos, sizeof(double) * 2 * np);
}

/* FIRE::minimizerRigidLoop, FIRE.cu:279-298, with checkRigidFIREParams (217-254) and
 * bendRigidVelocityTowardsForce (256-277): power and bending over the ring velocities and forces
 * only; a restart steps the centres and angles back by half the already-decreased step and
 * clears both velocities.  The last centres and angles are stored at the start of every
 * iteration, so the vertices follow that iteration's own move (the reference stores them only
 * in initRigidFIRE, so each later iteration would re-apply every earlier move). */
void orc_rigid_fire_loop(orc_state* s, long iterations) {
  long np = s->numParticles;
  for (long it = 0; it < iterations; it++) {
    memcpy(s->particleLastPos, s->particlePos, sizeof(double) * 2 * np);
    memcpy(s->particleLastAngle, s->particleAngle, sizeof(double) * np);
    double vDotF = 0;
    for (long i = 0; i < 2 * np; i++) vDotF += s->particleVel[i] * s->particleForce[i];
    s->lastP = vDotF;
    if (vDotF > 0) {
      s->numStep++;
      if (s->numStep > s->minStep) {
        s->fire_dt = fmin(s->fire_dt * s->f_inc, s->fire_dt_max);
        s->a *= s->f_a;
      }
    } else {
      s->numStep = 0;
      s->fire_dt = fmax(s->fire_dt * s->f_dec, s->fire_dt_max / 10);
      s->a = s->a_start;
      double h = 0.5 * s->fire_dt;
      for (long p = 0; p < np; p++) {
        for (int dim = 0; dim < 2; dim++) s->particlePos[p * 2 + dim] -= h * s->particleVel[p * 2 + dim];
        s->particleAngle[p] -= h * s->particleAngvel[p];
      }
      memset(s->particleVel, 0, sizeof(double) * 2 * np);
      memset(s->particleAngvel, 0, sizeof(double) * np);
    }
    rigid_kick(s, 0.5 * s->fire_dt, NULL);
    double vv = 0, ff = 0;
    for (long i = 0; i < 2 * np; i++) vv += s->particleVel[i] * s->particleVel[i];
    for (long i = 0; i < 2 * np; i++) ff += s->particleForce[i] * s->particleForce[i];
    if (ff > 0) {
      double ratio = sqrt(vv / ff), fa = s->a;
      for (long i = 0; i < 2 * np; i++)
        s->particleVel[i] = (1 - fa) * s->particleVel[i] + fa * s->particleForce[i] * ratio;
    }
    for (long p = 0; p < np; p++) {
      for (int dim = 0; dim < 2; dim++) s->particlePos[p * 2 + dim] += s->fire_dt * s->particleVel[p * 2 + dim];
      s->particleAngle[p] += s->fire_dt * s->particleAngvel[p];
    }
    rigid_move_vertices(s);
    orc_check_neighbors(s);
    orc_calc_rigid_force_energy(s);
    rigid_kick(s, 0.5 * s->fire_dt, NULL);
  }
}

/* getRigidMaxUnbalancedForce, DPM2D.cu:3643-3652: the larger of the square roots of the largest
 * squared ring force component and the largest squared ring torque */
double orc_get_rigid_max_unbalanced_force(const orc_state* s) {
  double mf = -1, mt = -1;
  for (long i = 0; i < 2 * s->numParticles; i++) mf = fmax(mf, s->particleForce[i] * s->particleForce[i]);
  for (long p = 0; p < s->numParticles; p++) mt = fmax(mt, s->particleTorque[p] * s->particleTorque[p]);
  return fmax(sqrt(mf), sqrt(mt));
}

/* getRigidKineticEnergy, DPM2D.cu:1149-1159: 0.5 sum v^2 + 0.5 sum (omega R)^2 with R the ring's
 * particle radius */
double orc_get_rigid_kinetic_energy(const orc_state* s) {
  double k = 0, r = 0;
  for (long i = 0; i < 2 * s->numParticles; i++) k += s->particleVel[i] * s->particleVel[i];
  for (long p = 0; p < s->numParticles; p++) {
    double w = s->particleAngvel[p] * s->particleRad[p];
    r += w * w;
  }
  return 0.5 * k + 0.5 * r;
}

/* ---------------------------------------------------------- observables --- */
/* getPotentialEnergy, DPM2D.cu:1062-1080 ('all' returns the real sum here, see A9) */
double orc_get_potential_energy(orc_state* s) {
  double e = 0;
  for (long v = 0; v < s->numVertices; v++) e += s->energy[v];
  if (s->interactionType == ORC_VERTEX_SMOOTH || s->interactionType == ORC_CELL_SMOOTH)
    for (long p = 0; p < s->numParticles; p++) e += s->particleEnergy[p];
  return e;
}

/* getMaxUnbalancedForce, DPM2D.cu:997-1005: sqrt of the largest squared component */
double orc_get_max_unbalanced_force(const orc_state* s) {
  double m = -1;
  for (long i = 0; i < 2 * s->numVertices; i++) {
    double f2 = s->force[i] * s->force[i];
    if (f2 > m) m = f2;
  }
  return sqrt(m);
}

/* getTotalForceMagnitude, DPM2D.cu:986-995 */
double orc_get_total_force_magnitude(const orc_state* s) {
  double t = 0;
  for (long i = 0; i < 2 * s->numVertices; i++) t += s->force[i] * s->force[i];
  return sqrt(t / (s->numVertices * 2));
}

/* vertex-disk correction of kernelCalcVertexArea, cudaKernel.cuh:1636-1642 */
static double vertex_area_sum(const orc_state* s) {
  double t = 0;
  for (long p = 0; p < s->numParticles; p++) {
    double r = s->rad[s->firstVertexInParticleId[p]];
    t += r * r * (0.5 * s->numVertexInParticleList[p] - 1);
  }
  return t;
}

/* getPhi, DPM2D.cu:1202-1220 */
double orc_get_phi(orc_state* s) {
  double phi = 0;
  for (long p = 0; p < s->numParticles; p++) phi += s->area[p];
  phi += PI * vertex_area_sum(s);
  return phi / (s->boxSize[0] * s->boxSize[1]);
}

/* getPreferredPhi, DPM2D.cu:1222-1238 */
double orc_get_preferred_phi(const orc_state* s) {
  double phi = 0;
  for (long p = 0; p < s->numParticles; p++) phi += s->a0[p];
  phi += PI * vertex_area_sum(s);
  return phi / (s->boxSize[0] * s->boxSize[1]);
}

/* getMeanParticleSize, DPM2D.cu:758-760 */
double orc_get_mean_particle_size(const orc_state* s) {
  double t = 0;
  for (long p = 0; p < s->numParticles; p++) t += s->a0[p];
  return 2 * sqrt(t / (PI * s->numParticles));
}

/* getMeanVertexRadius, DPM2D.cu:770-772 */
double orc_get_mean_vertex_radius(const orc_state* s) {
  double t = 0;
  for (long v = 0; v < s->numVertices; v++) t += s->rad[v];
  return t / s->numVertices;
}

/* adjustTemperature, DPM2D.cu:1185-1200 */
void orc_adjust_temperature(orc_state* s, double targetTemp) {
  double scale = sqrt(targetTemp / orc_get_temperature(s));
  for (long i = 0; i < 2 * s->numVertices; i++) s->vel[i] *= scale;
}

/* adjustKineticEnergy, DPM2D.cu:1161-1183: scale the velocities so that K + U returns to
 * prevEtot, when the kinetic energy covers the change; returns the scale, or 0 when no
 * adjustment is made (the reference prints and leaves the velocities) */
double orc_adjust_kinetic_energy(orc_state* s, double prevEtot) {
  double ekin = orc_get_kinetic_energy(s);
  double deltaEtot = orc_get_potential_energy(s) + ekin;
  deltaEtot -= prevEtot;
  if (ekin > deltaEtot) {
    double scale = sqrt((ekin - deltaEtot) / ekin);
    for (long i = 0; i < 2 * s->numVertices; i++) s->vel[i] *= scale;
    return scale;
  }
  return 0.;
}

/* scaleVertices, DPM2D.cu:1802-1811 with kernelScaleVertexPositions, cudaKernel.cuh:1644-1658:
 * positions are scaled about the stored particlePos (not recomputed, as in the reference). */
void orc_scale_vertices(orc_state* s, double scale) {
  for (long p = 0; p < s->numParticles; p++) {
    long first = s->firstVertexInParticleId[p];
    long last = first + s->numVertexInParticleList[p];
    for (long v = first; v < last; v++)
      for (int dim = 0; dim < 2; dim++) {
        double distance = s->pos[v * 2 + dim] - s->particlePos[p * 2 + dim];
        s->pos[v * 2 + dim] += (scale - 1) * distance;
      }
  }
  for (long p = 0; p < s->numParticles; p++) { s->a0[p] *= scale * scale; s->area[p] *= scale * scale; }
  for (long v = 0; v < s->numVertices; v++) { s->l0[v] *= scale; s->rad[v] *= scale; }
}

/* scaleBox, DPM2D.cu:1693-1719: ring centres and the box (the round box's radius) shrink by
 * 1/scale; translateVertices (DPM2D.cu:1834-1839, kernelTranslateVertices cudaKernel.cuh:1678-1693)
 * then moves each ring by the shift of its centre, a box distance in the new box.  The centres
 * are those of the last force evaluation. */
void orc_scale_box(orc_state* s, double scale) {
  long np = s->numParticles;
  double* last = (double*)malloc(sizeof(double) * 2 * np);
  memcpy(last, s->particlePos, sizeof(double) * 2 * np);
  for (long i = 0; i < 2 * np; i++) s->particlePos[i] /= scale;
  if (s->geometryType == ORC_GEOM_ROUND_BOX) {
    s->boxRadius /= scale;
  } else {
    s->boxSize[0] /= scale;
    s->boxSize[1] /= scale;
  }
  for (long p = 0; p < np; p++) {
    double delta[2];
    for (int dim = 0; dim < 2; dim++) delta[dim] = pbcd(s, s->particlePos[p * 2 + dim], last[p * 2 + dim], dim);
    long first = s->firstVertexInParticleId[p];
    long end = first + s->numVertexInParticleList[p];
    for (long v = first; v < end; v++)
      for (int dim = 0; dim < 2; dim++) s->pos[v * 2 + dim] += delta[dim];
  }
  free(last);
}

/* getParticleShapes, DPM2D.cu:739-756 */
void orc_get_particle_shapes(orc_state* s, double* shape) {
  for (long p = 0; p < s->numParticles; p++) {
    shape[p * 3] = s->area[p];
    shape[p * 3 + 1] = s->perimeter[p];
    shape[p * 3 + 2] = s->perimeter[p] * s->perimeter[p] / (4 * PI * s->area[p]);
  }
}

/* ---------------------------------------------------------- observables --- */
/* calcContacts + kernelCalcContacts, DPM2D.cu:3408-3428, cudaKernel.cuh:1791-1822: the particles
 * owning any entry of the particle's vertices' rows, in order of first appearance; the row width
 * doubles until every particle's contacts fit */
void orc_calc_contacts(orc_state* s) {
  long width = s->contactListSize > 0 ? s->contactListSize : 16;
  for (;;) {
    free(s->contactList);
    s->contactList = (long*)malloc(sizeof(long) * s->numParticles * width);
    for (long i = 0; i < s->numParticles * width; i++) s->contactList[i] = -1;
    int fits = 1;
    for (long p = 0; p < s->numParticles && fits; p++) {
      long added = 0, f = s->firstVertexInParticleId[p];
      for (long v = f; v < f + s->numVertexInParticleList[p]; v++) {
        for (long k = 0; k < s->maxNeighborList[v]; k++) {
          long other = s->neighborList[v * s->neighborListSize + k];
          if (other == v || other == -1) continue;
          long c = s->particleIdList[other];
          int isNew = 1;
          for (long q = 0; q < added; q++) if (s->contactList[p * width + q] == c) isNew = 0;
          if (!isNew) continue;
          if (added == width) { fits = 0; break; }
          s->contactList[p * width + added++] = c;
        }
        if (!fits) break;
      }
      s->maxContactList[p] = added;
    }
    if (fits) break;
    width *= 2;
  }
  s->contactListSize = width;
}

/* calcGeometricContacts + kernelCalcGeometricContacts, DPM2D.cu:3430-3450, cudaKernel.cuh:1823-1861:
 * the rings owning a row entry within a gap of `gap` radius sums of one of the ring's vertices
 * (-(1 - d / radSum) < gap), in order of first appearance, into rows of 4 * nDim = 8; a ring's
 * contacts beyond 8 are dropped (the reference's resize loop never triggers).  The row loop is
 * bounded by the vertex's own row count (the reference bounds it by the ring's). */
void orc_calc_geometric_contacts(orc_state* s, double gap) {
  const long width = 8;
  free(s->contactList);
  s->contactList = (long*)malloc(sizeof(long) * s->numParticles * width);
  for (long i = 0; i < s->numParticles * width; i++) s->contactList[i] = -1;
  for (long p = 0; p < s->numParticles; p++) {
    long added = 0, f = s->firstVertexInParticleId[p];
    for (long v = f; v < f + s->numVertexInParticleList[p]; v++) {
      for (long k = 0; k < s->maxNeighborList[v]; k++) {
        long other = s->neighborList[v * s->neighborListSize + k];
        if (other == v || other == -1) continue;
        double radSum = s->rad[v] + s->rad[other];
        double d = calc_distance(s, &s->pos[v * 2], &s->pos[other * 2]);
        if (!(-(1 - d / radSum) < gap)) continue;
        if (added >= width) continue;
        long c = s->particleIdList[other];
        int isNew = 1;
        for (long q = 0; q < width; q++) if (s->contactList[p * width + q] == c) isNew = 0;
        if (isNew) s->contactList[p * width + added++] = c;
      }
    }
    s->maxContactList[p] = added;
  }
  s->contactListSize = width;
}

long orc_contact_list_size(const orc_state* s) { return s->contactListSize; }

void orc_get_contacts(const orc_state* s, long* list, long* counts) {
  if (list) memcpy(list, s->contactList, sizeof(long) * s->numParticles * s->contactListSize);
  if (counts) memcpy(counts, s->maxContactList, sizeof(long) * s->numParticles);
}

/* getParticleHexaticOrder + kernelCalcHexaticOrderParameter, DPM2D.cu:1505-1510,
 * cudaKernel.cuh:1951-1976: per particle |<exp(6 i theta)>| over its contacts (0 without), mean */
double orc_get_hexatic_order(orc_state* s) {
  orc_calc_particle_positions(s);
  double sum = 0;
  for (long p = 0; p < s->numParticles; p++) {
    double re = 0, im = 0;
    long n = s->maxContactList[p];
    if (n > 0) {
      for (long q = 0; q < n; q++) {
        long o = s->contactList[p * s->contactListSize + q];
        if (o == p || o == -1) continue;
        double dx = pbcd(s, s->particlePos[2 * p], s->particlePos[2 * o], 0);
        double dy = pbcd(s, s->particlePos[2 * p + 1], s->particlePos[2 * o + 1], 1);
        double a = atan2(dy, dx);
        re += cos(6.0 * a);
        im += sin(6.0 * a);
      }
      re /= n;
      im /= n;
    }
    sum += hypot(re, im);
  }
  return sum / s->numParticles;
}

/* setInitialPositions + setInitialParticlePositions, DPM2D.cu:921-930 */
void orc_set_initial_positions(orc_state* s) {
  memcpy(s->initialPos, s->pos, sizeof(double) * 2 * s->numVertices);
  orc_calc_particle_positions(s);
  memcpy(s->particleInitPos, s->particlePos, sizeof(double) * 2 * s->numParticles);
}

/* getVertexMSD / getParticleMSD, DPM2D.cu:1284-1298 (normalised by the box area) */
double orc_get_vertex_msd(orc_state* s) {
  double sum = 0;
  for (long v = 0; v < s->numVertices; v++)
    for (int d = 0; d < 2; d++) {
      double dd = pbcd(s, s->pos[2 * v + d], s->initialPos[2 * v + d], d);
      sum += dd * dd;
    }
  return sum / (s->numVertices * s->boxSize[0] * s->boxSize[1]);
}

double orc_get_particle_msd(orc_state* s) {
  orc_calc_particle_positions(s);
  double sum = 0;
  for (long p = 0; p < s->numParticles; p++)
    for (int d = 0; d < 2; d++) {
      double dd = pbcd(s, s->particlePos[2 * p + d], s->particleInitPos[2 * p + d], d);
      sum += dd * dd;
    }
  return sum / (s->numParticles * s->boxSize[0] * s->boxSize[1]);
}

/* getVertexISF / getParticleISF, DPM2D.cu:1486-1503, cudaKernel.cuh:1935-1949: isotropic
 * sin(k d) / (k d), vertex wave number pi / (2 rad[0]) */
double orc_get_vertex_isf(orc_state* s) {
  double k = PI / (2 * s->rad[0]), sum = 0;
  for (long v = 0; v < s->numVertices; v++) {
    double d = calc_distance(s, &s->pos[2 * v], &s->initialPos[2 * v]);
    sum += sin(k * d) / (k * d);
  }
  return sum / s->numVertices;
}

double orc_get_particle_isf(orc_state* s, double k) {
  orc_calc_particle_positions(s);
  double sum = 0;
  for (long p = 0; p < s->numParticles; p++) {
    double d = calc_distance(s, &s->particlePos[2 * p], &s->particleInitPos[2 * p]);
    sum += sin(k * d) / (k * d);
  }
  return sum / s->numParticles;
}

/* kernelCalcPerParticleStressTensor, cudaKernel.cuh:1525-1565 (harmonic contacts): for each
 * overlapping listed pair of the particle's vertices, (lever) (x) f with f = ec overlap / radSum
 * * delta / (distance radSum), delta = other - this, and the lever = the vertex position
 * relative to the particle centre plus delta / 2 of this and every earlier overlap of the same
 * vertex (the reference accumulates it).  The rows are the vertices' own. */
void orc_calc_per_particle_stress(orc_state* s, double* out) {
  orc_calc_particle_positions(s);
  for (long p = 0; p < s->numParticles; p++) {
    double* st = out + 4 * p;
    st[0] = st[1] = st[2] = st[3] = 0;
    long f = s->firstVertexInParticleId[p];
    for (long v = f; v < f + s->numVertexInParticleList[p]; v++) {
      double rel[2] = {pbcd(s, s->pos[2 * v], s->particlePos[2 * p], 0),
                       pbcd(s, s->pos[2 * v + 1], s->particlePos[2 * p + 1], 1)};
      for (long k = 0; k < s->maxNeighborList[v]; k++) {
        long o = s->neighborList[v * s->neighborListSize + k];
        if (o == v || o == -1) continue;
        double radSum = s->rad[v] + s->rad[o];
        double distance = calc_distance(s, &s->pos[2 * v], &s->pos[2 * o]);
        double overlap = 1 - distance / radSum;
        if (overlap > 0) {
          double gm = s->ec * overlap / radSum, delta[2], force[2];
          for (int d = 0; d < 2; d++) {
            delta[d] = pbcd(s, s->pos[2 * o + d], s->pos[2 * v + d], d);
            force[d] = gm * delta[d] / (distance * radSum);
            rel[d] += delta[d] * 0.5;
          }
          st[0] += rel[0] * force[0];
          st[3] += rel[1] * force[1];
          st[1] += rel[0] * force[1];
          st[2] += rel[1] * force[0];
        }
      }
    }
  }
}

/* calcStressTensor, DPM2D.cu:2181-2187: the sum of the per-particle tensors (the reference's
 * kernelCalcStressTensor adds particle 0's entries once per particle, without atomics); and
 * getPressure, DPM2D.cu:976-983: trace / (nDim numVertices) */
void orc_calc_stress_tensor(orc_state* s, double* out4) {
  double* per = (double*)malloc(sizeof(double) * 4 * s->numParticles);
  orc_calc_per_particle_stress(s, per);
  out4[0] = out4[1] = out4[2] = out4[3] = 0;
  for (long p = 0; p < s->numParticles; p++)
    for (int c = 0; c < 4; c++) out4[c] += per[4 * p + c];
  free(per);
}

double orc_get_pressure(orc_state* s) {
  double st[4];
  orc_calc_stress_tensor(s, st);
  return (st[0] + st[3]) / (2 * s->numVertices);
}
