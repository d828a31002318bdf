// HIP kernels of the LS-DEM timestep for gfx950 (CDNA4, wave64).
//
// Per step (launch order, one stream):
//   k_kick_drift      half-kick, drift, reference-grid key, search cell, periodic wrap + ghost images,
//                     per-body motion record for the rebuild test and the separation certificates
//   k_broad_phase     on rebuild steps only (else it returns at once): counting sort of particles and
//                     ghost images by search-grid cell, FP32 bounding-sphere / grid-box candidate pairs,
//                     pair list in master order and the reaction CSR, phases joined by grid barriers
//                     (the same phases also run as separate gated launches)
//   k_pair_check      per candidate pair: partner this step, separation certificate against the motion
//                     since it was issued; the pairs to evaluate are compacted
//   k_candidates      one wavefront per batch of evaluated pairs: FP32 cluster cull and node prefilter
//                     (conservative), survivors written to a contiguous range per pair, then the exact
//                     FP64 node-vs-SDF test, tangential history and linear spring-dashpot per survivor
//   k_bonds           parallel bonds
//   k_reduce_kick     per-body segmented sum of contact and bond contributions + second half-kick
#include "device.h"
#include "surrogate.h"

namespace lsd {

// ------------------------------------------------------------------------------------------
// Reference grid (spatialGrid) and search grid helpers
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ i3 refCell(const GridGeom& G, d3 p) {
  return i3{int((p.x - G.mn.x) * G.inv.x), int((p.y - G.mn.y) * G.inv.y), int((p.z - G.mn.z) * G.inv.z)};
}
__device__ __forceinline__ int refKey(const GridGeom& G, d3 p) {
  if (p.x < G.mn.x) return -1; else if (p.x >= G.mx.x) return -1;
  if (p.y < G.mn.y) return -1; else if (p.y >= G.mx.y) return -1;
  if (p.z < G.mn.z) return -1; else if (p.z >= G.mx.z) return -1;
  return lin3(refCell(G, p), G.size);
}
// Squared distance from p to the reference domain box (0 inside).
__device__ __forceinline__ double domainDist2(const GridGeom& G, d3 p) {
  const double dx = fmax(0.0, fmax(G.mn.x - p.x, p.x - G.mx.x));
  const double dy = fmax(0.0, fmax(G.mn.y - p.y, p.y - G.mx.y));
  const double dz = fmax(0.0, fmax(G.mn.z - p.z, p.z - G.mx.z));
  return dx * dx + dy * dy + dz * dz;
}
// Ghost-image eligibility of a source at post-wrap position p, widened by hs (half the skin) so
// that it holds for every position within hs of p.  Set s: 0 = +x, 1 = +y, 2 = +x+y.
__device__ __forceinline__ bool potInColumn(const GridGeom& G, d3 p, int s, double hs) {
  const bool cx = (p.x - G.mn.x) * G.inv.x < 1.0 + hs * G.inv.x + 1e-9;
  const bool cy = (p.y - G.mn.y) * G.inv.y < 1.0 + hs * G.inv.y + 1e-9;
  return s == 0 ? cx : (s == 1 ? cy : (cx && cy));
}
__device__ __forceinline__ bool potOutColumn(const GridGeom& G, d3 p, int s, double hs) {
  const bool nx = (p.x - G.mn.x) * G.inv.x >= 1.0 - hs * G.inv.x - 1e-9;
  const bool ny = (p.y - G.mn.y) * G.inv.y >= 1.0 - hs * G.inv.y - 1e-9;
  return s == 0 ? nx : (s == 1 ? ny : (nx || ny));
}
// Can the image key of an unshifted image be valid (the source outside the ghost domain)?
__device__ __forceinline__ bool potOutOfDomain(const Params& P, d3 p, double hs) {
  const GridGeom& G = P.gridP;
  const double zmax = (P.quirks & 2) ? G.mx.y : G.mx.z;
  bool out = p.z < G.mn.z + hs || p.z >= zmax - hs;
  if (!P.periodicX) out = out || p.x < G.mn.x + hs || p.x >= G.mx.x - hs;
  if (!P.periodicY) out = out || p.y < G.mn.y + hs || p.y >= G.mx.y - hs;
  return out;
}

// Clamped cell and +-1 window of a node (the reference's 27-cell traversal bounds).
__device__ __forceinline__ void refWindow(const GridGeom& G, d3 p, i3& c, i3& lo, i3& hi) {
  c = refCell(G, p);
  lo = i3{-1, -1, -1};
  hi = i3{1, 1, 1};
  if (c.x <= 0) { c.x = 0; lo.x = 0; }
  if (c.x >= G.size.x - 1) { c.x = G.size.x - 1; hi.x = 0; }
  if (c.y <= 0) { c.y = 0; lo.y = 0; }
  if (c.y >= G.size.y - 1) { c.y = G.size.y - 1; hi.y = 0; }
  if (c.z <= 0) { c.z = 0; lo.z = 0; }
  if (c.z >= G.size.z - 1) { c.z = G.size.z - 1; hi.z = 0; }
}
__device__ __forceinline__ bool keyInWindow(const GridGeom& G, int key, d3 p) {
  i3 c, lo, hi;
  refWindow(G, p, c, lo, hi);
  const int kx = key % G.size.x, ky = (key / G.size.x) % G.size.y, kz = key / (G.size.x * G.size.y);
  return kx >= c.x + lo.x && kx <= c.x + hi.x && ky >= c.y + lo.y && ky <= c.y + hi.y &&
         kz >= c.z + lo.z && kz <= c.z + hi.z;
}
__device__ __forceinline__ int clampCell(double v, int n) {
  v = fmin(fmax(v, 0.0), double(n - 1));
  return int(v);
}
__device__ __forceinline__ i3 searchCell(const SearchGrid& S, d3 p) {
  return i3{clampCell(floor((p.x - S.mn.x) * S.invCell), S.size.x),
            clampCell(floor((p.y - S.mn.y) * S.invCell), S.size.y),
            clampCell(floor((p.z - S.mn.z) * S.invCell), S.size.z)};
}
__device__ __forceinline__ int ghostKey(const Params& P, d3 p) {
  const GridGeom& G = P.gridP;
  i3 c = refCell(G, p);
  bool out = false;
  if (p.x < G.mn.x) { c.x = 0; out = true; }
  else if (p.x >= G.mx.x) { c.x = G.size.x - 1; out = true; }
  if (p.y < G.mn.y) { c.y = 0; out = true; }
  else if (p.y >= G.mx.y) { c.y = G.size.y - 1; out = true; }
  const double zmax = (P.quirks & 2) ? G.mx.y : G.mx.z;
  if (p.z < G.mn.z) { c.z = 0; out = true; }
  else if (p.z >= zmax) { c.z = G.size.z - 1; out = true; }
  return out ? lin3(c, G.size) : -1;
}

// ------------------------------------------------------------------------------------------
// Small utilities
// ------------------------------------------------------------------------------------------
// Broad-phase kernels run only on steps that rebuild the candidate pairs: `gate` points at the
// step's rebuild flag (nullptr = always run).
#define LSD_GATE(g) \
  if ((g) != nullptr && *(g) == 0) return

// Raise error flag `bit` (kErr*): the call halts (every later kernel returns at entry).  The rebuild
// gates of both step buffers are cleared so that gated broad-phase launches skip as well; the index
// of the failing step is kept for the host.
__device__ __forceinline__ void raiseError(const Params& P, int bit) {
  atomicOr(&P.devInt[3], bit);
  atomicMax(&P.devInt[kErrStepIdx], P.callStep);
  atomicExch(&P.devInt[rebuildFlagIdx(0)], 0);
  atomicExch(&P.devInt[rebuildFlagIdx(1)], 0);
}
// An error flag is up: this kernel has nothing to do.
#define LSD_HALT(P) \
  if ((P).devInt[3] != 0) return

__global__ void k_fill_int(int* a, long long n, int v, const int* gate) {
  LSD_GATE(gate);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    a[i] = v;
}
__global__ void k_fill_u64(unsigned long long* a, long long n, unsigned long long v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    a[i] = v;
}

// Exclusive scan of n ints (n + 1 outputs, out[n] = total) in two phases over tiles.
constexpr int kScanThreads = 256;   // block of the separate scan launches
constexpr int kScanLds = 1024 / 64 + 1;  // LDS ints of blockExclusiveScan (blocks of up to 1024)
constexpr int kScanItems = 8;

__device__ __forceinline__ int blockExclusiveScan(int v, int* lds, int& total) {
  // any block size that is a multiple of 64 up to 1024; lds needs kScanLds ints
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int incl = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  if (lane == 63) lds[wid] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int w = 0; w < int(blockDim.x >> 6); ++w) {
      int t = lds[w];
      lds[w] = acc;
      acc += t;
    }
    lds[blockDim.x >> 6] = acc;
  }
  __syncthreads();
  total = lds[blockDim.x >> 6];
  int r = lds[wid] + incl - v;
  __syncthreads();
  return r;
}

// Grid-stride loop bounds of one phase: this thread's index and the number of threads.
struct Span {
  int tid, nth;
};
__device__ __forceinline__ Span gridSpan() {
  return Span{int(blockIdx.x * blockDim.x + threadIdx.x), int(gridDim.x * blockDim.x)};
}

// Phase 1 of the scan: the sum of each tile of blockDim.x x kScanItems inputs (blocks loop over
// tiles).
__device__ void scanReduceTiles(const int* in, long long n, int* tileSum, int* lds) {
  const long long tile = (long long)blockDim.x * kScanItems;
  const long long nTiles = (n + tile - 1) / tile;
  for (long long t = blockIdx.x; t < nTiles; t += gridDim.x) {
    const long long base = t * tile + (long long)threadIdx.x * kScanItems;
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
      if (base + k < n) s += in[base + k];
    int total;
    blockExclusiveScan(s, lds, total);
    if (threadIdx.x == 0) tileSum[t] = total;
  }
}
// Phase 2: each tile adds the sums of the tiles before it and writes its exclusive prefix; the
// last tile writes the total to out[n].  n == 0 writes out[0] = 0.
__device__ void scanApplyTiles(const int* in, long long n, const int* tileSum, int* out, int* lds) {
  const long long tile = (long long)blockDim.x * kScanItems;
  const long long nTiles = (n + tile - 1) / tile;
  if (n == 0 && blockIdx.x == 0 && threadIdx.x == 0) out[0] = 0;
  for (long long t = blockIdx.x; t < nTiles; t += gridDim.x) {
    int pre = 0;
    for (long long u = threadIdx.x; u < t; u += blockDim.x) pre += tileSum[u];
    int preTotal;
    blockExclusiveScan(pre, lds, preTotal);
    const long long base = t * tile + (long long)threadIdx.x * kScanItems;
    int v[kScanItems];
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      v[k] = (base + k < n) ? in[base + k] : 0;
      s += v[k];
    }
    int total;
    int ex = blockExclusiveScan(s, lds, total) + preTotal;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      if (base + k < n) out[base + k] = ex;
      ex += v[k];
    }
    if (t == nTiles - 1 && threadIdx.x == 0) out[n] = preTotal + total;
  }
}
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const int* in, long long n, int* tileSum,
                                                               const int* gate) {
  LSD_GATE(gate);
  __shared__ int lds[kScanLds];
  scanReduceTiles(in, n, tileSum, lds);
}
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const int* in, long long n, const int* tileSum, int* out,
                                                              const int* gate) {
  LSD_GATE(gate);
  __shared__ int lds[kScanLds];
  scanApplyTiles(in, n, tileSum, out, lds);
}

// ------------------------------------------------------------------------------------------
// Integration: first half-kick + drift (+ keys, wrap, ghost images)
// ------------------------------------------------------------------------------------------
// Chord 2 sin(angle / 2) of the rotation from unit quaternion b to unit quaternion q: the largest
// displacement of a unit vector between the two orientations (|R(q) - R(b)| in the 2-norm).
__device__ __forceinline__ double relChord(q4 b, q4 q) {
  const double vx = b.q0 * q.q1 - q.q0 * b.q1 - (b.q2 * q.q3 - b.q3 * q.q2);
  const double vy = b.q0 * q.q2 - q.q0 * b.q2 - (b.q3 * q.q1 - b.q1 * q.q3);
  const double vz = b.q0 * q.q3 - q.q0 * b.q3 - (b.q1 * q.q2 - b.q2 * q.q1);
  return 2.0 * sqrt(vx * vx + vy * vy + vz * vz);
}

// conj(b) q in FP32: the rotation from b to q as a quaternion.
__device__ __forceinline__ float4 relQuat(q4 b, q4 q) {
  const double w = b.q0 * q.q0 + b.q1 * q.q1 + b.q2 * q.q2 + b.q3 * q.q3;
  const double vx = b.q0 * q.q1 - q.q0 * b.q1 - (b.q2 * q.q3 - b.q3 * q.q2);
  const double vy = b.q0 * q.q2 - q.q0 * b.q2 - (b.q3 * q.q1 - b.q1 * q.q3);
  const double vz = b.q0 * q.q3 - q.q0 * b.q3 - (b.q1 * q.q2 - b.q2 * q.q1);
  return make_float4(float(w), float(vx), float(vy), float(vz));
}
// Chord between the rotations of two FP32 quaternions (as relChord), plus an allowance for their
// FP32 rounding.
__device__ __forceinline__ float relChordF(float4 b, float4 q) {
  const float vx = b.x * q.y - q.x * b.y - (b.z * q.w - b.w * q.z);
  const float vy = b.x * q.z - q.x * b.z - (b.w * q.y - b.y * q.w);
  const float vz = b.x * q.w - q.x * b.w - (b.y * q.z - b.z * q.y);
  return 2.f * sqrtf(vx * vx + vy * vy + vz * vz) * (1.f + 1e-5f) + 1e-6f;
}

// Which ghost images of a particle at (wrapped) position p lie a period away from it: bit 0 the
// +x image (first reference-grid column), bit 1 the +y image, bit 2 the corner image.
// (SEC: quarter-sector periodicity, known at compile time in the per-step kernels)
template <bool SEC>
__device__ __forceinline__ int imgShiftMask(const Params& P, d3 p) {
  const i3 c = refCell(P.gridP, p);
  if (SEC)  // turned images: +90 (first y cell), 180 (first x and y cells), -90 (first x cell)
    return (c.y == 0 ? 1 : 0) | (c.x == 0 && c.y == 0 ? 2 : 0) | (c.x == 0 ? 4 : 0);
  return (c.x == 0 ? 1 : 0) | (c.y == 0 ? 2 : 0) | (c.x == 0 && c.y == 0 ? 4 : 0);
}
// Lanes of the wavefront below this one.
__device__ __forceinline__ unsigned long long laneLtMask() {
  const int lane = int(__lane_id());
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Members of Params chosen by an image set known only at run time, read through selects: indexing
// a by-value kernel argument with a run-time index makes the compiler copy the whole argument block
// into per-lane scratch memory.
template <class T>
__device__ __forceinline__ T pick3(int s, T a, T b, T c) { return s == 0 ? a : (s == 1 ? b : c); }

// (an image position array indexed directly compiles to one load of the pointer when no other
// member is indexed at run time: the per-step kernels without sector periodicity keep that form)
template <bool SEC>
__device__ __forceinline__ d3* imgPos(const Params& P, int s) {
  return SEC ? pick3(s, P.gpos[0], P.gpos[1], P.gpos[2]) : P.gpos[s];
}
__device__ __forceinline__ q4* imgQuat(const Params& P, int s) { return pick3(s, P.gquat[0], P.gquat[1], P.gquat[2]); }
__device__ __forceinline__ d3* imgVel(const Params& P, int s) { return pick3(s, P.gvel[0], P.gvel[1], P.gvel[2]); }
// v turned by the rotation of sector image set s (about the z axis through the origin)
__device__ __forceinline__ d3 secTurn(const Params& P, int s, d3 v) {
  return rodrigues(v, pick3(s, P.secW[0], P.secW[1], P.secW[2]), pick3(s, P.secA[0], P.secA[1], P.secA[2]),
                   pick3(s, P.secB[0], P.secB[1], P.secB[2]));
}
__device__ __forceinline__ q4 secQuat(const Params& P, int s) { return pick3(s, P.secR[0], P.secR[1], P.secR[2]); }

// Column predicate (potInColumn / potOutColumn set) of image set s: with sector periodicity the sets
// turned by +90, 180 and -90 degrees take the sources of the first y cell, of the corner and of the
// first x cell.
template <bool SEC>
__device__ __forceinline__ int imageColumn(int s) { return SEC ? (s == 0 ? 1 : (s == 1 ? 2 : 0)) : s; }
// A body that left the quarter sector across x = min (y = min) turns by -90 (+90) degrees about the
// z axis through the corner: position, velocity and orientation (updateVelocityPositionOrientation-
// OutOfSectorKernel, mySpatialGrid.cu:58-85).  Returns whether it turned.
__device__ __forceinline__ bool sectorWrap(const Params& P, d3& p, d3& v, q4& q) {
  const d3 mn = P.gridP.mn;
  bool turned = false;
  if (p.x < mn.x && p.y >= mn.y) {
    v = rodrigues(v, P.secW[2], P.secA[2], P.secB[2]);
    p = mn + rodrigues(p - mn, P.secW[2], P.secA[2], P.secB[2]);
    q = qrotate(P.secR[2], q);
    turned = true;
  }
  if (p.y < mn.y && p.x >= mn.x) {
    v = rodrigues(v, P.secW[0], P.secA[0], P.secB[0]);
    p = mn + rodrigues(p - mn, P.secW[0], P.secA[0], P.secB[0]);
    q = qrotate(P.secR[0], q);
    turned = true;
  }
  return turned;
}

__device__ __forceinline__ void halfKick(const Params& P, int i, d3 F, d3 T) {
  const double invM = P.invM[i];
  if (isZero(invM)) return;
  const double h = 0.5 * P.dt;
  P.vel[i] += (F * invM + P.g) * h;
  P.angvel[i] += (mul(rotInvInertia(P.quat[i], P.invI[i]), T)) * h;
}

__device__ void wallDrift(const Params& P, int w, int buf);

// Thread 0 clears the step's counters and raises the host's rebuild request; all threads empty the
// entries of the history table this step fills (left there two steps ago; their slots were
// recorded at insertion).
__device__ __forceinline__ void stepHead(const Params& P, const StepHead& H, int i, int nth) {
  if (i == 0) {
    if (H.forceRebuild) atomicOr(&P.devInt[rebuildFlagIdx(H.buf)], 1);
    P.devInt[1] = 0;   // contact slots allocated
  }
  if (i < 8) P.evalCount[i * kBarrierLine] = 0;  // map records of the pair check's block groups
  const int nClear = P.devInt[10 + H.buf];
  for (int c = i; c < nClear; c += nth) {
    const int slot = H.hashSlot[c];
    if (slot >= 0) H.hashKey[slot] = kHashEmpty;
  }
}

// Half-kick(s) with (F, T), drift, reference-grid key, search cell, periodic wrap and ghost images,
// the rebuild test and the motion record of particle i.  kicks = 2 applies the previous step's
// second half-kick and this step's first, the same increment twice (the force of a step is used by
// both), exactly as two separate half-kicks round.
// Everything the kick and drift of particle i read, loaded before any of its stores (the stores
// could alias later loads as far as the compiler knows, which would serialise the round trips).
// (the inverse inertia is read by kickDrift only for a body with a torque)
struct BodyIn {
  d3 vel, w, pos, xBuild, xBuildW;
  q4 quat, qBuild;
  double invM;
  int imgBuild;
};
template <bool SEC>
__device__ __forceinline__ BodyIn loadBody(const Params& P, int i) {
  BodyIn b;
  b.vel = P.vel[i];
  b.w = P.angvel[i];
  b.quat = P.quat[i];
  b.invM = P.invM[i];
  b.pos = P.pos[i];
  b.xBuild = P.xBuild[i];
  b.xBuildW = P.xBuildW[i];
  b.qBuild = P.qBuild[i];
  b.imgBuild = (P.periodicX || P.periodicY || SEC) ? P.imgBuild[i] : 0;
  return b;
}
// SEC: quarter-sector periodicity (a compile-time switch, so that simulations without it run the code
// they ran before it existed)
template <bool SEC>
__device__ __forceinline__ void kickDrift(const Params& P, int i, const BodyIn& in, d3 F, d3 T, int kicks, int buf) {
  d3 vel = in.vel, w = in.w;
  const q4 q0 = in.quat;
  const double invM = in.invM;
  // a body without a torque keeps its angular velocity: R I^-1 R^T 0 adds only zeros, so neither its
  // inverse inertia (48 bytes) is read nor its angular velocity written back
  const bool torqued = T.x != 0.0 || T.y != 0.0 || T.z != 0.0;
  const s6 invI = torqued ? P.invI[i] : s6{};
  if (!isZero(invM) && !P.prime) {
    const double h = 0.5 * P.dt;
    const d3 dv = (F * invM + P.g) * h;
    vel += dv;
    if (kicks == 2) vel += dv;
    P.vel[i] = vel;
    if (torqued) {
      const d3 dw = (mul(rotInvInertia(q0, invI), T)) * h;
      w += dw;
      if (kicks == 2) w += dw;
      P.angvel[i] = w;
    }
  }
  d3 p = in.pos;
  q4 qn = q0;
  if (!P.prime) {  // (the pre-loop build detects at the current poses)
    p += P.dt * vel;
    qn = qintegrate(q0, w, P.dt);
    P.quat[i] = qn;
  }
  constexpr bool sector = SEC;
  const bool periodic = P.periodicX || P.periodicY || sector;
  const bool wrapAfter = (P.quirks & 4) != 0;
  const GridGeom& G = P.gridP;
  const d3 ox = mk3(G.mx.x - G.mn.x, 0., 0.), oy = mk3(0., G.mx.y - G.mn.y, 0.);
  bool turned = false;  // (sector periodicity: the body left the sector and turned this step)
  if (periodic && !wrapAfter) {
    if (P.periodicX) { if (p.x < G.mn.x) p += ox; else if (p.x >= G.mx.x) p -= ox; }
    if (P.periodicY) { if (p.y < G.mn.y) p += oy; else if (p.y >= G.mx.y) p -= oy; }
    if (sector) turned = sectorWrap(P, p, vel, qn);
  }
  P.posDet[i] = p;
  const d3 pDet = p;
  const q4 qDet = qn;  // orientation of particle-particle detection (before a turn after the build)
  if (sector) P.quatDet[i] = qn;
  const int keyRef = refKey(G, p);
  int gk[3] = {-1, -1, -1};  // image keys; an image's position is stored only when it exists
  int unshifted = 0;  // valid images at the particle's own position (particle outside the domain)
  if (periodic) {
    d3 img;
    if (P.periodicX) {
      if (wrapAfter) { if (p.x < G.mn.x) p += ox; else if (p.x >= G.mx.x) p -= ox; }
      img = p;
      if (refCell(G, p).x == 0) img += ox;
      const int k = ghostKey(P, img);
      if (k >= 0) P.gpos[0][i] = img;
      gk[0] = k;
      unshifted += (k >= 0 && refCell(G, p).x != 0);
    }
    if (P.periodicY) {
      if (wrapAfter) { if (p.y < G.mn.y) p += oy; else if (p.y >= G.mx.y) p -= oy; }
      img = p;
      if (refCell(G, p).y == 0) img += oy;
      const int k = ghostKey(P, img);
      if (k >= 0) P.gpos[1][i] = img;
      gk[1] = k;
      unshifted += (k >= 0 && refCell(G, p).y != 0);
    }
    if (P.periodicX && P.periodicY) {
      img = p;
      const i3 c = refCell(G, p);
      if (c.x == 0 && c.y == 0) img += ox + oy;
      const int k = ghostKey(P, img);
      if (k >= 0) P.gpos[2][i] = img;
      gk[2] = k;
      unshifted += (k >= 0 && !(c.x == 0 && c.y == 0));
    }
    if (sector) {
      // (mySpatialGrid.cu:154-212) images turned by +90 degrees for bodies in the first y cell, 180 for
      // those in the first x and y cells, -90 for those in the first x cell; other copies keep the pose
      if (wrapAfter) turned = sectorWrap(P, p, vel, qn) || turned;
      const i3 c = refCell(G, p);
      for (int s = 0; s < 3; ++s) {
        const bool on = s == 0 ? c.y == 0 : (s == 1 ? (c.x == 0 && c.y == 0) : c.x == 0);
        const d3 im = on ? G.mn + secTurn(P, s, p - G.mn) : p;
        const int k = ghostKey(P, im);
        if (k >= 0) {
          imgPos<SEC>(P, s)[i] = im;
          imgQuat(P, s)[i] = on ? qrotate(secQuat(P, s), qn) : qn;
          imgVel(P, s)[i] = on ? secTurn(P, s, vel) : vel;
        }
        gk[s] = k;
        unshifted += (k >= 0 && !on);
      }
    }
  }
  if (turned) {
    P.vel[i] = vel;
    P.quat[i] = qn;
  }
  P.pos[i] = p;
  // sources that may hold an unshifted image before the next rebuild
  if (periodic && P.halfSkin > 0.0) {
    unshifted = 0;
    if (potOutOfDomain(P, p, P.halfSkin))
      for (int s = 0; s < 3; ++s) unshifted += potOutColumn(G, p, imageColumn<SEC>(s), P.halfSkin);
  }
  if (unshifted) atomicAdd(&P.devInt[ghostCountIdx(buf)], unshifted);
  // rebuild the candidate pairs when a body has moved half the skin since the last build (a
  // periodic wrap always does)
  const double hs2 = P.halfSkin * P.halfSkin;
  const d3 dd = pDet - in.xBuild, dw = p - in.xBuildW;
  // (with sector periodicity the pairs are rebuilt every step: a turned body's images and
  // detection pose change by a quarter turn)
  const bool moved = sector || len2(dd) > hs2 || len2(dw) > hs2;
  if (__ballot(moved) && (threadIdx.x & 63) == __builtin_ctzll(__ballot(moved))) atomicOr(&P.devInt[rebuildFlagIdx(buf)], 1);
  // motion since the build for the separation certificates (a step that wraps rebuilds, so posDet
  // and pos move alike on every step that reads this)
  BodyStep st;
  const q4 qb = in.qBuild;
  st.motion = make_float4(float(dd.x), float(dd.y), float(dd.z), float(relChord(qb, qDet)));
  P.qrel[i] = relQuat(qb, qDet);
  st.key = make_int4(keyRef, -1, -1, -1);
  if (periodic) {
    const int flip = imgShiftMask<SEC>(P, p) ^ in.imgBuild;
    int* sk = &st.key.y;
    for (int g = 0; g < 3; ++g) sk[g] = gk[g] >= 0 ? (gk[g] | (((flip >> g) & 1) << kFlipBit)) : -1;
  }
  P.step[i] = st;
}

// One thread per particle, then one per wall.
template <bool SEC>
__global__ void k_kick_drift(Params P, int forceRebuild) {
  LSD_HALT(P);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  BodyIn in;
  d3 F, T;
  if (i < P.nP) {
    in = loadBody<SEC>(P, i);
    F = P.force[i];
    T = P.torque[i];
  }
  stepHead(P, StepHead{P.hashKey, P.hashSlot, P.curBuf, forceRebuild}, i, int(gridDim.x * blockDim.x));
  if (i >= P.nP) {
    if (i - P.nP < P.nW) wallDrift(P, i - P.nP, P.curBuf);
    return;
  }
  kickDrift<SEC>(P, i, in, F, T, 1, P.curBuf);
}

__device__ void wallDrift(const Params& P, int w, int buf) {
  d3 p = P.wpos[w];
  q4 q = P.wquat[w];
  if (!P.prime) {
    p += P.dt * P.wvel[w];
    P.wpos[w] = p;
    q = qintegrate(q, P.wangvel[w], P.dt);
    P.wquat[w] = q;
  }
  P.wkey[w] = refKey(P.gridW, p);
  // displacement bound of any point of the wall's grid box since the last build: translation plus
  // 2 R sin(angle / 2) for the rotation
  const double chord = relChord(P.wBuildQ[w], q);
  const double R = P.shapes[P.wshape[w]].boxR;
  const d3 dp = p - P.wBuildPos[w];
  const double disp = len(dp) + R * chord;
  if (disp > P.halfSkin) atomicOr(&P.devInt[rebuildFlagIdx(buf)], 1);
  P.wMotion[w] = make_float4(float(dp.x), float(dp.y), float(dp.z), float(chord));
  P.wQrel[w] = relQuat(P.wBuildQ[w], q);
  P.wSph[w] = make_float4(float(p.x), float(p.y), float(p.z), float(R));
}

// Record the poses the candidate pairs were built from (rebuild steps only).
__device__ void markBuildOne(const Params& P, int i);
__device__ void markBuild(const Params& P, Span g) {
  for (int i = g.tid; i < max(P.nP, P.nW); i += g.nth) markBuildOne(P, i);
}
__device__ void markBuildOne(const Params& P, int i) {
  if (i < P.nP) {
    P.xBuild[i] = P.posDet[i];
    P.xBuildW[i] = P.pos[i];
    P.qBuild[i] = P.periodicSector ? P.quatDet[i] : P.quat[i];
    BodyStep st = P.step[i];
    st.motion = make_float4(0.f, 0.f, 0.f, 0.f);
    P.qrel[i] = make_float4(1.f, 0.f, 0.f, 0.f);
    st.key.y = st.key.y >= 0 ? (st.key.y & ~(1 << kFlipBit)) : -1;
    st.key.z = st.key.z >= 0 ? (st.key.z & ~(1 << kFlipBit)) : -1;
    st.key.w = st.key.w >= 0 ? (st.key.w & ~(1 << kFlipBit)) : -1;
    P.step[i] = st;
    if (P.periodicX || P.periodicY || P.periodicSector)
      P.imgBuild[i] = P.periodicSector ? imgShiftMask<true>(P, P.pos[i]) : imgShiftMask<false>(P, P.pos[i]);
  }
  if (i < P.nW) {
    P.wBuildPos[i] = P.wpos[i];
    P.wBuildQ[i] = P.wquat[i];
    P.wMotion[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    P.wQrel[i] = make_float4(1.f, 0.f, 0.f, 0.f);
  }
}

// ------------------------------------------------------------------------------------------
// Deterministic counting sort by bucket: count (atomic rank) -> scan -> scatter -> per-bucket
// insertion sort by item index.
// ------------------------------------------------------------------------------------------
__device__ void bucketCount(const int* bucket, int n, int* count, int* rank, Span g) {
  for (int i = g.tid; i < n; i += g.nth) {
    const int b = bucket[i];
    rank[i] = b >= 0 ? atomicAdd(&count[b], 1) : -1;
  }
}
__device__ void bucketScatter(const int* bucket, int n, const int* start, const int* rank, int* out, Span g) {
  for (int i = g.tid; i < n; i += g.nth) {
    const int b = bucket[i];
    const int d = b >= 0 ? start[b] + rank[i] : -1;
    if (d >= 0 && d < n) out[d] = i;  // (always in range when the counts and the scan are complete)
  }
}
// Buckets of up to 16 items are sorted in registers (odd-even transposition network) from
// `tmp` into `items`; larger ones are listed for bucketSortBig (*bigCount starts at 0).
__device__ void bucketSortSmall(const int* start, int nb, const int* tmp, int* items, int* bigList, int* bigCount,
                                Span g, int nItems) {
  for (int b = g.tid; b < nb; b += g.nth) {
    const int s = start[b], e = start[b + 1];
    const int n = e - s;
    if (n <= 0 || s < 0 || e > nItems) continue;  // (never with a complete scan of the counts)
    if (n == 1) {
      items[s] = tmp[s];
      continue;
    }
    if (n > 16) {
      const int q = atomicAdd(bigCount, 1);
      if (q < nb) bigList[q] = b;
      continue;
    }
    int v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = k < n ? tmp[s + k] : 0x7fffffff;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
#pragma unroll
      for (int k = (r & 1); k + 1 < 16; k += 2) {
        const int a = v[k], c = v[k + 1];
        v[k] = min(a, c);
        v[k + 1] = max(a, c);
      }
    }
#pragma unroll
    for (int k = 0; k < 16; ++k)
      if (k < n) items[s + k] = v[k];
  }
}
// One wavefront per listed bucket: each item's place is the number of smaller items (items are
// distinct indices).  `wave` / `nWaves`: this wavefront's index and the number of wavefronts.
__device__ void bucketSortBig(const int* start, const int* tmp, int* items, const int* bigList, const int* bigCount,
                              int wave, int nWaves, int nb, int nItems) {
  const int lane = threadIdx.x & 63;
  const int nBig = min(*bigCount, nb);
  for (int q = wave; q < nBig; q += nWaves) {
    const int b = bigList[q];
    const int s = start[b], e = start[b + 1];
    if (s < 0 || e > nItems || s >= e) continue;  // (never with a complete scan of the counts)
    for (int c0 = s; c0 < e; c0 += 64) {
      const int v = c0 + lane < e ? tmp[c0 + lane] : 0x7fffffff;
      int rank = 0;
      for (int c1 = s; c1 < e; c1 += 64) {
        const int u = c1 + lane < e ? tmp[c1 + lane] : 0x7fffffff;
        for (int k = 0; k < 64; ++k) rank += __shfl(u, k, 64) < v ? 1 : 0;
      }
      if (c0 + lane < e) items[s + rank] = v;
    }
  }
}
__global__ void k_bucket_count(const int* bucket, int n, int* count, int* rank, const int* gate) {
  LSD_GATE(gate);
  bucketCount(bucket, n, count, rank, gridSpan());
}
__global__ void k_bucket_scatter(const int* bucket, int n, const int* start, const int* rank, int* out,
                                 int* bigCount, const int* gate) {
  LSD_GATE(gate);
  if (blockIdx.x == 0 && threadIdx.x == 0) *bigCount = 0;  // list of large buckets of k_bucket_sort
  bucketScatter(bucket, n, start, rank, out, gridSpan());
}
__global__ void k_bucket_sort(const int* start, int nb, const int* tmp, int* items, int* bigList, int* bigCount,
                              const int* gate, int nItems) {
  LSD_GATE(gate);
  bucketSortSmall(start, nb, tmp, items, bigList, bigCount, gridSpan(), nItems);
}
__global__ void k_bucket_sort_big(const int* start, int nb, const int* tmp, int* items, const int* bigList,
                                  const int* bigCount, const int* gate, int nItems) {
  LSD_GATE(gate);
  const Span g = gridSpan();
  bucketSortBig(start, tmp, items, bigList, bigCount, g.tid >> 6, g.nth >> 6, nb, nItems);
}

// Search-grid cell of a particle's post-wrap centre: ghost images are searched through their
// sources, and a particle that wrapped this step is outside the domain before the wrap (never a pp
// partner).
__device__ __forceinline__ int searchCellOf(const Params& P, int i) {
  return lin3(searchCell(P.sg, P.pos[i]), P.sg.size);
}

// Gather particle (or ghost-image) data into the sorted order of the search-grid table so
// that the candidate loops stream contiguous memory.
__device__ void gatherParticle(const Params& P, int k) {
  const int i = P.sorted[k];
  const d3 x = P.posDet[i], w = P.pos[i];
  const float r = float(P.radius[i]);
  // a pp partner must be inside the domain when the pair is used; one that can get there before
  // the next rebuild (within half the skin) stays a candidate
  const bool eligible = P.step[i].key.x >= 0 || domainDist2(P.gridP, x) <= P.halfSkin * P.halfSkin;
  P.sPR[k] = make_float4(float(x.x), float(x.y), float(x.z), eligible ? r : -1.f);
  P.sPRw[k] = make_float4(float(w.x), float(w.y), float(w.z), r);
  P.sBoxR[k] = float(P.scale[i] * P.shapes[P.shape[i]].boxR);
}
__global__ void k_search_cells(Params P) {
  LSD_GATE(P.devInt + rebuildFlagIdx(P.curBuf));
  const Span g = gridSpan();
  for (int i = g.tid; i < P.nP; i += g.nth) P.scell[i] = searchCellOf(P, i);
}
__global__ void k_gather_particles(Params P) {
  LSD_GATE(P.devInt + rebuildFlagIdx(P.curBuf));
  const Span g = gridSpan();
  for (int k = g.tid; k < P.nP; k += g.nth) gatherParticle(P, k);
}

// ------------------------------------------------------------------------------------------
// Candidate pairs
// ------------------------------------------------------------------------------------------
// Squared distance from x to the search grid's box (0 inside).
__device__ __forceinline__ double searchBoxDist2(const SearchGrid& S, d3 x) {
  const double ex = S.size.x * S.cell, ey = S.size.y * S.cell, ez = S.size.z * S.cell;
  const double dx = fmax(0.0, fmax(S.mn.x - x.x, x.x - (S.mn.x + ex)));
  const double dy = fmax(0.0, fmax(S.mn.y - x.y, x.y - (S.mn.y + ey)));
  const double dz = fmax(0.0, fmax(S.mn.z - x.z, x.z - (S.mn.z + ez)));
  return dx * dx + dy * dy + dz * dz;
}

// Visit every item of a search-grid table lying in cells that can hold a point within `reach`
// of x.  Rows (y, z) are pruned by their distance to x; the x-run of cells of a row is
// contiguous in the sorted order, so each row costs one range lookup.  The lookups of four rows
// are issued before their items are visited, so four rows cost one round trip to memory (the
// order of the visits is the same as row by row).
// The x-run of search-grid row (cy, cz) within reach of x, given the row's remaining squared
// reach rem >= 0: its first and one-past-last sorted positions.
__device__ __forceinline__ int2 rowRun(const SearchGrid& S, const int* cellStart, d3 x, int cy, int cz, double rem) {
  const double hw = sqrt(rem);
  const int x0 = clampCell(floor((x.x - hw - S.mn.x) * S.invCell), S.size.x);
  const int x1 = clampCell(floor((x.x + hw - S.mn.x) * S.invCell), S.size.x);
  const int row = (cz * S.size.y + cy) * S.size.x;
  return make_int2(cellStart[row + x0], cellStart[row + x1 + 1]);
}
// r2 minus the squared distance from x to row cy of a plane at squared distance dz2.
__device__ __forceinline__ double rowRem(const SearchGrid& S, d3 x, int cy, double r2, double dz2) {
  const double y0 = cy == 0 ? -1e300 : S.mn.y + cy * S.cell;
  const double y1 = cy == S.size.y - 1 ? 1e300 : S.mn.y + (cy + 1) * S.cell;
  const double dy = fmax(0.0, fmax(y0 - x.y, x.y - y1));
  return r2 - dy * dy - dz2;
}

template <class F>
__device__ __forceinline__ void visitNear(const SearchGrid& S, const int* cellStart, d3 x, double reach, F&& f) {
  const i3 c = searchCell(S, x);
  const int R = max(1, int(ceil(reach * S.invCell)));
  const double r2 = reach * reach;
  for (int cz = max(0, c.z - R); cz <= min(S.size.z - 1, c.z + R); ++cz) {
    const double z0 = cz == 0 ? -1e300 : S.mn.z + cz * S.cell;
    const double z1 = cz == S.size.z - 1 ? 1e300 : S.mn.z + (cz + 1) * S.cell;
    const double dz = fmax(0.0, fmax(z0 - x.z, x.z - z1));
    const double dz2 = dz * dz;
    if (dz2 > r2) continue;
    const int cyLo = max(0, c.y - R), cyHi = min(S.size.y - 1, c.y + R);
    // rows in groups of four: a group's range lookups are issued together, then its items are
    // visited (four named runs picked by selects: an indexed array would go to scratch memory, and
    // unrolling the visits would copy f four times)
    for (int g0 = cyLo; g0 <= cyHi; g0 += 4) {
      auto load = [&](int cy) {
        const double rem = cy <= cyHi ? rowRem(S, x, cy, r2, dz2) : -1.0;
        return rem >= 0.0 ? rowRun(S, cellStart, x, cy, cz, rem) : make_int2(0, 0);
      };
      const int2 q0 = load(g0), q1 = load(g0 + 1), q2 = load(g0 + 2), q3 = load(g0 + 3);
      for (int u = 0; u < 4; ++u) {
        const int2 r = u == 0 ? q0 : u == 1 ? q1 : u == 2 ? q2 : q3;
        for (int k = r.x; k < r.y; ++k) f(k);
      }
    }
  }
}

__device__ __forceinline__ double triAt(const Params& P, const ShapeDev& S, double sc, d3 loc, d3 go, double invSp,
                                        bool& inside, double& cmax) {
  // trilinear SDF value at a local point (for pair-level culls), inside = within the grid box,
  // cmax = largest |corner| (bounds the rounding of the value)
  const double gx = (loc.x - go.x) * invSp, gy = (loc.y - go.y) * invSp, gz = (loc.z - go.z) * invSp;
  const int i0 = (int)floor(gx), j0 = (int)floor(gy), k0 = (int)floor(gz);
  inside = !(i0 < 0 || j0 < 0 || k0 < 0 || i0 >= S.nx - 1 || j0 >= S.ny - 1 || k0 >= S.nz - 1);
  if (!inside) return 0.0;
  const i3 n{S.nx, S.ny, S.nz};
  const float* g = P.sdf + S.sdfOff;
  double p[8];
  p[0] = double(g[lin3(i3{i0, j0, k0}, n)]) * sc;
  p[1] = double(g[lin3(i3{i0 + 1, j0, k0}, n)]) * sc;
  p[2] = double(g[lin3(i3{i0, j0 + 1, k0}, n)]) * sc;
  p[3] = double(g[lin3(i3{i0 + 1, j0 + 1, k0}, n)]) * sc;
  p[4] = double(g[lin3(i3{i0, j0, k0 + 1}, n)]) * sc;
  p[5] = double(g[lin3(i3{i0 + 1, j0, k0 + 1}, n)]) * sc;
  p[6] = double(g[lin3(i3{i0, j0 + 1, k0 + 1}, n)]) * sc;
  p[7] = double(g[lin3(i3{i0 + 1, j0 + 1, k0 + 1}, n)]) * sc;
  cmax = 0.0;
  for (int c = 0; c < 8; ++c) cmax = fmax(cmax, fabs(p[c]));
  return triValue(gx - double(i0), gy - double(j0), gz - double(k0), p);
}

// Distance from a point to the grid box [o, o + (n-1) h] (all in the same local frame).
__device__ __forceinline__ double boxDist2(d3 loc, d3 o, double h, const ShapeDev& S) {
  const double ex = (S.nx - 1) * h, ey = (S.ny - 1) * h, ez = (S.nz - 1) * h;
  const double dx = fmax(0.0, fmax(o.x - loc.x, loc.x - (o.x + ex)));
  const double dy = fmax(0.0, fmax(o.y - loc.y, loc.y - (o.y + ey)));
  const double dz = fmax(0.0, fmax(o.z - loc.z, loc.z - (o.z + ez)));
  return dx * dx + dy * dy + dz * dz;
}

// FP32 sphere-overlap test with a margin covering the FP32 rounding of both centres and radii:
// accepts every pair whose FP64 bounding spheres overlap (a superset is harmless, every pair is
// settled by the exact node tests).
__device__ __forceinline__ bool spheresMayOverlap(float4 a, float4 b, float rb) {
  const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  const float ma = fmaxf(fmaxf(fabsf(a.x), fabsf(a.y)), fabsf(a.z));
  const float mb = fmaxf(fmaxf(fabsf(b.x), fabsf(b.y)), fabsf(b.z));
  const float rr = (a.w + rb) * (1.f + 1e-6f) + 3e-7f * (ma + mb) + 1e-30f;
  return dx * dx + dy * dy + dz * dz <= rr * rr * (1.f + 1e-6f);
}

// Candidate pairs of the master in sorted position t (A = sorted[t]), in a fixed order:
// particle partners, walls, then ghost images set by set.  DIRECT writes them to the pair list
// at `out`; otherwise the first kPairSlab go to the master's slab column and all are counted.
constexpr int kPairSlab = kPairSlabHost;
template <bool DIRECT>
__device__ __forceinline__ int pairsOf(const Params& P, int t, int A, long long out) {
  int cnt = 0;
  auto emit = [&](uint32_t type, int partner) {
    const uint32_t info = (type << 28) | uint32_t(partner);
    if (DIRECT) {
      if (out < P.pairCap) {
        P.pairA[out] = A;
        P.pairInfo[out] = info;
      }
      ++out;
    } else if (cnt < kPairSlab) {
      P.pairSlab[(long long)cnt * P.nP + t] = info;
    }
    ++cnt;
  };
  const float4 a = P.sPR[t];
  const float skin = float(2.0 * P.halfSkin);
  // particle-particle: B > A, B inside the domain, bounding spheres within the skin
  {
    const float ma = fmaxf(fmaxf(fabsf(a.x), fabsf(a.y)), fabsf(a.z));
    const double reach = (double(a.w) + P.rMax + 2.0 * P.halfSkin) * (1.0 + 2e-6) + 1e-6 * double(ma) + 1e-12;
    visitNear(P.sg, P.cellStart, mk3(a.x, a.y, a.z), reach, [&](int k) {
      const float4 b = P.sPR[k];
      if (b.w < 0.f || !spheresMayOverlap(a, b, b.w + skin)) return;
      const int B = P.sorted[k];
      if (B <= A) return;
      emit(PT_PP, B);
    });
  }
  // walls: sphere vs grid box, then a Lipschitz bound on the wall SDF at the centre
  if (P.nW > 0) {
    const double rA = P.radius[A];
    const d3 xA = P.posDet[A];
    const double rs = rA + 2.0 * P.halfSkin;  // the wall and the particle each move < half the skin
    const double rr = rs * (1.0 + 1e-9) + 1e-12;
    // every point of a wall's grid box lies within boxR of the wall's origin: walls far away (the
    // other drums of an ensemble) are dropped on an FP32 sphere test (its margin covers the FP32
    // roundings of the centres, of boxR and of the test) before the shape record is read and the
    // centre turned into the wall's frame.  The searching lanes of the wavefront load the spheres of
    // as many walls at once, one each, and pass them round: one memory round trip per that many
    // walls instead of one per wall.
    const float fx = float(xA.x), fy = float(xA.y), fz = float(xA.z);
    const float fr = float(rr);
    const float mA = fmaxf(fmaxf(fabsf(fx), fabsf(fy)), fabsf(fz));
    const unsigned long long act = __ballot(1);
    const int nAct = __popcll(act);
    const int myRank = __popcll(act & laneLtMask());
    for (int w0 = 0; w0 < P.nW; w0 += nAct) {
      const int wl = w0 + myRank;
      const float4 mine = wl < P.nW ? P.wSph[wl] : make_float4(0.f, 0.f, 0.f, 0.f);
      const int cnt = min(nAct, P.nW - w0);
      unsigned long long rest = act;
      for (int u = 0; u < cnt; ++u) {  // a wall's key is checked each step when the pair is used
        const int src = __ffsll((long long)rest) - 1;  // the lane holding wall w0 + u
        rest &= rest - 1;
        const float4 sw = make_float4(__shfl(mine.x, src, 64), __shfl(mine.y, src, 64), __shfl(mine.z, src, 64),
                                      __shfl(mine.w, src, 64));
        const int w = w0 + u;
        const float dx = fx - sw.x, dy = fy - sw.y, dz = fz - sw.z;
        const float ms = fmaxf(fmaxf(fabsf(sw.x), fabsf(sw.y)), fabsf(sw.z));
        const float reach = (sw.w + fr) * (1.f + 1e-5f) + 1e-6f * (mA + ms) + 1e-30f;
        if (dx * dx + dy * dy + dz * dz > reach * reach * (1.f + 1e-5f)) continue;
        const d3 dw = xA - P.wpos[w];
        const ShapeDev S = P.shapes[P.wshape[w]];
        const d3 loc = rotInv(dw, P.wquat[w]);
        const d3 o = mk3(S.ox, S.oy, S.oz);
        if (boxDist2(loc, o, S.h, S) > rr * rr) continue;
        bool inside;
        double cmax;
        const double phi = triAt(P, S, 1.0, loc, o, 1. / S.h, inside, cmax);
        if (inside && phi > S.lip * rs * (1.0 + 1e-6) + 1e-9 * S.h + 1e-12 * cmax) continue;
        emit(PT_PW, w);
      }
    }
  }
  // periodic ghost images: any node inside the image's grid box is tested, so the reach is
  // the image's grid-box radius.  Images are found through their sources in the particle grid:
  // an image shifted by the set's offset lies near x_A - offset; images left at the source's own
  // position (a source outside the domain) are searched around x_A only when such images exist.
  if (P.periodicX || P.periodicY || P.periodicSector) {
    const float4 g = P.sPRw[t];  // post-wrap centre (images are built from it)
    const float mg = fmaxf(fmaxf(fabsf(g.x), fabsf(g.y)), fabsf(g.z));
    const double reach = (double(g.w) + P.boxRMax + 2.0 * P.halfSkin) * (1.0 + 2e-6) + 1e-6 * double(mg) + 1e-12;
    const GridGeom& G = P.gridP;
    const double hs = P.halfSkin;
    const float lx = float(G.mx.x - G.mn.x), ly = float(G.mx.y - G.mn.y);
    const bool anyUnshifted = P.devInt[ghostCountIdx(P.curBuf)] > 0;
    const bool sector = P.periodicSector != 0;
    for (int s = 0; s < 3; ++s) {
      const bool on = sector || (s == 0 && P.periodicX) || (s == 1 && P.periodicY) ||
                      (s == 2 && P.periodicX && P.periodicY);
      if (!on) continue;
      const float offx = s != 1 ? lx : 0.f, offy = s != 0 ? ly : 0.f;
      const int col = sector ? imageColumn<true>(s) : imageColumn<false>(s);
      // the image of a source at b (shifted by the period, or turned about the corner's z axis)
      auto image = [&](float4 b) {
        if (!sector) return make_float4(b.x + offx, b.y + offy, b.z, 0.f);
        const d3 r = G.mn + secTurn(P, s, mk3(b.x, b.y, b.z) - G.mn);
        return make_float4(float(r.x), float(r.y), float(r.z), 0.f);
      };
      for (int pass = 0; pass < 2; ++pass) {
        const bool shifted = pass == 0;
        if (!shifted && !anyUnshifted) break;
        // sources whose image lies near x_A lie near the inverse image of x_A (a turn the other way)
        const int inv = 2 - s;
        const d3 c = !shifted ? mk3(g.x, g.y, g.z)
                     : (sector ? G.mn + secTurn(P, inv, mk3(g.x, g.y, g.z) - G.mn)
                               : mk3(double(g.x) - double(offx), double(g.y) - double(offy), double(g.z)));
        if (searchBoxDist2(P.sg, c) > reach * reach) continue;
        visitNear(P.sg, P.cellStart, c, reach, [&](int k) {
          const float4 b = P.sPRw[k];
          const float rb = P.sBoxR[k] + skin;
          const float4 im = shifted ? image(b) : b;
          if (!spheresMayOverlap(g, im, rb)) return;
          const int j = P.sorted[k];
          const d3 pj = P.pos[j];
          if (shifted) {
            if (!potInColumn(G, pj, col, hs)) return;
          } else {
            if (!potOutOfDomain(P, pj, hs) || !potOutColumn(G, pj, col, hs)) return;
            // one pair per (A, image): the shifted pass already took it
            if (potInColumn(G, pj, col, hs) && spheresMayOverlap(g, image(b), rb)) return;
          }
          emit(PT_GX + s, j);
        });
      }
    }
  }
  return cnt;
}

// Broad phase, one thread per master in spatial order: count and keep the first candidates.
__device__ void pairsFind(const Params& P, Span g) {
  for (int t = g.tid; t < P.nP; t += g.nth) {
    const int A = P.sorted[t];
    P.pairCount[A] = pairsOf<false>(P, t, A, 0);
  }
}
__global__ void k_pairs_find(Params P) {
  LSD_GATE(P.devInt + rebuildFlagIdx(P.curBuf));
  pairsFind(P, gridSpan());
}
// Pair list in master order: copy the slab, or search again for masters with more candidates.
__device__ void pairsEmit(const Params& P, Span g) {
  for (int t = g.tid; t < P.nP; t += g.nth) {
    const int A = P.sorted[t];
    const int n = P.pairCount[A];
    const long long out = P.pairStart[A];
    if (n > kPairSlab) {
      pairsOf<true>(P, t, A, out);
      continue;
    }
    for (int k = 0; k < n; ++k) {
      if (out + k >= P.pairCap) break;
      P.pairA[out + k] = A;
      P.pairInfo[out + k] = P.pairSlab[(long long)k * P.nP + t];
    }
  }
}
// Keep the pair list being replaced (per-master starts, entries, certificates) for the
// certificates the new list inherits; runs before the new starts and entries are written.
__device__ void backupPairs(const Params& P, Span g) {
  const int n = min(P.pairStart[P.nP], (int)P.pairCap);
  for (int i = g.tid; i <= P.nP; i += g.nth) P.oldStart[i] = P.pairStart[i];
  for (int p = g.tid; p < n; p += g.nth) {
    P.oldInfo[p] = P.pairInfo[p];
    P.oldCert[p] = P.cert[p];
  }
}
__global__ void k_backup_pairs(Params P) {
  LSD_GATE(P.devInt + rebuildFlagIdx(P.curBuf));
  backupPairs(P, gridSpan());
}
// A particle-particle pair of the new list that was in the previous one keeps its certificate,
// rebased to this build: the room its fast check leaves now (bound from the old motion records,
// read before they restart), with the rotation coefficients of this build's poses (|d| + r_A and
// r_A).  A node's motion since the old evaluation is at most the old bound now plus the new bound
// later, so the new bound staying below that room keeps the pair separated.  Other pairs start
// without a certificate.
__device__ void inheritCerts(const Params& P, Span g) {
  const int n = min(P.pairStart[P.nP], (int)P.pairCap);
  for (int p = g.tid; p < n; p += g.nth) {
    float4 nc = make_float4(-1.f, 0.f, 0.f, -1.f);
    const uint32_t info = P.pairInfo[p];
    if (P.useCert && (info >> 28) == PT_PP) {
      const int A = P.pairA[p], j = int(info & kPartnerMask);
      const int q1 = min(P.oldStart[A + 1], (int)P.pairCap);
      for (int q = P.oldStart[A]; q < q1; ++q) {
        if (P.oldInfo[q] != info) continue;
        const float4 c = P.oldCert[q];
        if (c.x > 0.f) {
          const float4 mA = P.step[A].motion, mB = P.step[j].motion;
          const float ex = mA.x - mB.x, ey = mA.y - mB.y, ez = mA.z - mB.z;
          const float bound = (sqrtf(ex * ex + ey * ey + ez * ez) + mB.w * c.y + mA.w * fabsf(c.z)) * (1.f + 1e-5f);
          const float rem = (c.x - bound) * (1.f - 1e-5f) - 1e-6f * c.x;  // the roundings of both checks
          if (rem > 0.f) {
            const double cA = P.scale[A] * P.shapes[P.shape[A]].nodeR * (1.0 + 1e-6);
            const double cB = (len(P.posDet[A] - P.posDet[j]) + cA) * (1.0 + 1e-6);
            nc = make_float4(rem, float(cB * (1.0 + 1e-6)), copysignf(float(cA * (1.0 + 1e-6)), c.z), -1.f);
          }
        }
        break;
      }
    }
    P.cert[p] = nc;
  }
}
__global__ void k_inherit_certs(Params P) {
  LSD_GATE(P.devInt + rebuildFlagIdx(P.curBuf));
  inheritCerts(P, gridSpan());
}
// The new list does not fit: raise the error (the rest of the call halts) instead of writing it.
__device__ __forceinline__ bool pairOverflow(const Params& P) {
  const int np = P.pairStart[P.nP];
  if ((long long)np <= P.pairCap) return false;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    atomicMax(&P.devInt[4], np);
    raiseError(P, kErrPairs);
  }
  return true;
}
__global__ void k_pairs_emit(Params P) {
  LSD_GATE(P.devInt + rebuildFlagIdx(P.curBuf));
  if (pairOverflow(P)) return;
  pairsEmit(P, gridSpan());
}
__global__ void k_mark_build(Params P) {
  LSD_GATE(P.devInt + rebuildFlagIdx(P.curBuf));
  markBuild(P, gridSpan());
}

// Pairs whose partner particle receives the reaction (particle-particle, ghosts with j != A).
__device__ void rcCountPairs(const Params& P, Span g) {
  const int n = min(P.pairStart[P.nP], (int)P.pairCap);
  for (int p = g.tid; p < n; p += g.nth) {
    const uint32_t info = P.pairInfo[p];
    const uint32_t type = info >> 28;
    const int partner = int(info & kPartnerMask);
    int r = -1;
    if (type == PT_PP || (type >= PT_GX && partner != P.pairA[p])) r = atomicAdd(&P.rcCount[partner], 1);
    P.rcRank[p] = r;
  }
}
__device__ void rcScatterPairs(const Params& P, Span g) {
  const int n = min(P.pairStart[P.nP], (int)P.pairCap);
  for (int p = g.tid; p < n; p += g.nth) {
    const int r = P.rcRank[p];
    if (r < 0) continue;
    const int partner = int(P.pairInfo[p] & kPartnerMask);
    const int d = P.rcStart[partner] + r;
    if (d < P.pairCap) P.rcTmp[d] = p;
  }
}
__global__ void k_rc_count(Params P) {
  LSD_GATE(P.devInt + rebuildFlagIdx(P.curBuf));
  rcCountPairs(P, gridSpan());
}
__global__ void k_rc_scatter(Params P) {
  LSD_GATE(P.devInt + rebuildFlagIdx(P.curBuf));
  if (blockIdx.x == 0 && threadIdx.x == 0) P.devInt[2] = 0;  // large-bucket list of the following sort
  rcScatterPairs(P, gridSpan());
}

// ------------------------------------------------------------------------------------------
// History table
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long histKey(uint32_t type, int gnode, int partner) {
  return ((unsigned long long)type << 59) | ((unsigned long long)(unsigned)gnode << 28) | (unsigned long long)partner;
}
__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ull;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebull;
  x ^= x >> 31;
  return x;
}
// Insert a contact of this step into this step's table (open addressing, linear probing); returns
// the slot, recorded so that the entry can be removed when the table is reused two steps later.
// The table holds at least twice the contact capacity, so a probe finds a free slot long before it
// has seen every slot; a probe that does see every slot raises kErrHash and gives up.
__device__ __forceinline__ int histInsert(const Params& P, unsigned long long key, int c) {
  unsigned long long slot = mix64(key) & P.hashMask;
  for (unsigned long long probe = 0; probe <= P.hashMask; ++probe) {
    const unsigned long long old = atomicCAS(&P.hashKey[slot], kHashEmpty, key);
    if (old == kHashEmpty || old == key) {
      P.hashVal[slot] = c;
      return int(slot);
    }
    slot = (slot + 1) & P.hashMask;
  }
  raiseError(P, kErrHash);
  return -1;
}

// The previous step's spring of a contact key (zero when it had none), from the previous table,
// given the key k0 and value v0 already read at the key's first slot slot0.
__device__ __forceinline__ d3 histLookup(const Params& P, unsigned long long key, unsigned long long slot0,
                                         unsigned long long k0, int v0) {
  if (k0 == key) return P.prevSpring[v0];
  unsigned long long slot = slot0, k = k0;
  for (unsigned long long probe = 0; probe <= P.hashMask; ++probe) {
    if (k == key) return P.prevSpring[P.prevHashVal[slot]];
    if (k == kHashEmpty) return mk3(0., 0., 0.);
    slot = (slot + 1) & P.hashMask;
    k = P.prevHashKey[slot];
  }
  return mk3(0., 0., 0.);  // (a full table: the insert that filled it raised kErrHash)
}

// ------------------------------------------------------------------------------------------
// Narrow phase: node-vs-SDF detection, one wavefront per batch of 8 candidate pairs
// ------------------------------------------------------------------------------------------
constexpr int kNarrowBlock = 256;

// Stat slot `slot` of accumulator family f (one 128-byte line per slot).
__device__ __forceinline__ unsigned long long* statSlot(const Params& P, int f, int slot) {
  return P.statSlots + (f * kStatSlots + slot) * kStatStride;
}

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Detection-side view of a pair (everything the node tests need).
struct PairView {
  int pair, A, partner, nodeBase, valid;
  uint32_t type, info;
  d3 xA;        // master position used for detection
  q4 qA;
  double sA;
  double rA;    // largest distance of a node of A from A's origin (scaled)
  int clusterOff, nClusters, nodeOff, shapeB;
  d3 xB;        // partner pose used for detection
  q4 qB;
  double rB;    // particle-particle radius cull (0 for fixed partners)
  double scB, invSpB;
  d3 goB;
  int keyB;
};

// (info and A as in pairInfo[p], pairA[p], when the caller already holds them)
template <bool SEC>
__device__ __forceinline__ PairView loadPair(const Params& P, int p, uint32_t info, int A) {
  PairView v;
  v.pair = p;
  v.info = info;
  v.type = v.info >> 28;
  v.partner = int(v.info & kPartnerMask);
  v.A = A;
  const bool ghost = v.type >= PT_GX;
  v.xA = ghost ? P.pos[v.A] : P.posDet[v.A];
  v.qA = (SEC && !ghost) ? P.quatDet[v.A] : P.quat[v.A];
  v.sA = P.scale[v.A];
  const ShapeDev SA = P.shapes[P.shape[v.A]];
  v.clusterOff = SA.clusterOff;
  v.nClusters = SA.nClusters;
  v.nodeOff = SA.nodeOff;
  v.rA = v.sA * SA.nodeR;
  v.nodeBase = P.nodeStart[v.A];
  const int j = v.partner;
  if (v.type == PT_PW) {
    v.xB = P.wpos[j]; v.qB = P.wquat[j]; v.rB = 0.; v.shapeB = P.wshape[j]; v.scB = 1.0;
    const ShapeDev SB = P.shapes[v.shapeB];
    v.invSpB = 1. / SB.h; v.goB = mk3(SB.ox, SB.oy, SB.oz); v.keyB = P.wkey[j];
  } else {
    // keys from the partner's step record (an image key carries its shift-flip bit there)
    const int* keys = reinterpret_cast<const int*>(P.step + j) + 4;  // BodyStep::key as 4 ints
    const int k = keys[v.type == PT_PP ? 0 : 1 + int(v.type - PT_GX)];
    if (v.type == PT_PP) {
      v.xB = P.posDet[j]; v.keyB = k;
      v.qB = SEC ? P.quatDet[j] : P.quat[j];
    } else {
      v.xB = imgPos<SEC>(P, int(v.type - PT_GX))[j];
      v.keyB = k >= 0 ? (k & ~(1 << kFlipBit)) : -1;
      v.qB = SEC ? imgQuat(P, int(v.type - PT_GX))[j] : P.quat[j];
    }
    v.rB = P.radius[j]; v.shapeB = P.shape[j]; v.scB = P.scale[j];
    v.invSpB = P.invSp[j]; v.goB = P.gorig[j];
  }
  v.valid = 1;
  return v;
}

constexpr int kContactChunk = 1024;  // candidate slots reserved per atomic

// The step record's key of a pair's partner of type t (the body's own key, or its image's): a
// select, since a runtime index into the record's fields would go through scratch memory.
__device__ __forceinline__ int partnerKey(int4 key, uint32_t t) {
  return t == PT_PP ? key.x : (t == PT_GX ? key.y : (t == PT_GY ? key.z : key.w));
}

struct NodeHit {
  d3 point, normal;
  double ov;
  int gnode;
};

// Exact node-vs-SDF test of global node `ni` (cluster-ordered node array) in the reference's
// arithmetic: node transform, radius cull or cell window, grid bounds, trilinear overlap >= 0.
// (nB and sdfOff: the partner's grid size and its offset in the SDF pool)
__device__ __forceinline__ bool exactTest(const Params& P, const PairView& v, const NodeDev& nd, i3 nB,
                                          long long sdfOff, bool full, NodeHit& h, unsigned long long& tests,
                                          unsigned long long& queries) {
  h.gnode = v.nodeBase + nd.orig;
  const d3 xw = rot(v.qA, mk3(nd.x * v.sA, nd.y * v.sA, nd.z * v.sA)) + v.xA;
  tests++;
  if (v.type == PT_PP) {
    if (len2(xw - v.xB) > v.rB * v.rB) return false;
  } else {
    if (!keyInWindow(v.type == PT_PW ? P.gridW : P.gridP, v.keyB, xw)) return false;
  }
  const d3 loc = rotInv(xw - v.xB, v.qB);
  const double gx = (loc.x - v.goB.x) * v.invSpB, gy = (loc.y - v.goB.y) * v.invSpB,
               gz = (loc.z - v.goB.z) * v.invSpB;
  const int i0 = (int)floor(gx), j0 = (int)floor(gy), k0 = (int)floor(gz);
  if (i0 < 0 || j0 < 0 || k0 < 0 || i0 >= nB.x - 1 || j0 >= nB.y - 1 || k0 >= nB.z - 1) return false;
  const int i1 = i0 + 1, j1 = j0 + 1, k1 = k0 + 1;
  const double x = gx - double(i0), y = gy - double(j0), z = gz - double(k0);
  const float* g = P.sdf + sdfOff;
  double cc[8];
  cc[0] = double(g[lin3(i3{i0, j0, k0}, nB)]) * v.scB;
  cc[1] = double(g[lin3(i3{i1, j0, k0}, nB)]) * v.scB;
  cc[2] = double(g[lin3(i3{i0, j1, k0}, nB)]) * v.scB;
  cc[3] = double(g[lin3(i3{i1, j1, k0}, nB)]) * v.scB;
  cc[4] = double(g[lin3(i3{i0, j0, k1}, nB)]) * v.scB;
  cc[5] = double(g[lin3(i3{i1, j0, k1}, nB)]) * v.scB;
  cc[6] = double(g[lin3(i3{i0, j1, k1}, nB)]) * v.scB;
  cc[7] = double(g[lin3(i3{i1, j1, k1}, nB)]) * v.scB;
  queries++;
  h.ov = -triValue(x, y, z, cc);
  if (!(h.ov >= 0.)) return false;
  if (full) {
    d3 n = triGradient((P.quirks & 1) ? double(int(x)) : x, y, z, cc);
    n = rot(v.qB, n);
    n = normalize(n);
    h.normal = n;
    h.point = xw + 0.5 * h.ov * n;
  }
  return true;
}

// FP32 map of A's unscaled node coordinates n into the partner's grid index space,
// g = M n + t (g = (R_B^T (x_world - x_B) - origin_B) / spacing_B), with conservative acceptance
// regions: every margin is far above the FP32 rounding of the map and of the interpolation.
struct alignas(16) PairF32 {
  float m[9], t[3];
  float c[3], r2;     // pp: partner centre in g and squared bounding radius (g units); r2 < 0: no sphere test
  float hi[3], lo;    // grid box in g: [lo, hi] per axis
  float phiMargin;    // SDF prefilter margin for FP32 map rounding (unscaled SDF units)
  float rhoScale;     // A's unscaled cluster radius -> g units
  float lipH;         // Lipschitz bound of the partner's SDF per g unit
  int nx, ny, nz, sdfOff;
  int nodeOff, clusterOff, nClusters;  // A's shape
  int pair;           // candidate pair index
  float gToWorld;     // world length of one g unit (the partner's scaled spacing)
  float o[3], h;      // partner's unscaled grid origin and spacing (g -> local point)
  int slot;           // partner's surrogate family slot, -1 if none (walls)
  float certD0;       // motion bound of the pair since the build, at this step (inf: none)
  float certCB, certCA;  // certificate coefficients of the partner's and the master's rotation
};
static_assert(sizeof(PairF32) == 160, "pair map record is 160 bytes");

__device__ __forceinline__ void makeF32(const Params& P, const PairView& v, PairF32& f) {
  const ShapeDev SB = P.shapes[v.shapeB];
  const double is = v.invSpB;  // 1 / (scaled spacing)
  const d3 c0 = rotInv(rot(v.qA, mk3(v.sA, 0., 0.)), v.qB) * is;
  const d3 c1 = rotInv(rot(v.qA, mk3(0., v.sA, 0.)), v.qB) * is;
  const d3 c2 = rotInv(rot(v.qA, mk3(0., 0., v.sA)), v.qB) * is;
  const d3 t = (rotInv(v.xA - v.xB, v.qB) - v.goB) * is;
  f.m[0] = float(c0.x); f.m[1] = float(c1.x); f.m[2] = float(c2.x);
  f.m[3] = float(c0.y); f.m[4] = float(c1.y); f.m[5] = float(c2.y);
  f.m[6] = float(c0.z); f.m[7] = float(c1.z); f.m[8] = float(c2.z);
  f.t[0] = float(t.x); f.t[1] = float(t.y); f.t[2] = float(t.z);
  // |g| of any node that can pass is below reach; FP32 error of g is ~1e-7 reach
  const double cg = len(v.goB) * is;
  const double reach = len(t) + 8.0 * v.sA * is + cg + double(SB.nx + SB.ny + SB.nz);
  const double abs = 1e-5 * reach + 1e-6;
  f.c[0] = float(-v.goB.x * is); f.c[1] = float(-v.goB.y * is); f.c[2] = float(-v.goB.z * is);
  if (v.type == PT_PP) {
    const double r = v.rB * is * (1.0 + 1e-5) + abs;
    f.r2 = float(r * r);
  } else {
    f.r2 = -1.f;
  }
  f.lo = float(-abs);
  f.hi[0] = float(SB.nx - 1 + abs);
  f.hi[1] = float(SB.ny - 1 + abs);
  f.hi[2] = float(SB.nz - 1 + abs);
  f.lipH = float(SB.lip * SB.h * (1.0 + 1e-5));
  f.phiMargin = float(SB.lip * SB.h * abs * 2.0 + 1e-30);
  f.rhoScale = float(v.sA * is * (1.0 + 1e-5));
  f.nx = SB.nx; f.ny = SB.ny; f.nz = SB.nz;
  f.sdfOff = int(SB.sdfOff);
  f.nodeOff = v.nodeOff;
  f.clusterOff = v.clusterOff;
  f.nClusters = v.nClusters;
  f.pair = v.pair;
  f.gToWorld = float(1.0 / is);
  f.o[0] = float(SB.ox); f.o[1] = float(SB.oy); f.o[2] = float(SB.oz);
  f.h = float(SB.h);
  f.slot = (P.surSlot && v.type != PT_PW) ? P.surSlot[v.shapeB] : -1;
  // Separation-certificate coefficients.  A node u of A (|u| <= rA) sits at
  // g = R_B^T (d + R_A u) / spacing, d = x_A - x_B.  Between the build and now it moved at most
  //   |dA - dB| + chord_B |d_build| + (chord_A + chord_B) rA,   |d_build| <= |d| + |dA - dB|,
  // with dA, dB the displacements and chord the rotation chords since the build.
  const float4 mA = P.step[v.A].motion;
  float4 mB;
  bool flip = false;  // an image that changed its shift state has no motion bound
  if (v.type == PT_PW) mB = P.wMotion[v.partner];
  else {
    const BodyStep sb = P.step[v.partner];
    mB = sb.motion;
    if (v.type >= PT_GX) flip = ((partnerKey(sb.key, v.type) >> kFlipBit) & 1) != 0;
  }
  const double ex = double(mA.x) - double(mB.x), ey = double(mA.y) - double(mB.y), ez = double(mA.z) - double(mB.z);
  const double del = sqrt(ex * ex + ey * ey + ez * ez);
  const double cA = v.rA * (1.0 + 1e-6);
  const double cB = (len(v.xA - v.xB) + del + cA) * (1.0 + 1e-6);
  const double d0 = del + double(mB.w) * cB + double(mA.w) * cA;
  f.certD0 = flip ? __builtin_huge_valf() : float(d0 * (1.0 + 1e-5));
  f.certCB = float(cB * (1.0 + 1e-6));
  f.certCA = float(cA * (1.0 + 1e-6));
}

// Which pairs the narrow phase evaluates this step, one thread per candidate pair.  A pair is
// skipped when its partner does not exist now (pp: outside the domain; wall: its key; ghost: the
// image) or when its separation certificate still holds: the motion bound of its nodes since the
// build stays below the slack recorded at its last evaluation (all its nodes were culled then
// with that much room).  Skipped pairs have no contacts.  full != 0 (an audit step of the
// surrogate) also evaluates every live pair whose certificate rests on the surrogate's slack.
// A block compacts a tile of 1024 pairs in LDS, appends it with one atomic, then writes the FP32
// map records of its evaluated pairs (FP64 poses in, one 160-byte record out, in evaluation
// order).  A block group g (blockIdx mod 8) checks one contiguous eighth of the tiles and writes the
// records of its evaluated pairs from the first pair of its first tile on, with its own counter
// (evalCount[g]): 8 counters on 8 lines instead of one, since returning atomics on one address are
// served one after another (~11 ns each) and every tile takes one.
constexpr int kCheckTile = 1024;

// The evaluated pairs' map records, numbered densely group after group (record a of the narrow
// phase) and found at their place in the groups' regions of pairMap.
struct EvalRecords {
  int n;  // records this step
  int pre[8], base[8];
  // position in pairMap of record a (0 <= a < n): the last group whose first record is at or below a
  // (a group without records has the same first record as the next one)
  __device__ __forceinline__ int at(int a) const {
    int pos = a;
#pragma unroll
    for (int g = 0; g < 8; ++g)
      if (a >= pre[g]) pos = base[g] + (a - pre[g]);
    return pos;
  }
};
__device__ __forceinline__ int checkTiles(const Params& P) {
  const int nPairs = min(P.pairStart[P.nP], (int)P.pairCap);
  return (nPairs + kCheckTile - 1) / kCheckTile;
}
__device__ __forceinline__ int groupFirstTile(int nTiles, int g) { return int((long long)nTiles * g / 8); }
__device__ __forceinline__ EvalRecords evalRecords(const Params& P) {
  EvalRecords e;
  const int nTiles = checkTiles(P);
  int sum = 0;
#pragma unroll
  for (int g = 0; g < 8; ++g) {
    e.pre[g] = sum;
    e.base[g] = groupFirstTile(nTiles, g) * kCheckTile;
    sum += P.evalCount[g * kBarrierLine];
  }
  e.n = sum;
  return e;
}
template <bool SEC>
__global__ __launch_bounds__(256) void k_pair_check(Params P, int full) {
  __shared__ int sList[kCheckTile], sA[kCheckTile];
  __shared__ uint32_t sInfo[kCheckTile];
  PairF32* out = reinterpret_cast<PairF32*>(P.pairMap);
  __shared__ int sCount, sBase;
  LSD_HALT(P);  // (a failed broad phase leaves no pair list to read)
  // the broad phase of this step is over: its grid barrier's words start the next one at zero
  if (blockIdx.x == 0 && threadIdx.x < 10) P.barrier[threadIdx.x * kBarrierLine] = 0;
  const int nPairs = min(P.pairStart[P.nP], (int)P.pairCap);
  const bool rebuilt = P.devInt[rebuildFlagIdx(P.curBuf)] != 0;  // a new pair list this step
  const bool fresh = !P.useCert;  // (a new list holds the certificates it inherited)
  const int lane = threadIdx.x & 63;
  const unsigned long long ltMask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  // XCD-aware tiles: workgroups go to the 8 XCDs in turn (blockIdx mod 8), so each XCD takes one
  // contiguous eighth of the pair list and its partners' data stays in that XCD's L2
  const int nTiles = (nPairs + kCheckTile - 1) / kCheckTile;
  const int xcd = int(blockIdx.x) & 7, perX = (int(gridDim.x) - xcd + 7) >> 3;  // blocks on this XCD
  const int tLo = groupFirstTile(nTiles, xcd), tHi = groupFirstTile(nTiles, xcd + 1);
  const int region = tLo * kCheckTile;  // this group's records: at most its pairs, from its first pair on
  unsigned int evNodes = 0;  // boundary nodes of the masters of this thread's evaluated pairs
  for (int t = tLo + (int(blockIdx.x) >> 3); t < tHi; t += perX) {
    const int t0 = t * kCheckTile;
    if (threadIdx.x == 0) sCount = 0;
    __syncthreads();
    // the tile's pairs of this thread, with the loads of one kind issued together
    constexpr int R = kCheckTile / 256;
    // (vector types: an element picked by a runtime index stays in registers, where an array's
    // would be read from scratch)
    typedef int IntR __attribute__((ext_vector_type(R)));
    typedef uint32_t UintR __attribute__((ext_vector_type(R)));
    IntR pv, Av, jv;
    UintR tv;
    int key[R];
    float4 cv[R], mB[R], mA[R];
#pragma clang loop unroll(full)
    for (int r = 0; r < R; ++r) {
      pv[r] = t0 + r * 256 + int(threadIdx.x);
      const bool in = pv[r] < nPairs;
      const uint32_t info = in ? P.pairInfo[pv[r]] : (uint32_t(PT_PW) << 28);
      tv[r] = info >> 28;
      jv[r] = int(info & kPartnerMask);
      Av[r] = in ? P.pairA[pv[r]] : 0;
      cv[r] = (in && !fresh) ? P.cert[pv[r]] : make_float4(-1.f, 0.f, 0.f, -1.f);
    }
#pragma clang loop unroll(full)
    for (int r = 0; r < R; ++r) {
      // the partner's step record (liveness key with the image's flip, motion): one 32-byte read
      if (pv[r] >= nPairs) {
        key[r] = -1;
        mB[r] = make_float4(0.f, 0.f, 0.f, 0.f);
      } else if (tv[r] == PT_PW) {
        key[r] = P.wkey[jv[r]];
        mB[r] = P.wMotion[jv[r]];
      } else {
        const BodyStep sb = P.step[jv[r]];
        key[r] = partnerKey(sb.key, tv[r]);
        mB[r] = sb.motion;
      }
      mA[r] = (cv[r].x > 0.f || cv[r].w > 0.f) ? P.step[Av[r]].motion : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    // A pair whose bound against the motion since the build is spent (c.x <= 0) is checked against
    // the motion since its evaluation whatever the partner's motion: the inputs of that bound are
    // read now, with the partner's record, for the first such pair of the thread (about 1 % of the
    // pairs are; a thread's others read them one round trip later).
    int early = -1;
#pragma clang loop unroll(full)
    for (int r = R - 1; r >= 0; --r)
      if (pv[r] < nPairs && cv[r].x <= 0.f && cv[r].w > 0.f) early = r;
    float4 e0e, qa0e, qb0e, qae, qbe;
    if (early >= 0) {
      const int p = pv[early], A = Av[early], j = jv[early];
      e0e = P.certE[p];
      qa0e = P.certQA[p];
      qb0e = P.certQB[p];
      qae = P.qrel[A];
      qbe = tv[early] == PT_PW ? P.wQrel[j] : P.qrel[j];
    }
    bool actv[R];
#pragma clang loop unroll(full)
    for (int r = 0; r < R; ++r) {
      const int p = pv[r];
      const bool live = key[r] >= 0;
      bool act = live;
      // an audit step of the surrogate (full) trusts only certificates of the exact culls; the
      // turned images of sector periodicity are evaluated every step (their motion bound would need
      // the turn)
      if (live && !fresh && (cv[r].x > 0.f || cv[r].w > 0.f) && !(full && signbit(cv[r].z)) &&
          !(SEC && tv[r] >= PT_GX)) {
        float4 c = cv[r];
        c.z = fabsf(c.z);
        const bool flip = tv[r] >= PT_GX && ((key[r] >> kFlipBit) & 1);
        const float ex = mA[r].x - mB[r].x, ey = mA[r].y - mB[r].y, ez = mA[r].z - mB[r].z;
        if (!flip && c.x > 0.f) {
          // against the motion since the build (the slack net of the motion at the evaluation); a
          // pair that fails it is evaluated, which issues it a new certificate (checking it against
          // the motion since the evaluation instead would cost one more round trip)
          const float bound = (sqrtf(ex * ex + ey * ey + ez * ez) + mB[r].w * c.y + mA[r].w * c.z) * (1.f + 1e-5f);
          act = !(bound < c.x);
        } else if (c.x <= 0.f && c.w > 0.f) {
          // against the motion since the evaluation: the same bound with the displacement and the
          // rotations taken from that step
          float4 e0, qa0, qb0, qa, qb;
          if (r == early) {
            e0 = e0e; qa0 = qa0e; qb0 = qb0e; qa = qae; qb = qbe;
          } else {
            e0 = P.certE[p];
            qa0 = P.certQA[p];
            qb0 = P.certQB[p];
            qa = P.qrel[Av[r]];
            qb = tv[r] == PT_PW ? P.wQrel[jv[r]] : P.qrel[jv[r]];
          }
          if ((e0.w != 0.f) == flip) {
            const float fx = ex - e0.x, fy = ey - e0.y, fz = ez - e0.z;
            const float chA = relChordF(qa0, qa), chB = relChordF(qb0, qb);
            const float bound = (sqrtf(fx * fx + fy * fy + fz * fz) + chB * c.y + chA * c.z) * (1.f + 1e-5f);
            act = !(bound < c.w);
          }
        }
      }
      if (p < nPairs) {
        if (!live && rebuilt) P.cert[p] = make_float4(-1.f, 0.f, 0.f, -1.f);
        // a pair skipped under a certificate has had no contacts since the evaluation that issued
        // it (a pair with candidate nodes gets no certificate); others drop last step's count
        // (on a new list every count is the old list's)
        if (!act && (rebuilt || !(cv[r].x > 0.f || cv[r].w > 0.f))) P.pairCcnt[p] = 0;
      }
      actv[r] = act && p < nPairs;
    }
#pragma clang loop unroll(full)
    for (int r = 0; r < R; ++r) {
      const unsigned long long m = __ballot(actv[r]);
      int first = 0;
      if (lane == 0 && m) first = atomicAdd(&sCount, __popcll(m));
      first = __shfl(first, 0, 64);
      if (actv[r]) {
        const int slot = first + __popcll(m & ltMask);
        sList[slot] = pv[r];
        sInfo[slot] = (tv[r] << 28) | uint32_t(jv[r]);
        sA[slot] = Av[r];
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) sBase = sCount ? region + atomicAdd(&P.evalCount[xcd * kBarrierLine], sCount) : 0;
    __syncthreads();
    for (int k = threadIdx.x; k < sCount; k += 256) {
      const int p = sList[k], a = sBase + k;
      const PairView v = loadPair<SEC>(P, p, sInfo[k], sA[k]);
      // the motion state of this evaluation, for the pair's certificate against later motion
      // (loaded before the stores below)
      const BodyStep sa = P.step[v.A];
      const float4 qA = P.qrel[v.A];
      float4 mB, qB;
      float flip = 0.f;
      if (v.type == PT_PW) {
        mB = P.wMotion[v.partner];
        qB = P.wQrel[v.partner];
      } else {
        const BodyStep sb = P.step[v.partner];
        mB = sb.motion;
        qB = P.qrel[v.partner];
        if (v.type >= PT_GX) flip = float((partnerKey(sb.key, v.type) >> kFlipBit) & 1);
      }
      PairF32 f;
      makeF32(P, v, f);
      out[a] = f;
      P.pairRec[p] = a;
      evNodes += unsigned(P.nodeStart[v.A + 1] - v.nodeBase);
      P.certE[p] = make_float4(sa.motion.x - mB.x, sa.motion.y - mB.y, sa.motion.z - mB.z, flip);
      P.certQA[p] = qA;
      P.certQB[p] = qB;
    }
    __syncthreads();
  }
  // (only threads that wrote map records count: a few thousand atomics per step at 200k particles)
  if (__ballot(evNodes != 0)) {  // (one atomic per wavefront)
    const unsigned long long sum = (unsigned long long)wsum(double(evNodes));
    if (lane == 0) atomicAdd(statSlot(P, 4, blockIdx.x & (kStatSlots - 1)), sum);
  }
}

// The FP32 prefilters may contract a * b + c into one rounding: their margins bound the rounding
// of separate operations, and a fused operation rounds less.  (The FP64 membership tests keep the
// file-wide -ffp-contract=off.)
#define FP32_CONTRACT _Pragma("clang fp contract(fast)")

__device__ __forceinline__ void mapF32(const PairF32& f, float nx, float ny, float nz, float& gx, float& gy, float& gz) {
  FP32_CONTRACT
  gx = f.m[0] * nx + f.m[1] * ny + f.m[2] * nz + f.t[0];
  gy = f.m[3] * nx + f.m[4] * ny + f.m[5] * nz + f.t[1];
  gz = f.m[6] * nx + f.m[7] * ny + f.m[8] * nz + f.t[2];
}

// FP32 trilinear SDF value (unscaled) at grid coordinates g, clamped into the grid; a prefilter
// only.  Its rounding error is below 1e-6 of the largest |corner|, which the caller adds (kRelPhi).
constexpr float kRelPhi = 4e-6f;
__device__ __forceinline__ float phiF32(const float* __restrict__ sdf, const PairF32& f, float gx, float gy, float gz,
                                        float& cmax) {
  FP32_CONTRACT
  const int i0 = min(max(int(floorf(gx)), 0), f.nx - 2), j0 = min(max(int(floorf(gy)), 0), f.ny - 2),
            k0 = min(max(int(floorf(gz)), 0), f.nz - 2);
  const float x = fminf(fmaxf(gx - float(i0), 0.f), 1.f), y = fminf(fmaxf(gy - float(j0), 0.f), 1.f),
              z = fminf(fmaxf(gz - float(k0), 0.f), 1.f);
  const int sx = 1, sy = f.nx, sz = f.nx * f.ny;
  const float* g = sdf + f.sdfOff + i0 + j0 * sy + k0 * sz;
  const float c000 = g[0], c100 = g[sx], c010 = g[sy], c110 = g[sx + sy];
  const float c001 = g[sz], c101 = g[sx + sz], c011 = g[sy + sz], c111 = g[sx + sy + sz];
  cmax = fmaxf(fmaxf(fmaxf(fabsf(c000), fabsf(c100)), fmaxf(fabsf(c010), fabsf(c110))),
               fmaxf(fmaxf(fabsf(c001), fabsf(c101)), fmaxf(fabsf(c011), fabsf(c111))));
  const float a = (c000 * (1.f - x) + c100 * x) * (1.f - y) + (c010 * (1.f - x) + c110 * x) * y;
  const float b = (c001 * (1.f - x) + c101 * x) * (1.f - y) + (c011 * (1.f - x) + c111 * x) * y;
  return a * (1.f - z) + b * z;
}

// Cluster acceptance (conservative): the cluster's bounding sphere must reach the partner's
// bounding sphere (particle partners) and grid box, and, when its centre lies in the box, the
// partner's SDF at the centre must be within Lipschitz reach of zero.
//
// Every test below is a function of the g coordinates of A's nodes alone (the partner's sphere,
// box and SDF are fixed in its frame).  When a test culls, `slack` (g units) is how far the
// cluster's radius could grow with the same test still culling it: while no node has moved
// further than that in the partner's frame, the test still rules out every node of the cluster.
// Slacks are rounded down; a cull that holds by less than the rounding gives a slack <= 0.

// A slack is rounded down by kSlackRound, which covers the FP32 rounding of computing it (a few
// ulps of the distances involved).  The rounding of the map itself needs no allowance: every
// threshold already contains it (the margins abs, phiMargin), so a test that fails with a radius
// grown by the slack fails for every exact point of the grown ball.
constexpr float kSlackRound = 1.f - 2e-6f;

// Geometric part: bounding spheres and grid box.  `inside` reports whether the centre lies in the
// grid box (then an SDF bound at the centre can cull the cluster).
__device__ __forceinline__ bool clusterGeomF32(const PairF32& f, const ClusterF32& cl, float& gx, float& gy, float& gz,
                                               float& rho, bool& inside, float& slack) {
  FP32_CONTRACT
  mapF32(f, cl.cx, cl.cy, cl.cz, gx, gy, gz);
  rho = cl.rho * f.rhoScale;
  inside = false;
  slack = 0.f;
  if (f.r2 >= 0.f) {
    const float dx = gx - f.c[0], dy = gy - f.c[1], dz = gz - f.c[2];
    const float d2 = dx * dx + dy * dy + dz * dz;
    const float r = sqrtf(f.r2) + rho;
    if (d2 > r * r * 1.0001f) {
      slack = sqrtf(d2) * kSlackRound - r * 1.00005f;
      return false;
    }
  }
  const float dx = fmaxf(0.f, fmaxf(f.lo - gx, gx - f.hi[0]));
  const float dy = fmaxf(0.f, fmaxf(f.lo - gy, gy - f.hi[1]));
  const float dz = fmaxf(0.f, fmaxf(f.lo - gz, gz - f.hi[2]));
  const float d2 = dx * dx + dy * dy + dz * dz;
  if (d2 > rho * rho * 1.0001f) {
    slack = sqrtf(d2) * kSlackRound - rho * 1.00005f;
    return false;
  }
  inside = dx == 0.f && dy == 0.f && dz == 0.f;
  return true;
}
// Radius growth that a Lipschitz bound phi > lipH * 1.001 * radius + lim still allows.
__device__ __forceinline__ float lipSlack(const PairF32& f, float phi, float lim, float rho) {
  FP32_CONTRACT
  // hardware reciprocal (1 ulp), covered by kSlackRound
  return f.lipH > 0.f ? (phi - lim) * __builtin_amdgcn_rcpf(f.lipH * 1.001f) * kSlackRound - rho : 0.f;
}
// SDF part with the exact grid: phi(centre) beyond Lipschitz reach of zero culls the cluster.
__device__ __forceinline__ bool clusterSdfF32(const float* __restrict__ sdf, const PairF32& f, float gx, float gy,
                                              float gz, float rho, float& slack) {
  FP32_CONTRACT
  float cmax;
  const float phi = phiF32(sdf, f, gx, gy, gz, cmax);
  const float lim = f.phiMargin + kRelPhi * cmax;
  if (phi > f.lipH * rho * 1.001f + lim) {
    slack = lipSlack(f, phi, lim, rho);
    return false;
  }
  return true;
}
__device__ __forceinline__ bool clusterF32(const float* __restrict__ sdf, const PairF32& f, const ClusterF32& cl,
                                           bool sdfCull) {
  float gx, gy, gz, rho, slack;
  bool inside;
  if (!clusterGeomF32(f, cl, gx, gy, gz, rho, inside, slack)) return false;
  return !inside || !sdfCull || clusterSdfF32(sdf, f, gx, gy, gz, rho, slack);
}
// Surrogate part: the same bound with the learned distance, widened by the band (twice the
// surrogate's measured maximum error).  `d` is the network output in normalised units.
__device__ __forceinline__ bool clusterSurKeep(const Params& P, const PairF32& f, float d, float rho) {
  const float phi = d * (1.0f / P.sur.invLen);
  return !(phi - P.surBand > f.lipH * rho * 1.001f + f.phiMargin);
}
__device__ __forceinline__ float clusterSurSlack(const Params& P, const PairF32& f, float d, float rho) {
  return lipSlack(f, d * (1.0f / P.sur.invLen), P.surBand + f.phiMargin, rho);
}
// Lane's surrogate query for a cluster centre at grid coordinates g (normalised local point).
__device__ __forceinline__ void surQueryPoint(const Params& P, const PairF32& f, float gx, float gy, float gz,
                                              float& ux, float& uy, float& uz) {
  ux = (f.o[0] + gx * f.h) * P.sur.invLen;
  uy = (f.o[1] + gy * f.h) * P.sur.invLen;
  uz = (f.o[2] + gz * f.h) * P.sur.invLen;
}
// Node acceptance (conservative superset of the exact FP64 test); slack as for clusters, with
// the node as a cluster of radius 0.
__device__ __forceinline__ bool nodeF32(const float* __restrict__ sdf, const PairF32& f, float4 n, float& slack) {
  FP32_CONTRACT
  float gx, gy, gz;
  mapF32(f, n.x, n.y, n.z, gx, gy, gz);
  slack = 0.f;
  if (f.r2 >= 0.f) {
    const float dx = gx - f.c[0], dy = gy - f.c[1], dz = gz - f.c[2];
    const float d2 = dx * dx + dy * dy + dz * dz;
    if (d2 > f.r2) {
      slack = sqrtf(d2) * kSlackRound - sqrtf(f.r2);
      return false;
    }
  }
  const float out = fmaxf(fmaxf(fmaxf(f.lo - gx, gx - f.hi[0]), fmaxf(f.lo - gy, gy - f.hi[1])),
                          fmaxf(f.lo - gz, gz - f.hi[2]));
  if (out > 0.f) {  // outside the grid box on some axis, by `out`
    slack = out * kSlackRound;
    return false;
  }
  float cmax;
  const float phi = phiF32(sdf, f, gx, gy, gz, cmax);
  const float lim = f.phiMargin + kRelPhi * cmax;
  if (phi <= lim) return true;
  slack = lipSlack(f, phi, lim, 0.f);
  return false;
}

// nodeF32 in three stages, so that the grid reads of several nodes can be in flight together:
// the sphere and box tests (slack when they cull), the eight corner reads, the SDF test.
__device__ __forceinline__ bool nodeGeomF32(const PairF32& f, float4 n, float& gx, float& gy, float& gz, float& slack) {
  FP32_CONTRACT
  mapF32(f, n.x, n.y, n.z, gx, gy, gz);
  slack = 0.f;
  if (f.r2 >= 0.f) {
    const float dx = gx - f.c[0], dy = gy - f.c[1], dz = gz - f.c[2];
    const float d2 = dx * dx + dy * dy + dz * dz;
    if (d2 > f.r2) {
      slack = sqrtf(d2) * kSlackRound - sqrtf(f.r2);
      return false;
    }
  }
  const float out = fmaxf(fmaxf(fmaxf(f.lo - gx, gx - f.hi[0]), fmaxf(f.lo - gy, gy - f.hi[1])),
                          fmaxf(f.lo - gz, gz - f.hi[2]));
  if (out > 0.f) {
    slack = out * kSlackRound;
    return false;
  }
  return true;
}
__device__ __forceinline__ void sdfCorners(const float* __restrict__ sdf, const PairF32& f, float gx, float gy, float gz,
                                           float (&c)[8], float& x, float& y, float& z) {
  FP32_CONTRACT
  const int i0 = min(max(int(floorf(gx)), 0), f.nx - 2), j0 = min(max(int(floorf(gy)), 0), f.ny - 2),
            k0 = min(max(int(floorf(gz)), 0), f.nz - 2);
  x = fminf(fmaxf(gx - float(i0), 0.f), 1.f);
  y = fminf(fmaxf(gy - float(j0), 0.f), 1.f);
  z = fminf(fmaxf(gz - float(k0), 0.f), 1.f);
  const int sx = 1, sy = f.nx, sz = f.nx * f.ny;
  const float* g = sdf + f.sdfOff + i0 + j0 * sy + k0 * sz;
  c[0] = g[0]; c[1] = g[sx]; c[2] = g[sy]; c[3] = g[sx + sy];
  c[4] = g[sz]; c[5] = g[sx + sz]; c[6] = g[sy + sz]; c[7] = g[sx + sy + sz];
}
// Same arithmetic as phiF32 and the SDF part of nodeF32.
__device__ __forceinline__ bool nodeSdfF32(const PairF32& f, const float (&c)[8], float x, float y, float z,
                                           float& slack) {
  FP32_CONTRACT
  const float cmax = fmaxf(fmaxf(fmaxf(fabsf(c[0]), fabsf(c[1])), fmaxf(fabsf(c[2]), fabsf(c[3]))),
                           fmaxf(fmaxf(fabsf(c[4]), fabsf(c[5])), fmaxf(fabsf(c[6]), fabsf(c[7]))));
  const float a = (c[0] * (1.f - x) + c[1] * x) * (1.f - y) + (c[2] * (1.f - x) + c[3] * x) * y;
  const float b = (c[4] * (1.f - x) + c[5] * x) * (1.f - y) + (c[6] * (1.f - x) + c[7] * x) * y;
  const float phi = a * (1.f - z) + b * z;
  const float lim = f.phiMargin + kRelPhi * cmax;
  if (phi <= lim) return true;
  slack = lipSlack(f, phi, lim, 0.f);
  return false;
}

// Exact test and contact law of one candidate slot: the reference's node-vs-SDF membership in
// FP64, then (for contacts) tangential history and the linear spring-dashpot.
// (p, node, pInfo, pA: the slot's pair and node, and that pair's pairInfo and pairA)
// grid: the partner's grid size and SDF offset when the caller holds the pair's map record (which
// spares the read of the partner's shape record), else x < 0.  (Values, not a pointer to the record:
// a pointer that is either to LDS or to global memory makes flat loads.)
template <bool SEC>
__device__ void exactSlot(const Params& P, int c, int p, int node, uint32_t pInfo, int pA, int4 grid,
                          unsigned long long& statQueries, unsigned long long& statContacts) {
  if (p < 0) {
    P.cInfo[c] = kInvalidInfo;
    P.hashSlot[c] = -1;
    return;
  }
  const PairView v = loadPair<SEC>(P, p, pInfo, pA);
  // everything the contact law and the tangential history need is read here, before the exact
  // test, so that its round trips overlap the test's: a candidate that is no contact wastes these
  // reads, a contact saves their latency
  const NodeDev nd = P.nodes[node];
  const uint32_t type = v.type;
  const int j = v.partner;
  const int A = v.A;
  const d3 riA = P.pos[A];
  const d3 vA = P.vel[A], wA = P.angvel[A];
  const double invMA = P.invM[A], knA = P.kn[A], ksA = P.ks[A], muA = P.mu[A], resA = P.res[A];
  d3 xf, vB, wB;
  double knB, ksB, invMB, muB, resB;
  if (type == PT_PW) {
    xf = P.wpos[j]; vB = P.wvel[j]; wB = P.wangvel[j];
    knB = ksB = invMB = 0.;
    muB = P.wmu[j];
    resB = P.wres[j];
  } else {
    xf = (type == PT_PP) ? P.pos[j] : imgPos<SEC>(P, int(type - PT_GX))[j];
    // an image's velocity (a turned image's is turned), the real body's angular velocity
    // (contactKernel.cu:236-238)
    vB = (type != PT_PP && SEC) ? imgVel(P, int(type - PT_GX))[j] : P.vel[j];
    wB = P.angvel[j];
    knB = P.kn[j];
    ksB = P.ks[j];
    invMB = P.invM[j];
    muB = P.mu[j];
    resB = P.res[j];
  }
  // tangential history of the same (list, node, partner) from the previous step (its first slot
  // read now); this step's contact goes into this step's table for the next one
  const unsigned long long key = histKey(type, v.nodeBase + nd.orig, j);
  const unsigned long long slot0 = mix64(key) & P.hashMask;
  const unsigned long long k0 = P.prevHashKey[slot0];
  const int v0 = P.prevHashVal[slot0];
  NodeHit h;
  unsigned long long tests = 0;
  i3 nB;
  long long sdfOff;
  if (grid.x >= 0) {
    nB = i3{grid.x, grid.y, grid.z};
    sdfOff = grid.w;
  } else {
    const ShapeDev SB = P.shapes[v.shapeB];
    nB = i3{SB.nx, SB.ny, SB.nz};
    sdfOff = SB.sdfOff;
  }
  const bool hit = exactTest(P, v, nd, nB, sdfOff, true, h, tests, statQueries);
  if (!hit) {
    P.cInfo[c] = kInvalidInfo;
    P.hashSlot[c] = -1;
    return;
  }
  statContacts++;
  double kn = 0., ks = 0., effM = 0., mu, res;
  if (type == PT_PW) {
    kn = knA; ks = ksA;
    if (invMA > 0.) effM = 1. / invMA;
    mu = fmin(muA, muB);
    res = fmin(resA, resB);
  } else {
    if (knA > 0. && knB > 0.) kn = knA * knB / (knA + knB);
    if (ksA > 0. && ksB > 0.) ks = ksA * ksB / (ksA + ksB);
    if (invMA > 0. || invMB > 0.) effM = 1. / (invMA + invMB);
    res = fmin(resA, resB);
    mu = fmin(muA, muB);
  }
  d3 spring = histLookup(P, key, slot0, k0, v0);
  // (the slot's stores after every load above)
  P.cInfo[c] = v.info;
  P.cNode[c] = h.gnode;
  P.cPoint[c] = h.point;
  P.cNormal[c] = h.normal;
  P.cOverlap[c] = h.ov;
  P.hashSlot[c] = histInsert(P, key, c);
  const d3 vc = vA + cross(wA, h.point - riA) - (vB + cross(wB, h.point - xf));
  d3 F = mk3(0., 0., 0.);
  if (P.prime) {
    // the pre-loop build only lists the contact with the spring it found
    // (LSParticleContactDetectionKernel.cu:280-289); the force kernels do not run
    P.cSpring[c] = spring;
    P.cEn[c] = 0.;
    P.cEs[c] = 0.;
    P.cF[c] = F;
    P.cTA[c] = F;
    P.cTB[c] = F;
    return;
  }
  const ContactLaw law = linearContact(spring, vc, h.normal, h.ov, P.dt, kn, ks, mu, res, effM);
  F = law.F;
  spring = law.spring;
  P.cSpring[c] = spring;
  P.cEn[c] = 0.5 * h.ov * h.ov * kn;
  P.cEs[c] = 0.5 * len2(spring) * ks;
  P.cF[c] = F;
  P.cTA[c] = cross(h.point - riA, F);
  P.cTB[c] = cross(h.point - xf, -F);
}

constexpr int kBatch = 8;  // candidate pairs per wave batch
constexpr int kMaxSub = 4;  // sub-clusters per cluster (64 / 16)
constexpr int kSlotStage = 128;  // candidate slots of a batch handed to the exact test through LDS

// Per-wave LDS of k_candidates: pair maps, surviving (pair, cluster) entries, surviving
// (pair, sub-cluster) entries with their 16-bit node masks, per-pair counts, offsets and slacks.
__host__ __device__ inline int candWaveBytes(int maxClusters) {
  const int nEnt = kBatch * maxClusters, nSub = kMaxSub * nEnt;
  const int b = int(sizeof(PairF32)) * kBatch + 8 * nEnt + 8 * nSub + 4 * nSub + 12 * kBatch + 16 + 8 * kBatch +
                8 * kSlotStage;
  return (b + 15) & ~15;
}
size_t candidatesLdsBytes(int maxClusters) { return size_t(candWaveBytes(maxClusters)) * (kNarrowBlock / 64); }

// Value of another lane of the same 16-lane row (every lane active).
template <int CTRL>
__device__ __forceinline__ float dppMove(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
// Minimum over aligned groups of W lanes (4, 8 or 16).
template <int W>
__device__ __forceinline__ float groupMin(float v) {
  // lane exchanges within 16-lane rows (DPP): quad permutations, half-row mirror, row rotations
  static_assert(W == 4 || W == 8 || W == 16, "group of 4, 8 or 16 lanes");
  v = fminf(v, dppMove<0xB1>(v));  // quad_perm [1,0,3,2]
  v = fminf(v, dppMove<0x4E>(v));  // quad_perm [2,3,0,1]: min of the quad
  if (W == 8) v = fminf(v, dppMove<0x141>(v));  // row_half_mirror: the other quad of the 8
  if (W == 16) {
    v = fminf(v, dppMove<0x124>(v));  // row_ror 4
    v = fminf(v, dppMove<0x128>(v));  // row_ror 8
  }
  return v;
}

// Candidate nodes: one wavefront per batch of 8 evaluated pairs, three levels of conservative
// FP32 culls.
//   1. lanes = (pair, cluster of <= 64 nodes): bounding spheres, grid box, SDF Lipschitz bound at
//      the cluster centre (or the surrogate's decision); survivors become entries;
//   2. lanes = (entry, sub-cluster of <= 16 nodes), 16 entries per pass: the same tests with the
//      sub-cluster's smaller sphere; survivors become sub-entries;
//   3. lanes = (sub-entry, node), 4 sub-entries per pass: the node prefilter.
// The batch reserves its exact candidate count with one atomic and writes each pair's candidates
// to a contiguous range, in (cluster, sub-cluster, node) order.  Every node of a pair is ruled
// out by exactly one cull (cluster, sub-cluster or node) or becomes a candidate; the smallest
// slack of those culls, less the rounding of the map, is the pair's separation certificate
// (none when a node is a candidate).
template <bool SUR, bool SEC>
__global__ __launch_bounds__(kNarrowBlock) void k_candidates(Params P) {
  extern __shared__ __align__(16) unsigned char candLds[];
  const int lane = threadIdx.x & 63, wl = threadIdx.x >> 6;
  const int maxC = P.maxClusters;
  const int entCap = kBatch * maxC, subCap = kMaxSub * entCap;
  unsigned char* wb = candLds + wl * candWaveBytes(maxC);
  PairF32* sF = reinterpret_cast<PairF32*>(wb);
  int2* sEnt = reinterpret_cast<int2*>(wb + sizeof(PairF32) * kBatch);  // (first sub-cluster, count | pair j << 4)
  int2* sSub = sEnt + entCap;                                            // (first node, count | j << 16)
  unsigned int* sMask = reinterpret_cast<unsigned int*>(sSub + subCap);  // 16-bit node masks
  int* sCnt = reinterpret_cast<int*>(sMask + subCap);
  int* sOff = sCnt + kBatch;
  unsigned int* sSlack = reinterpret_cast<unsigned int*>(sOff + kBatch);  // bits of a float >= 0
  int* sBase4 = reinterpret_cast<int*>(sSlack + kBatch);  // write offsets of the 4 sub-entries of a pass
  uint32_t* sPInfo = reinterpret_cast<uint32_t*>(sBase4 + 4);  // pairInfo and pairA of the batch's pairs
  int* sPA = reinterpret_cast<int*>(sPInfo + kBatch);
  int2* sStage = reinterpret_cast<int2*>(sPA + kBatch);  // (batch pair, node) of the first slots
  const int q = lane >> 4, b16 = lane & 15;
  const unsigned int lt16 = (1u << b16) - 1u;
  const float kInf = __builtin_huge_valf();
  LSD_HALT(P);
  unsigned long long statCand = 0, statEnt = 0, statQueries = 0, statContacts = 0;
  const EvalRecords rec = evalRecords(P);
  const int nAct = rec.n;
  const int wave = __builtin_amdgcn_readfirstlane(int(blockIdx.x * (kNarrowBlock / 64) + wl));
  const int nWaves = int(gridDim.x * (kNarrowBlock / 64));
  if (wave == 0 && lane == 0) P.devStat[13] += (unsigned long long)nAct;  // evaluated pairs, summed over steps
  const int iters = (maxC + 7) >> 3;
  const int j = lane >> 3, k = lane & 7;
  const unsigned long long ltMask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  // batch bi takes the evaluated pairs bi, bi + NB, bi + 2 NB, ...: pairs near contact come in
  // spatial runs, and spreading them over the batches keeps the slowest wavefront short
  // pairs per batch: when few pairs are evaluated, small batches spread the slowest pairs (those
  // near contact, with many surviving clusters) over more wavefronts (one pair per wavefront when
  // there are at least eight wavefronts per pair, as in small simulations)
  const int nb = P.narrowBatch > 0 ? P.narrowBatch
                                   : (nAct <= nWaves / 8 ? 1 : (nAct <= nWaves / 2 ? 2 : (nAct <= nWaves ? 4 : kBatch)));
  const int NB = (nAct + nb - 1) / nb;
  for (int bi = wave; bi < NB; bi += nWaves) {
    // ---- level 1: clusters ----
    const int a = bi + j * NB;  // map record
    const bool own = j < nb && a < nAct;
    // the 8 lanes of a pair copy its 160-byte map record into LDS, 20 bytes each
    if (own) {
      const unsigned int* src = reinterpret_cast<const unsigned int*>(P.pairMap) + (long long)rec.at(a) * 40 + 5 * k;
      unsigned int* dst = reinterpret_cast<unsigned int*>(&sF[j]) + 5 * k;
#pragma unroll
      for (int w = 0; w < 5; ++w) dst[w] = src[w];
    }
    if (lane < kBatch) sCnt[lane] = 0;
    __builtin_amdgcn_wave_barrier();
    const PairF32& f = sF[j];
    const int clusterOff = own ? f.clusterOff : 0, nCl = own ? f.nClusters : 0;
    // the pair's list entries for the exact test, loaded now and first used at its end
    uint32_t myInfo = 0;
    int myA = 0;
    if (own && k == 0) {
      myInfo = P.pairInfo[f.pair];
      myA = P.pairA[f.pair];
    }
    int E = 0;
    float lmin = kInf;  // smallest slack of this lane's culled clusters
    for (int it = 0; it < iters; ++it) {
      const int c = it * 8 + k;
      bool keep = false;
      int subFirst = 0, nSub = 0;
      if (own && c < nCl) {
        const ClusterF32 cl = P.clusters32[clusterOff + c];
        subFirst = cl.subFirst;
        nSub = cl.nSub;
        float gx, gy, gz, rho, slack;
        bool inside;
        keep = clusterGeomF32(f, cl, gx, gy, gz, rho, inside, slack);
        if (keep && inside) {
          if (SUR && f.slot >= 0) keep = (P.surMask[a] >> c) & 1ull;  // slack: surSlack
          else if (P.sdfCull) keep = clusterSdfF32(P.sdf, f, gx, gy, gz, rho, slack);
          if (!keep && SUR && f.slot >= 0) slack = kInf;
        }
        if (!keep) lmin = fminf(lmin, fmaxf(slack, 0.f));
      }
      const unsigned long long m = __ballot(keep);
      // (the entry carries the cluster's sub-cluster range: level 2 reads no cluster record)
      if (keep) sEnt[E + __popcll(m & ltMask)] = make_int2(subFirst, nSub | (j << 4));
      E += __popcll(m);
    }
    lmin = groupMin<8>(lmin);
    if (k == 0) sSlack[j] = __float_as_uint(lmin);
    statEnt += (lane == 0) ? (unsigned long long)E : 0ull;
    __builtin_amdgcn_wave_barrier();
    // ---- level 2: sub-clusters, 16 entries x 4 per pass ----
    int NS = 0;
    for (int e0 = 0; e0 < E; e0 += 16) {
      const int e = e0 + (lane >> 2), u = lane & 3;
      bool keep = false;
      float smin = kInf;
      int first = 0, count = 0, jj = 0;
      if (e < E) {
        const int2 en = sEnt[e];
        jj = en.y >> 4;
        if (u < (en.y & 15)) {
          const ClusterF32 sc = P.subs32[en.x + u];
          const PairF32& fj = sF[jj];
          float gx, gy, gz, rho, slack;
          bool inside;
          keep = clusterGeomF32(fj, sc, gx, gy, gz, rho, inside, slack);
          // with the surrogate the network has decided on the whole cluster; the sub-cluster keeps
          // only its geometric tests then
          if (keep && inside && !(SUR && fj.slot >= 0) && P.sdfCull)
            keep = clusterSdfF32(P.sdf, fj, gx, gy, gz, rho, slack);
          if (!keep) smin = fmaxf(slack, 0.f);
          first = fj.nodeOff + sc.first;
          count = sc.count;
        }
      }
      smin = groupMin<4>(smin);  // the 4 lanes of an entry share its pair
      if (e < E && (lane & 3) == 0 && smin < kInf) atomicMin(&sSlack[jj], __float_as_uint(smin));
      const unsigned long long m = __ballot(keep);
      if (keep) sSub[NS + __popcll(m & ltMask)] = make_int2(first, count | (jj << 16));
      NS += __popcll(m);
    }
    __builtin_amdgcn_wave_barrier();
    // ---- level 3: nodes, 4 sub-entries x 16 per pass, two passes per iteration: the node loads,
    // then the grid reads of the nodes that pass the sphere and box tests, are issued for both
    // passes before either is used; the next iteration's node loads are issued before this
    // iteration's tests ----
    int total = 0;
    int jjn[2];
    bool liven[2];
    float4 ndn[2];
    auto loadNodes = [&](int e0) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int e = e0 + 4 * h + q;
        jjn[h] = 0;
        liven[h] = false;
        ndn[h] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (e < NS) {
          const int2 se = sSub[e];
          jjn[h] = se.y >> 16;
          liven[h] = b16 < (se.y & 0xFFFF);
          if (liven[h]) ndn[h] = P.nodes32[se.x + b16];
        }
      }
    };
    if (NS > 0) loadNodes(0);
    for (int e0 = 0; e0 < NS; e0 += 8) {
      int jjv[2];
      bool live[2], geo[2];
      float4 nd[2];
      float gx[2], gy[2], gz[2], sl[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        jjv[h] = jjn[h];
        live[h] = liven[h];
        nd[h] = ndn[h];
      }
      if (e0 + 8 < NS) loadNodes(e0 + 8);
#pragma unroll
      for (int h = 0; h < 2; ++h) geo[h] = live[h] && nodeGeomF32(sF[jjv[h]], nd[h], gx[h], gy[h], gz[h], sl[h]);
      float cc[2][8], fx[2], fy[2], fz[2];
#pragma unroll
      for (int h = 0; h < 2; ++h)
        if (geo[h]) sdfCorners(P.sdf, sF[jjv[h]], gx[h], gy[h], gz[h], cc[h], fx[h], fy[h], fz[h]);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int e = e0 + 4 * h + q;
        bool cand = false;
        float smin = kInf;
        if (live[h]) {
          float slack = sl[h];
          if (geo[h]) cand = nodeSdfF32(sF[jjv[h]], cc[h], fx[h], fy[h], fz[h], slack);
          smin = cand ? 0.f : fmaxf(slack, 0.f);
        }
        smin = groupMin<16>(smin);  // the 16 lanes of a sub-entry share its pair
        const unsigned long long m = __ballot(cand);
        const unsigned int mq = unsigned((m >> (16 * q)) & 0xFFFFull);
        if (b16 == 0 && e < NS) {
          sMask[e] = mq;
          atomicAdd(&sCnt[jjv[h]], __popc(mq));  // LDS: 4 groups may share a pair
          if (smin < kInf) atomicMin(&sSlack[jjv[h]], __float_as_uint(smin));
        }
        total += __popcll(m);
      }
    }
    // the batch's pair entries (loaded at level 1) for the contact flags and the exact test
    if (k == 0 && j < kBatch) {
      sPInfo[j] = myInfo;
      sPA[j] = myA;
    }
    __builtin_amdgcn_wave_barrier();
    // ---- one reservation per batch, exact size (its result is first needed after the
    // certificates) ----
    int first = 0;
    if (total > 0 && lane == 0) first = atomicAdd(&P.devInt[1], total);
    if (lane < nb && bi + lane * NB < nAct) {
      const PairF32& fl = sF[lane];
      // separation certificate in world units, net of the motion since the build
      float sg = __uint_as_float(sSlack[lane]);
      bool surDep = false;  // the surrogate's slack is the smaller one
      if (SUR) {
        const float ss = P.surSlack[bi + lane * NB];
        if (ss < sg) {
          sg = ss;
          surDep = true;
        }
      }
      float sw = -1.f, s0 = -1.f;
      const float sgr = sg * (1.f - 1e-5f) - 1e-6f;  // g units; the map's rounding is in the thresholds
      if (sgr > 0.f && P.useCert) {
        s0 = sgr * fl.gToWorld * (1.f - 1e-5f);
        sw = (sgr * fl.gToWorld - fl.certD0) * (1.f - 1e-5f);
      }
      // (the sign of .z marks a certificate that rests on the surrogate's slack)
      P.cert[fl.pair] = make_float4(sw > 0.f ? sw : -1.f, fl.certCB, surDep ? -fl.certCA : fl.certCA,
                                    s0 > 0.f ? s0 : -1.f);
    }
    if (total > 0) {
      first = __builtin_amdgcn_readfirstlane(__shfl(first, 0, 64));
      if ((long long)first + total > P.contactCap) {
        // the call halts with a capacity error; slots reserved below the capacity stay empty
        if (lane == 0) {
          atomicMax(&P.devInt[5], first + total);
          raiseError(P, kErrContacts);
        }
        for (int c = first + lane; c < (int)P.contactCap; c += 64) {
          if (c >= first + total) break;
          P.cInfo[c] = kInvalidInfo;
          P.hashSlot[c] = -1;
        }
        total = 0;
      }
    }
    if (lane < nb && bi + lane * NB < nAct) {
      const int pr = sF[lane].pair;
      int off = first;
      for (int jj = 0; jj < lane; ++jj) off += sCnt[jj];
      P.pairCoff[pr] = off;
      P.pairCcnt[pr] = total ? sCnt[lane] : 0;
      if (total && sCnt[lane] > 0) {
        // both bodies have contact terms to sum this step
        P.bodyHit[sPA[lane]] = 1;
        const uint32_t info = sPInfo[lane];
        if ((info >> 28) != PT_PW) P.bodyHit[info & kPartnerMask] = 1;
      }
      sOff[lane] = off;  // running write offset of each pair
    }
    __builtin_amdgcn_wave_barrier();
    // ---- write candidate slots; a pair's candidates are contiguous ----
    if (total) {
      for (int e0 = 0; e0 < NS; e0 += 4) {
        if (lane == 0) {
          for (int t = 0; t < 4 && e0 + t < NS; ++t) {
            const int jj = sSub[e0 + t].y >> 16;
            sBase4[t] = sOff[jj];
            sOff[jj] += __popc(sMask[e0 + t]);
          }
        }
        __builtin_amdgcn_wave_barrier();
        const int e = e0 + q;
        if (e < NS) {
          const unsigned int mq = sMask[e];
          if ((mq >> b16) & 1u) {
            const int2 se = sSub[e];
            const int dst = sBase4[q] + __popc(mq & lt16);
            P.cPair[dst] = sF[se.y >> 16].pair;
            P.cNode[dst] = se.x + b16;  // cluster-ordered node index until the exact test
            if (dst - first < kSlotStage) sStage[dst - first] = make_int2(se.y >> 16, se.x + b16);
          }
        }
        __builtin_amdgcn_wave_barrier();
      }
      // the exact test and contact law of this batch's candidates, by the whole wavefront
      __threadfence_block();
      __builtin_amdgcn_wave_barrier();
      for (int c = first + lane; c < first + total; c += 64) {
        if (c - first < kSlotStage) {
          const int2 st = sStage[c - first];
          const PairF32& fs = sF[st.x];
          exactSlot<SEC>(P, c, fs.pair, st.y, sPInfo[st.x], sPA[st.x], make_int4(fs.nx, fs.ny, fs.nz, fs.sdfOff),
                         statQueries, statContacts);
        } else {
          const int pc = P.cPair[c];
          exactSlot<SEC>(P, c, pc, P.cNode[c], P.pairInfo[pc], P.pairA[pc], make_int4(-1, 0, 0, 0), statQueries,
                         statContacts);
        }
      }
    }
    statCand += (lane == 0) ? (unsigned long long)total : 0ull;
  }
  const unsigned long long t = (unsigned long long)wsum(double(statCand));
  const unsigned long long te = (unsigned long long)wsum(double(statEnt));
  const unsigned long long tq = (unsigned long long)wsum(double(statQueries));
  const unsigned long long tc = (unsigned long long)wsum(double(statContacts));
  if (lane == 0 && (t | te)) {
    const int slot = (blockIdx.x + wl) & (kStatSlots - 1);
    atomicAdd(statSlot(P, 0, slot), t);
    atomicAdd(statSlot(P, 1, slot), tq);
    atomicAdd(statSlot(P, 2, slot), tc);
    atomicAdd(statSlot(P, 3, slot), te);
  }
}

// Surrogate cluster decisions, one wavefront per batch of 8 candidate pairs (lanes = pair x
// cluster, as stage 1 of k_candidates): bit c of surMask[p] is 0 when the network culls cluster c
// of pair p.  Only clusters whose centre lies in a family partner's grid box are queried, packed
// 16 per network pass.
__global__ __launch_bounds__(kNarrowBlock) void k_sur_masks(Params P) {
  LSD_HALT(P);
  __shared__ int sQ[kNarrowBlock / 64][64];
  const int lane = threadIdx.x & 63, wl = threadIdx.x >> 6;
  const EvalRecords rec = evalRecords(P);
  const int nAct = rec.n;
  const int wave = __builtin_amdgcn_readfirstlane(int(blockIdx.x * (kNarrowBlock / 64) + wl));
  const int nWaves = int(gridDim.x * (kNarrowBlock / 64));
  const int iters = (P.maxClusters + 7) >> 3;
  const int j = lane >> 3, k = lane & 7;
  for (int base = wave * kBatch; base < nAct; base += nWaves * kBatch) {
    const int a = base + j;  // map record
    const bool own = a < nAct;
    PairF32 f;
    if (own) f = reinterpret_cast<const PairF32*>(P.pairMap)[rec.at(a)];
    unsigned long long mask = ~0ull;
    float smin = __builtin_huge_valf();  // smallest slack of the clusters the network culls
    for (int it = 0; it < iters; ++it) {
      const int c = it * 8 + k;
      float gx = 0.f, gy = 0.f, gz = 0.f, rho = 0.f, ux = 0.f, uy = 0.f, uz = 0.f;
      int qslot = -1;
      if (own && c < f.nClusters && f.slot >= 0) {
        const ClusterF32 cl = P.clusters32[f.clusterOff + c];
        bool inside;
        float slack;
        if (clusterGeomF32(f, cl, gx, gy, gz, rho, inside, slack) && inside) {
          qslot = f.slot;
          surQueryPoint(P, f, gx, gy, gz, ux, uy, uz);
        }
      }
      const float d = surInferPacked(P.sur, ux, uy, uz, qslot, lane, sQ[wl]);
      const bool cull = qslot >= 0 && !clusterSurKeep(P, f, d, rho);
      if (cull) smin = fminf(smin, fmaxf(clusterSurSlack(P, f, d, rho), 0.f));
      // collect this group's 8 decisions of pair j into its mask
      const unsigned long long b = __ballot(cull);
      const unsigned long long mine = (b >> (8 * j)) & 0xFFull;
      mask &= ~(mine << (8 * it));
    }
    smin = groupMin<8>(smin);
    if (own && k == 0) {
      P.surMask[a] = mask;
      P.surSlack[a] = smin;
    }
  }
}

// Exact audit of the surrogate: every contact of this (plain) step whose partner is in the
// family must lie in a cluster that the surrogate keeps; a culled one is a missed contact.
// One lane per contact slot, 64 per wave through the network.
__global__ __launch_bounds__(kNarrowBlock) void k_sur_audit(Params P) {
  LSD_HALT(P);
  __shared__ int sQ[kNarrowBlock / 64][64];
  const int lane = threadIdx.x & 63;
  const int n = min(P.devInt[1], (int)P.contactCap);
  const int nWaves = int(gridDim.x * (kNarrowBlock / 64));
  unsigned long long checked = 0, missed = 0;
  for (int base = int(blockIdx.x * (kNarrowBlock / 64) + (threadIdx.x >> 6)) * 64; base < n; base += nWaves * 64) {
    const int c = base + lane;
    bool q = false;
    PairF32 f;
    float gx = 0.f, gy = 0.f, gz = 0.f, rho = 0.f, ux = 0.f, uy = 0.f, uz = 0.f;
    if (c < n && P.cInfo[c] != kInvalidInfo) {
      const int p = P.cPair[c];
      f = reinterpret_cast<const PairF32*>(P.pairMap)[P.pairRec[p]];
      if (f.slot >= 0) {
        const int A = P.pairA[p];
        const int orig = P.cNode[c] - P.nodeStart[A];
        const int cluster = P.nodeCluster[f.nodeOff + orig];
        const ClusterF32 cl = P.clusters32[f.clusterOff + cluster];
        bool inside;
        float slack;
        const bool geom = clusterGeomF32(f, cl, gx, gy, gz, rho, inside, slack);
        q = geom && inside;  // the only clusters the surrogate decides on
        if (q) surQueryPoint(P, f, gx, gy, gz, ux, uy, uz);
        checked++;
      }
    }
    const float d = surInferPacked(P.sur, ux, uy, uz, q ? f.slot : -1, lane, sQ[threadIdx.x >> 6]);
    if (q && !clusterSurKeep(P, f, d, rho)) missed++;
  }
  const unsigned long long tc = (unsigned long long)wsum(double(checked));
  const unsigned long long tm = (unsigned long long)wsum(double(missed));
  if (lane == 0 && (tc | tm)) {
    atomicAdd(&P.devStat[10], tc);
    atomicAdd(&P.devStat[11], tm);
  }
}

// ------------------------------------------------------------------------------------------
// Bonds (one thread per bond)
// ------------------------------------------------------------------------------------------
__global__ void k_bonds(Params P) {
  LSD_HALT(P);
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P.nBonds) return;
  double* out = P.bOut + 6ll * k;
  if (P.bActive[k] == 0) {
    for (int t = 0; t < 6; ++t) out[t] = 0.;
    P.bF[k] = mk3(0, 0, 0);
    P.bMi[k] = mk3(0, 0, 0);
    P.bMj[k] = mk3(0, 0, 0);
    return;
  }
  const double* C = P.bConst + 9ll * k;
  const double B1 = C[0], B2 = C[1], B3 = C[2], B4 = C[3], rad = C[4], L0 = C[5], sT = C[6], coh = C[7],
               bmu = C[8];
  const d3* L = P.bLocal + 8ll * k;
  const int i = P.bI[k], j = P.bJ[k];
  const q4 qi = P.quat[i], qj = P.quat[j];
  const d3 n1i = rot(qi, L[0]), n2i = rot(qi, L[1]), n3i = rot(qi, L[2]);
  const d3 n1j = rot(qj, L[4]), n2j = rot(qj, L[5]), n3j = rot(qj, L[6]);
  const d3 ri = P.pos[i], rj = P.pos[j];
  const d3 rbi = rot(qi, L[3]) + ri, rbj = rot(qj, L[7]) + rj;
  P.bPoint[k] = 0.5 * (rbi + rbj);
  d3 F = mk3(0, 0, 0), Mij = mk3(0, 0, 0), Mji = mk3(0, 0, 0);
  int active = 1;
  const d3 rbij = rbi - rbj;
  const double D = len(rbij);
  if (!isZero(D)) {
    const d3 nij = normalize(rbij);
    const d3 e = -nij;
    out[0] = 0.5 * B1 * (D - L0) * (D - L0);
    out[1] = B2 * (0.5 * dot((n1j - n1i), e) - 0.25 * dot(n1i, n1j) + 0.75);
    out[2] = (0.25 * B2 + B3 + 0.5 * B4) * (dot(n1i, n1j) + 1.);
    out[3] = -0.5 * B4 * (dot(n1i, n1j) + dot(n2i, n2j) + dot(n3i, n3j) - 1);
    F = B1 * (D - L0) * e + B2 / (2 * D) * ((n1j - n1i) - dot((n1j - n1i), e) * e);
    const d3 MTB = B3 * cross(n1j, n1i) - 0.5 * B4 * (cross(n2j, n2i) + cross(n3j, n3i));
    Mij = -0.5 * B2 * cross(e, n1i) + MTB;
    Mji = 0.5 * B2 * cross(e, n1j) - MTB;
    const double Fn = dot(F, nij);
    const d3 Fs = F - Fn * nij;
    const double Tt = dot(MTB, nij);
    const d3 Tb = MTB - Tt * nij;
    const double Ab = rad * rad * kPi;
    const double Ib = Ab * rad * rad / 4.;
    const double Jb = 2 * Ib;
    out[4] = -Fn / Ab + len(Tb) / Ib * rad;
    out[5] = len(Fs) / Ab + fabs(Tt) / Jb * rad;
    if (sT > 0. && out[4] > sT) active = 0;
    else if (coh > 0. && out[5] > coh - bmu * out[4]) active = 0;
  }
  P.bActive[k] = active;
  P.bF[k] = F;
  P.bMi[k] = Mij + cross(rbi - ri, F);
  P.bMj[k] = Mji + cross(rbj - rj, -F);
}

// ------------------------------------------------------------------------------------------
// Per-body segmented reduction + second half-kick
// ------------------------------------------------------------------------------------------
__device__ void stepEnd(const Params& P);
__device__ void stepEndWave(const Params& P, int l);

// Contact and bond sums on body b in a fixed order: its master pairs contact by contact, then the
// reactions of the pairs where it is the partner, then its bonds.
// A body none of whose pairs holds candidate nodes this step has no contact terms (its pairs'
// contact counts are all 0) and skips the walk over its pairs.  Returns whether any term was
// summed (false: F and T are exactly zero).
__device__ __forceinline__ bool reduceBody(const Params& P, int b, d3& F, d3& T) {
  F = mk3(0., 0., 0.);
  T = mk3(0., 0., 0.);
  const bool hit = P.bodyHit[b] != 0;
  if (hit) P.bodyHit[b] = 0;
  const long long cap = P.pairCap;
  for (int p = P.pairStart[b]; hit && p < P.pairStart[b + 1] && p < cap; ++p) {
    const int c0 = P.pairCoff[p], n = P.pairCcnt[p];
    for (int c = c0; c < c0 + n; ++c) {
      if (P.cInfo[c] == kInvalidInfo) continue;
      F += P.cF[c];
      T += P.cTA[c];
    }
  }
  for (int q = P.rcStart[b]; hit && q < P.rcStart[b + 1]; ++q) {
    const int p = P.rcPair[q];
    const int c0 = P.pairCoff[p], n = P.pairCcnt[p];
    for (int c = c0; c < c0 + n; ++c) {
      if (P.cInfo[c] == kInvalidInfo) continue;
      F += -P.cF[c];
      T += P.cTB[c];
    }
  }
  bool any = hit;
  if (P.nBonds > 0) {
    for (int q = P.bcStart[b]; q < P.bcStart[b + 1]; ++q) {
      const int it = P.bcItem[q];
      const int k = it >> 1;
      if (it & 1) { F += -P.bF[k]; T += P.bMj[k]; }
      else { F += P.bF[k]; T += P.bMi[k]; }
      any = true;
    }
  }
  return any;
}

// One thread per body.  Thread 0 also closes the step's bookkeeping.
__global__ void k_reduce_kick(Params P, int kick) {
  LSD_HALT(P);
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b == 0) stepEnd(P);
  if (b >= P.nP) return;
  d3 F, T;
  reduceBody(P, b, F, T);
  P.force[b] = F;
  P.torque[b] = T;
  P.forceSet[b] = 1;  // an external-force hook may add to the stored sums
  if (kick) halfKick(P, b, F, T);
}

// The end of one step and the start of the next in one pass over the bodies: sums, second
// half-kick, then the next step's first half-kick and drift (one thread per particle, then one per
// wall).  H describes the next step; P is this step's.  Nothing here reads another body's state,
// and this step's flags and counters are closed by thread 0 before it clears the next step's.
template <bool SEC>
__global__ void k_reduce_drift(Params P, StepHead H) {
  LSD_HALT(P);
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  BodyIn in;
  bool set = false;
  if (b < P.nP) {
    in = loadBody<SEC>(P, b);
    set = P.forceSet[b] != 0;
  }
  if (b == 0) stepEnd(P);
  stepHead(P, H, b, int(gridDim.x * blockDim.x));
  if (b >= P.nP) {
    if (b - P.nP < P.nW) wallDrift(P, b - P.nP, H.buf);
    return;
  }
  d3 F, T;
  const bool any = reduceBody(P, b, F, T);
  kickDrift<SEC>(P, b, in, F, T, 2, H.buf);
  // the stored sums of a body without terms are already zero unless written otherwise since
  if (any || set) {
    P.force[b] = F;
    P.torque[b] = T;
    P.forceSet[b] = any ? 1 : 0;
  }
}
// The same for a small simulation (up to kWaveEndBodies bodies, no sector periodicity), whose
// counters the first wavefront closes (stepEndWave): its body pass is short, and thread 0's rounds of
// counter reads were its last wavefront's (256 drums of 19k bodies in all: 0.0472 -> 0.0433 ms per
// step).  In the 200k-body pass thread 0's version is the faster, and the kernel above is left as it
// compiles alone (any change to it shifts its code generation).
__global__ void k_reduce_drift_small(Params P, StepHead H) {
  LSD_HALT(P);
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  BodyIn in;
  bool set = false;
  if (b < P.nP) {
    in = loadBody<false>(P, b);
    set = P.forceSet[b] != 0;
  }
  if (blockIdx.x == 0 && threadIdx.x < 64) stepEndWave(P, int(threadIdx.x));
  stepHead(P, H, b, int(gridDim.x * blockDim.x));
  if (b >= P.nP) {
    if (b - P.nP < P.nW) wallDrift(P, b - P.nP, H.buf);
    return;
  }
  d3 F, T;
  const bool any = reduceBody(P, b, F, T);
  kickDrift<false>(P, b, in, F, T, 2, H.buf);
  if (any || set) {
    P.force[b] = F;
    P.torque[b] = T;
    P.forceSet[b] = any ? 1 : 0;
  }
}

// Close of the pre-loop build: step bookkeeping, and the contact flags of the candidate kernel
// cleared (no sums follow; the stored forces stay those of the last step).
__global__ void k_prime_end(Params P) {
  LSD_HALT(P);
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b == 0) stepEnd(P);
  if (b < P.nP) P.bodyHit[b] = 0;
}

// ------------------------------------------------------------------------------------------
// Call snapshot (restart of a call that halted on an error)
// ------------------------------------------------------------------------------------------
__global__ void k_snap_save(Params P, SnapDev S) {
  const Span g = gridSpan();
  for (int i = g.tid; i < P.nP; i += g.nth) {
    S.pos[i] = P.pos[i]; S.vel[i] = P.vel[i]; S.ang[i] = P.angvel[i];
    S.force[i] = P.force[i]; S.torque[i] = P.torque[i]; S.quat[i] = P.quat[i]; S.forceSet[i] = P.forceSet[i];
  }
  for (int i = g.tid; i < P.nW; i += g.nth) { S.wpos[i] = P.wpos[i]; S.wquat[i] = P.wquat[i]; }
  for (int i = g.tid; i < P.nBonds; i += g.nth) S.bActive[i] = P.bActive[i];
  const int n = int(min((long long)P.devInt[10 + S.prevBuf], S.hCap));
  for (int c = g.tid; c < n; c += g.nth) {
    S.hNode[c] = S.prevNode[c]; S.hInfo[c] = S.prevInfo[c]; S.hSpring[c] = S.prevSpring[c];
  }
  if (g.tid == 0) S.meta[0] = n;
  if (g.tid < kDevStats) S.devStat[g.tid] = P.devStat[g.tid];
}

// Back to the call's start; the history tables must be empty (filled by the host before).  The
// previous pair list is dropped (no certificate is inherited from it) and the step flags cleared.
__global__ void k_snap_restore(Params P, SnapDev S) {
  const Span g = gridSpan();
  for (int i = g.tid; i < P.nP; i += g.nth) {
    P.pos[i] = S.pos[i]; P.posDet[i] = S.pos[i]; P.vel[i] = S.vel[i]; P.angvel[i] = S.ang[i];
    P.force[i] = S.force[i]; P.torque[i] = S.torque[i]; P.quat[i] = S.quat[i]; P.forceSet[i] = S.forceSet[i];
    P.bodyHit[i] = 0;
  }
  for (int i = g.tid; i <= P.nP; i += g.nth) P.pairStart[i] = 0;
  for (int i = g.tid; i < P.nW; i += g.nth) { P.wpos[i] = S.wpos[i]; P.wquat[i] = S.wquat[i]; }
  for (int i = g.tid; i < P.nBonds; i += g.nth) P.bActive[i] = S.bActive[i];
  const int n = S.meta[0];
  for (int c = g.tid; c < n; c += g.nth) {
    S.prevNode[c] = S.hNode[c]; S.prevInfo[c] = S.hInfo[c]; S.prevSpring[c] = S.hSpring[c];
  }
  if (g.tid < kDevStats) P.devStat[g.tid] = S.devStat[g.tid];
  if (g.tid < 18) P.barrier[g.tid * kBarrierLine] = 0;  // (the grid barrier's words and evalCount)
  if (g.tid == 0) {
    P.devInt[10 + S.prevBuf] = n;  // the restored slots, cleared from their table two steps on
    P.devInt[11 - S.prevBuf] = 0;  // the other table is empty
    P.devInt[1] = 0;
    P.devInt[rebuildFlagIdx(0)] = 0; P.devInt[rebuildFlagIdx(1)] = 0;
    P.devInt[ghostCountIdx(0)] = 0; P.devInt[ghostCountIdx(1)] = 0;
    P.devInt[3] = 0;
    P.devInt[kErrStepIdx] = -1;
  }
  for (int k = g.tid; k < 5 * kStatSlots; k += g.nth) P.statSlots[k * kStatStride] = 0;
}

// The history table of the restored slots: each contact slot's key inserted again (the table was
// emptied first), its table slot recorded for the clearing two steps on.
__global__ void k_hist_reinsert(Params P, SnapDev S, unsigned long long* key, int* val, int* slotOf) {
  const Span g = gridSpan();
  const int n = S.meta[0];
  for (int c = g.tid; c < n; c += g.nth) {
    const uint32_t info = S.prevInfo[c];
    int at = -1;
    if (info != kInvalidInfo) {
      const unsigned long long k = histKey(info >> 28, S.prevNode[c], int(info & kPartnerMask));
      unsigned long long slot = mix64(k) & P.hashMask;
      for (unsigned long long probe = 0; probe <= P.hashMask; ++probe) {
        const unsigned long long old = atomicCAS(&key[slot], kHashEmpty, k);
        if (old == kHashEmpty || old == k) {
          val[slot] = c;
          at = int(slot);
          break;
        }
        slot = (slot + 1) & P.hashMask;
      }
    }
    slotOf[c] = at;
  }
}

// Second half-kick from the stored sums (after an external-force hook has added to them).
__global__ void k_kick(Params P) {
  LSD_HALT(P);
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= P.nP) return;
  halfKick(P, b, P.force[b], P.torque[b]);
}

// ------------------------------------------------------------------------------------------
// Step bookkeeping
// ------------------------------------------------------------------------------------------
// stepEnd by the first wavefront of block 0 (lanes l < min(64, blockDim.x)): the lanes read the stat
// slots (a line each) together and sum them across lanes (in FP64, exact for counts below 2^53), where
// thread 0 alone takes 8 dependent rounds of scalar loads.  Totals, flags and counters as stepEnd.
__device__ void stepEndWave(const Params& P, int l) {
  const int nl = min(64, int(blockDim.x));
  double s[5] = {0., 0., 0., 0., 0.};
  for (int k = l; k < kStatSlots; k += nl) {
#pragma unroll
    for (int f = 0; f < 5; ++f) s[f] += double(*statSlot(P, f, k));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      const double o = __shfl(s[f], l + off, 64);
      if (l + off < nl) s[f] += o;
    }
  }
  if (l == 0) {
    const int np = P.pairStart[P.nP];
    P.devInt[4] = max(P.devInt[4], np);
    P.devInt[5] = max(P.devInt[5], P.devInt[1]);
    P.devInt[10 + P.curBuf] = min(P.devInt[1], (int)P.contactCap);
    const unsigned long long t = (unsigned long long)s[0], q = (unsigned long long)s[1],
                             c = (unsigned long long)s[2], e = (unsigned long long)s[3],
                             mn = (unsigned long long)s[4];
    P.devStat[12] += mn;
    P.devStat[14] += P.devInt[rebuildFlagIdx(P.curBuf)] != 0 ? 1ull : 0ull;
    P.devStat[0] = t;
    P.devStat[1] = q;
    P.devStat[2] = c;
    P.devStat[3] += (unsigned long long)np;
    P.devStat[4] += t;
    P.devStat[5] += q;
    P.devStat[6] += c;
    P.devStat[8] += e;
    P.devInt[rebuildFlagIdx(P.curBuf)] = 0;
    P.devInt[ghostCountIdx(P.curBuf)] = 0;
  }
  // (after every lane's loads: their sums were exchanged above)
  for (int k = l; k < 5 * kStatSlots; k += nl) P.statSlots[k * kStatStride] = 0;
}

// Step totals into the counters, then the per-step counters that the drift kernel accumulates into
// cleared for the next step.
__device__ void stepEnd(const Params& P) {
  {
    const int np = P.pairStart[P.nP];  // (an overflowing list halted the call in the broad phase)
    P.devInt[4] = max(P.devInt[4], np);
    P.devInt[5] = max(P.devInt[5], P.devInt[1]);
    P.devInt[10 + P.curBuf] = min(P.devInt[1], (int)P.contactCap);  // slots to clear two steps on
    unsigned long long t = 0, q = 0, c = 0, e = 0, mn = 0;
    for (int k = 0; k < kStatSlots; ++k) {
      t += *statSlot(P, 0, k);
      q += *statSlot(P, 1, k);
      c += *statSlot(P, 2, k);
      e += *statSlot(P, 3, k);
      mn += *statSlot(P, 4, k);
    }
    P.devStat[12] += mn;
    P.devStat[14] += P.devInt[rebuildFlagIdx(P.curBuf)] != 0 ? 1ull : 0ull;
    P.devStat[0] = t;
    P.devStat[1] = q;
    P.devStat[2] = c;
    // sums over steps (profiling)
    P.devStat[3] += (unsigned long long)np;
    P.devStat[4] += t;
    P.devStat[5] += q;
    P.devStat[6] += c;
    P.devStat[8] += e;
  }
  P.devInt[rebuildFlagIdx(P.curBuf)] = 0;  // rebuild flag (the drift kernel raises it)
  P.devInt[ghostCountIdx(P.curBuf)] = 0;  // valid ghost images at their source's own position
  for (int k = 0; k < 5 * kStatSlots; ++k) P.statSlots[k * kStatStride] = 0;
  // (the contact slot count [1] and the pair check's evalCount stay readable after the step; the
  // drift kernel clears them)
}


// ------------------------------------------------------------------------------------------
// The whole broad phase in one launch (rebuild steps only; every block returns at once
// otherwise).  Phases are separated by grid barriers; every block runs every phase, so all blocks
// reach every barrier.  The host sizes the grid so that all its blocks are resident together.
// ------------------------------------------------------------------------------------------
// Blocks of the single-launch broad phase: large, so that few blocks meet at each barrier (the
// cost of a barrier grows with the number of blocks) while many threads work between them.
constexpr int kBroadThreads = 1024;

// Grid barrier in two levels (its words are zeroed by the pair check, the launch that follows every
// broad phase, and by the host after a failed call): a block arrives on
// its group's counter (group = blockIdx mod 8, the XCD the block runs on when the dispatcher deals
// blocks round-robin over the XCDs); the last arrival of a group arrives on the groups' counter, and
// the last of those writes the release word, which the blocks poll.  A barrier then serialises about
// nBlocks / 8 + 8 atomics on any one word instead of nBlocks, and the polls read a word no arrival
// writes.  The wait is bounded (P.barrierSpin polls, about a second by default): a block that gives
// up raises kErrBarrier.  A block also stops waiting as soon as any error flag is up, and after
// every barrier all blocks read the flags: sync() returns true when the launch must be abandoned,
// so every block leaves at the same barrier and no phase runs on the incomplete results of an
// earlier one.
struct GridBarrier {
  const Params& P;
  int nBlocks, passed;
  __device__ bool sync() {
    __shared__ int abandon;
    __syncthreads();
    if (threadIdx.x == 0) {
      ++passed;
      const int g = int(blockIdx.x) & 7;
      const int nGroups = min(nBlocks, 8);
      const int inGroup = (nBlocks - g + 7) >> 3;  // blocks b < nBlocks with b mod 8 == g
      int* groupCount = P.barrier + g * kBarrierLine;
      int* groupsDone = P.barrier + 8 * kBarrierLine;
      int* release = P.barrier + 9 * kBarrierLine;
      __threadfence();  // this block's writes, before the arrival
      if (atomicAdd(groupCount, 1) == passed * inGroup - 1) {
        __threadfence();  // (the group's arrivals, before the group's)
        if (atomicAdd(groupsDone, 1) == passed * nGroups - 1) {
          __threadfence();
          __hip_atomic_store(release, passed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      // relaxed polling (an acquire per poll would invalidate the caches every time), then one
      // acquire fence
      for (unsigned int it = 0;
           __hip_atomic_load(release, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < passed; ++it) {
        if (__hip_atomic_load(P.devInt + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
        if (it >= unsigned(P.barrierSpin)) {
          raiseError(P, kErrBarrier);
          break;
        }
        __builtin_amdgcn_s_sleep(10);  // ~640 cycles between polls of the release word
      }
      __threadfence();  // the other blocks' writes, after the wait
      abandon = __hip_atomic_load(P.devInt + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
    }
    __syncthreads();
    return abandon != 0;
  }
};

__global__ __launch_bounds__(kBroadThreads) void k_broad_phase(Params P) {
  LSD_HALT(P);
  if (P.devInt[rebuildFlagIdx(P.curBuf)] == 0) return;  // written by the drift kernel before this launch: grid-uniform
  __shared__ int lds[kScanLds];
  GridBarrier grid{P, int(gridDim.x), 0};
  const Span g = gridSpan();
  const int wave = g.tid >> 6, nWaves = g.nth >> 6;
  const int ncell = P.sg.ncell;
  // counting sort of the particles by search cell
  for (int i = g.tid; i < ncell; i += g.nth) P.cellCount[i] = 0;
  for (int i = g.tid; i < P.nP; i += g.nth) {
    P.rcCount[i] = 0;
    P.scell[i] = searchCellOf(P, i);
  }
  if (g.tid == 0) P.devInt[2] = 0;  // large-bucket list
  if (grid.sync()) return;
  bucketCount(P.scell, P.nP, P.cellCount, P.rank, g);
  if (grid.sync()) return;
  scanReduceTiles(P.cellCount, ncell, P.scanTiles, lds);
  if (grid.sync()) return;
  scanApplyTiles(P.cellCount, ncell, P.scanTiles, P.cellStart, lds);
  if (grid.sync()) return;
  bucketScatter(P.scell, P.nP, P.cellStart, P.rank, P.sortTmp, g);
  if (grid.sync()) return;
  bucketSortSmall(P.cellStart, ncell, P.sortTmp, P.sorted, P.sortBig, P.devInt + 2, g, P.nP);
  if (grid.sync()) return;
  bucketSortBig(P.cellStart, P.sortTmp, P.sorted, P.sortBig, P.devInt + 2, wave, nWaves, ncell, P.nP);
  if (grid.sync()) return;
  // particle data in sorted order; candidate pairs per master, prefix, pair list
  for (int k = g.tid; k < P.nP; k += g.nth) gatherParticle(P, k);
  backupPairs(P, g);
  if (g.tid == 0) P.devInt[2] = 0;  // large-bucket list of the reaction sort
  if (grid.sync()) return;
  pairsFind(P, g);
  if (grid.sync()) return;
  scanReduceTiles(P.pairCount, P.nP, P.scanTiles, lds);
  if (grid.sync()) return;
  scanApplyTiles(P.pairCount, P.nP, P.scanTiles, P.pairStart, lds);
  if (grid.sync()) return;
  // every block reads the same total after the barrier: all leave together when it does not fit
  if (pairOverflow(P)) return;
  pairsEmit(P, g);
  if (grid.sync()) return;
  inheritCerts(P, g);  // reads the motion records before they restart
  if (grid.sync()) return;
  markBuild(P, g);
  if (grid.sync()) return;
  // reaction CSR: pairs by partner, in pair order
  rcCountPairs(P, g);
  if (grid.sync()) return;
  scanReduceTiles(P.rcCount, P.nP, P.scanTiles, lds);
  if (grid.sync()) return;
  scanApplyTiles(P.rcCount, P.nP, P.scanTiles, P.rcStart, lds);
  if (grid.sync()) return;
  rcScatterPairs(P, g);
  if (grid.sync()) return;
  const int nRc = min(P.rcStart[P.nP], int(P.pairCap));
  bucketSortSmall(P.rcStart, P.nP, P.rcTmp, P.rcPair, P.sortBig, P.devInt + 2, g, nRc);
  if (grid.sync()) return;
  bucketSortBig(P.rcStart, P.rcTmp, P.rcPair, P.sortBig, P.devInt + 2, wave, nWaves, P.nP, nRc);
}

template __global__ void k_candidates<false, false>(Params P);
template __global__ void k_candidates<true, false>(Params P);
template __global__ void k_candidates<false, true>(Params P);
template __global__ void k_candidates<true, true>(Params P);
template __global__ void k_kick_drift<false>(Params P, int forceRebuild);
template __global__ void k_kick_drift<true>(Params P, int forceRebuild);
template __global__ void k_reduce_drift<false>(Params P, StepHead H);
template __global__ void k_reduce_drift<true>(Params P, StepHead H);
template __global__ void k_pair_check<false>(Params P, int full);
template __global__ void k_pair_check<true>(Params P, int full);

}  // namespace lsd
