// Device data layout of one LS-DEM context.  All arrays live in HBM; the struct is passed by
// value to every kernel.
#pragma once
#include "dmath.h"
#include "surrogate.h"

namespace lsd {

enum PairType : uint32_t { PT_PP = 0, PT_PW = 1, PT_GX = 2, PT_GY = 3, PT_GXY = 4 };
constexpr uint32_t kPartnerMask = 0x0FFFFFFFu;
constexpr uint64_t kHashEmpty = ~0ull;

// Shape: FP32 SDF grid (x fastest) + boundary nodes grouped into clusters of <= 64 nodes.
struct ShapeDev {
  int nx, ny, nz, nNodes;
  int nodeOff, nClusters, clusterOff, pad0;
  long long sdfOff;
  double ox, oy, oz, h;  // unscaled grid origin and spacing
  double boxR;           // max |corner| of the grid box (unscaled), used for fixed-partner culls
  double lip;            // Lipschitz bound of the trilinear interpolant over cells with a positive
                         // corner (scale invariant)
  double nodeR;          // max |node| (unscaled): how far a node lies from the body origin
};
struct ClusterDev {
  double cx, cy, cz, rho;  // bounding sphere of the cluster's nodes (unscaled)
  int first, count;        // node range in the shape's cluster-ordered nodes
  int subFirst, nSub;      // sub-clusters (<= 16 nodes each)
};
struct ClusterF32 {
  float cx, cy, cz, rho;  // FP32 copy of the bounding sphere, rho rounded up
  int first, count;
  int subFirst, nSub;     // clusters: global index of the first sub-cluster and their number
};
struct NodeDev {
  double x, y, z;  // unscaled local position
  int orig;        // index of the node in the shape's node list as given by the caller
  int pad;
};

struct GridGeom {  // geometry of the reference spatial grid (spatialGrid::set)
  d3 mn, mx, inv;
  i3 size;
  int pad;
};
struct SearchGrid {  // this library's own broad-phase grid
  d3 mn;
  double cell, invCell;
  i3 size;
  int ncell;
};

// A particle's per-step record for the pair check:
//   motion: displacement of posDet since the build (x, y, z) and the chord 2 sin(angle / 2) of the
//           rotation since the build (w);
//   key:    reference-grid key (x; -1 outside the domain) and the keys of ghost images 0..2 (y, z, w;
//           -1: no image), each with bit kFlipBit set when the image's shift state differs from the
//           build's.
// (32 bytes: one read per partner.)  The rotation since the build as a quaternion is kept apart
// (qrel), read only when a certificate is checked against the motion since an evaluation.
struct alignas(16) BodyStep {
  float4 motion;
  int4 key;
};
constexpr int kFlipBit = 30;
static_assert(sizeof(BodyStep) == 32 && alignof(BodyStep) == 16, "step record: motion then key, 16 B each");

struct Params {
  // ---- sizes and scalars ----
  int nP, nW, nShapes, nBonds;
  d3 g;
  double dt;
  int quirks, periodicX, periodicY, sdfCull;
  int narrowBatch;    // pairs per wavefront batch of k_candidates (0: chosen on the device)
  int prime;          // 1: the pre-loop build of a solve phase (detection at the current poses: no
                      // kick, no drift, no contact law; springs carried over as found)
  int callStep;       // index of this step within the lsdem_step call (recorded with an error flag)
  int barrierSpin;    // poll limit of a broad-phase grid barrier before it gives up (time-out error)
  int periodicSector; // quarter-sector rotational periodicity about the z axis through gridP.mn
  GridGeom gridP, gridW;
  SearchGrid sg;
  double rMax;        // max particle bounding radius
  double halfSkin;    // candidate pairs are reused until a body has moved this far since they were built
  double boxRMax;     // max scaled grid-box radius over particles (ghost search reach)
  long long pairCap, contactCap;
  unsigned long long hashMask;

  // ---- shapes ----
  const ShapeDev* shapes;
  const ClusterDev* clusters;
  const NodeDev* nodes;
  const ClusterF32* clusters32;
  const ClusterF32* subs32;  // sub-clusters of <= 16 nodes
  float4* pairMap;        // per evaluated pair (list position): 160-byte FP32 map record
  const float4* nodes32;  // same nodes in FP32 (x, y, z, 0) for the conservative prefilter
  const float* sdf;
  int maxClusters;        // largest cluster count of any shape

  // ---- particles ----
  d3* pos;      // state position (after any periodic wrap)
  d3* posDet;   // position used by the particle-particle and wall detection of this step
  q4* quat;
  d3* vel;
  d3* angvel;
  d3* force;
  d3* torque;
  const s6* invI;
  const double* invM;
  const double* radius;
  const double* kn;
  const double* ks;
  const double* mu;
  const double* res;
  const double* scale;
  const double* invSp;   // 1 / (scale * spacing)
  const d3* gorig;       // scale * grid origin
  const int* shape;
  const int* nodeStart;  // N+1 prefix of boundary-node counts
  int* scell;            // search-grid cell of pos (post-wrap), rebuild steps only
  d3* xBuild;            // posDet when the candidate pairs were built
  d3* xBuildW;           // pos when the candidate pairs were built
  q4* qBuild;            // quat when the candidate pairs were built
  int* imgBuild;         // image shift states at the build (bit g: image g lies a period away)
  BodyStep* step;        // per particle, written every step: motion since the build and keys
  float4* qrel;          // per particle: conj(q_build) q in FP32 (rotation since the build)

  // ---- walls ----
  d3* wpos;
  q4* wquat;
  d3* wvel;
  d3* wangvel;
  const double* wmu;
  const double* wres;
  const int* wshape;
  int* wkey;
  d3* wBuildPos;         // wall pose when the candidate pairs were built
  q4* wBuildQ;
  float4* wMotion;       // as motion, for walls
  float4* wSph;          // FP32 bounding sphere of each wall's grid box at its current pose (centre,
                         // boxR), written at every drift: the pair search's first wall test
  float4* wQrel;         // as qrel, for walls

  // ---- periodic ghost images (3 sets: +x, +y, +x+y; with sector periodicity, turned by +90, 180
  // and -90 degrees about the z axis through gridP.mn) ----
  d3* gpos[3];       // image positions, written only for images that exist (key >= 0 in the step
                     // record); the keys live in the step record
  q4* gquat[3];      // sector images only: orientation and velocity of each existing image
  d3* gvel[3];
  q4* quatDet;       // sector periodicity: orientation used for particle-particle detection (before
                     // this step's turn of a body that left the sector)
  d3 secW[3];        // rotation vectors (0, 0, angle) of the three sector image sets
  double secA[3], secB[3];  // their Rodrigues coefficients, computed on the host
  q4 secR[3];        // their unit quaternions, computed on the host

  // ---- search grid of particles ----
  int* cellCount;   // ncell (scratch)
  int* cellStart;   // ncell + 1
  int* sorted;      // nP: particle ids in (cell, id) order
  int* rank;        // nP (scratch)
  // particle data gathered into sorted order (contiguous runs per search-grid row)
  float4* sPR;      // FP32 (posDet, bounding radius; radius -1 when the particle has no reference key)
  float4* sPRw;     // FP32 (pos after the periodic wrap, bounding radius)
  float* sBoxR;     // scaled grid-box radius (ghost-image reach)
  int* rcTmp;       // reaction pairs before the per-partner sort
  int* sortTmp;     // counting-sort scratch (items before the per-bucket sort)
  int* sortBig;     // counting-sort scratch (list of buckets of more than 16 items)
  int* scanTiles;   // scan scratch (tile sums)
  uint32_t* pairSlab;  // kPairSlab x nP: first candidate pairs of each master (column per master)

  // ---- candidate pairs (current step) ----
  int* pairCount;   // nP (+1) counts then exclusive prefix pairStart
  int* pairStart;   // nP + 1
  int* pairA;       // pairCap
  uint32_t* pairInfo;  // (type << 28) | partner
  int* pairCoff;    // first contact slot of the pair (its contacts are contiguous)
  int* pairCcnt;    // contacts of the pair
  // separation certificates: a pair whose nodes were all culled with slack s (world units, less
  // the motion bound at that step) is skipped while the motion bound since the build stays below s
  int* oldStart;           // the previous pair list kept through a rebuild: per-master starts (nP + 1),
  uint32_t* oldInfo;       // pair entries and certificates, for the certificates the new list inherits
  float4* oldCert;
  float4* cert;     // pairCap: (s_build or -1, cB, cA, s or -1): s_build against the motion since the
                    // build, s against the motion since the pair's last evaluation (below)
  float4* certE;    // pairCap: at the last evaluation, dA - dB (x, y, z) and the image's flip state (w)
  float4* certQA;   // pairCap: at the last evaluation, the master's qrel
  float4* certQB;   // pairCap: at the last evaluation, the partner's qrel
  int* pairRec;     // pair -> position of its map record in pairMap this step
  int useCert;
  // reaction CSR (pairs whose partner is a particle that receives the reaction)
  int* rcCount;     // nP
  int* rcStart;     // nP + 1
  int* rcPair;      // pairCap
  int* rcRank;      // pairCap (scratch)

  // ---- contacts (current step) ----
  int* cNode;
  uint32_t* cInfo;      // (type << 28) | partner, kInvalidInfo for an unused slot
  int* cPair;           // candidate pair of the contact
  d3* cF;               // force on the master particle
  d3* cTA;              // torque on the master particle
  d3* cTB;              // torque on the partner
  d3* cSpring;
  d3* cPoint;
  d3* cNormal;
  double* cOverlap;
  double* cEn;
  double* cEs;
  // history: two hash tables keyed by (list, node, partner), used in alternate steps; this step
  // inserts its contacts into one and looks up the previous step's springs in the other
  const d3* prevSpring;
  unsigned long long* hashKey;        // this step's table
  int* hashVal;
  int* hashSlot;                      // slot taken by each contact slot (-1: none), per table
  const unsigned long long* prevHashKey;  // the previous step's table
  const int* prevHashVal;
  int curBuf;                         // which of the two tables and contact buffers is this step's

  // ---- bonds ----
  const int* bI;
  const int* bJ;
  int* bActive;
  const double* bConst;  // 9 per bond: B1 B2 B3 B4 radius L0 sigmaT cohesion mu
  const d3* bLocal;      // 8 per bond: n1i n2i n3i pi n1j n2j n3j pj
  d3* bPoint;
  double* bOut;          // 6 per bond: Un Us Ub Ut sigmaMax tauMax
  d3* bF;                // force on i
  d3* bMi;               // torque on i (moment + arm x F)
  d3* bMj;               // torque on j
  const int* bcStart;    // nP + 1 CSR of bond ends per particle
  const int* bcItem;     // 2*nBonds: (bond << 1) | (0 = i end, 1 = j end)

  // ---- device counters ----
  // devInt: [1] contact slots allocated this step, [3] error flags (kErr*: any flag halts every later
  //         kernel of the call),
  //         [4] peak pairs, [5] peak candidate slots, [6] slots scanned by the history insert,
  //         [7] images at their source's own position, [8] rebuild flag of the step,
  //         [9] (unused), [10 + b] contact slots of the
  //         last step that filled history table b, [12] grid-barrier arrivals of the broad phase,
  //         [20] callStep of the step that raised the first error flag (-1: none)
  int* devInt;
  int* barrier;          // grid barrier of the single-launch broad phase: kBarrierLine-int lines, one per
                         // block group (blockIdx mod 8), then the groups' counter, then the release word;
                         // then 8 lines of evalCount
  int* evalCount;        // word g * kBarrierLine: map records the pair check's block group g (blockIdx
                         // mod 8) wrote this step, in its own region of pairMap
  // devStat: [0] node tests, [1] sdf queries, [2] contacts (this step);
  //          [3] pairs, [4] node tests, [5] sdf queries, [6] contacts, [8] cluster entries
  //          (summed while profiling); [9] audits, [10] audited contacts, [11] missed contacts
  //          (cumulative); [12] master nodes of the candidate pairs, [13] evaluated pairs,
  //          [14] rebuild steps (summed while profiling)
  unsigned long long* devStat;
  unsigned long long* statSlots;  // 5 families x kStatSlots spread accumulators of the narrow phase,
                                  // kStatStride words apart

  // ---- learned culling surrogate ----
  SurNet sur;              // network (valid when surSlot != nullptr)
  const int* surSlot;      // family slot of each shape id, -1 outside the family; nullptr: untrained
  float surBand;           // cull band, unscaled SDF units (2 x measured max error)
  const int* nodeCluster;  // per shape node (caller order, at the shape's node offset): its cluster
  unsigned long long* surMask;  // per evaluated pair (list position): bit c = 0 when the network
                                // culls cluster c
  float* surSlack;              // per evaluated pair: smallest slack of the clusters it culled
  unsigned char* bodyHit;       // per particle: 1 when a pair of it holds candidate nodes this step
                                // (set by the narrow phase, cleared by the sums that read it)
  unsigned char* forceSet;      // per particle: 0 when its stored force and torque are exactly zero
};

// What opens a step besides the bodies: the table of tangential history this step fills, its slot
// list from two steps ago, the step's buffer (flag set, clear count) and the host's rebuild request.
struct StepHead {
  unsigned long long* hashKey;
  const int* hashSlot;
  int buf;
  int forceRebuild;
};

// Error flags in devInt[3].  Any of them halts the call: every kernel launched after the flag is up
// returns at once (the gated broad-phase launches find their gate cleared), so the host can restore
// the state of the call's start, grow the buffers and run the call again.
constexpr int kErrPairs = 1;      // candidate pairs beyond pairCap
constexpr int kErrContacts = 2;   // candidate slots beyond contactCap
constexpr int kErrBarrier = 4;    // a grid barrier of the single-launch broad phase gave up waiting
constexpr int kErrHash = 8;       // a history table probe found no free slot
constexpr int kErrStepIdx = 20;   // devInt index of the failing step

// Copy of the state at the start of an lsdem_step call, so that a call that halted on an error can
// be run again from there (after the buffers grew): the bodies' state, the walls' poses, the bonds'
// flags and the contact slots of the step before the call (their springs are the next step's
// history; the table that finds them is rebuilt from the slots).
struct SnapDev {
  d3 *pos, *vel, *ang, *force, *torque, *wpos;
  q4 *quat, *wquat;
  int* bActive;
  unsigned char* forceSet;
  unsigned long long* devStat;  // kDevStats counters
  int* meta;                    // [0] history slots kept
  int* hNode;                   // history slots: node, info, spring
  uint32_t* hInfo;
  d3* hSpring;
  long long hCap;
  // the live contact buffer of the step before the call (1 - cur at the call's start)
  int* prevNode;
  uint32_t* prevInfo;
  d3* prevSpring;
  int prevBuf;
};

constexpr int kStatSlots = 32;
// bodies up to which the drift after a step closes the step's counters with a wavefront (k_reduce_drift_small)
constexpr int kWaveEndBodies = 65536;
// u64 words per stat slot: each slot has a 128-byte line of its own, since atomics on one line are
// served one after another (~10 ns each), whatever words of the line they touch
constexpr int kStatStride = 16;
constexpr int kStatWords = 5 * kStatSlots * kStatStride;
constexpr int kDevInts = 32;   // devInt entries
constexpr int kBarrierLine = 32;                  // ints per barrier word (one 128-byte line each)
constexpr int kBarrierInts = 18 * kBarrierLine;   // 8 groups, the groups' counter, the release word, evalCount
// The step's rebuild flag and its count of ghost images left at their source's position, two sets
// used in alternate steps (buffer = Params::curBuf): the kernel that closes one step may already
// raise the next step's.
__host__ __device__ inline int rebuildFlagIdx(int buf) { return 16 + buf; }
__host__ __device__ inline int ghostCountIdx(int buf) { return 18 + buf; }
constexpr int kDevStats = 16;  // devStat entries
constexpr int kPairSlabHost = 16;
constexpr int kBroadThreadsHost = 1024;  // block of the single-launch broad phase (kBroadThreads)  // slab depth of the broad phase (kPairSlab in the kernels)
constexpr uint32_t kInvalidInfo = 0xFFFFFFFFu;  // unused contact slot

}  // namespace lsd
