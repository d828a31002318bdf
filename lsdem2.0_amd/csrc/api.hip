// Host side of liblsdem.so: context, uploads, per-step launch sequence and the C ABI of
// include/lsdem.h.  No CPU fallback exists: every step runs on the GPU.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/lsdem.h"
#include "device.h"
#include "surrogate_train.h"

namespace lsd {
// kernels.hip
__global__ void k_fill_int(int* a, long long n, int v, const int* gate);
__global__ void k_fill_u64(unsigned long long* a, long long n, unsigned long long v);
__global__ void k_scan_reduce(const int* in, long long n, int* blockSum, const int* gate);
__global__ void k_scan_apply(const int* in, long long n, const int* blockSum, int* out, const int* gate);
template <bool SEC> __global__ void k_kick_drift(Params P, int forceRebuild);
__global__ void k_search_cells(Params P);
template <bool SEC> __global__ void k_reduce_drift(Params P, StepHead H);
__global__ void k_reduce_drift_small(Params P, StepHead H);
__global__ void k_broad_phase(Params P);
__global__ void k_bucket_count(const int* bucket, int n, int* count, int* rank, const int* gate);
__global__ void k_bucket_scatter(const int* bucket, int n, const int* start, const int* rank, int* out,
                                 int* bigCount, const int* gate);
__global__ void k_bucket_sort(const int* start, int nb, const int* tmp, int* items, int* bigList, int* bigCount,
                              const int* gate, int nItems);
__global__ void k_bucket_sort_big(const int* start, int nb, const int* tmp, int* items, const int* bigList,
                                  const int* bigCount, const int* gate, int nItems);
__global__ void k_mark_build(Params P);
__global__ void k_pairs_find(Params P);
__global__ void k_pairs_emit(Params P);
__global__ void k_backup_pairs(Params P);
__global__ void k_inherit_certs(Params P);
__global__ void k_rc_count(Params P);
__global__ void k_rc_scatter(Params P);
__global__ void k_gather_particles(Params P);
template <bool SEC> __global__ void k_pair_check(Params P, int full);
template <bool SUR, bool SEC> __global__ void k_candidates(Params P);
__global__ void k_sur_audit(Params P);
__global__ void k_sur_masks(Params P);
size_t candidatesLdsBytes(int maxClusters);
__global__ void k_bonds(Params P);
__global__ void k_reduce_kick(Params P, int kick);
__global__ void k_kick(Params P);
__global__ void k_prime_end(Params P);
__global__ void k_snap_save(Params P, SnapDev S);
__global__ void k_snap_restore(Params P, SnapDev S);
__global__ void k_hist_reinsert(Params P, SnapDev S, unsigned long long* key, int* val, int* slotOf);
}  // namespace lsd

using namespace lsd;

namespace {

thread_local std::string g_err;
int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess)                                                          \
      return fail(LSDEM_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct HostShape {
  int nx, ny, nz;
  d3 origin;
  double h;
  std::vector<float> sdf;
  std::vector<d3> nodes;
  std::vector<NodeDev> perm;        // nodes in cluster order
  std::vector<ClusterDev> clusters;
  std::vector<ClusterDev> subs;     // sub-clusters of <= 16 nodes
  double boxR = 0, lip = 0;
};

// Longest-axis median splits of idx[first, first + count) into leaves of <= leaf items, left
// parts kept at multiples of `leaf` where possible; leaves returned in index order.
struct Seg { int first, count; };
std::vector<Seg> splitLeaves(const std::vector<d3>& pts, std::vector<int>& idx, Seg root, int leaf) {
  std::vector<Seg> work{root}, leaves;
  while (!work.empty()) {
    Seg s = work.back();
    work.pop_back();
    if (s.count <= leaf) { if (s.count > 0) leaves.push_back(s); continue; }
    d3 lo = pts[idx[s.first]], hi = lo;
    for (int k = s.first; k < s.first + s.count; ++k) {
      const d3 p = pts[idx[k]];
      lo = mk3(std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z));
      hi = mk3(std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z));
    }
    const d3 ext = hi - lo;
    const int axis = (ext.x >= ext.y && ext.x >= ext.z) ? 0 : (ext.y >= ext.z ? 1 : 2);
    int left = ((s.count + 1) / 2 + leaf - 1) / leaf * leaf;
    left = std::min(left, s.count - 1);
    auto coord = [&](int i) { const d3 p = pts[i]; return axis == 0 ? p.x : axis == 1 ? p.y : p.z; };
    std::nth_element(idx.begin() + s.first, idx.begin() + s.first + left, idx.begin() + s.first + s.count,
                     [&](int a2, int b2) { return coord(a2) < coord(b2) || (coord(a2) == coord(b2) && a2 < b2); });
    work.push_back({s.first + left, s.count - left});
    work.push_back({s.first, left});
  }
  std::sort(leaves.begin(), leaves.end(), [](const Seg& x, const Seg& y) { return x.first < y.first; });
  return leaves;
}

// Bounding sphere (box centre, max distance) of the nodes idx[s].
ClusterDev boundSeg(const std::vector<d3>& pts, const std::vector<int>& idx, Seg s) {
  d3 lo = pts[idx[s.first]], hi = lo;
  for (int k = s.first; k < s.first + s.count; ++k) {
    const d3 p = pts[idx[k]];
    lo = mk3(std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z));
    hi = mk3(std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z));
  }
  const d3 c = 0.5 * (lo + hi);
  double rho = 0;
  for (int k = s.first; k < s.first + s.count; ++k) rho = std::max(rho, len(pts[idx[k]] - c));
  ClusterDev cl{};
  cl.cx = c.x; cl.cy = c.y; cl.cz = c.z;
  cl.rho = rho * (1.0 + 1e-12);
  cl.first = s.first;
  cl.count = s.count;
  return cl;
}

// Node hierarchy of a shape: clusters of <= 64 nodes, each split into sub-clusters of <= 16,
// both spatially compact, with bounding spheres.  Nodes are stored in cluster (and sub-cluster)
// order; a cluster's sub-clusters are consecutive (subFirst local to the shape, nSub).
void buildClusters(HostShape& S) {
  std::vector<int> idx(S.nodes.size());
  std::iota(idx.begin(), idx.end(), 0);
  const std::vector<Seg> leaves = splitLeaves(S.nodes, idx, Seg{0, int(idx.size())}, 64);
  S.perm.clear();
  S.clusters.clear();
  S.subs.clear();
  for (const Seg& s : leaves) {
    const std::vector<Seg> subs = splitLeaves(S.nodes, idx, s, 16);
    ClusterDev cl = boundSeg(S.nodes, idx, s);
    cl.subFirst = int(S.subs.size());
    cl.nSub = int(subs.size());
    for (const Seg& u : subs) S.subs.push_back(boundSeg(S.nodes, idx, u));
    S.clusters.push_back(cl);
  }
  for (size_t k = 0; k < idx.size(); ++k) {
    const d3 p = S.nodes[idx[k]];
    NodeDev nd{};
    nd.x = p.x; nd.y = p.y; nd.z = p.z; nd.orig = idx[k];
    S.perm.push_back(nd);
  }
}

// Grid-box radius and a Lipschitz bound of the trilinear interpolant over the cells with a
// positive corner (the only cells a path from a point with phi > 0 to the first zero of phi
// crosses, so phi(x) <= lip * |x - y| for the nearest y with phi(y) <= 0).  Inside one cell each
// partial derivative is a bilinear blend of the 4 edge differences, so it is bounded by the
// largest of them.
void shapeBounds(HostShape& S) {
  double br = 0;
  for (int c = 0; c < 8; ++c) {
    const d3 p = mk3(S.origin.x + ((c & 1) ? (S.nx - 1) * S.h : 0.), S.origin.y + ((c & 2) ? (S.ny - 1) * S.h : 0.),
                     S.origin.z + ((c & 4) ? (S.nz - 1) * S.h : 0.));
    br = std::max(br, len(p));
  }
  S.boxR = br * (1.0 + 1e-12);
  auto at = [&](int i, int j, int k) { return double(S.sdf[(size_t(k) * S.ny + j) * S.nx + i]); };
  double L = 0;
  for (int k = 0; k + 1 < S.nz; ++k)
    for (int j = 0; j + 1 < S.ny; ++j)
      for (int i = 0; i + 1 < S.nx; ++i) {
        bool pos = false;
        for (int c = 0; c < 8 && !pos; ++c) pos = at(i + (c & 1), j + ((c >> 1) & 1), k + (c >> 2)) > 0.0;
        if (!pos) continue;
        double gx = 0, gy = 0, gz = 0;
        for (int a = 0; a < 2; ++a)
          for (int b = 0; b < 2; ++b) {
            gx = std::max(gx, std::fabs(at(i + 1, j + a, k + b) - at(i, j + a, k + b)));
            gy = std::max(gy, std::fabs(at(i + a, j + 1, k + b) - at(i + a, j, k + b)));
            gz = std::max(gz, std::fabs(at(i + a, j + b, k + 1) - at(i + a, j + b, k)));
          }
        L = std::max(L, std::sqrt(gx * gx + gy * gy + gz * gz) / S.h);
      }
  S.lip = L * (1.0 + 1e-9);
}

template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  hipError_t alloc(size_t count) {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = count;
    if (count == 0) return hipSuccess;
    return hipMalloc(&p, count * sizeof(T));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

}  // namespace

struct lsdem_ctx {
  lsdem_config cfg;
  lsdem_external_force_fn hookFn = nullptr;
  void* hookUser = nullptr;
  double time = 0.0;
  hipStream_t stream = nullptr;
  bool initialized = false;
  int numCU = 256;
  int narrowBlocks = 2048;
  int coopBlocks = 0;  // grid of the single-launch broad phase (0: separate launches)
  // host staging
  std::vector<HostShape> shapes;
  std::vector<int> pShape;
  std::vector<double> pScale, pInvM, pKn, pKs, pMu, pRes, pRadius;
  std::vector<d3> pPos, pVel, pAng;
  std::vector<q4> pQuat;
  std::vector<s6> pInvI;
  std::vector<int> wShape;
  std::vector<double> wRadius, wMu, wRes;
  std::vector<d3> wPos, wVel, wAng;
  std::vector<q4> wQuat;
  // bonds (host)
  std::vector<int> bI, bJ, bActive;
  std::vector<double> bConst;
  std::vector<d3> bLocal, bPoint;

  // device
  DBuf<ShapeDev> dShapes;
  DBuf<ClusterDev> dClusters;
  DBuf<NodeDev> dNodes;
  DBuf<float4> dNodes32, dPairMap;
  DBuf<unsigned long long> dSurMask;
  DBuf<ClusterF32> dClusters32, dSubs32;
  DBuf<d3> dXBuild, dXBuildW, dWBuildPos;
  DBuf<q4> dQBuild;
  DBuf<int> dImgBuild, dPairRec;
  DBuf<BodyStep> dStep;
  DBuf<float4> dQrel, dWMotion, dWSph, dWQrel, dCert, dCertE, dCertQA, dCertQB;
  DBuf<float> dSurSlack;
  DBuf<unsigned char> dBodyHit, dForceSet;
  DBuf<q4> dWBuildQ;
  bool forceRebuild = true;
  bool primePending = false;  // the pre-loop build runs before the next step
  bool haveContacts = false;  // a detection has filled the contact buffers
  int phaseNext = 0;          // next lsdem_phase_* call of a step driven phase by phase (0: none begun)
  // the state at the start of the running call (see SnapDev), with the host's own step bookkeeping
  struct Snap {
    DBuf<d3> pos, vel, ang, force, torque, wpos, hSpring;
    DBuf<q4> quat, wquat;
    DBuf<int> bActive, hNode, meta;
    DBuf<uint32_t> hInfo;
    DBuf<unsigned char> forceSet;
    DBuf<unsigned long long> devStat;
    long long hCap = 0;
    long long stepCount = 0, audits = 0, profSeen = 0;
    double time = 0.0;
    int cur = 0, profSteps = 0;
    bool primePending = false, forceRebuild = false, haveContacts = false;
    void release() {
      pos.release(); vel.release(); ang.release(); force.release(); torque.release(); wpos.release(); hSpring.release();
      quat.release(); wquat.release(); bActive.release(); hNode.release(); meta.release(); hInfo.release();
      forceSet.release(); devStat.release();
    }
  } snap;
  int recoveries = 0;  // calls run again after an error (buffers grown or broad phase switched)
  DBuf<int> dNodeCluster;
  // learned culling surrogate
  struct Sur {
    bool trained = false, enabled = false;
    int nShapes = 0;
    std::vector<int> famShape;  // slot -> shape id
    float invLen = 1.f, clamp = 0.6f, band = 0.f;
    double maxErr = 0, meanErr = 0, loss = 0, trainMs = 0;
    long long iterations = 0, audits = 0;
    int auditEvery = 100;
    DBuf<float> param, embF, m, v, partial, loss1, sx, st, vx, vt, out, errSum;
    DBuf<int> sslot, vslot, famShapeD, slotOfShape;
    DBuf<sur_v4s> wF, wB;
    DBuf<uint16_t> emb;
    DBuf<unsigned int> errBits;
    void release() {
      param.release(); embF.release(); m.release(); v.release(); partial.release(); loss1.release();
      sx.release(); st.release(); vx.release(); vt.release(); out.release(); errSum.release();
      sslot.release(); vslot.release(); famShapeD.release(); slotOfShape.release();
      wF.release(); wB.release(); emb.release(); errBits.release();
    }
  } sur;
  DBuf<float> dSdf;
  DBuf<d3> dPos, dPosDet, dVel, dAng, dForce, dTorque, dGorig;
  DBuf<q4> dQuat;
  DBuf<s6> dInvI;
  DBuf<double> dInvM, dRadius, dKn, dKs, dMu, dRes, dScale, dInvSp;
  DBuf<int> dShape, dNodeStart, dScell;
  DBuf<d3> dWPos, dWVel, dWAng;
  DBuf<q4> dWQuat;
  DBuf<double> dWMu, dWRes;
  DBuf<int> dWShape, dWKey;
  DBuf<d3> dGPos[3];
  DBuf<q4> dGQuat[3];  // sector images: orientations and velocities
  DBuf<d3> dGVel[3];
  DBuf<q4> dQuatDet;   // sector periodicity: detection orientation of the particles
  DBuf<int> dCellCount, dCellStart, dSorted, dRank;
  DBuf<float4> dSPR, dSPRw;
  DBuf<float> dSBoxR;
  DBuf<uint32_t> dPairSlab;
  DBuf<int> dPairCount, dPairStart, dPairA, dPairCoff, dPairCcnt, dRcCount, dRcStart, dRcPair, dRcRank, dRcTmp;
  DBuf<int> dSortTmp, dSortBig;
  DBuf<uint32_t> dPairInfo, dOldInfo;
  DBuf<int> dOldStart;
  DBuf<float4> dOldCert;
  DBuf<int> dCNode[2], dCPair;
  DBuf<d3> dCF, dCTA, dCTB;
  DBuf<uint32_t> dCInfo[2];
  DBuf<d3> dCSpring[2], dCPoint[2], dCNormal[2];
  DBuf<double> dCOverlap[2], dCEn[2], dCEs[2];
  DBuf<unsigned long long> dHashKey[2];
  DBuf<int> dHashVal[2], dHashSlot[2];
  DBuf<unsigned long long> dStatSlots;
  DBuf<int> dBI, dBJ, dBActive, dBcStart, dBcItem;
  DBuf<double> dBConst, dBOut;
  DBuf<d3> dBLocal, dBPoint, dBF, dBMi, dBMj;
  DBuf<int> dDevInt, dScanBlocks, dBarrier;
  DBuf<unsigned long long> dDevStat;
  long long hashSize = 0;
  int scanBlocksCap = 0;
  Params P{};
  long long stepCount = 0;
  int cur = 0;  // contact buffer written by the next step
  // profiling: kPhaseEvents events per profiled step
  std::vector<hipEvent_t> events;
  std::vector<int> eventOf;  // per profiled step and phase mark: the event slot recorded for it
  int profMax = 0, profSteps = 0;
  int profEvery = 1;
  long long profSeen = 0;  // steps since lsdem_profile_begin
  bool profiling = false;
  // steps captured as one HIP graph (lsdem_step_graph), for graphN steps per launch
  hipGraph_t graph = nullptr;
  hipGraphExec_t graphExec = nullptr;
  int graphN = 0;
  void dropGraph() {
    if (graphExec) (void)hipGraphExecDestroy(graphExec);
    if (graph) (void)hipGraphDestroy(graph);
    graphExec = nullptr;
    graph = nullptr;
    graphN = 0;
  }

  ~lsdem_ctx() {
    dropGraph();
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
    releaseAll();
    if (stream) (void)hipStreamDestroy(stream);
  }
  void releaseAll();
};

void lsdem_ctx::releaseAll() {
  dShapes.release(); dClusters.release(); dNodes.release(); dNodes32.release(); dPairMap.release(); dSurMask.release(); dClusters32.release(); dSubs32.release(); dSdf.release();
  dCPair.release(); dCF.release(); dCTA.release(); dCTB.release();
  dPos.release(); dPosDet.release(); dVel.release(); dAng.release(); dForce.release(); dTorque.release();
  dGorig.release(); dQuat.release(); dInvI.release();
  dInvM.release(); dRadius.release(); dKn.release(); dKs.release(); dMu.release(); dRes.release();
  dScale.release(); dInvSp.release(); dShape.release(); dNodeStart.release(); dScell.release();
  dWPos.release(); dWVel.release(); dWAng.release(); dWQuat.release(); dWMu.release(); dWRes.release();
  dWShape.release(); dWKey.release();
  for (int s = 0; s < 3; ++s) {
    dGPos[s].release();
    dGQuat[s].release();
    dGVel[s].release();
  }
  dCellCount.release(); dCellStart.release(); dSorted.release(); dRank.release();
  dSPR.release(); dSPRw.release(); dPairSlab.release(); dStatSlots.release();
  dSBoxR.release();
  dPairCount.release(); dPairStart.release(); dPairA.release(); dPairCoff.release(); dPairCcnt.release();
  dRcCount.release(); dRcStart.release(); dRcPair.release(); dRcRank.release(); dRcTmp.release();
  dSortTmp.release(); dSortBig.release();
  dPairInfo.release(); dOldInfo.release(); dOldStart.release(); dOldCert.release();
  for (int b = 0; b < 2; ++b) {
    dCNode[b].release(); dCInfo[b].release(); dCSpring[b].release(); dCPoint[b].release();
    dCNormal[b].release(); dCOverlap[b].release(); dCEn[b].release(); dCEs[b].release();
  }
  for (int b = 0; b < 2; ++b) {
    dHashKey[b].release(); dHashVal[b].release(); dHashSlot[b].release();
  }
  dBI.release(); dBJ.release(); dBActive.release(); dBcStart.release(); dBcItem.release();
  dBConst.release(); dBOut.release(); dBLocal.release(); dBPoint.release(); dBF.release(); dBMi.release();
  dBMj.release();
  dBarrier.release(); dDevInt.release(); dScanBlocks.release(); dDevStat.release();
  dNodeCluster.release();
  dXBuild.release(); dXBuildW.release(); dWBuildPos.release(); dWBuildQ.release();
  dQBuild.release(); dImgBuild.release(); dStep.release(); dPairRec.release();
  dQrel.release(); dWQrel.release(); dCertE.release(); dCertQA.release(); dCertQB.release();
  dWMotion.release(); dWSph.release(); dCert.release(); dSurSlack.release(); dBodyHit.release(); dForceSet.release();
  sur.release();
  snap.release();
}

namespace {

GridGeom makeGrid(d3 mn, d3 mx, double cell) {  // spatialGrid::set
  GridGeom G{};
  if (mx.x <= mn.x) mx.x = mn.x + cell;
  if (mx.y <= mn.y) mx.y = mn.y + cell;
  if (mx.z <= mn.z) mx.z = mn.z + cell;
  G.mn = mn;
  G.mx = mx;
  G.size = i3{2, 2, 2};
  G.inv = mk3(0, 0, 0);
  if (isZero(cell)) return G;
  const d3 d = mx - mn;
  G.size.x = d.x > 2. * cell ? int(d.x / cell) : 2;
  G.size.y = d.y > 2. * cell ? int(d.y / cell) : 2;
  G.size.z = d.z > 2. * cell ? int(d.z / cell) : 2;
  G.inv.x = (d.x > 0.0) ? (double(G.size.x) / d.x) : 0.0;
  G.inv.y = (d.y > 0.0) ? (double(G.size.y) / d.y) : 0.0;
  G.inv.z = (d.z > 0.0) ? (double(G.size.z) / d.z) : 0.0;
  return G;
}

template <class T>
int upload(DBuf<T>& b, const std::vector<T>& v, hipStream_t s) {
  HIPCHK(b.alloc(std::max<size_t>(v.size(), 1)));
  if (!v.empty()) HIPCHK(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
  return LSDEM_OK;
}

inline unsigned blocksFor(long long n, int bs) { return unsigned(std::max<long long>(1, (n + bs - 1) / bs)); }

// Exclusive scan (out[n] = total).  `gate`: run only when *gate != 0 (nullptr: always).
int scan(lsdem_ctx* c, const int* in, long long n, int* out, const int* gate) {
  const int nb = int(std::max<long long>(1, (n + 2047) / 2048));
  if (nb > c->scanBlocksCap) return fail(LSDEM_ERR_STATE, "scan buffer too small");
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(256), 0, c->stream, in, n, c->dScanBlocks.p, gate);
  hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(256), 0, c->stream, in, n, (const int*)c->dScanBlocks.p, out, gate);
  return LSDEM_OK;
}

// Deterministic counting sort of `n` items by bucket id (ids < 0 skipped), gated like scan.
int bucketSort(lsdem_ctx* c, const int* bucket, int n, int nb, int* count, int* start, int* rank, int* out,
               const int* gate) {
  const int bs = 256;
  hipLaunchKernelGGL(k_fill_int, dim3(blocksFor(nb, bs)), dim3(bs), 0, c->stream, count, (long long)nb, 0, gate);
  if (n > 0)
    hipLaunchKernelGGL(k_bucket_count, dim3(blocksFor(n, bs)), dim3(bs), 0, c->stream, bucket, n, count, rank, gate);
  int rc = scan(c, count, nb, start, gate);
  if (rc) return rc;
  if (n > 0) {
    int* bigCount = c->P.devInt + 2;
    hipLaunchKernelGGL(k_bucket_scatter, dim3(blocksFor(n, bs)), dim3(bs), 0, c->stream, bucket, n,
                       (const int*)start, (const int*)rank, c->dSortTmp.p, bigCount, gate);
    hipLaunchKernelGGL(k_bucket_sort, dim3(blocksFor(nb, bs)), dim3(bs), 0, c->stream, (const int*)start, nb,
                       (const int*)c->dSortTmp.p, out, c->dSortBig.p, bigCount, gate, n);
    hipLaunchKernelGGL(k_bucket_sort_big, dim3(unsigned(c->numCU)), dim3(256), 0, c->stream, (const int*)start, nb,
                       (const int*)c->dSortTmp.p, out, (const int*)c->dSortBig.p, (const int*)bigCount, gate, n);
  }
  return LSDEM_OK;
}

void setCurrentBuffers(lsdem_ctx* c) {
  Params& P = c->P;
  const int b = c->cur, o = 1 - c->cur;
  P.cNode = c->dCNode[b].p; P.cInfo = c->dCInfo[b].p; P.cSpring = c->dCSpring[b].p;
  P.cPoint = c->dCPoint[b].p; P.cNormal = c->dCNormal[b].p; P.cOverlap = c->dCOverlap[b].p;
  P.cEn = c->dCEn[b].p; P.cEs = c->dCEs[b].p;
  P.prevSpring = c->dCSpring[o].p;
  P.hashKey = c->dHashKey[b].p; P.hashVal = c->dHashVal[b].p; P.hashSlot = c->dHashSlot[b].p;
  P.prevHashKey = c->dHashKey[o].p; P.prevHashVal = c->dHashVal[o].p;
  P.curBuf = b;
}

constexpr int kPhaseEvents = LSDEM_NUM_PHASES + 1;

void deviceView(lsdem_ctx* c, lsdem_device_view* v) {
  const Params& P = c->P;
  v->force = reinterpret_cast<double*>(P.force);
  v->torque = reinterpret_cast<double*>(P.torque);
  v->velocity = reinterpret_cast<double*>(P.vel);
  v->angular_velocity = reinterpret_cast<double*>(P.angvel);
  v->position = reinterpret_cast<const double*>(P.pos);
  v->quaternion = reinterpret_cast<const double*>(P.quat);
  v->num = P.nP;
  v->block_dim = c->cfg.block_size > 0 ? c->cfg.block_size : 256;
  v->grid_dim = (P.nP + v->block_dim - 1) / v->block_dim;
  v->stream = c->stream;
}

// The broad phase as separate launches, each gated on the device rebuild flag: search cells,
// counting sort and cell table, then the candidate pairs and reaction lists.
int hashLaunches(lsdem_ctx* c, const int* gate) {
  Params& P = c->P;
  hipStream_t s = c->stream;
  hipLaunchKernelGGL(k_search_cells, dim3(blocksFor(P.nP, 256)), dim3(256), 0, s, P);
  int rc = bucketSort(c, P.scell, P.nP, P.sg.ncell, P.cellCount, P.cellStart, P.rank, P.sorted, gate);
  if (rc) return rc;
  hipLaunchKernelGGL(k_gather_particles, dim3(blocksFor(P.nP, 256)), dim3(256), 0, s, P);
  return LSDEM_OK;
}

template <class Phase>
int pairLaunches(lsdem_ctx* c, const int* gate, Phase& phase) {
  Params& P = c->P;
  hipStream_t s = c->stream;
  int rc;
  hipLaunchKernelGGL(k_pairs_find, dim3(blocksFor(P.nP, 128)), dim3(128), 0, s, P);
  hipLaunchKernelGGL(k_backup_pairs, dim3(unsigned(c->numCU * 4)), dim3(256), 0, s, P);
  rc = scan(c, P.pairCount, P.nP, P.pairStart, gate);
  if (rc) return rc;
  hipLaunchKernelGGL(k_pairs_emit, dim3(blocksFor(P.nP, 128)), dim3(128), 0, s, P);
  hipLaunchKernelGGL(k_inherit_certs, dim3(unsigned(c->numCU * 4)), dim3(256), 0, s, P);
  hipLaunchKernelGGL(k_mark_build, dim3(blocksFor(std::max(P.nP, P.nW), 256)), dim3(256), 0, s, P);
  phase();
  // reaction CSR
  hipLaunchKernelGGL(k_fill_int, dim3(blocksFor(P.nP, 256)), dim3(256), 0, s, P.rcCount, (long long)P.nP, 0, gate);
  const unsigned gsPairs = unsigned(c->numCU * 8);
  hipLaunchKernelGGL(k_rc_count, dim3(gsPairs), dim3(256), 0, s, P);
  rc = scan(c, P.rcCount, P.nP, P.rcStart, gate);
  if (rc) return rc;
  hipLaunchKernelGGL(k_rc_scatter, dim3(gsPairs), dim3(256), 0, s, P);
  const int rcItems = int(std::min<long long>(P.pairCap, (1ll << 31) - 1));
  hipLaunchKernelGGL(k_bucket_sort, dim3(blocksFor(P.nP, 256)), dim3(256), 0, s, (const int*)P.rcStart, P.nP,
                     (const int*)P.rcTmp, P.rcPair, c->dSortBig.p, P.devInt + 2, gate, rcItems);
  hipLaunchKernelGGL(k_bucket_sort_big, dim3(unsigned(c->numCU)), dim3(256), 0, s, (const int*)P.rcStart, P.nP,
                     (const int*)P.rcTmp, P.rcPair, (const int*)c->dSortBig.p, (const int*)(P.devInt + 2), gate,
                     rcItems);
  return LSDEM_OK;
}

template <class Phase>
int broadPhaseLaunches(lsdem_ctx* c, const int* gate, Phase& phase) {
  const int rc = hashLaunches(c, gate);
  if (rc) return rc;
  phase();
  return pairLaunches(c, gate, phase);
}

// The narrow phase: pair check, then the exact node tests and the contact law (behind the
// surrogate's masks when it is on, except on audit steps, which run the exact path on every live
// pair that holds no certificate of the exact culls and then check the surrogate's decisions).
// Launches of the kernels instantiated for sector periodicity or not (a compile-time switch in the
// kernels, so that simulations without it run the code they ran before it existed).
void launchKickDrift(lsdem_ctx* c, int force) {
  const Params& P = c->P;
  const int bs = c->cfg.block_size > 0 ? c->cfg.block_size : 256;
  if (P.periodicSector)
    hipLaunchKernelGGL(k_kick_drift<true>, dim3(blocksFor(P.nP + P.nW, bs)), dim3(bs), 0, c->stream, P, force);
  else
    hipLaunchKernelGGL(k_kick_drift<false>, dim3(blocksFor(P.nP + P.nW, bs)), dim3(bs), 0, c->stream, P, force);
}
void launchReduceDrift(lsdem_ctx* c, const StepHead& H) {
  const Params& P = c->P;
  const int bs = c->cfg.block_size > 0 ? c->cfg.block_size : 256;
  if (P.periodicSector)
    hipLaunchKernelGGL(k_reduce_drift<true>, dim3(blocksFor(P.nP + P.nW, bs)), dim3(bs), 0, c->stream, P, H);
  else if (P.nP <= kWaveEndBodies)
    hipLaunchKernelGGL(k_reduce_drift_small, dim3(blocksFor(P.nP + P.nW, bs)), dim3(bs), 0, c->stream, P, H);
  else
    hipLaunchKernelGGL(k_reduce_drift<false>, dim3(blocksFor(P.nP + P.nW, bs)), dim3(bs), 0, c->stream, P, H);
}
template <bool SUR>
void launchCandidatesKernel(lsdem_ctx* c) {
  const Params& P = c->P;
  const unsigned lds = unsigned(candidatesLdsBytes(P.maxClusters));
  if (P.periodicSector)
    hipLaunchKernelGGL((k_candidates<SUR, true>), dim3(unsigned(c->narrowBlocks)), dim3(256), lds, c->stream, P);
  else
    hipLaunchKernelGGL((k_candidates<SUR, false>), dim3(unsigned(c->narrowBlocks)), dim3(256), lds, c->stream, P);
}

bool auditStep(lsdem_ctx* c) {
  const bool useSur = c->sur.enabled && c->sur.trained;
  return useSur && c->sur.auditEvery > 0 && (c->stepCount % c->sur.auditEvery == 0);
}
void launchPairCheck(lsdem_ctx* c, bool audit) {
  // (a multiple of 8 workgroups: the kernel gives each XCD an eighth of the pairs)
  const dim3 grid(unsigned((c->numCU * 4 + 7) / 8 * 8));
  if (c->P.periodicSector)
    hipLaunchKernelGGL(k_pair_check<true>, grid, dim3(256), 0, c->stream, c->P, audit ? 1 : 0);
  else
    hipLaunchKernelGGL(k_pair_check<false>, grid, dim3(256), 0, c->stream, c->P, audit ? 1 : 0);
}
void launchCandidates(lsdem_ctx* c, bool audit) {
  const Params& P = c->P;
  hipStream_t s = c->stream;
  const bool useSur = c->sur.enabled && c->sur.trained;
  if (useSur && !audit) hipLaunchKernelGGL(k_sur_masks, dim3(unsigned(c->narrowBlocks)), dim3(256), 0, s, P);
  if (useSur && !audit)
    launchCandidatesKernel<true>(c);
  else
    launchCandidatesKernel<false>(c);
  if (audit) {
    hipLaunchKernelGGL(k_sur_audit, dim3(unsigned(c->numCU * 2)), dim3(256), 0, s, P);
    c->sur.audits++;
  }
}

// The end of a step that is not fused with the next one: per-body sums and the second half-kick
// (split around the external-force hook when one is set).
void launchReduceKick(lsdem_ctx* c) {
  Params& P = c->P;
  const int bs = c->cfg.block_size > 0 ? c->cfg.block_size : 256;
  hipStream_t s = c->stream;
  if (c->hookFn) {
    hipLaunchKernelGGL(k_reduce_kick, dim3(blocksFor(P.nP, bs)), dim3(bs), 0, s, P, 0);
    lsdem_device_view v;
    deviceView(c, &v);
    c->hookFn(c->hookUser, c->time, &v);
    hipLaunchKernelGGL(k_kick, dim3(blocksFor(P.nP, bs)), dim3(bs), 0, s, P);
  } else {
    hipLaunchKernelGGL(k_reduce_kick, dim3(blocksFor(P.nP, bs)), dim3(bs), 0, s, P, 1);
  }
}

int launchBroadPhase(lsdem_ctx* c);
int scratchOn(int dev, unsigned char** out);

// One step.  headDone: the previous launch already ran this step's kick and drift.  fuseNext: the
// step ends with the next step's kick and drift (another step follows with no host work between).
int runStep(lsdem_ctx* c, bool headDone, bool fuseNext, int callStep) {
  Params& P = c->P;
  setCurrentBuffers(c);
  P.callStep = callStep;
  const int bs = c->cfg.block_size > 0 ? c->cfg.block_size : 256;
  hipStream_t s = c->stream;
  bool prof = false;
  if (c->profiling) {
    prof = c->profSteps < c->profMax && c->profSeen % c->profEvery == 0;
    c->profSeen++;
  }
  hipEvent_t* ev = prof ? &c->events[size_t(c->profSteps) * kPhaseEvents] : nullptr;
  int* evOf = prof ? &c->eventOf[size_t(c->profSteps) * kPhaseEvents] : nullptr;
  int mark = 0;
  // phase(false): the phase had no work; its boundary reuses the previous event (an event
  // record costs GPU time of its own)
  auto phase = [&](bool work = true) {
    if (prof && mark < kPhaseEvents) {
      if (work || mark == 0) {
        (void)hipEventRecord(ev[mark], s);
        evOf[mark] = mark;
      } else {
        evOf[mark] = evOf[mark - 1];
      }
    }
    ++mark;
  };
  phase();
  const int* gate = P.devInt + rebuildFlagIdx(P.curBuf);  // broad phase only on rebuild steps
  if (!headDone) {
    const int force = c->forceRebuild ? 1 : 0;
    c->forceRebuild = false;
    launchKickDrift(c, force);
  }
  phase(!headDone);
  if (c->coopBlocks > 0) {
    // one launch with grid barriers; all of its blocks return at once on steps that keep the pairs
    const int rc = launchBroadPhase(c);
    if (rc) return rc;
    phase();        // sort (the whole broad phase)
    phase(false);   // candidate pairs
  } else {
    const int rc = broadPhaseLaunches(c, gate, phase);
    if (rc) return rc;
  }
  phase(c->coopBlocks == 0);  // reaction lists
  // narrow phase + contact law
  const bool audit = auditStep(c);
  launchPairCheck(c, audit);
  phase();
  launchCandidates(c, audit);
  phase();  // (this step's contacts went into this step's history table inside k_candidates)
  if (P.nBonds > 0) hipLaunchKernelGGL(k_bonds, dim3(blocksFor(P.nBonds, bs)), dim3(bs), 0, s, P);
  phase(P.nBonds > 0);
  if (fuseNext && !c->hookFn) {
    const int nb = 1 - c->cur;  // the next step's buffers
    const StepHead H{c->dHashKey[nb].p, c->dHashSlot[nb].p, nb, c->forceRebuild ? 1 : 0};
    c->forceRebuild = false;
    launchReduceDrift(c, H);
  } else {
    launchReduceKick(c);
  }
  phase();
  if (prof) c->profSteps++;
  c->cur = 1 - c->cur;
  c->stepCount++;
  c->time += P.dt;
  return LSDEM_OK;
}

// The single-launch broad phase (its blocks meet at grid barriers).
int launchBroadPhase(lsdem_ctx* c) {
  Params& P = c->P;
  if (c->cfg.broad_phase_coop) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIPCHK(hipStreamIsCapturing(c->stream, &cs));
    if (cs == hipStreamCaptureStatusNone) {
      void* args[] = {&P};
      HIPCHK(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_broad_phase), dim3(unsigned(c->coopBlocks)),
                                        dim3(kBroadThreadsHost), args, 0, c->stream));
      return LSDEM_OK;
    }
  }
  hipLaunchKernelGGL(k_broad_phase, dim3(unsigned(c->coopBlocks)), dim3(kBroadThreadsHost), 0, c->stream, P);
  return LSDEM_OK;
}

// The pre-loop build of a solve phase (LSSolver.h:996-998, 1029-1031, 1061-1066, 1097-1104):
// keys, candidate pairs and contact lists at the current poses, the periodic wrap and the ghost
// images, with springs carried over; no integration, no forces, no step counted.
int runPrime(lsdem_ctx* c) {
  Params& P = c->P;
  setCurrentBuffers(c);
  const int bs = c->cfg.block_size > 0 ? c->cfg.block_size : 256;
  hipStream_t s = c->stream;
  P.prime = 1;
  P.callStep = 0;
  c->forceRebuild = false;
  launchKickDrift(c, 1);
  int rc = LSDEM_OK;
  if (c->coopBlocks > 0) {
    rc = launchBroadPhase(c);
  } else {
    auto none = [](bool = true) {};
    rc = broadPhaseLaunches(c, P.devInt + rebuildFlagIdx(P.curBuf), none);
  }
  if (rc == LSDEM_OK) {
    launchPairCheck(c, false);
    launchCandidatesKernel<false>(c);
    hipLaunchKernelGGL(k_prime_end, dim3(blocksFor(P.nP, bs)), dim3(bs), 0, s, P);
  }
  P.prime = 0;
  if (rc) return rc;
  c->cur = 1 - c->cur;
  c->primePending = false;
  c->haveContacts = true;
  return LSDEM_OK;
}

d3 rd3(const double* p, size_t i) { return mk3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
void wr3(double* p, size_t i, d3 v) { p[3 * i] = v.x; p[3 * i + 1] = v.y; p[3 * i + 2] = v.z; }
q4 rq4(const double* p, size_t i) { return q4{p[4 * i], p[4 * i + 1], p[4 * i + 2], p[4 * i + 3]}; }

s6 inverse6(s6 A) {
  const double det = A.xx * (A.yy * A.zz - A.yz * A.yz) + A.xy * (A.yz * A.xz - A.xy * A.zz) +
                     A.xz * (A.xy * A.yz - A.yy * A.xz);
  if (std::fabs(det) < 1e-30) return s6{0, 0, 0, 0, 0, 0};
  const double di = 1.0 / det;
  return s6{(A.yy * A.zz - A.yz * A.yz) * di,  (A.xx * A.zz - A.xz * A.xz) * di,
            (A.xx * A.yy - A.xy * A.xy) * di,  -(A.xy * A.zz - A.xz * A.yz) * di,
            (A.xy * A.yz - A.xz * A.yy) * di,  -(A.xx * A.yz - A.xz * A.xy) * di};
}

// Pair-sized arrays (candidate pairs, certificates, map records, reaction lists).  Contents are
// not kept: a new capacity comes with a rebuild of the pair list.
int allocPairs(lsdem_ctx* c, long long pairCap) {
  HIPCHK(c->dPairA.alloc(pairCap));
  HIPCHK(c->dPairInfo.alloc(pairCap));
  HIPCHK(c->dPairCoff.alloc(pairCap));
  HIPCHK(c->dPairCcnt.alloc(pairCap));
  HIPCHK(c->dSurMask.alloc(pairCap));
  HIPCHK(c->dSurSlack.alloc(pairCap));
  HIPCHK(c->dCert.alloc(pairCap));
  HIPCHK(c->dOldInfo.alloc(pairCap));
  HIPCHK(c->dOldCert.alloc(pairCap));
  HIPCHK(c->dCertE.alloc(pairCap));
  HIPCHK(c->dCertQA.alloc(pairCap));
  HIPCHK(c->dCertQB.alloc(pairCap));
  HIPCHK(c->dPairRec.alloc(pairCap));
  HIPCHK(c->dPairMap.alloc(pairCap * 10));
  HIPCHK(c->dRcPair.alloc(pairCap));
  HIPCHK(c->dRcTmp.alloc(pairCap));
  HIPCHK(c->dRcRank.alloc(pairCap));
  c->P.pairCap = pairCap;
  return LSDEM_OK;
}

// Contact-slot arrays and the two history tables (at least twice the slots, a power of two).
// Contents are not kept; the call snapshot restores the history of the step before the call.
int allocContacts(lsdem_ctx* c, long long contactCap) {
  hipStream_t s = c->stream;
  HIPCHK(c->dCPair.alloc(contactCap));
  HIPCHK(c->dCF.alloc(contactCap));
  HIPCHK(c->dCTA.alloc(contactCap));
  HIPCHK(c->dCTB.alloc(contactCap));
  for (int b = 0; b < 2; ++b) {
    HIPCHK(c->dCNode[b].alloc(contactCap));
    HIPCHK(c->dCInfo[b].alloc(contactCap));
    HIPCHK(c->dCSpring[b].alloc(contactCap));
    HIPCHK(c->dCPoint[b].alloc(contactCap));
    HIPCHK(c->dCNormal[b].alloc(contactCap));
    HIPCHK(c->dCOverlap[b].alloc(contactCap));
    HIPCHK(c->dCEn[b].alloc(contactCap));
    HIPCHK(c->dCEs[b].alloc(contactCap));
    HIPCHK(hipMemsetAsync(c->dCSpring[b].p, 0, contactCap * sizeof(d3), s));
  }
  long long hs = 1024;
  while (hs < 2 * contactCap) hs <<= 1;
  c->hashSize = hs;
  for (int b = 0; b < 2; ++b) {
    HIPCHK(c->dHashKey[b].alloc(hs));
    HIPCHK(c->dHashVal[b].alloc(hs));
    HIPCHK(c->dHashSlot[b].alloc(contactCap));
    hipLaunchKernelGGL(k_fill_u64, dim3(unsigned(c->numCU * 4)), dim3(256), 0, s, c->dHashKey[b].p, hs, kHashEmpty);
  }
  c->P.hashMask = (unsigned long long)(hs - 1);
  c->P.contactCap = contactCap;
  return LSDEM_OK;
}

// Snapshot buffers: the bodies, walls and bonds, and room for one step's contact slots.
int snapAlloc(lsdem_ctx* c) {
  auto& S = c->snap;
  const int nP = std::max(c->P.nP, 1), nW = std::max(c->P.nW, 1), nB = std::max(c->P.nBonds, 1);
  if (S.pos.n != size_t(nP)) {
    HIPCHK(S.pos.alloc(nP)); HIPCHK(S.vel.alloc(nP)); HIPCHK(S.ang.alloc(nP)); HIPCHK(S.force.alloc(nP));
    HIPCHK(S.torque.alloc(nP)); HIPCHK(S.quat.alloc(nP)); HIPCHK(S.forceSet.alloc(nP));
    HIPCHK(S.wpos.alloc(nW)); HIPCHK(S.wquat.alloc(nW)); HIPCHK(S.bActive.alloc(nB));
    HIPCHK(S.meta.alloc(4)); HIPCHK(S.devStat.alloc(kDevStats));
  }
  if (S.hCap != c->P.contactCap) {
    HIPCHK(S.hNode.alloc(c->P.contactCap)); HIPCHK(S.hInfo.alloc(c->P.contactCap));
    HIPCHK(S.hSpring.alloc(c->P.contactCap));
    S.hCap = c->P.contactCap;
  }
  return LSDEM_OK;
}

SnapDev snapDev(lsdem_ctx* c, int cur) {
  auto& S = c->snap;
  const int o = 1 - cur;
  return SnapDev{S.pos.p, S.vel.p, S.ang.p, S.force.p, S.torque.p, S.wpos.p, S.quat.p, S.wquat.p, S.bActive.p,
                 S.forceSet.p, S.devStat.p, S.meta.p, S.hNode.p, S.hInfo.p, S.hSpring.p, S.hCap,
                 c->dCNode[o].p, c->dCInfo[o].p, c->dCSpring[o].p, o};
}

unsigned snapBlocks(lsdem_ctx* c) { return unsigned(c->numCU * 4); }

// The state at the start of a call (enqueued on the context's stream).
int snapSave(lsdem_ctx* c) {
  auto& S = c->snap;
  hipLaunchKernelGGL(k_snap_save, dim3(snapBlocks(c)), dim3(256), 0, c->stream, c->P, snapDev(c, c->cur));
  HIPCHK(hipGetLastError());
  S.stepCount = c->stepCount; S.time = c->time; S.cur = c->cur; S.audits = c->sur.audits;
  S.profSteps = c->profSteps; S.profSeen = c->profSeen;
  S.primePending = c->primePending; S.forceRebuild = c->forceRebuild; S.haveContacts = c->haveContacts;
  return LSDEM_OK;
}

// Back to the state of the last snapSave (buffers may have been reallocated since); the next step
// rebuilds the candidate pairs.
int snapRestore(lsdem_ctx* c) {
  auto& S = c->snap;
  c->stepCount = S.stepCount; c->time = S.time; c->cur = S.cur; c->sur.audits = S.audits;
  c->profSteps = S.profSteps; c->profSeen = S.profSeen;
  c->primePending = S.primePending; c->haveContacts = S.haveContacts;
  c->forceRebuild = true;
  setCurrentBuffers(c);
  hipStream_t s = c->stream;
  for (int b = 0; b < 2; ++b)
    hipLaunchKernelGGL(k_fill_u64, dim3(unsigned(c->numCU * 4)), dim3(256), 0, s, c->dHashKey[b].p,
                       (long long)c->hashSize, kHashEmpty);
  const SnapDev D = snapDev(c, c->cur);
  hipLaunchKernelGGL(k_snap_restore, dim3(snapBlocks(c)), dim3(256), 0, s, c->P, D);
  const int o = 1 - c->cur;
  hipLaunchKernelGGL(k_hist_reinsert, dim3(snapBlocks(c)), dim3(256), 0, s, c->P, D, c->dHashKey[o].p,
                     c->dHashVal[o].p, c->dHashSlot[o].p);
  HIPCHK(hipGetLastError());
  return LSDEM_OK;
}

// Rotation by `angle` about the z axis as the reference builds it from the rotation vector
// (0, 0, angle): the Rodrigues coefficients of rotateVector (myVec.h:143-184) and the unit
// quaternion of rotateQuaternion (myQua.h:131-175), host branches, so that the device applies the
// same numbers with the same expressions.
void sectorRotation(double angle, d3& w, double& A, double& B, q4& r) {
  w = mk3(0.0, 0.0, angle);
  const double theta2 = dot(w, w);
  const double invTheta = 1.0 / std::sqrt(theta2);
  const double theta = theta2 * invTheta;
  A = std::sin(theta) * invTheta;
  B = (1.0 - std::cos(theta)) * (invTheta * invTheta);
  const double t = std::sqrt(w.x * w.x + w.y * w.y + w.z * w.z);
  const double it = 1.0 / t;
  const double c = std::cos(0.5 * t), sn = std::sin(0.5 * t);
  r = q4{c, sn * (w.x * it), sn * (w.y * it), sn * (w.z * it)};
}

// Device pointers of every context array into the Params table.
void bindPointers(lsdem_ctx* c) {
  Params& P = c->P;
  P.shapes = c->dShapes.p; P.clusters = c->dClusters.p; P.nodes = c->dNodes.p; P.sdf = c->dSdf.p;
  P.nodes32 = c->dNodes32.p; P.clusters32 = c->dClusters32.p; P.subs32 = c->dSubs32.p; P.nodeCluster = c->dNodeCluster.p;
  P.cPair = c->dCPair.p; P.cF = c->dCF.p; P.cTA = c->dCTA.p; P.cTB = c->dCTB.p;
  P.pos = c->dPos.p; P.posDet = c->dPosDet.p; P.quat = c->dQuat.p; P.vel = c->dVel.p; P.angvel = c->dAng.p;
  P.force = c->dForce.p; P.torque = c->dTorque.p; P.invI = c->dInvI.p; P.invM = c->dInvM.p;
  P.radius = c->dRadius.p; P.kn = c->dKn.p; P.ks = c->dKs.p; P.mu = c->dMu.p; P.res = c->dRes.p;
  P.scale = c->dScale.p; P.invSp = c->dInvSp.p; P.gorig = c->dGorig.p; P.shape = c->dShape.p;
  P.nodeStart = c->dNodeStart.p; P.scell = c->dScell.p;
  P.xBuild = c->dXBuild.p; P.xBuildW = c->dXBuildW.p; P.wBuildPos = c->dWBuildPos.p; P.wBuildQ = c->dWBuildQ.p;
  P.qBuild = c->dQBuild.p; P.imgBuild = c->dImgBuild.p; P.step = c->dStep.p;
  P.wMotion = c->dWMotion.p; P.wSph = c->dWSph.p; P.cert = c->dCert.p; P.oldInfo = c->dOldInfo.p; P.oldCert = c->dOldCert.p; P.oldStart = c->dOldStart.p; P.pairRec = c->dPairRec.p;
  P.qrel = c->dQrel.p; P.wQrel = c->dWQrel.p; P.certE = c->dCertE.p; P.certQA = c->dCertQA.p; P.certQB = c->dCertQB.p;
  P.surSlack = c->dSurSlack.p;
  P.bodyHit = c->dBodyHit.p;
  P.forceSet = c->dForceSet.p;
  P.wpos = c->dWPos.p; P.wquat = c->dWQuat.p; P.wvel = c->dWVel.p; P.wangvel = c->dWAng.p;
  P.wmu = c->dWMu.p; P.wres = c->dWRes.p; P.wshape = c->dWShape.p; P.wkey = c->dWKey.p;
  for (int g = 0; g < 3; ++g) {
    P.gpos[g] = c->dGPos[g].p;
    P.gquat[g] = c->dGQuat[g].p;
    P.gvel[g] = c->dGVel[g].p;
  }
  P.quatDet = c->dQuatDet.p;
  P.cellCount = c->dCellCount.p; P.cellStart = c->dCellStart.p; P.sorted = c->dSorted.p; P.rank = c->dRank.p;
  P.sPR = c->dSPR.p; P.sPRw = c->dSPRw.p; P.pairSlab = c->dPairSlab.p;
  P.sBoxR = c->dSBoxR.p;
  P.statSlots = c->dStatSlots.p;
  P.pairCount = c->dPairCount.p; P.pairStart = c->dPairStart.p; P.pairA = c->dPairA.p;
  P.pairInfo = c->dPairInfo.p; P.pairCoff = c->dPairCoff.p; P.pairCcnt = c->dPairCcnt.p; P.pairMap = c->dPairMap.p; P.surMask = c->dSurMask.p;
  P.rcCount = c->dRcCount.p; P.rcStart = c->dRcStart.p; P.rcPair = c->dRcPair.p; P.rcRank = c->dRcRank.p;
  P.rcTmp = c->dRcTmp.p; P.sortTmp = c->dSortTmp.p; P.sortBig = c->dSortBig.p; P.scanTiles = c->dScanBlocks.p;
  P.bI = c->dBI.p; P.bJ = c->dBJ.p; P.bActive = c->dBActive.p; P.bConst = c->dBConst.p; P.bLocal = c->dBLocal.p;
  P.bPoint = c->dBPoint.p; P.bOut = c->dBOut.p; P.bF = c->dBF.p; P.bMi = c->dBMi.p; P.bMj = c->dBMj.p;
  P.bcStart = c->dBcStart.p; P.bcItem = c->dBcItem.p;
  P.barrier = c->dBarrier.p; P.evalCount = c->dBarrier.p + 10 * kBarrierLine;
  P.devInt = c->dDevInt.p; P.devStat = c->dDevStat.p;
  setCurrentBuffers(c);
}

}  // namespace

extern "C" {

void lsdem_config_default(lsdem_config* cfg) {
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->domain_max[0] = cfg->domain_max[1] = cfg->domain_max[2] = 1.0;
  cfg->dt = 1e-5;
  cfg->quirks = LSDEM_QUIRK_ALL;
  cfg->block_size = 256;
  cfg->skin = -1.0;
}

const char* lsdem_last_error(void) { return g_err.c_str(); }

int lsdem_create(const lsdem_config* cfg, lsdem_ctx** out) {
  if (!cfg || !out) return fail(LSDEM_ERR_ARG, "null argument");
  *out = nullptr;
  const d3 ext = mk3(cfg->domain_max[0] - cfg->domain_min[0], cfg->domain_max[1] - cfg->domain_min[1],
                     cfg->domain_max[2] - cfg->domain_min[2]);
  if (!(ext.x > 0.0 && ext.y > 0.0 && ext.z > 0.0)) return fail(LSDEM_ERR_ARG, "invalid simulation domain size");
  if (!(cfg->dt > 0.0)) return fail(LSDEM_ERR_ARG, "invalid timeStep");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(LSDEM_ERR_HIP, "no HIP device available");
  if (cfg->device < 0 || cfg->device >= ndev) return fail(LSDEM_ERR_ARG, "device ordinal out of range");
  if (cfg->periodic_sector && (cfg->periodic_x || cfg->periodic_y))
    return fail(LSDEM_ERR_ARG, "sector periodicity cannot be combined with x/y periodicity");
  HIPCHK(hipSetDevice(cfg->device));
  lsdem_ctx* c = new lsdem_ctx();
  c->cfg = *cfg;
  // a blocking stream: work an external-force hook puts on the null stream stays ordered with the step
  if (hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return fail(LSDEM_ERR_HIP, "hipStreamCreate failed");
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, cfg->device) == hipSuccess && prop.multiProcessorCount > 0)
    c->numCU = prop.multiProcessorCount;
  *out = c;
  return LSDEM_OK;
}

void lsdem_destroy(lsdem_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->cfg.device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  delete ctx;
}

int lsdem_add_shape(lsdem_ctx* c, const float* sdf, int32_t nx, int32_t ny, int32_t nz, const double origin[3],
                    double spacing, const double* nodes, int32_t n_nodes, int32_t* shape_id) {
  if (!c || !sdf || !origin) return fail(LSDEM_ERR_ARG, "null argument");
  if (c->initialized) return fail(LSDEM_ERR_STATE, "shapes must be added before lsdem_initialize");
  if (nx < 2 || ny < 2 || nz < 2) return fail(LSDEM_ERR_ARG, "invalid level-set grid size (< 2)");
  if (!(spacing > 0.0)) return fail(LSDEM_ERR_ARG, "invalid level-set grid spacing");
  if (n_nodes < 0 || (n_nodes > 0 && !nodes)) return fail(LSDEM_ERR_ARG, "invalid boundary nodes");
  HostShape S;
  S.nx = nx; S.ny = ny; S.nz = nz;
  S.origin = mk3(origin[0], origin[1], origin[2]);
  S.h = spacing;
  S.sdf.assign(sdf, sdf + size_t(nx) * ny * nz);
  for (int i = 0; i < n_nodes; ++i) S.nodes.push_back(rd3(nodes, i));
  buildClusters(S);
  shapeBounds(S);
  c->shapes.push_back(std::move(S));
  if (shape_id) *shape_id = int(c->shapes.size()) - 1;
  return LSDEM_OK;
}

int lsdem_add_particles(lsdem_ctx* c, int32_t n, const int32_t* shape, const double* scale, const double* pos,
                        const double* quat, const double* vel, const double* angvel, const double* inv_mass,
                        const double* inv_inertia, const double* kn, const double* ks, const double* mu,
                        const double* restitution) {
  if (!c) return fail(LSDEM_ERR_ARG, "null context");
  if (c->initialized) return fail(LSDEM_ERR_STATE, "particles must be added before lsdem_initialize");
  if (n < 0 || (n > 0 && (!shape || !pos || !quat || !vel || !angvel || !inv_mass || !inv_inertia || !kn || !ks ||
                          !mu || !restitution)))
    return fail(LSDEM_ERR_ARG, "null particle array");
  for (int i = 0; i < n; ++i)
    if (shape[i] < 0 || shape[i] >= int(c->shapes.size())) return fail(LSDEM_ERR_ARG, "bad shape id");
  for (int i = 0; i < n; ++i) {
    const double s = scale ? scale[i] : 1.0;
    if (!(s > 0.0)) return fail(LSDEM_ERR_ARG, "scale must be > 0");
    const HostShape& S = c->shapes[shape[i]];
    double r = 0.;
    for (const d3& nd : S.nodes) r = std::max(len(nd * s), r);
    c->pShape.push_back(shape[i]);
    c->pScale.push_back(s);
    c->pRadius.push_back(r);
    c->pPos.push_back(rd3(pos, i));
    c->pVel.push_back(rd3(vel, i));
    c->pAng.push_back(rd3(angvel, i));
    c->pQuat.push_back(qnormalize(rq4(quat, i)));
    c->pInvM.push_back(inv_mass[i]);
    const double* I = inv_inertia + 6 * i;
    c->pInvI.push_back(s6{I[0], I[1], I[2], I[3], I[4], I[5]});
    c->pKn.push_back(kn[i] > 0. ? kn[i] : 0.);
    c->pKs.push_back(ks[i] > 0. ? ks[i] : 0.);
    c->pMu.push_back(mu[i] > 0. ? mu[i] : 0.);
    double e = restitution[i];
    if (e <= 0.0 || e > 1.0) e = 1.0;
    c->pRes.push_back(e);
  }
  return LSDEM_OK;
}

int lsdem_add_walls(lsdem_ctx* c, int32_t n, const int32_t* shape, const double* bound_radius, const double* pos,
                    const double* quat, const double* vel, const double* angvel, const double* mu,
                    const double* restitution) {
  if (!c) return fail(LSDEM_ERR_ARG, "null context");
  if (c->initialized) return fail(LSDEM_ERR_STATE, "walls must be added before lsdem_initialize");
  if (n < 0 || (n > 0 && (!shape || !bound_radius || !pos || !quat || !vel || !angvel || !mu || !restitution)))
    return fail(LSDEM_ERR_ARG, "null wall array");
  for (int i = 0; i < n; ++i)
    if (shape[i] < 0 || shape[i] >= int(c->shapes.size())) return fail(LSDEM_ERR_ARG, "bad shape id");
  for (int i = 0; i < n; ++i) {
    c->wShape.push_back(shape[i]);
    c->wRadius.push_back(bound_radius[i]);
    c->wPos.push_back(rd3(pos, i));
    c->wQuat.push_back(qnormalize(rq4(quat, i)));
    c->wVel.push_back(rd3(vel, i));
    c->wAng.push_back(rd3(angvel, i));
    c->wMu.push_back(mu[i] > 0. ? mu[i] : 0.);
    double e = restitution[i];
    if (e <= 0.0 || e > 1.0) e = 1.0;
    c->wRes.push_back(e);
  }
  return LSDEM_OK;
}

int lsdem_add_bond(lsdem_ctx* c, int32_t i, int32_t j, const double point[3], const double normal[3], double r,
                   double L, double E, double nu, double sigmaT, double coh, double bmu) {
  if (!c || !point || !normal) return fail(LSDEM_ERR_ARG, "null argument");
  if (c->initialized) return fail(LSDEM_ERR_STATE, "bonds must be added before lsdem_initialize");
  const int np = int(c->pPos.size());
  if (i < 0 || j < 0 || i >= np || j >= np) return fail(LSDEM_ERR_ARG, "bond particle id out of range");
  const d3 nrm = mk3(normal[0], normal[1], normal[2]);
  if (isZero(len2(nrm))) return fail(LSDEM_ERR_ARG, "invalid bond normal");
  if (r <= 0.) return fail(LSDEM_ERR_ARG, "invalid bond radius");
  if (L <= 0.) return fail(LSDEM_ERR_ARG, "invalid bond length");
  if (E <= 0.) return fail(LSDEM_ERR_ARG, "invalid Young's modulus");
  if (nu <= -1.0 || nu >= 0.5) return fail(LSDEM_ERR_ARG, "invalid Poisson ratio");
  // beam constants of the energy-based bond
  const double G = E / (2. * (1. + nu));
  const double A = kPi * r * r;
  const double I = A * r * r / 4.;
  const double J = 2. * I;
  const double Kn = E * A / L, Kt = G * J / L, Ks = 12. * E * I / (L * L * L), Kb = E * I / L;
  const double B1 = Kn, B2 = L * L * Ks, B4 = Kt, B3 = Kb - B2 / 4. - B4 / 2.;
  const d3 n = nrm / len(nrm);
  const d3 pt = mk3(point[0], point[1], point[2]);
  const d3 gpi = pt + 0.5 * L * n, gpj = pt - 0.5 * L * n;
  const q4 qi = c->pQuat[i], qj = c->pQuat[j];
  const d3 N1i = -n;
  const d3 ref = (std::fabs(N1i.x) < 0.9) ? mk3(1., 0., 0.) : mk3(0., 1., 0.);
  const d3 N2i = normalize(cross(ref, N1i));
  const d3 N3i = normalize(cross(N1i, N2i));
  const d3 N1j = -N1i, N2j = N2i, N3j = N3i;
  c->bI.push_back(i);
  c->bJ.push_back(j);
  c->bActive.push_back(1);
  const double cst[9] = {B1, B2, B3, B4, r, L, sigmaT, coh, bmu};
  c->bConst.insert(c->bConst.end(), cst, cst + 9);
  // local frames stored as rotated, without renormalising (VBondedInteraction.h:161-166)
  c->bLocal.push_back(rotInv(N1i, qi));
  c->bLocal.push_back(rotInv(N2i, qi));
  c->bLocal.push_back(rotInv(N3i, qi));
  c->bLocal.push_back(rotInv(gpi - c->pPos[i], qi));
  c->bLocal.push_back(rotInv(N1j, qj));
  c->bLocal.push_back(rotInv(N2j, qj));
  c->bLocal.push_back(rotInv(N3j, qj));
  c->bLocal.push_back(rotInv(gpj - c->pPos[j], qj));
  c->bPoint.push_back(pt);
  return LSDEM_OK;
}

int lsdem_initialize(lsdem_ctx* c) {
  if (!c) return fail(LSDEM_ERR_ARG, "null context");
  if (c->initialized) return fail(LSDEM_ERR_STATE, "already initialized");
  HIPCHK(hipSetDevice(c->cfg.device));
  {  // device address of the allocation-free entry point's scratch on this device
    unsigned char* sp = nullptr;
    if (int rc = scratchOn(c->cfg.device, &sp)) return rc;
  }
  hipStream_t s = c->stream;
  const int nP = int(c->pPos.size()), nW = int(c->wPos.size()), nS = int(c->shapes.size());
  const int nB = int(c->bI.size());
  Params& P = c->P;
  P = Params{};
  P.nP = nP; P.nW = nW; P.nShapes = nS; P.nBonds = nB;
  P.g = mk3(c->cfg.gravity[0], c->cfg.gravity[1], c->cfg.gravity[2]);
  P.dt = c->cfg.dt;
  P.quirks = c->cfg.quirks;
  P.sdfCull = c->cfg.no_sdf_cull ? 0 : 1;
  P.useCert = c->cfg.no_certificates ? 0 : 1;
  {
    const int nb = c->cfg.narrow_batch;
    if (nb != 0 && nb != 1 && nb != 2 && nb != 4 && nb != 8) return fail(LSDEM_ERR_ARG, "narrow_batch must be 0, 1, 2, 4 or 8");
    P.narrowBatch = nb;
  }
  P.periodicX = c->cfg.periodic_x != 0;
  P.periodicY = c->cfg.periodic_y != 0;
  P.periodicSector = c->cfg.periodic_sector != 0;
  for (int g = 0; g < 3; ++g) sectorRotation(g == 0 ? 0.5 * kPi : (g == 1 ? kPi : -0.5 * kPi), P.secW[g], P.secA[g],
                                             P.secB[g], P.secR[g]);
  const d3 dmn = mk3(c->cfg.domain_min[0], c->cfg.domain_min[1], c->cfg.domain_min[2]);
  const d3 dmx = mk3(c->cfg.domain_max[0], c->cfg.domain_max[1], c->cfg.domain_max[2]);

  // ---- shapes ----
  std::vector<ShapeDev> sd(std::max(nS, 1));
  std::vector<ClusterDev> cls;
  std::vector<NodeDev> nds;
  std::vector<float> sdf;
  for (int k = 0; k < nS; ++k) {
    const HostShape& S = c->shapes[k];
    ShapeDev& d = sd[k];
    d = ShapeDev{};
    d.nx = S.nx; d.ny = S.ny; d.nz = S.nz; d.nNodes = int(S.nodes.size());
    d.nodeOff = int(nds.size()); d.nClusters = int(S.clusters.size()); d.clusterOff = int(cls.size());
    d.sdfOff = (long long)sdf.size();
    d.ox = S.origin.x; d.oy = S.origin.y; d.oz = S.origin.z; d.h = S.h;
    d.boxR = S.boxR; d.lip = S.lip;
    for (const d3& n : S.nodes) d.nodeR = std::max(d.nodeR, len(n));
    d.nodeR *= 1.0 + 1e-12;
    if (d.nClusters > 64) return fail(LSDEM_ERR_ARG, "shape has more than 4096 boundary nodes");
    nds.insert(nds.end(), S.perm.begin(), S.perm.end());
    cls.insert(cls.end(), S.clusters.begin(), S.clusters.end());
    sdf.insert(sdf.end(), S.sdf.begin(), S.sdf.end());
  }
  int rc;
  if ((rc = upload(c->dShapes, sd, s))) return rc;
  if ((rc = upload(c->dClusters, cls, s))) return rc;
  if ((rc = upload(c->dNodes, nds, s))) return rc;
  {
    std::vector<float4> n32(nds.size());
    for (size_t k = 0; k < nds.size(); ++k) n32[k] = make_float4(float(nds[k].x), float(nds[k].y), float(nds[k].z), 0.f);
    if ((rc = upload(c->dNodes32, n32, s))) return rc;
    // FP32 copies; sub-cluster ranges of a cluster become global indices into the sub-cluster array
    std::vector<ClusterF32> c32(cls.size()), s32;
    std::vector<int> subBase(nS + 1, 0);
    for (int k = 0; k < nS; ++k) subBase[k + 1] = subBase[k] + int(c->shapes[k].subs.size());
    for (int k = 0; k < nS; ++k) {
      const HostShape& Sh = c->shapes[k];
      for (size_t u = 0; u < Sh.subs.size(); ++u) {
        const ClusterDev& d = Sh.subs[u];
        s32.push_back(ClusterF32{float(d.cx), float(d.cy), float(d.cz), float(d.rho * (1.0 + 1e-6)), d.first, d.count,
                                 0, 0});
      }
      for (size_t cc = 0; cc < Sh.clusters.size(); ++cc) {
        const ClusterDev& d = Sh.clusters[cc];
        c32[sd[k].clusterOff + cc] = ClusterF32{float(d.cx), float(d.cy), float(d.cz), float(d.rho * (1.0 + 1e-6)),
                                                d.first, d.count, subBase[k] + d.subFirst, d.nSub};
      }
    }
    if ((rc = upload(c->dClusters32, c32, s))) return rc;
    if ((rc = upload(c->dSubs32, s32, s))) return rc;
  }
  int maxClusters = 1;
  for (int k = 0; k < nS; ++k) maxClusters = std::max(maxClusters, sd[k].nClusters);
  {
    // cluster of every node, indexed by the shape's node offset + the caller's node index
    std::vector<int> nc(std::max<size_t>(nds.size(), 1), 0);
    for (int k = 0; k < nS; ++k) {
      const HostShape& S = c->shapes[k];
      for (size_t cc = 0; cc < S.clusters.size(); ++cc)
        for (int j = S.clusters[cc].first; j < S.clusters[cc].first + S.clusters[cc].count; ++j)
          nc[sd[k].nodeOff + S.perm[j].orig] = int(cc);
    }
    if ((rc = upload(c->dNodeCluster, nc, s))) return rc;
  }
  if ((rc = upload(c->dSdf, sdf, s))) return rc;

  // ---- particles ----
  std::vector<double> invSp(nP);
  std::vector<d3> gorig(nP), zero3(nP, mk3(0, 0, 0));
  std::vector<int> nodeStart(nP + 1, 0);
  double rMax = 0, boxRMax = 0;
  for (int i = 0; i < nP; ++i) {
    const HostShape& S = c->shapes[c->pShape[i]];
    invSp[i] = 1. / (S.h * c->pScale[i]);
    gorig[i] = S.origin * c->pScale[i];
    nodeStart[i + 1] = nodeStart[i] + int(S.nodes.size());
    rMax = std::max(rMax, c->pRadius[i]);
    boxRMax = std::max(boxRMax, c->pScale[i] * S.boxR);
  }
  if ((long long)nodeStart[nP] >= (1ll << 31) - 1) return fail(LSDEM_ERR_ARG, "too many boundary nodes");
  if ((rc = upload(c->dPos, c->pPos, s))) return rc;
  if ((rc = upload(c->dPosDet, c->pPos, s))) return rc;
  if ((rc = upload(c->dVel, c->pVel, s))) return rc;
  if ((rc = upload(c->dAng, c->pAng, s))) return rc;
  if ((rc = upload(c->dForce, zero3, s))) return rc;
  if ((rc = upload(c->dTorque, zero3, s))) return rc;
  if ((rc = upload(c->dQuat, c->pQuat, s))) return rc;
  if ((rc = upload(c->dInvI, c->pInvI, s))) return rc;
  if ((rc = upload(c->dInvM, c->pInvM, s))) return rc;
  if ((rc = upload(c->dRadius, c->pRadius, s))) return rc;
  if ((rc = upload(c->dKn, c->pKn, s))) return rc;
  if ((rc = upload(c->dKs, c->pKs, s))) return rc;
  if ((rc = upload(c->dMu, c->pMu, s))) return rc;
  if ((rc = upload(c->dRes, c->pRes, s))) return rc;
  if ((rc = upload(c->dScale, c->pScale, s))) return rc;
  if ((rc = upload(c->dInvSp, invSp, s))) return rc;
  if ((rc = upload(c->dGorig, gorig, s))) return rc;
  if ((rc = upload(c->dShape, c->pShape, s))) return rc;
  if ((rc = upload(c->dNodeStart, nodeStart, s))) return rc;
  HIPCHK(c->dXBuild.alloc(std::max(nP, 1)));
  HIPCHK(c->dXBuildW.alloc(std::max(nP, 1)));
  HIPCHK(c->dQBuild.alloc(std::max(nP, 1)));
  HIPCHK(c->dImgBuild.alloc(std::max(nP, 1)));
  HIPCHK(c->dStep.alloc(std::max(nP, 1)));
  HIPCHK(c->dQrel.alloc(std::max(nP, 1)));
  HIPCHK(c->dScell.alloc(std::max(nP, 1)));

  // ---- walls ----
  if ((rc = upload(c->dWPos, c->wPos, s))) return rc;
  if ((rc = upload(c->dWVel, c->wVel, s))) return rc;
  if ((rc = upload(c->dWAng, c->wAng, s))) return rc;
  if ((rc = upload(c->dWQuat, c->wQuat, s))) return rc;
  if ((rc = upload(c->dWMu, c->wMu, s))) return rc;
  if ((rc = upload(c->dWRes, c->wRes, s))) return rc;
  if ((rc = upload(c->dWShape, c->wShape, s))) return rc;
  HIPCHK(c->dWKey.alloc(std::max(nW, 1)));
  HIPCHK(c->dWBuildPos.alloc(std::max(nW, 1)));
  HIPCHK(c->dWBuildQ.alloc(std::max(nW, 1)));
  HIPCHK(c->dWMotion.alloc(std::max(nW, 1)));
  HIPCHK(c->dWSph.alloc(std::max(nW, 1)));
  HIPCHK(c->dWQrel.alloc(std::max(nW, 1)));

  // ---- grids ----
  double cellW = 0;
  for (double r : c->wRadius) cellW = std::max(cellW, r);
  P.gridP = makeGrid(dmn, dmx, nP ? rMax * 2.0 : 0.);
  P.gridW = makeGrid(dmn, dmx, nW ? cellW * 2.0 : 0.);
  P.rMax = rMax;
  P.boxRMax = boxRMax;
  // skin of the reused candidate pairs (rebuilt when a body has moved half of it)
  const double skin = c->cfg.skin >= 0.0 ? c->cfg.skin : 0.08 * rMax;
  P.halfSkin = 0.5 * skin;
  c->forceRebuild = true;
  {
    const d3 ext = dmx - dmn;
    double cell = rMax > 0 ? rMax + 0.5 * skin : std::max(ext.x, std::max(ext.y, ext.z));
    for (;;) {
      const double nc = std::ceil(ext.x / cell) * std::ceil(ext.y / cell) * std::ceil(ext.z / cell);
      if (nc <= 4.0e6) break;
      cell *= 1.25;
    }
    SearchGrid& S = P.sg;
    S.mn = dmn;
    S.cell = cell;
    S.invCell = 1.0 / cell;
    S.size = i3{std::max(1, int(std::ceil(ext.x / cell))), std::max(1, int(std::ceil(ext.y / cell))),
                std::max(1, int(std::ceil(ext.z / cell)))};
    S.ncell = S.size.x * S.size.y * S.size.z;
  }
  const int ncell = P.sg.ncell;
  HIPCHK(c->dCellCount.alloc(ncell));
  HIPCHK(c->dCellStart.alloc(ncell + 1));
  HIPCHK(c->dSorted.alloc(std::max(nP, 1)));
  HIPCHK(c->dSortTmp.alloc(std::max(nP, 1)));
  HIPCHK(c->dSortBig.alloc(std::max(ncell, nP) + 1));
  HIPCHK(c->dRank.alloc(std::max(nP, 1)));
  HIPCHK(c->dSPR.alloc(std::max(nP, 1)));
  HIPCHK(c->dSPRw.alloc(std::max(nP, 1)));
  HIPCHK(c->dSBoxR.alloc(std::max(nP, 1)));
  HIPCHK(c->dPairSlab.alloc(size_t(std::max(nP, 1)) * kPairSlabHost));
  const bool periodic = P.periodicX || P.periodicY || P.periodicSector;
  for (int g = 0; g < 3; ++g) {
    const bool on = (g == 0 && P.periodicX) || (g == 1 && P.periodicY) || (g == 2 && P.periodicX && P.periodicY) ||
                    P.periodicSector;
    const size_t n = on ? size_t(std::max(nP, 1)) : 1;
    HIPCHK(c->dGPos[g].alloc(n));
    HIPCHK(c->dGQuat[g].alloc(P.periodicSector ? n : 1));
    HIPCHK(c->dGVel[g].alloc(P.periodicSector ? n : 1));
  }
  HIPCHK(c->dQuatDet.alloc(P.periodicSector ? size_t(std::max(nP, 1)) : 1));

  // ---- pairs and contacts ----
  long long pairCap = c->cfg.pair_capacity > 0 ? c->cfg.pair_capacity : 64ll * nP + 4096 + (periodic ? 32ll * nP : 0);
  // narrow-phase grid: enough waves for ~16 candidate pairs per particle, at most 8 per SIMD
  c->narrowBlocks = int(std::min<long long>(c->numCU * 8ll, std::max<long long>(1, (16ll * nP + 8 * 4 - 1) / (8 * 4))));
  if (c->cfg.narrow_blocks > 0) c->narrowBlocks = int(c->cfg.narrow_blocks);
  // single-launch broad phase: a grid whose blocks can all be resident together
  c->coopBlocks = 0;
  if (!c->cfg.broad_phase_launches) {
    int perCU = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, reinterpret_cast<const void*>(k_broad_phase),
                                                        kBroadThreadsHost, 0));
    if (perCU < 1) return fail(LSDEM_ERR_HIP, "the single-launch broad phase does not fit a CU");
    // one block of 1024 threads per 512 particles or per 16,384 search cells (the cell loops of a
    // sparse domain, such as an ensemble of drums far apart), at least 8 and at most one per CU: a
    // small simulation leaves the rest of the GPU to others stepped concurrently, and its rebuilds
    // meet at barriers of few blocks
    const long long byCells = ((long long)P.sg.ncell + 16383) / 16384;
    c->coopBlocks = int(std::min<long long>(c->numCU, std::max<long long>({8, (nP + 511) / 512, byCells})));
    if (c->cfg.broad_phase_blocks > 0) c->coopBlocks = std::min(c->numCU, int(c->cfg.broad_phase_blocks));
    c->coopBlocks = std::min(c->coopBlocks, perCU * c->numCU);  // all resident together on an idle GPU
  }
  // the narrow phase's per-wavefront LDS grows with the largest node count of a shape
  {
    const size_t lds = candidatesLdsBytes(maxClusters);
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, c->cfg.device));
    if (lds > prop.sharedMemPerBlock) {
      const void* const kernels[4] = {reinterpret_cast<const void*>(k_candidates<false, false>),
                                      reinterpret_cast<const void*>(k_candidates<true, false>),
                                      reinterpret_cast<const void*>(k_candidates<false, true>),
                                      reinterpret_cast<const void*>(k_candidates<true, true>)};
      bool ok = true;
      for (const void* k : kernels)
        ok = ok && hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)) == hipSuccess;
      if (!ok)
        return fail(LSDEM_ERR_ARG, "shapes with " + std::to_string(64 * maxClusters) +
                                       " boundary nodes need " + std::to_string(lds) +
                                       " bytes of LDS per narrow-phase block, more than the device's " +
                                       std::to_string(prop.sharedMemPerBlock));
    }
  }
  // candidate nodes (slots of the exact test); reserved per batch at their exact count
  long long contactCap = c->cfg.contact_capacity > 0 ? c->cfg.contact_capacity : 64ll * nP + 16384;
  pairCap = std::min<long long>(pairCap, (1ll << 31) - 1);
  contactCap = std::min<long long>(contactCap, (1ll << 30));
  HIPCHK(c->dPairCount.alloc(std::max(nP, 1)));
  HIPCHK(c->dPairStart.alloc(nP + 1));
  HIPCHK(c->dBodyHit.alloc(std::max(nP, 1)));
  HIPCHK(c->dForceSet.alloc(std::max(nP, 1)));
  HIPCHK(c->dOldStart.alloc(nP + 1));
  HIPCHK(c->dRcCount.alloc(std::max(nP, 1)));
  HIPCHK(c->dRcStart.alloc(nP + 1));
  if ((rc = allocPairs(c, pairCap))) return rc;
  if ((rc = allocContacts(c, contactCap))) return rc;
  HIPCHK(c->dStatSlots.alloc(kStatWords));

  // ---- bonds (+ per-particle CSR of bond ends, in bond order) ----
  {
    std::vector<int> cnt(nP + 1, 0), item(2 * nB);
    for (int k = 0; k < nB; ++k) { cnt[c->bI[k] + 1]++; cnt[c->bJ[k] + 1]++; }
    for (int i = 0; i < nP; ++i) cnt[i + 1] += cnt[i];
    std::vector<int> fillp(cnt.begin(), cnt.end() - 1);
    for (int k = 0; k < nB; ++k) {
      item[fillp[c->bI[k]]++] = (k << 1);
      item[fillp[c->bJ[k]]++] = (k << 1) | 1;
    }
    if ((rc = upload(c->dBI, c->bI, s))) return rc;
    if ((rc = upload(c->dBJ, c->bJ, s))) return rc;
    if ((rc = upload(c->dBActive, c->bActive, s))) return rc;
    if ((rc = upload(c->dBConst, c->bConst, s))) return rc;
    if ((rc = upload(c->dBLocal, c->bLocal, s))) return rc;
    if ((rc = upload(c->dBPoint, c->bPoint, s))) return rc;
    if ((rc = upload(c->dBcStart, cnt, s))) return rc;
    if ((rc = upload(c->dBcItem, item, s))) return rc;
    HIPCHK(c->dBOut.alloc(std::max(6 * nB, 1)));
    HIPCHK(c->dBF.alloc(std::max(nB, 1)));
    HIPCHK(c->dBMi.alloc(std::max(nB, 1)));
    HIPCHK(c->dBMj.alloc(std::max(nB, 1)));
    HIPCHK(hipMemsetAsync(c->dBOut.p, 0, sizeof(double) * std::max(6 * nB, 1), s));
  }

  // ---- scratch ----
  long long maxScan = std::max<long long>({(long long)nP, (long long)ncell, 1ll});
  c->scanBlocksCap = int((maxScan + 2047) / 2048) + 1;
  HIPCHK(c->dScanBlocks.alloc(c->scanBlocksCap));
  HIPCHK(c->dDevInt.alloc(kDevInts));
  HIPCHK(c->dBarrier.alloc(kBarrierInts));
  HIPCHK(hipMemsetAsync(c->dBarrier.p, 0, kBarrierInts * sizeof(int), s));
  HIPCHK(c->dDevStat.alloc(kDevStats));
  HIPCHK(hipMemsetAsync(c->dStatSlots.p, 0, kStatWords * sizeof(unsigned long long), s));
  HIPCHK(hipMemsetAsync(c->dDevInt.p, 0, kDevInts * sizeof(int), s));
  HIPCHK(hipMemsetAsync(c->dDevStat.p, 0, kDevStats * sizeof(unsigned long long), s));
  HIPCHK(hipMemsetAsync(c->dPairStart.p, 0, (nP + 1) * sizeof(int), s));
  HIPCHK(hipMemsetAsync(c->dBodyHit.p, 0, std::max(nP, 1), s));
  HIPCHK(hipMemsetAsync(c->dForceSet.p, 1, std::max(nP, 1), s));

  P.surSlot = nullptr; P.surBand = 0.f; P.sur = SurNet{}; P.maxClusters = maxClusters;
  P.barrierSpin = c->cfg.barrier_spin > 0 ? c->cfg.barrier_spin : (c->cfg.barrier_spin < 0 ? 0 : (1 << 21));
  c->cur = 0;
  bindPointers(c);
  HIPCHK(hipMemsetAsync(c->dDevInt.p + kErrStepIdx, 0xFF, sizeof(int), s));
  if ((rc = snapAlloc(c))) return rc;
  HIPCHK(hipStreamSynchronize(s));
  HIPCHK(hipGetLastError());
  c->initialized = true;
  c->primePending = true;  // the first step is preceded by the pre-loop build, as in LSSolver::compute*
  c->haveContacts = false;
  return LSDEM_OK;
}

// Launches n steps on the context's stream (no wait).
int enqueueSteps(lsdem_ctx* c, int n) {
  if (n > 0 && c->primePending) {
    const int rc = runPrime(c);
    if (rc) return rc;
  }
  // consecutive steps share one launch at each boundary (reduce + next kick/drift) unless an
  // external-force hook runs between the second half-kick's sums and the kick
  const bool fuse = !c->hookFn && !c->cfg.no_fused_steps;
  for (int k = 0; k < n; ++k) {
    int rc = runStep(c, fuse && k > 0, fuse && k + 1 < n, k);
    if (rc) return rc;
  }
  if (n > 0) c->haveContacts = true;
  return LSDEM_OK;
}

// Waits for the stream and reports the device's error flags.
// Waits for the stream and reads the error flags of the call (kErr*) and the failing step.
int readFlags(lsdem_ctx* c, int& flags, int& step) {
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  int di[kErrStepIdx + 1];
  HIPCHK(hipMemcpy(di, c->dDevInt.p, sizeof di, hipMemcpyDeviceToHost));
  flags = di[3];
  step = di[kErrStepIdx];
  return LSDEM_OK;
}

int reportFlags(lsdem_ctx* c, int flags, int step) {
  int di[8];
  HIPCHK(hipMemcpy(di, c->dDevInt.p, sizeof di, hipMemcpyDeviceToHost));
  const std::string at = " (step " + std::to_string(step) + " of the call)";
  if (flags & kErrPairs)
    return fail(LSDEM_ERR_CAPACITY, "candidate-pair buffer overflow (needed " + std::to_string(di[4]) +
                                        " > capacity " + std::to_string(c->P.pairCap) + ")" + at);
  if (flags & kErrBarrier)
    return fail(LSDEM_ERR_STATE, "broad phase: a grid barrier timed out (blocks were not resident together)" + at);
  if (flags & kErrContacts)
    return fail(LSDEM_ERR_CAPACITY, "contact buffer overflow (needed " + std::to_string(di[5]) + " > capacity " +
                                        std::to_string(c->P.contactCap) + ")" + at);
  return fail(LSDEM_ERR_CAPACITY, "history table full" + at);
}

// Larger buffers (or the separate-launch broad phase) for what made the call halt.  Returns
// LSDEM_ERR_CAPACITY when a buffer is already at its largest size.
int growFor(lsdem_ctx* c, int flags) {
  int di[8];
  HIPCHK(hipMemcpy(di, c->dDevInt.p, sizeof di, hipMemcpyDeviceToHost));
  if (flags & kErrBarrier) c->coopBlocks = 0;  // separate gated launches need no co-residency
  int rc;
  if (flags & kErrPairs) {
    const long long maxCap = (1ll << 31) - 1;
    const long long cap = std::min(maxCap, std::max(2 * c->P.pairCap, (long long)di[4] + di[4] / 4 + 4096));
    if (cap <= c->P.pairCap) return reportFlags(c, flags, -1);
    if ((rc = allocPairs(c, cap))) return rc;
  }
  if (flags & (kErrContacts | kErrHash)) {
    const long long maxCap = 1ll << 30;
    const long long cap = std::min(maxCap, std::max(2 * c->P.contactCap, (long long)di[5] + di[5] / 4 + 16384));
    if (cap <= c->P.contactCap) return reportFlags(c, flags, -1);
    if ((rc = allocContacts(c, cap))) return rc;
  }
  bindPointers(c);
  c->dropGraph();
  c->recoveries++;
  return LSDEM_OK;
}

constexpr int kMaxRecoveries = 8;

// Runs `work` (which enqueues the call's launches) until it completes without an error flag: a call
// that halted is restored to its start, its buffers grown, and run again, so its results are those
// of a call that had the room from the start.  The external-force hook runs again for the repeated
// steps.  With cfg.no_recovery the error is returned at once and the state is undefined.
extern "C++" template <class Work>
int runCall(lsdem_ctx* c, Work&& work) {
  const bool recover = !c->cfg.no_recovery;
  int rc;
  if (recover && (rc = snapSave(c))) return rc;
  for (int attempt = 0;; ++attempt) {
    if ((rc = work())) return rc;
    int flags = 0, step = -1;
    if ((rc = readFlags(c, flags, step))) return rc;
    if (!flags) return LSDEM_OK;
    if (!recover || attempt >= kMaxRecoveries) {
      const int err = reportFlags(c, flags, step);
      // leave the context able to step again (from whatever state the halted call left)
      HIPCHK(hipMemset(c->dDevInt.p + 3, 0, sizeof(int)));
      HIPCHK(hipMemset(c->dDevInt.p + kErrStepIdx, 0xFF, sizeof(int)));
      HIPCHK(hipMemset(c->dBarrier.p, 0, kBarrierInts * sizeof(int)));
      c->forceRebuild = true;
      return err;
    }
    const long long hCap = c->snap.hCap;
    if ((rc = growFor(c, flags))) return rc;
    if ((rc = snapRestore(c))) return rc;
    if (c->P.contactCap != hCap) {  // room for the history of the larger buffers, filled again
      HIPCHK(hipStreamSynchronize(c->stream));  // (the restore still reads the old snapshot)
      if ((rc = snapAlloc(c))) return rc;
      if ((rc = snapSave(c))) return rc;
    }
  }
}

int lsdem_step(lsdem_ctx* c, int32_t n) {
  if (!c) return fail(LSDEM_ERR_ARG, "null context");
  if (!c->initialized) return fail(LSDEM_ERR_STATE, "lsdem_initialize has not been called");
  if (c->phaseNext) return fail(LSDEM_ERR_STATE, "a step driven by lsdem_phase_* calls is unfinished");
  if (n < 0) return fail(LSDEM_ERR_ARG, "negative step count");
  HIPCHK(hipSetDevice(c->cfg.device));
  if (c->P.nP == 0) { c->stepCount += n; return LSDEM_OK; }
  if (n == 0) return LSDEM_OK;
  return runCall(c, [&]() { return enqueueSteps(c, n); });
}

int lsdem_begin_phase(lsdem_ctx* c) {
  if (!c) return fail(LSDEM_ERR_ARG, "null context");
  if (!c->initialized) return fail(LSDEM_ERR_STATE, "lsdem_initialize has not been called");
  if (c->phaseNext) return fail(LSDEM_ERR_STATE, "a step driven by lsdem_phase_* calls is unfinished");
  HIPCHK(hipSetDevice(c->cfg.device));
  if (c->P.nP == 0) { c->primePending = false; return LSDEM_OK; }
  return runCall(c, [&]() { return runPrime(c); });
}

int lsdem_step_graph(lsdem_ctx* c, int32_t n) {
  if (!c) return fail(LSDEM_ERR_ARG, "null context");
  if (!c->initialized) return fail(LSDEM_ERR_STATE, "lsdem_initialize has not been called");
  if (c->phaseNext) return fail(LSDEM_ERR_STATE, "a step driven by lsdem_phase_* calls is unfinished");
  if (n < 2 || n % 2) return fail(LSDEM_ERR_ARG, "graph step count must be even and at least 2");
  if (c->hookFn || c->profiling || c->sur.enabled)
    return fail(LSDEM_ERR_STATE, "graph steps run without an external-force hook, profiling or the surrogate");
  HIPCHK(hipSetDevice(c->cfg.device));
  if (c->P.nP == 0) { c->stepCount += n; return LSDEM_OK; }
  // a pending rebuild request is a launch argument of the first step, and a pending pre-loop build
  // is launched before it: those steps run eagerly
  if (c->forceRebuild || c->primePending) return lsdem_step(c, n);
  if (c->graphN != n) {
    c->dropGraph();
    // capture the launches of n steps; the host's per-step bookkeeping is applied per replay
    // instead (n is even, so the contact buffers come back to the same parity)
    const int cur0 = c->cur;
    const long long step0 = c->stepCount;
    const double time0 = c->time;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueueSteps(c, n);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(c->stream, &g);
    c->cur = cur0;
    c->stepCount = step0;
    c->time = time0;
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (ec != hipSuccess) return fail(LSDEM_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ec));
    c->graph = g;
    HIPCHK(hipGraphInstantiate(&c->graphExec, c->graph, nullptr, nullptr, 0));
    c->graphN = n;
  }
  // a replay that halts on an error is restored and run again eagerly (with larger buffers)
  const bool recover = !c->cfg.no_recovery;
  int rc;
  if (recover && (rc = snapSave(c))) return rc;
  HIPCHK(hipGraphLaunch(c->graphExec, c->stream));
  c->stepCount += n;
  for (int k = 0; k < n; ++k) c->time += c->P.dt;  // as the eager steps add it
  c->haveContacts = true;
  int flags = 0, step = -1;
  if ((rc = readFlags(c, flags, step))) return rc;
  if (!flags) return LSDEM_OK;
  if (!recover) {
    const int err = reportFlags(c, flags, step);
    HIPCHK(hipMemset(c->dDevInt.p + 3, 0, sizeof(int)));
    HIPCHK(hipMemset(c->dDevInt.p + kErrStepIdx, 0xFF, sizeof(int)));
    HIPCHK(hipMemset(c->dBarrier.p, 0, kBarrierInts * sizeof(int)));
    c->forceRebuild = true;
    return err;
  }
  if ((rc = growFor(c, flags))) return rc;
  if ((rc = snapRestore(c))) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));  // (the restore still reads the old snapshot)
  if ((rc = snapAlloc(c))) return rc;
  return lsdem_step(c, n);
}

}  // extern "C"

namespace {

const char* const kPhaseNames[6] = {"integrate", "hash", "pairs", "narrow", "forces", "reduce"};

// Checks that phase k of a step driven phase by phase is the one due.
int phaseCheck(lsdem_ctx* c, int k) {
  if (!c) return fail(LSDEM_ERR_ARG, "null context");
  if (!c->initialized) return fail(LSDEM_ERR_STATE, "lsdem_initialize has not been called");
  if (c->phaseNext != k)
    return fail(LSDEM_ERR_STATE, std::string("lsdem_phase_") + kPhaseNames[k] + " called out of order (next: " +
                                     kPhaseNames[c->phaseNext] + ")");
  HIPCHK(hipSetDevice(c->cfg.device));
  return LSDEM_OK;
}

// Waits for the phase's launches and reports the device's error flags; an error abandons the step.
int phaseDone(lsdem_ctx* c, int next) {
  int flags = 0, step = -1, rc;
  if ((rc = readFlags(c, flags, step))) {
    c->phaseNext = 0;
    return rc;
  }
  if (flags) {
    const int err = reportFlags(c, flags, step);
    HIPCHK(hipMemset(c->dDevInt.p + 3, 0, sizeof(int)));
    HIPCHK(hipMemset(c->dDevInt.p + kErrStepIdx, 0xFF, sizeof(int)));
    HIPCHK(hipMemset(c->dBarrier.p, 0, kBarrierInts * sizeof(int)));
    c->forceRebuild = true;
    c->phaseNext = 0;
    return err;
  }
  c->phaseNext = next;
  return LSDEM_OK;
}

}  // namespace

extern "C" {

int lsdem_phase_integrate(lsdem_ctx* c) {
  int rc;
  if ((rc = phaseCheck(c, 0))) return rc;
  Params& P = c->P;
  if (P.nP == 0) { c->phaseNext = 1; return LSDEM_OK; }
  if (c->primePending) {
    if ((rc = runPrime(c))) return rc;
    if ((rc = phaseDone(c, 0))) return rc;
  }
  setCurrentBuffers(c);
  P.callStep = 0;
  c->forceRebuild = false;
  const int bs = c->cfg.block_size > 0 ? c->cfg.block_size : 256;
  // (a rebuild requested from the host: the hash and pairs phases run on every such step)
  launchKickDrift(c, 1);
  return phaseDone(c, 1);
}

int lsdem_phase_hash(lsdem_ctx* c) {
  int rc;
  if ((rc = phaseCheck(c, 1))) return rc;
  if (c->P.nP == 0) { c->phaseNext = 2; return LSDEM_OK; }
  if ((rc = hashLaunches(c, c->P.devInt + rebuildFlagIdx(c->P.curBuf)))) return rc;
  return phaseDone(c, 2);
}

int lsdem_phase_pairs(lsdem_ctx* c) {
  int rc;
  if ((rc = phaseCheck(c, 2))) return rc;
  if (c->P.nP == 0) { c->phaseNext = 3; return LSDEM_OK; }
  auto none = [](bool = true) {};
  if ((rc = pairLaunches(c, c->P.devInt + rebuildFlagIdx(c->P.curBuf), none))) return rc;
  return phaseDone(c, 3);
}

int lsdem_phase_narrow(lsdem_ctx* c) {
  int rc;
  if ((rc = phaseCheck(c, 3))) return rc;
  if (c->P.nP == 0) { c->phaseNext = 4; return LSDEM_OK; }
  const bool audit = auditStep(c);
  launchPairCheck(c, audit);
  launchCandidates(c, audit);
  return phaseDone(c, 4);
}

int lsdem_phase_forces(lsdem_ctx* c) {
  int rc;
  if ((rc = phaseCheck(c, 4))) return rc;
  const Params& P = c->P;
  if (P.nP == 0) { c->phaseNext = 5; return LSDEM_OK; }
  const int bs = c->cfg.block_size > 0 ? c->cfg.block_size : 256;
  if (P.nBonds > 0) hipLaunchKernelGGL(k_bonds, dim3(blocksFor(P.nBonds, bs)), dim3(bs), 0, c->stream, P);
  return phaseDone(c, 5);
}

int lsdem_phase_reduce(lsdem_ctx* c) {
  int rc;
  if ((rc = phaseCheck(c, 5))) return rc;
  if (c->P.nP == 0) {
    c->phaseNext = 0;
    c->stepCount++;
    return LSDEM_OK;
  }
  launchReduceKick(c);
  if ((rc = phaseDone(c, 0))) return rc;
  c->cur = 1 - c->cur;
  c->stepCount++;
  c->time += c->P.dt;
  c->haveContacts = true;
  return LSDEM_OK;
}

int lsdem_get_cell_table(lsdem_ctx* c, double origin[3], double* cell_edge, int32_t dims[3], int32_t* cell_of,
                         int32_t* sorted, int64_t start_capacity, int32_t* cell_start) {
  if (!c || !c->initialized) return fail(LSDEM_ERR_STATE, "not initialized");
  HIPCHK(hipSetDevice(c->cfg.device));
  const SearchGrid& S = c->P.sg;
  if (origin) { origin[0] = S.mn.x; origin[1] = S.mn.y; origin[2] = S.mn.z; }
  if (cell_edge) *cell_edge = S.cell;
  if (dims) { dims[0] = S.size.x; dims[1] = S.size.y; dims[2] = S.size.z; }
  if (cell_start && start_capacity < (int64_t)S.ncell + 1)
    return fail(LSDEM_ERR_CAPACITY, "cell_start needs " + std::to_string((long long)S.ncell + 1) + " entries");
  HIPCHK(hipStreamSynchronize(c->stream));
  const size_t n = size_t(c->P.nP);
  if (n && cell_of) HIPCHK(hipMemcpy(cell_of, c->P.scell, n * sizeof(int), hipMemcpyDeviceToHost));
  if (n && sorted) HIPCHK(hipMemcpy(sorted, c->P.sorted, n * sizeof(int), hipMemcpyDeviceToHost));
  if (cell_start)
    HIPCHK(hipMemcpy(cell_start, c->P.cellStart, (size_t(S.ncell) + 1) * sizeof(int), hipMemcpyDeviceToHost));
  return LSDEM_OK;
}

int lsdem_get_pairs(lsdem_ctx* c, int64_t capacity, int64_t* count, int32_t* master, int32_t* partner,
                    int32_t* type) {
  if (!c || !c->initialized) return fail(LSDEM_ERR_STATE, "not initialized");
  HIPCHK(hipSetDevice(c->cfg.device));
  HIPCHK(hipStreamSynchronize(c->stream));
  const Params& P = c->P;
  int total = 0;
  if (P.nP > 0) HIPCHK(hipMemcpy(&total, P.pairStart + P.nP, sizeof(int), hipMemcpyDeviceToHost));
  if (count) *count = total;
  const size_t m = size_t(std::max<int64_t>(0, std::min<int64_t>({capacity, (int64_t)total, (int64_t)P.pairCap})));
  if (!m) return LSDEM_OK;
  if (master) HIPCHK(hipMemcpy(master, P.pairA, m * sizeof(int), hipMemcpyDeviceToHost));
  if (partner || type) {
    std::vector<uint32_t> info(m);
    HIPCHK(hipMemcpy(info.data(), P.pairInfo, m * sizeof(uint32_t), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < m; ++k) {
      if (partner) partner[k] = int32_t(info[k] & 0x0fffffffu);
      if (type) type[k] = int32_t(info[k] >> 28);
    }
  }
  return LSDEM_OK;
}

int lsdem_get_particles(lsdem_ctx* c, double* pos, double* quat, double* vel, double* angvel, double* force,
                        double* torque) {
  if (!c || !c->initialized) return fail(LSDEM_ERR_STATE, "not initialized");
  HIPCHK(hipSetDevice(c->cfg.device));
  HIPCHK(hipStreamSynchronize(c->stream));
  const size_t n = size_t(c->P.nP);
  if (n == 0) return LSDEM_OK;
  if (pos) HIPCHK(hipMemcpy(pos, c->dPos.p, n * sizeof(d3), hipMemcpyDeviceToHost));
  if (quat) HIPCHK(hipMemcpy(quat, c->dQuat.p, n * sizeof(q4), hipMemcpyDeviceToHost));
  if (vel) HIPCHK(hipMemcpy(vel, c->dVel.p, n * sizeof(d3), hipMemcpyDeviceToHost));
  if (angvel) HIPCHK(hipMemcpy(angvel, c->dAng.p, n * sizeof(d3), hipMemcpyDeviceToHost));
  if (force) HIPCHK(hipMemcpy(force, c->dForce.p, n * sizeof(d3), hipMemcpyDeviceToHost));
  if (torque) HIPCHK(hipMemcpy(torque, c->dTorque.p, n * sizeof(d3), hipMemcpyDeviceToHost));
  return LSDEM_OK;
}

int lsdem_set_particles(lsdem_ctx* c, const double* pos, const double* quat, const double* vel, const double* angvel,
                        const double* force, const double* torque) {
  if (!c || !c->initialized) return fail(LSDEM_ERR_STATE, "not initialized");
  HIPCHK(hipSetDevice(c->cfg.device));
  HIPCHK(hipStreamSynchronize(c->stream));
  const size_t n = size_t(c->P.nP);
  if (n == 0) return LSDEM_OK;
  c->forceRebuild = true;  // poses changed outside the step: rebuild the candidate pairs
  if (pos) {
    HIPCHK(hipMemcpy(c->dPos.p, pos, n * sizeof(d3), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->dPosDet.p, pos, n * sizeof(d3), hipMemcpyHostToDevice));
  }
  if (quat) HIPCHK(hipMemcpy(c->dQuat.p, quat, n * sizeof(q4), hipMemcpyHostToDevice));
  if (vel) HIPCHK(hipMemcpy(c->dVel.p, vel, n * sizeof(d3), hipMemcpyHostToDevice));
  if (angvel) HIPCHK(hipMemcpy(c->dAng.p, angvel, n * sizeof(d3), hipMemcpyHostToDevice));
  if (force) HIPCHK(hipMemcpy(c->dForce.p, force, n * sizeof(d3), hipMemcpyHostToDevice));
  if (torque) HIPCHK(hipMemcpy(c->dTorque.p, torque, n * sizeof(d3), hipMemcpyHostToDevice));
  if (force || torque) HIPCHK(hipMemset(c->dForceSet.p, 1, n));  // stored sums no longer known to be zero
  return LSDEM_OK;
}

int lsdem_get_walls(lsdem_ctx* c, double* pos, double* quat, double* vel, double* angvel) {
  if (!c || !c->initialized) return fail(LSDEM_ERR_STATE, "not initialized");
  HIPCHK(hipSetDevice(c->cfg.device));
  HIPCHK(hipStreamSynchronize(c->stream));
  const size_t n = size_t(c->P.nW);
  if (n == 0) return LSDEM_OK;
  if (pos) HIPCHK(hipMemcpy(pos, c->dWPos.p, n * sizeof(d3), hipMemcpyDeviceToHost));
  if (quat) HIPCHK(hipMemcpy(quat, c->dWQuat.p, n * sizeof(q4), hipMemcpyDeviceToHost));
  if (vel) HIPCHK(hipMemcpy(vel, c->dWVel.p, n * sizeof(d3), hipMemcpyDeviceToHost));
  if (angvel) HIPCHK(hipMemcpy(angvel, c->dWAng.p, n * sizeof(d3), hipMemcpyDeviceToHost));
  return LSDEM_OK;
}

int lsdem_set_walls(lsdem_ctx* c, const double* pos, const double* quat, const double* vel, const double* angvel) {
  if (!c || !c->initialized) return fail(LSDEM_ERR_STATE, "not initialized");
  HIPCHK(hipSetDevice(c->cfg.device));
  HIPCHK(hipStreamSynchronize(c->stream));
  const size_t n = size_t(c->P.nW);
  if (n == 0) return LSDEM_OK;
  c->forceRebuild = true;  // poses changed outside the step: rebuild the candidate pairs
  if (pos) HIPCHK(hipMemcpy(c->dWPos.p, pos, n * sizeof(d3), hipMemcpyHostToDevice));
  if (quat) HIPCHK(hipMemcpy(c->dWQuat.p, quat, n * sizeof(q4), hipMemcpyHostToDevice));
  if (vel) HIPCHK(hipMemcpy(c->dWVel.p, vel, n * sizeof(d3), hipMemcpyHostToDevice));
  if (angvel) HIPCHK(hipMemcpy(c->dWAng.p, angvel, n * sizeof(d3), hipMemcpyHostToDevice));
  return LSDEM_OK;
}

int lsdem_get_contacts(lsdem_ctx* c, int32_t list, int64_t capacity, int64_t* count, int32_t* master,
                       int32_t* slave, double* point, double* normal, double* overlap, double* spring,
                       double* energy_n, double* energy_s) {
  if (!c || !c->initialized) return fail(LSDEM_ERR_STATE, "not initialized");
  if (list < 0 || list > 4) return fail(LSDEM_ERR_ARG, "bad list id");
  HIPCHK(hipSetDevice(c->cfg.device));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (count) *count = 0;
  if (!c->haveContacts || c->P.nP == 0) return LSDEM_OK;
  int di[8];
  HIPCHK(hipMemcpy(di, c->dDevInt.p, sizeof di, hipMemcpyDeviceToHost));
  const long long n = std::min<long long>(di[1], c->P.contactCap);
  const int b = 1 - c->cur;  // buffer written by the last step
  std::vector<int> node(n);
  std::vector<uint32_t> info(n);
  HIPCHK(hipMemcpy(node.data(), c->dCNode[b].p, n * sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(info.data(), c->dCInfo[b].p, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
  std::vector<long long> sel;
  for (long long k = 0; k < n; ++k)
    if (int(info[k] >> 28) == list) sel.push_back(k);
  std::sort(sel.begin(), sel.end(), [&](long long a, long long b2) {
    if (node[a] != node[b2]) return node[a] < node[b2];
    return (info[a] & kPartnerMask) < (info[b2] & kPartnerMask);
  });
  if (count) *count = (int64_t)sel.size();
  const long long m = std::min<long long>(capacity, (long long)sel.size());
  std::vector<d3> vp, vn, vs;
  std::vector<double> vo, ven, ves;
  if (point) { vp.resize(n); HIPCHK(hipMemcpy(vp.data(), c->dCPoint[b].p, n * sizeof(d3), hipMemcpyDeviceToHost)); }
  if (normal) { vn.resize(n); HIPCHK(hipMemcpy(vn.data(), c->dCNormal[b].p, n * sizeof(d3), hipMemcpyDeviceToHost)); }
  if (spring) { vs.resize(n); HIPCHK(hipMemcpy(vs.data(), c->dCSpring[b].p, n * sizeof(d3), hipMemcpyDeviceToHost)); }
  if (overlap) { vo.resize(n); HIPCHK(hipMemcpy(vo.data(), c->dCOverlap[b].p, n * sizeof(double), hipMemcpyDeviceToHost)); }
  if (energy_n) { ven.resize(n); HIPCHK(hipMemcpy(ven.data(), c->dCEn[b].p, n * sizeof(double), hipMemcpyDeviceToHost)); }
  if (energy_s) { ves.resize(n); HIPCHK(hipMemcpy(ves.data(), c->dCEs[b].p, n * sizeof(double), hipMemcpyDeviceToHost)); }
  for (long long k = 0; k < m; ++k) {
    const long long s = sel[k];
    if (master) master[k] = node[s];
    if (slave) slave[k] = int(info[s] & kPartnerMask);
    if (point) wr3(point, k, vp[s]);
    if (normal) wr3(normal, k, vn[s]);
    if (spring) wr3(spring, k, vs[s]);
    if (overlap) overlap[k] = vo[s];
    if (energy_n) energy_n[k] = ven[s];
    if (energy_s) energy_s[k] = ves[s];
  }
  return LSDEM_OK;
}

int lsdem_get_bonds(lsdem_ctx* c, int32_t* active, double* point, double* energies4, double* sigma_max,
                    double* tau_max) {
  if (!c || !c->initialized) return fail(LSDEM_ERR_STATE, "not initialized");
  HIPCHK(hipSetDevice(c->cfg.device));
  HIPCHK(hipStreamSynchronize(c->stream));
  const int nB = c->P.nBonds;
  if (nB == 0) return LSDEM_OK;
  if (active) HIPCHK(hipMemcpy(active, c->dBActive.p, nB * sizeof(int), hipMemcpyDeviceToHost));
  if (point) HIPCHK(hipMemcpy(point, c->dBPoint.p, nB * sizeof(d3), hipMemcpyDeviceToHost));
  if (energies4 || sigma_max || tau_max) {
    std::vector<double> o(6 * size_t(nB));
    HIPCHK(hipMemcpy(o.data(), c->dBOut.p, o.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int k = 0; k < nB; ++k) {
      if (energies4) for (int t = 0; t < 4; ++t) energies4[4 * k + t] = o[6 * k + t];
      if (sigma_max) sigma_max[k] = o[6 * k + 4];
      if (tau_max) tau_max[k] = o[6 * k + 5];
    }
  }
  return LSDEM_OK;
}

int lsdem_get_counters(lsdem_ctx* c, lsdem_counters* out) {
  if (!c || !out) return fail(LSDEM_ERR_ARG, "null argument");
  std::memset(out, 0, sizeof(*out));
  out->steps = c->stepCount;
  out->pair_capacity = c->P.pairCap;
  out->contact_capacity = c->P.contactCap;
  out->recoveries = c->recoveries;
  if (!c->initialized || c->P.nP == 0) return LSDEM_OK;
  HIPCHK(hipSetDevice(c->cfg.device));
  HIPCHK(hipStreamSynchronize(c->stream));
  int di[8];
  unsigned long long ds[8];
  int np = 0;
  HIPCHK(hipMemcpy(di, c->dDevInt.p, sizeof di, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(ds, c->dDevStat.p, sizeof ds, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(&np, c->dPairStart.p + c->P.nP, sizeof(int), hipMemcpyDeviceToHost));
  out->pairs = np;
  out->contacts = (int64_t)ds[2];
  out->node_tests = (int64_t)ds[0];
  out->sdf_queries = (int64_t)ds[1];
  return LSDEM_OK;
}

int lsdem_profile_begin(lsdem_ctx* c, int32_t max_steps) { return lsdem_profile_begin_sampled(c, max_steps, 1); }

int lsdem_profile_begin_sampled(lsdem_ctx* c, int32_t max_steps, int32_t every) {
  if (!c || !c->initialized) return fail(LSDEM_ERR_STATE, "not initialized");
  if (max_steps < 0) return fail(LSDEM_ERR_ARG, "negative max_steps");
  if (every < 1) return fail(LSDEM_ERR_ARG, "sampling interval below 1");
  HIPCHK(hipSetDevice(c->cfg.device));
  const size_t need = size_t(max_steps) * kPhaseEvents;
  while (c->events.size() < need) {
    hipEvent_t e;
    // timing only: no system-scope fence at each phase boundary (read after a stream sync)
    HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    c->events.push_back(e);
  }
  c->eventOf.assign(need, 0);
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemset(c->dDevStat.p + 3, 0, 6 * sizeof(unsigned long long)));
  HIPCHK(hipMemset(c->dDevStat.p + 12, 0, 3 * sizeof(unsigned long long)));
  c->profMax = max_steps;
  c->profSteps = 0;
  c->profEvery = every;
  c->profSeen = 0;
  c->profiling = true;
  return LSDEM_OK;
}

int lsdem_profile_end(lsdem_ctx* c, lsdem_profile* out) {
  if (!c || !out) return fail(LSDEM_ERR_ARG, "null argument");
  if (!c->profiling) return fail(LSDEM_ERR_STATE, "profiling not started");
  HIPCHK(hipSetDevice(c->cfg.device));
  HIPCHK(hipStreamSynchronize(c->stream));
  std::memset(out, 0, sizeof(*out));
  out->steps = c->profSeen;
  out->sampled_steps = c->profSteps;
  for (int st = 0; st < c->profSteps; ++st) {
    hipEvent_t* ev = &c->events[size_t(st) * kPhaseEvents];
    const int* of = &c->eventOf[size_t(st) * kPhaseEvents];
    for (int ph = 0; ph < LSDEM_NUM_PHASES; ++ph) {
      float ms = 0.f;
      if (of[ph] != of[ph + 1]) HIPCHK(hipEventElapsedTime(&ms, ev[of[ph]], ev[of[ph + 1]]));
      out->ms[ph] += ms;
    }
  }
  unsigned long long ds[kDevStats];
  HIPCHK(hipMemcpy(ds, c->dDevStat.p, sizeof ds, hipMemcpyDeviceToHost));
  out->cluster_tests = (int64_t)ds[8];
  out->master_nodes = (int64_t)ds[12];
  out->evaluated_pairs = (int64_t)ds[13];
  out->rebuilds = (int64_t)ds[14];
  out->pairs = (int64_t)ds[3];
  out->node_tests = (int64_t)ds[4];
  out->sdf_queries = (int64_t)ds[5];
  out->contacts = (int64_t)ds[6];
  c->profiling = false;
  return LSDEM_OK;
}

int lsdem_mass_properties(const double* sdf, int32_t nx, int32_t ny, int32_t nz, const double origin[3],
                          double h, double density, double* mass, double centroid[3], double* inertia6,
                          double* inv_inertia6) {
  if (!sdf || !origin || !mass || !centroid) return fail(LSDEM_ERR_ARG, "null argument");
  if (nx < 1 || ny < 1 || nz < 1 || !(h > 0.0)) return fail(LSDEM_ERR_ARG, "invalid grid");
  // smoothed Heaviside of half-width 1.5 cells
  auto H = [](double phi) -> double {
    const double eps = 1.5;
    if (phi < -eps) return 1.0;
    if (phi > eps) return 0.0;
    const double x = -phi / eps;
    return 0.5 * (1.0 + x + std::sin(kPi * x) / kPi);
  };
  const i3 n{nx, ny, nz};
  const double mg = density * h * h * h;
  double m = 0.;
  for (int x = 0; x < nx; x++)
    for (int y = 0; y < ny; y++)
      for (int z = 0; z < nz; z++) m += H(sdf[lin3(i3{x, y, z}, n)] / h);
  m *= mg;
  *mass = m;
  d3 c = mk3(0., 0., 0.);
  s6 I{0, 0, 0, 0, 0, 0};
  if (m > 0.) {
    const d3 o = mk3(origin[0], origin[1], origin[2]);
    for (int x = 0; x < nx; x++)
      for (int y = 0; y < ny; y++)
        for (int z = 0; z < nz; z++) {
          const double Hv = H(sdf[lin3(i3{x, y, z}, n)] / h);
          c.x += Hv * (o.x + double(x) * h);
          c.y += Hv * (o.y + double(y) * h);
          c.z += Hv * (o.z + double(z) * h);
        }
    c *= mg / m;
    for (int x = 0; x < nx; x++)
      for (int y = 0; y < ny; y++)
        for (int z = 0; z < nz; z++) {
          const double Hv = H(sdf[lin3(i3{x, y, z}, n)] / h);
          const d3 r = o + h * mk3(double(x), double(y), double(z)) - c;
          I.xx += Hv * (r.y * r.y + r.z * r.z) * mg;
          I.yy += Hv * (r.x * r.x + r.z * r.z) * mg;
          I.zz += Hv * (r.y * r.y + r.x * r.x) * mg;
          I.xy -= Hv * r.x * r.y * mg;
          I.xz -= Hv * r.x * r.z * mg;
          I.yz -= Hv * r.y * r.z * mg;
        }
  }
  centroid[0] = c.x; centroid[1] = c.y; centroid[2] = c.z;
  if (inertia6) { inertia6[0] = I.xx; inertia6[1] = I.yy; inertia6[2] = I.zz; inertia6[3] = I.xy; inertia6[4] = I.xz; inertia6[5] = I.yz; }
  if (inv_inertia6) {
    const s6 Ii = m > 0. ? inverse6(I) : s6{0, 0, 0, 0, 0, 0};
    inv_inertia6[0] = Ii.xx; inv_inertia6[1] = Ii.yy; inv_inertia6[2] = Ii.zz;
    inv_inertia6[3] = Ii.xy; inv_inertia6[4] = Ii.xz; inv_inertia6[5] = Ii.yz;
  }
  return LSDEM_OK;
}


int lsdem_set_external_force(lsdem_ctx* c, lsdem_external_force_fn fn, void* user) {
  if (!c) return fail(LSDEM_ERR_ARG, "null context");
  c->hookFn = fn;
  c->hookUser = user;
  return LSDEM_OK;
}

int lsdem_get_device_view(lsdem_ctx* c, lsdem_device_view* out) {
  if (!c || !out) return fail(LSDEM_ERR_ARG, "null argument");
  if (!c->initialized) return fail(LSDEM_ERR_STATE, "not initialized");
  deviceView(c, out);
  return LSDEM_OK;
}

int lsdem_get_wall_device_view(lsdem_ctx* c, lsdem_wall_device_view* out) {
  if (!c || !out) return fail(LSDEM_ERR_ARG, "null argument");
  if (!c->initialized) return fail(LSDEM_ERR_STATE, "lsdem_initialize has not been called");
  const Params& P = c->P;
  out->position = reinterpret_cast<const double*>(P.wpos);
  out->quaternion = reinterpret_cast<const double*>(P.wquat);
  out->velocity = reinterpret_cast<double*>(P.wvel);
  out->angular_velocity = reinterpret_cast<double*>(P.wangvel);
  out->num = P.nW;
  out->block_dim = c->cfg.block_size > 0 ? c->cfg.block_size : 256;
  out->grid_dim = (P.nW + out->block_dim - 1) / out->block_dim;
  out->stream = c->stream;
  return LSDEM_OK;
}

int lsdem_get_time(lsdem_ctx* c, double* time) {
  if (!c || !time) return fail(LSDEM_ERR_ARG, "null argument");
  *time = c->time;
  return LSDEM_OK;
}

int lsdem_set_time(lsdem_ctx* c, double time) {
  if (!c) return fail(LSDEM_ERR_ARG, "null context");
  c->time = time;
  return LSDEM_OK;
}


// ------------------------------------------------------------------------------------------
// Learned culling surrogate
// ------------------------------------------------------------------------------------------
void lsdem_surrogate_params_default(lsdem_surrogate_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->iterations = 3000;
  p->batch = 4096;
  p->samples = 2000000;
  p->validation = 262144;
  p->lr = 2e-3;
  p->lr_final = 1e-4;
  p->clamp = 0.6;
  p->band_factor = 2.0;
  p->audit_every = 100;
  p->seed = 20261206ull;
}

}  // extern "C"

namespace {

// The family is every shape used by a particle; slots follow shape ids.
int surEnsureFamily(lsdem_ctx* c) {
  auto& S = c->sur;
  if (S.nShapes > 0) return LSDEM_OK;
  std::vector<int> used(c->shapes.size(), 0);
  for (int sh : c->pShape) used[sh] = 1;
  std::vector<int> slotOf(std::max<size_t>(c->shapes.size(), 1), -1);
  double maxBox = 0;
  for (size_t k = 0; k < c->shapes.size(); ++k)
    if (used[k]) {
      slotOf[k] = int(S.famShape.size());
      S.famShape.push_back(int(k));
      maxBox = std::max(maxBox, c->shapes[k].boxR);
    }
  if (S.famShape.empty()) return fail(LSDEM_ERR_STATE, "no particle shapes to learn");
  S.nShapes = int(S.famShape.size());
  S.invLen = float(1.0 / maxBox);
  const int nAll = kSurNParam + S.nShapes * kSurEmb;
  HIPCHK(S.param.alloc(kSurNParam));
  HIPCHK(S.embF.alloc(size_t(S.nShapes) * kSurEmb));
  HIPCHK(S.m.alloc(nAll));
  HIPCHK(S.v.alloc(nAll));
  HIPCHK(S.wF.alloc(kSurNTiles * 64));
  HIPCHK(S.wB.alloc(kSurNTiles * 64));
  HIPCHK(S.emb.alloc(size_t(S.nShapes) * kSurEmb));
  HIPCHK(S.loss1.alloc(1));
  HIPCHK(S.errBits.alloc(1));
  HIPCHK(S.errSum.alloc(1));
  HIPCHK(S.famShapeD.alloc(S.nShapes));
  HIPCHK(S.slotOfShape.alloc(slotOf.size()));
  HIPCHK(hipMemcpy(S.famShapeD.p, S.famShape.data(), S.nShapes * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(S.slotOfShape.p, slotOf.data(), slotOf.size() * sizeof(int), hipMemcpyHostToDevice));
  return LSDEM_OK;
}

SurNet surNet(lsdem_ctx* c) {
  auto& S = c->sur;
  SurNet N;
  N.wF = S.wF.p;
  N.wB = S.wB.p;
  N.param = S.param.p;
  N.emb = S.emb.p;
  N.invLen = S.invLen;
  N.nShapes = S.nShapes;
  return N;
}

// Network in use by the step (pair maps carry the family slot from now on).
void surPublish(lsdem_ctx* c) {
  c->P.sur = surNet(c);
  c->P.surSlot = c->sur.slotOfShape.p;
  c->P.surBand = c->sur.band;
}

int surUpload(lsdem_ctx* c, const float* all) {
  auto& S = c->sur;
  HIPCHK(hipMemcpy(S.param.p, all, kSurNParam * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(S.embF.p, all + kSurNParam, size_t(S.nShapes) * kSurEmb * sizeof(float), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_sur_pack, dim3(80), dim3(256), 0, c->stream, (const float*)S.param.p, (const float*)S.embF.p,
                     S.nShapes, S.wF.p, S.wB.p, S.emb.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return LSDEM_OK;
}

int surSamples(lsdem_ctx* c, DBuf<float>& x, DBuf<int>& slot, DBuf<float>& t, long long n, unsigned long long seed) {
  auto& S = c->sur;
  HIPCHK(x.alloc(size_t(n) * 3));
  HIPCHK(slot.alloc(n));
  HIPCHK(t.alloc(n));
  SurFamily F{c->P.shapes, c->P.sdf, S.famShapeD.p, S.nShapes, S.invLen, S.clamp};
  SurSamples D{x.p, slot.p, t.p, n};
  hipLaunchKernelGGL(k_sur_sample, dim3(unsigned(c->numCU * 4)), dim3(256), 0, c->stream, F, D, seed);
  HIPCHK(hipGetLastError());
  return LSDEM_OK;
}

// Max and mean |out - target| over a sample set (normalised units).
int surMeasure(lsdem_ctx* c, DBuf<float>& x, DBuf<int>& slot, DBuf<float>& t, long long n, double& maxErr,
               double& meanErr) {
  auto& S = c->sur;
  HIPCHK(hipMemsetAsync(S.errBits.p, 0, sizeof(unsigned int), c->stream));
  HIPCHK(hipMemsetAsync(S.errSum.p, 0, sizeof(float), c->stream));
  SurSamples D{x.p, slot.p, t.p, n};
  hipLaunchKernelGGL(k_sur_eval, dim3(unsigned(c->numCU * 2)), dim3(256), 0, c->stream, surNet(c), D, (float*)nullptr,
                     S.errBits.p, S.errSum.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  unsigned int bits = 0;
  float sum = 0.f;
  HIPCHK(hipMemcpy(&bits, S.errBits.p, sizeof bits, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(&sum, S.errSum.p, sizeof sum, hipMemcpyDeviceToHost));
  float m;
  std::memcpy(&m, &bits, sizeof m);
  maxErr = m;
  meanErr = n > 0 ? double(sum) / double(n) : 0.0;
  return LSDEM_OK;
}

long long coprimeStride(long long n) {
  long long s = (2654435761ll % n) | 1;
  auto gcd = [](long long a, long long b) {
    while (b) { const long long t = a % b; a = b; b = t; }
    return a;
  };
  while (gcd(s, n) != 1) s += 2;
  return s;
}

}  // namespace

extern "C" {

int lsdem_surrogate_set_weights(lsdem_ctx* c, const float* w, int64_t count) {
  if (!c || !w) return fail(LSDEM_ERR_ARG, "null argument");
  if (!c->initialized) return fail(LSDEM_ERR_STATE, "not initialized");
  HIPCHK(hipSetDevice(c->cfg.device));
  int rc = surEnsureFamily(c);
  if (rc) return rc;
  if (count != kSurNParam + int64_t(c->sur.nShapes) * kSurEmb)
    return fail(LSDEM_ERR_ARG, "weight count must be " + std::to_string(kSurNParam + c->sur.nShapes * kSurEmb));
  if ((rc = surUpload(c, w))) return rc;
  c->sur.trained = true;
  surPublish(c);
  return LSDEM_OK;
}

int lsdem_surrogate_get_weights(lsdem_ctx* c, float* w, int64_t count) {
  if (!c || !w) return fail(LSDEM_ERR_ARG, "null argument");
  if (c->sur.nShapes == 0) return fail(LSDEM_ERR_STATE, "no surrogate");
  if (count != kSurNParam + int64_t(c->sur.nShapes) * kSurEmb) return fail(LSDEM_ERR_ARG, "bad weight count");
  HIPCHK(hipSetDevice(c->cfg.device));
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(w, c->sur.param.p, kSurNParam * sizeof(float), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(w + kSurNParam, c->sur.embF.p, size_t(c->sur.nShapes) * kSurEmb * sizeof(float),
                   hipMemcpyDeviceToHost));
  return LSDEM_OK;
}

int lsdem_surrogate_train(lsdem_ctx* c, const lsdem_surrogate_params* prm) {
  if (!c) return fail(LSDEM_ERR_ARG, "null context");
  if (!c->initialized) return fail(LSDEM_ERR_STATE, "not initialized");
  lsdem_surrogate_params p;
  if (prm) p = *prm;
  else lsdem_surrogate_params_default(&p);
  if (p.iterations < 0 || p.batch <= 0 || p.batch % kSurRows || p.samples <= 0 || p.validation <= 0 || p.clamp <= 0)
    return fail(LSDEM_ERR_ARG, "bad surrogate parameters (batch must be a positive multiple of 64)");
  HIPCHK(hipSetDevice(c->cfg.device));
  int rc = surEnsureFamily(c);
  if (rc) return rc;
  auto& S = c->sur;
  S.clamp = float(p.clamp);
  S.auditEvery = p.audit_every;
  const int nAll = kSurNParam + S.nShapes * kSurEmb;
  // initial weights: He-normal (Box-Muller over a splitmix stream), zero biases
  {
    std::vector<float> w(nAll, 0.f);
    unsigned long long st = p.seed;
    auto next = [&]() {
      st += 0x9e3779b97f4a7c15ull;
      unsigned long long z = st;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
      return double((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0);
    };
    auto normal = [&]() {
      const double u1 = std::max(next(), 1e-300), u2 = next();
      return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    for (int k = 0; k < kSurIn * kSurHid; ++k) w[kSurOffW[0] + k] = float(normal() * std::sqrt(2.0 / kSurIn));
    for (int l = 1; l < 4; ++l)
      for (int k = 0; k < kSurHid * kSurHid; ++k) w[kSurOffW[l] + k] = float(normal() * std::sqrt(2.0 / kSurHid));
    for (int k = 0; k < kSurHid; ++k) w[kSurOffW5 + k] = float(normal() * std::sqrt(1.0 / kSurHid));
    for (int k = 0; k < S.nShapes * kSurEmb; ++k) w[kSurNParam + k] = float(normal() * 0.5);
    if ((rc = surUpload(c, w.data()))) return rc;
  }
  HIPCHK(hipMemsetAsync(S.m.p, 0, nAll * sizeof(float), c->stream));
  HIPCHK(hipMemsetAsync(S.v.p, 0, nAll * sizeof(float), c->stream));
  if ((rc = surSamples(c, S.sx, S.sslot, S.st, p.samples, p.seed ^ 0x51ull))) return rc;
  if ((rc = surSamples(c, S.vx, S.vslot, S.vt, p.validation, p.seed ^ 0xA7ull))) return rc;
  const int blocks = p.batch / kSurRows;
  HIPCHK(S.partial.alloc(size_t(blocks) * (nAll + 1)));
  const long long stride = coprimeStride(p.samples);
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, c->stream));
  const double b1 = 0.9, b2 = 0.999;
  for (int it = 0; it < p.iterations; ++it) {
    const double lr = p.lr_final + 0.5 * (p.lr - p.lr_final) * (1.0 + std::cos(3.141592653589793 * it / p.iterations));
    SurTrainArgs T{surNet(c), SurSamples{S.sx.p, S.sslot.p, S.st.p, p.samples},
                   (long long)((unsigned long long)it * (unsigned long long)p.batch % (unsigned long long)p.samples),
                   stride, p.batch, S.partial.p};
    hipLaunchKernelGGL(k_sur_train, dim3(unsigned(blocks)), dim3(256), 0, c->stream, T);
    SurAdamArgs Ad{S.param.p, S.embF.p, S.m.p, S.v.p, S.partial.p, blocks, S.nShapes, float(lr), float(b1), float(b2),
                   1e-8f, float(1.0 / (1.0 - std::pow(b1, it + 1))), float(1.0 / (1.0 - std::pow(b2, it + 1))),
                   S.wF.p, S.wB.p, S.emb.p, S.loss1.p};
    hipLaunchKernelGGL(k_sur_adam, dim3(unsigned((nAll + 255) / 256)), dim3(256), 0, c->stream, Ad);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(e1, c->stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  float loss = 0.f;
  HIPCHK(hipMemcpy(&loss, S.loss1.p, sizeof loss, hipMemcpyDeviceToHost));
  S.loss = loss;
  S.trainMs = ms;
  S.iterations = p.iterations;
  if ((rc = surMeasure(c, S.vx, S.vslot, S.vt, p.validation, S.maxErr, S.meanErr))) return rc;
  S.band = float(p.band_factor * S.maxErr / S.invLen);
  S.trained = true;
  surPublish(c);
  return LSDEM_OK;
}

int lsdem_surrogate_enable(lsdem_ctx* c, int32_t on) {
  if (!c) return fail(LSDEM_ERR_ARG, "null context");
  if (on && !c->sur.trained) return fail(LSDEM_ERR_STATE, "surrogate not trained");
  c->sur.enabled = on != 0;
  return LSDEM_OK;
}

int lsdem_surrogate_set_band(lsdem_ctx* c, double band, int32_t audit_every) {
  if (!c) return fail(LSDEM_ERR_ARG, "null context");
  if (!(band >= 0.0)) return fail(LSDEM_ERR_ARG, "band must be >= 0");
  c->sur.band = float(band);
  c->sur.auditEvery = audit_every;
  if (c->sur.trained) surPublish(c);
  return LSDEM_OK;
}

int lsdem_surrogate_get_report(lsdem_ctx* c, lsdem_surrogate_report* r) {
  if (!c || !r) return fail(LSDEM_ERR_ARG, "null argument");
  std::memset(r, 0, sizeof(*r));
  const auto& S = c->sur;
  r->trained = S.trained;
  r->enabled = S.enabled;
  r->n_shapes = S.nShapes;
  r->n_params = kSurNParam + S.nShapes * kSurEmb;
  r->iterations = S.iterations;
  r->final_loss = S.loss;
  r->max_error = S.maxErr / S.invLen;
  r->mean_error = S.meanErr / S.invLen;
  r->band = S.band;
  r->train_ms = S.trainMs;
  r->length_scale = 1.0 / S.invLen;
  r->audits = S.audits;
  if (c->initialized) {
    HIPCHK(hipSetDevice(c->cfg.device));
    HIPCHK(hipStreamSynchronize(c->stream));
    unsigned long long ds[12];
    HIPCHK(hipMemcpy(ds, c->dDevStat.p, sizeof ds, hipMemcpyDeviceToHost));
    r->audited_contacts = (int64_t)ds[10];
    r->missed_contacts = (int64_t)ds[11];
  }
  return LSDEM_OK;
}

int lsdem_surrogate_eval(lsdem_ctx* c, int64_t n, const float* points, const int32_t* shape_ids, float* out) {
  if (!c || (n > 0 && (!points || !shape_ids || !out))) return fail(LSDEM_ERR_ARG, "null argument");
  if (c->sur.nShapes == 0) return fail(LSDEM_ERR_STATE, "no surrogate");
  HIPCHK(hipSetDevice(c->cfg.device));
  auto& S = c->sur;
  std::vector<int> slot(n);
  std::vector<int> slotOf(c->shapes.size(), -1);
  for (int k = 0; k < S.nShapes; ++k) slotOf[S.famShape[k]] = k;
  for (int64_t i = 0; i < n; ++i) {
    if (shape_ids[i] < 0 || shape_ids[i] >= int(slotOf.size()) || slotOf[shape_ids[i]] < 0)
      return fail(LSDEM_ERR_ARG, "shape outside the surrogate family");
    slot[i] = slotOf[shape_ids[i]];
  }
  DBuf<float> x, o;
  DBuf<int> sl;
  HIPCHK(x.alloc(size_t(n) * 3));
  HIPCHK(o.alloc(n));
  HIPCHK(sl.alloc(n));
  HIPCHK(hipMemcpy(x.p, points, size_t(n) * 3 * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(sl.p, slot.data(), size_t(n) * sizeof(int), hipMemcpyHostToDevice));
  SurSamples D{x.p, sl.p, nullptr, n};
  hipLaunchKernelGGL(k_sur_eval, dim3(unsigned(c->numCU * 2)), dim3(256), 0, c->stream, surNet(c), D, o.p,
                     (unsigned int*)nullptr, (float*)nullptr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  std::vector<float> r(n);
  HIPCHK(hipMemcpy(r.data(), o.p, size_t(n) * sizeof(float), hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < n; ++i) out[i] = r[i] / S.invLen;
  return LSDEM_OK;
}

int lsdem_surrogate_gradient(lsdem_ctx* c, int64_t n, const float* points, const int32_t* shape_ids,
                             const float* targets, float* grad, double* loss) {
  if (!c || !points || !shape_ids || !targets || !grad) return fail(LSDEM_ERR_ARG, "null argument");
  if (n <= 0 || n % kSurRows) return fail(LSDEM_ERR_ARG, "n must be a positive multiple of 64");
  if (c->sur.nShapes == 0) return fail(LSDEM_ERR_STATE, "no surrogate");
  HIPCHK(hipSetDevice(c->cfg.device));
  auto& S = c->sur;
  std::vector<int> slotOf(c->shapes.size(), -1), slot(n);
  for (int k = 0; k < S.nShapes; ++k) slotOf[S.famShape[k]] = k;
  for (int64_t i = 0; i < n; ++i) {
    if (shape_ids[i] < 0 || shape_ids[i] >= int(slotOf.size()) || slotOf[shape_ids[i]] < 0)
      return fail(LSDEM_ERR_ARG, "shape outside the surrogate family");
    slot[i] = slotOf[shape_ids[i]];
  }
  const int blocks = int(n / kSurRows);
  const int nAll = kSurNParam + S.nShapes * kSurEmb;
  DBuf<float> x, t, part;
  DBuf<int> sl;
  HIPCHK(x.alloc(size_t(n) * 3));
  HIPCHK(t.alloc(n));
  HIPCHK(sl.alloc(n));
  HIPCHK(part.alloc(size_t(blocks) * (nAll + 1)));
  HIPCHK(hipMemcpy(x.p, points, size_t(n) * 3 * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(t.p, targets, size_t(n) * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(sl.p, slot.data(), size_t(n) * sizeof(int), hipMemcpyHostToDevice));
  SurTrainArgs T{surNet(c), SurSamples{x.p, sl.p, t.p, n}, 0, 1, int(n), part.p};
  hipLaunchKernelGGL(k_sur_train, dim3(unsigned(blocks)), dim3(256), 0, c->stream, T);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  std::vector<float> h(size_t(blocks) * (nAll + 1));
  HIPCHK(hipMemcpy(h.data(), part.p, h.size() * sizeof(float), hipMemcpyDeviceToHost));
  double l = 0;
  for (int k = 0; k < nAll; ++k) {
    double g = 0;
    for (int b = 0; b < blocks; ++b) g += h[size_t(b) * (nAll + 1) + k];
    grad[k] = float(g);
  }
  for (int b = 0; b < blocks; ++b) l += h[size_t(b) * (nAll + 1) + nAll];
  if (loss) *loss = l;
  return LSDEM_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------
// Allocation-free variant of lsdem_surrogate_gradient
// ------------------------------------------------------------------------------------------
namespace {

// Scratch of lsdem_surrogate_gradient_noalloc: an error word (first 256 bytes), the family slot
// of each sample, then the per-block partial gradients and losses that k_sum_partials adds up in
// block order.  Every GPU the library runs on holds its own copy until the process ends; host code
// reaches a device's copy only through the address hipGetSymbolAddress gives on that device.
constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratchMutex;
unsigned char* g_scratchPtr[kMaxScratchDevices] = {};

// Device address of `scratch` on device `dev`, which must be current; looked up once per device.
int scratchOn(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kMaxScratchDevices) return fail(LSDEM_ERR_ARG, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratchMutex);
  if (!g_scratchPtr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      return fail(LSDEM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratchPtr[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratchPtr[dev];
  return LSDEM_OK;
}

__global__ void k_clear_word(int* w) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *w = 0;
}

// Family slot of each sample's shape id; an id outside the family raises *bad and takes slot 0
// (so the kernels that follow stay in bounds; the call then reports the error).
__global__ void k_sur_slots(const int32_t* shapeIds, long long n, const int* slotOfShape, int nShapeIds, int* slot,
                            int* bad) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const int id = shapeIds[i];
    int sl = (id >= 0 && id < nShapeIds) ? slotOfShape[id] : -1;
    if (sl < 0) {
      atomicOr(bad, 1);
      sl = 0;
    }
    slot[i] = sl;
  }
}

// grad[k] = sum over blocks of part[b][k] in block order, in FP64 then rounded to FP32 (k < nAll);
// k = nAll: the loss, kept in FP64.  The host sum of lsdem_surrogate_gradient, term for term.
__global__ void k_sum_partials(const float* part, int blocks, int nAll, float* grad, double* loss) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > nAll) return;
  double g = 0;
  for (int b = 0; b < blocks; ++b) g += part[size_t(b) * (nAll + 1) + k];
  if (k < nAll) grad[k] = float(g);
  else if (loss) *loss = g;
}

}  // namespace

extern "C" {

int lsdem_surrogate_gradient_noalloc(lsdem_ctx* c, int64_t n, const float* points, const int32_t* shape_ids,
                                     const float* targets, float* grad, double* loss) {
  if (!c || !points || !shape_ids || !targets || !grad) return fail(LSDEM_ERR_ARG, "null argument");
  if (n <= 0 || n % kSurRows) return fail(LSDEM_ERR_ARG, "n must be a positive multiple of 64");
  if (c->sur.nShapes == 0) return fail(LSDEM_ERR_STATE, "no surrogate");
  HIPCHK(hipSetDevice(c->cfg.device));
  unsigned char* base = nullptr;
  if (int rc = scratchOn(c->cfg.device, &base)) return rc;
  auto& S = c->sur;
  const int blocks = int(n / kSurRows);
  const int nAll = kSurNParam + S.nShapes * kSurEmb;
  const size_t slotOff = 256;
  const size_t partOff = slotOff + (size_t(n) * sizeof(int) + 255) / 256 * 256;
  const size_t perBlock = size_t(nAll + 1) * sizeof(float);
  if (partOff + size_t(blocks) * perBlock > kNoAllocScratchBytes) {
    // largest multiple of 64 samples that fits: 256 + 256 * b + b * perBlock <= size (b blocks)
    const long long maxBlocks = (long long)((kNoAllocScratchBytes - 256) / (256 + perBlock));
    return fail(LSDEM_ERR_CAPACITY, "n beyond the scratch of the allocation-free gradient (" +
                                        std::to_string(kNoAllocScratchBytes >> 20) + " MiB): at most " +
                                        std::to_string(maxBlocks * kSurRows) + " samples with " +
                                        std::to_string(S.nShapes) + " shapes");
  }
  int* bad = reinterpret_cast<int*>(base);
  int* slot = reinterpret_cast<int*>(base + slotOff);
  float* part = reinterpret_cast<float*>(base + partOff);
  hipStream_t s = c->stream;
  hipLaunchKernelGGL(k_clear_word, dim3(1), dim3(64), 0, s, bad);
  hipLaunchKernelGGL(k_sur_slots, dim3(unsigned(std::min<long long>((n + 255) / 256, 1024))), dim3(256), 0, s,
                     shape_ids, (long long)n, (const int*)S.slotOfShape.p, int(c->shapes.size()), slot, bad);
  SurTrainArgs T{surNet(c), SurSamples{const_cast<float*>(points), slot, const_cast<float*>(targets), n}, 0, 1,
                 int(n), part};
  hipLaunchKernelGGL(k_sur_train, dim3(unsigned(blocks)), dim3(256), 0, s, T);
  hipLaunchKernelGGL(k_sum_partials, dim3(unsigned((nAll + 1 + 255) / 256)), dim3(256), 0, s, (const float*)part,
                     blocks, nAll, grad, loss);
  HIPCHK(hipGetLastError());
  int badHost = 0;
  HIPCHK(hipMemcpyAsync(&badHost, bad, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (badHost) return fail(LSDEM_ERR_ARG, "shape outside the surrogate family");
  return LSDEM_OK;
}

}  // extern "C"
