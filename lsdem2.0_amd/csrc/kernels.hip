// HIP kernels of the LS-DEM timestep for gfx950 (CDNA4, wave64).
//
// Per step (launch order, one stream):
//   k_kick_drift      half-kick, drift, reference-grid key, search cell, periodic wrap + ghost images
//   k_walls           wall drift + wall keys
//   counting sort     particles (and ghost images) by search-grid cell, deterministic order
//   k_pairs (x2)      bounding-sphere / grid-box candidate pairs per master particle (count, write)
//   reaction CSR      pairs grouped by the partner that receives the reaction force
//   k_narrow          one wavefront per candidate pair: cluster cull, node-vs-SDF, contact law,
//                     tangential history, per-pair force/torque sums
//   k_hash_*          history table of this step's contacts for the next step
//   k_bonds           parallel bonds
//   k_reduce_kick     per-body segmented sum of pair and bond contributions + second half-kick
#include "device.h"

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
__global__ void k_fill_int(int* a, long long n, int v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    a[i] = v;
}
__global__ void k_fill_u64(unsigned long long* a, long long n, unsigned long long v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    a[i] = v;
}

// Exclusive scan of n ints (n + 1 outputs, out[n] = total), three kernels.
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ int blockExclusiveScan(int v, int* lds, int& total) {
  // lds needs kScanThreads/64 + 1 ints
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
    for (int w = 0; w < kScanThreads / 64; ++w) {
      int t = lds[w];
      lds[w] = acc;
      acc += t;
    }
    lds[kScanThreads / 64] = acc;
  }
  __syncthreads();
  total = lds[kScanThreads / 64];
  int r = lds[wid] + incl - v;
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const int* in, long long n, int* blockSum) {
  __shared__ int lds[kScanThreads / 64 + 1];
  const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanItems;
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) s += in[base + k];
  int total;
  blockExclusiveScan(s, lds, total);
  if (threadIdx.x == 0) blockSum[blockIdx.x] = total;
}
__global__ __launch_bounds__(kScanThreads) void k_scan_top(int* blockSum, int nb, int* totalOut) {
  __shared__ int lds[kScanThreads / 64 + 1];
  int carry = 0;
  for (int base = 0; base < nb; base += kScanThreads) {
    const int i = base + threadIdx.x;
    const int v = i < nb ? blockSum[i] : 0;
    int total;
    const int ex = blockExclusiveScan(v, lds, total);
    if (i < nb) blockSum[i] = ex + carry;
    carry += total;
  }
  if (threadIdx.x == 0) *totalOut = carry;
}
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const int* in, long long n, const int* blockSum, int* out) {
  __shared__ int lds[kScanThreads / 64 + 1];
  const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanItems;
  int v[kScanItems];
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0;
    s += v[k];
  }
  int total;
  int ex = blockExclusiveScan(s, lds, total) + blockSum[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) out[base + k] = ex;
    ex += v[k];
  }
}

// ------------------------------------------------------------------------------------------
// Integration: first half-kick + drift (+ keys, wrap, ghost images)
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ void halfKick(const Params& P, int i, d3 F, d3 T) {
  const double invM = P.invM[i];
  if (isZero(invM)) return;
  const double h = 0.5 * P.dt;
  P.vel[i] += (F * invM + P.g) * h;
  P.angvel[i] += (mul(rotInvInertia(P.quat[i], P.invI[i]), T)) * h;
}

__global__ void k_kick_drift(Params P) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.nP) return;
  halfKick(P, i, P.force[i], P.torque[i]);
  d3 p = P.pos[i];
  p += P.dt * P.vel[i];
  P.quat[i] = qintegrate(P.quat[i], P.angvel[i], P.dt);
  const bool periodic = P.periodicX || P.periodicY;
  const bool wrapAfter = (P.quirks & 4) != 0;
  const GridGeom& G = P.gridP;
  const d3 ox = mk3(G.mx.x - G.mn.x, 0., 0.), oy = mk3(0., G.mx.y - G.mn.y, 0.);
  if (periodic && !wrapAfter) {
    if (P.periodicX) { if (p.x < G.mn.x) p += ox; else if (p.x >= G.mx.x) p -= ox; }
    if (P.periodicY) { if (p.y < G.mn.y) p += oy; else if (p.y >= G.mx.y) p -= oy; }
  }
  P.posDet[i] = p;
  P.keyRef[i] = refKey(G, p);
  P.scell[i] = lin3(searchCell(P.sg, p), P.sg.size);
  if (periodic) {
    d3 img;
    if (P.periodicX) {
      if (wrapAfter) { if (p.x < G.mn.x) p += ox; else if (p.x >= G.mx.x) p -= ox; }
      img = p;
      if (refCell(G, p).x == 0) img += ox;
      const int k = ghostKey(P, img);
      P.gpos[0][i] = img;
      P.gkey[0][i] = k;
      P.gcell[0][i] = k >= 0 ? lin3(searchCell(P.sg, img), P.sg.size) : -1;
    }
    if (P.periodicY) {
      if (wrapAfter) { if (p.y < G.mn.y) p += oy; else if (p.y >= G.mx.y) p -= oy; }
      img = p;
      if (refCell(G, p).y == 0) img += oy;
      const int k = ghostKey(P, img);
      P.gpos[1][i] = img;
      P.gkey[1][i] = k;
      P.gcell[1][i] = k >= 0 ? lin3(searchCell(P.sg, img), P.sg.size) : -1;
    }
    if (P.periodicX && P.periodicY) {
      img = p;
      const i3 c = refCell(G, p);
      if (c.x == 0 && c.y == 0) img += ox + oy;
      const int k = ghostKey(P, img);
      P.gpos[2][i] = img;
      P.gkey[2][i] = k;
      P.gcell[2][i] = k >= 0 ? lin3(searchCell(P.sg, img), P.sg.size) : -1;
    }
  }
  P.pos[i] = p;
}

__global__ void k_walls(Params P) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= P.nW) return;
  d3 p = P.wpos[w];
  p += P.dt * P.wvel[w];
  P.wpos[w] = p;
  P.wquat[w] = qintegrate(P.wquat[w], P.wangvel[w], P.dt);
  P.wkey[w] = refKey(P.gridW, p);
}

// ------------------------------------------------------------------------------------------
// Deterministic counting sort by bucket: count (atomic rank) -> scan -> scatter -> per-bucket
// insertion sort by item index.
// ------------------------------------------------------------------------------------------
__global__ void k_bucket_count(const int* bucket, int n, int* count, int* rank) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = bucket[i];
  rank[i] = b >= 0 ? atomicAdd(&count[b], 1) : -1;
}
__global__ void k_bucket_scatter(const int* bucket, int n, const int* start, const int* rank, int* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = bucket[i];
  if (b >= 0) out[start[b] + rank[i]] = i;
}
__global__ void k_bucket_sort(const int* start, int nb, int* items) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int s = start[b], e = start[b + 1];
  for (int k = s + 1; k < e; ++k) {
    const int v = items[k];
    int j = k - 1;
    while (j >= s && items[j] > v) {
      items[j + 1] = items[j];
      --j;
    }
    items[j + 1] = v;
  }
}

// ------------------------------------------------------------------------------------------
// Candidate pairs
// ------------------------------------------------------------------------------------------
// Visit every item of a search-grid table lying in cells that can hold a point within `reach`
// of x.  Rows (y, z) are pruned by their distance to x; the x-run of cells of a row is
// contiguous in the sorted order, so each row costs one range lookup.
template <class F>
__device__ __forceinline__ void visitNear(const SearchGrid& S, const int* cellStart, const int* sorted, d3 x,
                                          double reach, F&& f) {
  const i3 c = searchCell(S, x);
  const int R = max(1, int(ceil(reach * S.invCell)));
  const double r2 = reach * reach;
  for (int cz = max(0, c.z - R); cz <= min(S.size.z - 1, c.z + R); ++cz) {
    const double z0 = cz == 0 ? -1e300 : S.mn.z + cz * S.cell;
    const double z1 = cz == S.size.z - 1 ? 1e300 : S.mn.z + (cz + 1) * S.cell;
    const double dz = fmax(0.0, fmax(z0 - x.z, x.z - z1));
    if (dz * dz > r2) continue;
    for (int cy = max(0, c.y - R); cy <= min(S.size.y - 1, c.y + R); ++cy) {
      const double y0 = cy == 0 ? -1e300 : S.mn.y + cy * S.cell;
      const double y1 = cy == S.size.y - 1 ? 1e300 : S.mn.y + (cy + 1) * S.cell;
      const double dy = fmax(0.0, fmax(y0 - x.y, x.y - y1));
      const double rem = r2 - dy * dy - dz * dz;
      if (rem < 0.0) continue;
      const double hw = sqrt(rem);
      const int x0 = clampCell(floor((x.x - hw - S.mn.x) * S.invCell), S.size.x);
      const int x1 = clampCell(floor((x.x + hw - S.mn.x) * S.invCell), S.size.x);
      const int row = (cz * S.size.y + cy) * S.size.x;
      const int s = cellStart[row + x0], e = cellStart[row + x1 + 1];
      for (int k = s; k < e; ++k) f(sorted[k]);
    }
  }
}

__device__ __forceinline__ double triAt(const Params& P, const ShapeDev& S, double sc, d3 loc, d3 go, double invSp,
                                        bool& inside) {
  // trilinear SDF value at a local point (for pair-level culls), inside = within the grid box
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

template <bool WRITE>
__global__ void k_pairs(Params P) {
  const int A = blockIdx.x * blockDim.x + threadIdx.x;
  if (A >= P.nP) return;
  int cnt = 0;
  long long out = WRITE ? P.pairStart[A] : 0;
  auto emit = [&](uint32_t type, int partner) {
    if (WRITE) {
      if (out < P.pairCap) {
        P.pairA[out] = A;
        P.pairInfo[out] = (type << 28) | uint32_t(partner);
      }
      ++out;
    }
    ++cnt;
  };
  const double rA = P.radius[A];
  // particle-particle: B > A, B inside the domain, bounding spheres overlap
  {
    const d3 xA = P.posDet[A];
    const double reach = (rA + P.rMax) * (1.0 + 2e-9) + 2e-12;
    visitNear(P.sg, P.cellStart, P.sorted, xA, reach, [&](int B) {
      if (B <= A || P.keyRef[B] < 0) return;
      const double rr = (rA + P.radius[B]) * (1.0 + 1e-9) + 1e-12;
      if (len2(xA - P.posDet[B]) > rr * rr) return;
      emit(PT_PP, B);
    });
  }
  // walls: sphere vs grid box, then a Lipschitz bound on the wall SDF at the centre
  if (P.nW > 0) {
    const d3 xA = P.posDet[A];
    for (int w = 0; w < P.nW; ++w) {
      if (P.wkey[w] < 0) continue;
      const ShapeDev S = P.shapes[P.wshape[w]];
      const d3 loc = rotInv(xA - P.wpos[w], P.wquat[w]);
      const d3 o = mk3(S.ox, S.oy, S.oz);
      const double rr = rA * (1.0 + 1e-9) + 1e-12;
      if (boxDist2(loc, o, S.h, S) > rr * rr) continue;
      bool inside;
      const double phi = triAt(P, S, 1.0, loc, o, 1. / S.h, inside);
      if (inside && phi > S.lip * rA * (1.0 + 1e-6) + 1e-9 * S.h) continue;
      emit(PT_PW, w);
    }
  }
  // periodic ghost images: any node inside the image's grid box is tested, so the reach is
  // the image's grid-box radius
  if (P.periodicX || P.periodicY) {
    const d3 xA = P.pos[A];
    const double reach = (rA + P.boxRMax) * (1.0 + 2e-9) + 2e-12;
    for (int s = 0; s < 3; ++s) {
      const bool on = (s == 0 && P.periodicX) || (s == 1 && P.periodicY) || (s == 2 && P.periodicX && P.periodicY);
      if (!on) continue;
      const d3* gp = P.gpos[s];
      visitNear(P.sg, P.gCellStart[s], P.gSorted[s], xA, reach, [&](int j) {
        const double bj = P.scale[j] * P.shapes[P.shape[j]].boxR;
        const double rr = (rA + bj) * (1.0 + 1e-9) + 1e-12;
        if (len2(xA - gp[j]) > rr * rr) return;
        emit(PT_GX + s, j);
      });
    }
  }
  if (!WRITE) P.pairCount[A] = cnt;
}

// Pairs whose partner particle receives the reaction (particle-particle, ghosts with j != A).
__global__ void k_rc_count(Params P) {
  const int n = min(P.pairStart[P.nP], (int)P.pairCap);
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    const uint32_t info = P.pairInfo[p];
    const uint32_t type = info >> 28;
    const int partner = int(info & kPartnerMask);
    int r = -1;
    if (type == PT_PP || (type >= PT_GX && partner != P.pairA[p])) r = atomicAdd(&P.rcCount[partner], 1);
    P.rcRank[p] = r;
  }
}
__global__ void k_rc_scatter(Params P) {
  const int n = min(P.pairStart[P.nP], (int)P.pairCap);
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    const int r = P.rcRank[p];
    if (r < 0) continue;
    const int partner = int(P.pairInfo[p] & kPartnerMask);
    P.rcPair[P.rcStart[partner] + r] = p;
  }
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
__global__ void k_hash_insert(Params P) {
  const int n = min(P.devInt[1], (int)P.contactCap);
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    const uint32_t info = P.cInfo[c];
    const unsigned long long key = histKey(info >> 28, P.cNode[c], int(info & kPartnerMask));
    unsigned long long slot = mix64(key) & P.hashMask;
    for (;;) {
      const unsigned long long old = atomicCAS(&P.hashKey[slot], kHashEmpty, key);
      if (old == kHashEmpty || old == key) {
        P.hashVal[slot] = c;
        break;
      }
      slot = (slot + 1) & P.hashMask;
    }
  }
}
__device__ __forceinline__ d3 histLookup(const Params& P, unsigned long long key) {
  unsigned long long slot = mix64(key) & P.hashMask;
  for (;;) {
    const unsigned long long k = P.hashKey[slot];
    if (k == key) return P.prevSpring[P.hashVal[slot]];
    if (k == kHashEmpty) return mk3(0., 0., 0.);
    slot = (slot + 1) & P.hashMask;
  }
}

// ------------------------------------------------------------------------------------------
// Narrow phase + contact law, one wavefront per candidate pair
// ------------------------------------------------------------------------------------------
struct Partner {
  d3 x;       // pose used for detection
  q4 q;
  d3 xf;      // position used by the force law (torque arm)
  d3 v, w;
  double invM, kn, ks, mu, res;
  double r;   // particle bounding radius (particle-particle cull)
  double sc, invSp;
  d3 go;
  int shape, key;
};

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

constexpr int kNarrowBlock = 256;

__global__ __launch_bounds__(kNarrowBlock) void k_narrow(Params P) {
  const int lane = threadIdx.x & 63;
  const unsigned long long ltMask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  unsigned long long statTests = 0, statQueries = 0;
  const int nPairs = min(P.pairStart[P.nP], (int)P.pairCap);
  for (;;) {
    int p = 0;
    if (lane == 0) p = atomicAdd(&P.devInt[2], 1);
    p = __builtin_amdgcn_readfirstlane(__shfl(p, 0, 64));
    if (p >= nPairs) break;

    const uint32_t info = P.pairInfo[p];
    const uint32_t type = info >> 28;
    const int partner = int(info & kPartnerMask);
    const int A = P.pairA[p];
    const bool ghostPair = type >= PT_GX;
    // master particle
    const d3 xA = ghostPair ? P.pos[A] : P.posDet[A];
    const q4 qA = P.quat[A];
    const double sA = P.scale[A];
    const ShapeDev SA = P.shapes[P.shape[A]];
    const int nodeBase = P.nodeStart[A];
    const d3 riA = P.pos[A];
    const d3 vA = P.vel[A], wA = P.angvel[A];
    const double invMA = P.invM[A], knA = P.kn[A], ksA = P.ks[A], muA = P.mu[A], resA = P.res[A];
    // partner
    Partner B;
    if (type == PT_PW) {
      B.x = P.wpos[partner]; B.q = P.wquat[partner]; B.xf = B.x;
      B.v = P.wvel[partner]; B.w = P.wangvel[partner];
      B.invM = 0.; B.kn = 0.; B.ks = 0.; B.mu = P.wmu[partner]; B.res = P.wres[partner];
      B.r = 0.; B.shape = P.wshape[partner]; B.sc = 1.0;
      const ShapeDev SB = P.shapes[B.shape];
      B.invSp = 1. / SB.h; B.go = mk3(SB.ox, SB.oy, SB.oz);
      B.key = P.wkey[partner];
    } else {
      const int j = partner;
      if (type == PT_PP) { B.x = P.posDet[j]; B.xf = P.pos[j]; B.v = P.vel[j]; B.key = P.keyRef[j]; }
      else { const int s = int(type - PT_GX); B.x = P.gpos[s][j]; B.xf = B.x; B.v = P.vel[j]; B.key = P.gkey[s][j]; }
      B.q = P.quat[j]; B.w = P.angvel[j];
      B.invM = P.invM[j]; B.kn = P.kn[j]; B.ks = P.ks[j]; B.mu = P.mu[j]; B.res = P.res[j];
      B.r = P.radius[j]; B.shape = P.shape[j]; B.sc = P.scale[j]; B.invSp = P.invSp[j]; B.go = P.gorig[j];
    }
    const ShapeDev SB = P.shapes[B.shape];
    // pair-level material constants
    double kn = 0., ks = 0., effM = 0., mu, res;
    if (type == PT_PW) {
      kn = knA; ks = ksA;
      if (invMA > 0.) effM = 1. / invMA;
    } else {
      if (knA > 0. && B.kn > 0.) kn = knA * B.kn / (knA + B.kn);
      if (ksA > 0. && B.ks > 0.) ks = ksA * B.ks / (ksA + B.ks);
      if (invMA > 0. || B.invM > 0.) effM = 1. / (invMA + B.invM);
    }
    res = fmin(resA, B.res);
    mu = fmin(muA, B.mu);

    // cluster cull (conservative): lane k tests cluster k
    unsigned long long cmask;
    {
      bool keep = false;
      if (lane < SA.nClusters) {
        const ClusterDev cl = P.clusters[SA.clusterOff + lane];
        const d3 cw = rot(qA, mk3(cl.cx, cl.cy, cl.cz) * sA) + xA;  // cluster centre, world
        const double rho = cl.rho * sA * (1.0 + 1e-9) + 1e-12;
        if (type == PT_PP) {
          const double rr = rho + B.r * (1.0 + 1e-9) + 1e-12;
          keep = len2(cw - B.x) <= rr * rr;
        } else {
          const d3 loc = rotInv(cw - B.x, B.q);
          keep = boxDist2(loc, B.go, SB.h * B.sc, SB) <= rho * rho;
          if (keep && type == PT_PW) {
            bool inside;
            const double phi = triAt(P, SB, 1.0, loc, B.go, B.invSp, inside);
            if (inside && phi > SB.lip * rho * (1.0 + 1e-6) + 1e-9 * SB.h) keep = false;
          }
        }
      }
      cmask = __ballot(keep);
    }

    d3 FA = mk3(0., 0., 0.), TA = mk3(0., 0., 0.), TB = mk3(0., 0., 0.);
    int pairContacts = 0;
    const float* sdfB = P.sdf + SB.sdfOff;
    const i3 nB{SB.nx, SB.ny, SB.nz};
    const GridGeom& G = (type == PT_PW) ? P.gridW : P.gridP;
    while (cmask) {
      const int k = __builtin_ctzll(cmask);
      cmask &= cmask - 1;
      const ClusterDev cl = P.clusters[SA.clusterOff + k];
      bool hit = false;
      int gnode = 0;
      d3 point, normal, spring, F;
      double ov = 0., En = 0., Es = 0.;
      if (lane < cl.count) {
        const NodeDev nd = P.nodes[SA.nodeOff + cl.first + lane];
        gnode = nodeBase + nd.orig;
        const d3 xw = rot(qA, mk3(nd.x * sA, nd.y * sA, nd.z * sA)) + xA;
        bool eligible = true;
        statTests++;
        if (type == PT_PP) {
          const d3 rel = xw - B.x;
          eligible = !(len2(rel) > B.r * B.r);
        } else {
          eligible = keyInWindow(G, B.key, xw);
        }
        if (eligible) {
          const d3 loc = rotInv(xw - B.x, B.q);
          const double gx = (loc.x - B.go.x) * B.invSp, gy = (loc.y - B.go.y) * B.invSp,
                       gz = (loc.z - B.go.z) * B.invSp;
          const int i0 = (int)floor(gx), j0 = (int)floor(gy), k0 = (int)floor(gz);
          if (!(i0 < 0 || j0 < 0 || k0 < 0 || i0 >= nB.x - 1 || j0 >= nB.y - 1 || k0 >= nB.z - 1)) {
            const int i1 = i0 + 1, j1 = j0 + 1, k1 = k0 + 1;
            const double x = gx - double(i0), y = gy - double(j0), z = gz - double(k0);
            double c[8];
            c[0] = double(sdfB[lin3(i3{i0, j0, k0}, nB)]) * B.sc;
            c[1] = double(sdfB[lin3(i3{i1, j0, k0}, nB)]) * B.sc;
            c[2] = double(sdfB[lin3(i3{i0, j1, k0}, nB)]) * B.sc;
            c[3] = double(sdfB[lin3(i3{i1, j1, k0}, nB)]) * B.sc;
            c[4] = double(sdfB[lin3(i3{i0, j0, k1}, nB)]) * B.sc;
            c[5] = double(sdfB[lin3(i3{i1, j0, k1}, nB)]) * B.sc;
            c[6] = double(sdfB[lin3(i3{i0, j1, k1}, nB)]) * B.sc;
            c[7] = double(sdfB[lin3(i3{i1, j1, k1}, nB)]) * B.sc;
            statQueries++;
            ov = -triValue(x, y, z, c);
            if (ov >= 0.) {
              hit = true;
              d3 n = triGradient((P.quirks & 1) ? double(int(x)) : x, y, z, c);
              n = rot(B.q, n);
              n = normalize(n);
              normal = n;
              point = xw + 0.5 * ov * n;
            }
          }
        }
      }
      if (hit) {
        // tangential history of the same (list, node, partner) from the previous step
        spring = histLookup(P, histKey(type, gnode, partner));
        const d3 vc = vA + cross(wA, point - riA) - (B.v + cross(B.w, point - B.xf));
        F = mk3(0., 0., 0.);
        linearContact(F, spring, vc, normal, ov, P.dt, kn, ks, mu, res, effM);
        En = 0.5 * ov * ov * kn;
        Es = 0.5 * len2(spring) * ks;
        FA += F;
        TA += cross(point - riA, F);
        TB += cross(point - B.xf, -F);
      }
      const unsigned long long hm = __ballot(hit);
      if (hm) {
        const int cnt = __popcll(hm);
        int base = 0;
        if (lane == 0) base = atomicAdd(&P.devInt[1], cnt);
        base = __shfl(base, 0, 64);
        pairContacts += cnt;
        if ((long long)base + cnt > P.contactCap) {
          if (lane == 0) atomicOr(&P.devInt[3], 2);
        } else if (hit) {
          const int o = base + __popcll(hm & ltMask);
          P.cNode[o] = gnode;
          P.cInfo[o] = info;
          P.cSpring[o] = spring;
          P.cPoint[o] = point;
          P.cNormal[o] = normal;
          P.cOverlap[o] = ov;
          P.cEn[o] = En;
          P.cEs[o] = Es;
        }
      }
    }
    if (pairContacts > 0) {
      const double s0 = wsum(FA.x), s1 = wsum(FA.y), s2 = wsum(FA.z);
      const double s3 = wsum(TA.x), s4 = wsum(TA.y), s5 = wsum(TA.z);
      const double s6v = wsum(TB.x), s7 = wsum(TB.y), s8 = wsum(TB.z);
      if (lane == 0) {
        double* o = P.pairSum + 9ll * p;
        o[0] = s0; o[1] = s1; o[2] = s2; o[3] = s3; o[4] = s4; o[5] = s5; o[6] = s6v; o[7] = s7; o[8] = s8;
      }
    }
    if (lane == 0) P.pairFlag[p] = pairContacts;
  }
  // statistics, one atomic per wave
  const unsigned long long t = (unsigned long long)wsum(double(statTests));
  const unsigned long long q = (unsigned long long)wsum(double(statQueries));
  if (lane == 0) {
    atomicAdd(&P.devStat[0], t);
    atomicAdd(&P.devStat[1], q);
  }
}

// ------------------------------------------------------------------------------------------
// Bonds (one thread per bond)
// ------------------------------------------------------------------------------------------
__global__ void k_bonds(Params P) {
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
__global__ void k_reduce_kick(Params P) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= P.nP) return;
  d3 F = mk3(0., 0., 0.), T = mk3(0., 0., 0.);
  const long long cap = P.pairCap;
  for (int p = P.pairStart[b]; p < P.pairStart[b + 1] && p < cap; ++p) {
    if (P.pairFlag[p] == 0) continue;
    const double* s = P.pairSum + 9ll * p;
    F += mk3(s[0], s[1], s[2]);
    T += mk3(s[3], s[4], s[5]);
  }
  for (int q = P.rcStart[b]; q < P.rcStart[b + 1]; ++q) {
    const int p = P.rcPair[q];
    if (P.pairFlag[p] == 0) continue;
    const double* s = P.pairSum + 9ll * p;
    F += -mk3(s[0], s[1], s[2]);
    T += mk3(s[6], s[7], s[8]);
  }
  if (P.nBonds > 0) {
    for (int q = P.bcStart[b]; q < P.bcStart[b + 1]; ++q) {
      const int it = P.bcItem[q];
      const int k = it >> 1;
      if (it & 1) { F += -P.bF[k]; T += P.bMj[k]; }
      else { F += P.bF[k]; T += P.bMi[k]; }
    }
  }
  P.force[b] = F;
  P.torque[b] = T;
  halfKick(P, b, F, T);
}

// ------------------------------------------------------------------------------------------
// Step bookkeeping
// ------------------------------------------------------------------------------------------
__global__ void k_step_begin(Params P) {
  if (threadIdx.x == 0) {
    P.devInt[1] = 0;  // contacts written
    P.devInt[2] = 0;  // narrow-phase work counter
    P.devStat[0] = 0;
    P.devStat[1] = 0;
  }
}
__global__ void k_step_end(Params P) {
  if (threadIdx.x == 0) {
    const int np = P.pairStart[P.nP];
    if ((long long)np > P.pairCap) P.devInt[3] |= 1;
    P.devInt[4] = max(P.devInt[4], np);
    P.devInt[5] = max(P.devInt[5], P.devInt[1]);
  }
}

template __global__ void k_pairs<false>(Params P);
template __global__ void k_pairs<true>(Params P);

}  // namespace lsd
