// LS-DEM timestep oracle: a serial C++ restatement of FunDEM's level-set DEM step.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or calls this
// code; it is the checker that the GPU path is compared against (tests/, the smoke test
// and the CPU-baseline leg of bench.py).
//
// Arithmetic follows the reference kernels expression by expression (FP64, host branches
// of the reference math headers: 1/sqrt instead of rsqrt, std::sin/std::cos), so that the
// contact set is reproduced bit for bit.  Built with -ffp-contract=off.
//
// Reference files (relative to the reference checkout):
//   FunDEM/kernel/CUDAKernelFunction/myUtility/myVec.h, myQua.h, myMat.h   (math)
//   FunDEM/kernel/CUDAKernelFunction/particleIntegrationKernel.cu        (kick, drift)
//   FunDEM/kernel/CUDAKernelFunction/myUtility/mySpatialGrid.{h,cu}       (keys, wrap, ghosts)
//   FunDEM/kernel/CUDAKernelFunction/myUtility/buildHashStartEnd.cu       (sort, cell table)
//   FunDEM/kernel/CUDAKernelFunction/LSParticleContactDetectionKernel.{cu,cuh} (narrow phase)
//   FunDEM/kernel/CUDAKernelFunction/contactKernel.{cu,cuh}               (contact law, bonds)
//   FunDEM/kernel/LSSolver.h, LSParticle.h, SolidInteraction.h, PeriodicBoundary.h,
//   VBondedInteraction.h                                                  (step order, setup)
//
// Pinning status of each function is listed in the header of tests/test_oracle.py; the
// functions below whose results depend on the reference's CUDA device math (rsqrt, log)
// are pinned only up to tolerance.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <unordered_map>
#include <vector>

namespace lsdem_oracle {

// ---------------------------------------------------------------------------------------
// Math (myVec.h:13-235, myQua.h:10-129, myMat.h:15-195; host branches)
// ---------------------------------------------------------------------------------------
struct V3 { double x, y, z; };
struct Q4 { double q0, q1, q2, q3; };
struct S6 { double xx, yy, zz, xy, xz, yz; };
struct I3 { int x, y, z; };

constexpr double kPi = 3.14159265358979323846;  // myVec.h:13-16

inline V3 v3(double x, double y, double z) { return V3{x, y, z}; }
inline bool isZero(double x) { return std::fabs(x) <= 1.0e-20; }  // myVec.h:22-25
inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator*(V3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
inline V3 operator*(double s, V3 a) { return v3(a.x * s, a.y * s, a.z * s); }
inline V3 operator/(V3 a, double s) { double inv = 1.0 / s; return v3(a.x * inv, a.y * inv, a.z * inv); }
inline V3& operator+=(V3& a, V3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
inline V3& operator-=(V3& a, V3 b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
inline V3& operator*=(V3& a, double s) { a.x *= s; a.y *= s; a.z *= s; return a; }
inline V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline double lengthSquared(V3 v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
inline double length(V3 v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
inline V3 normalize(V3 v) {  // myVec.h:121-137, host branch
  double lenSq = v.x * v.x + v.y * v.y + v.z * v.z;
  if (isZero(lenSq)) return v3(0.0, 0.0, 0.0);
  double invLen = 1.0 / std::sqrt(lenSq);
  return v3(v.x * invLen, v.y * invLen, v.z * invLen);
}
inline int linearIndex3D(I3 ijk, I3 dims) {  // myVec.h:232-235
  return ijk.z * dims.y * dims.x + ijk.y * dims.x + ijk.x;
}

inline Q4 operator+(Q4 a, Q4 b) { return Q4{a.q0 + b.q0, a.q1 + b.q1, a.q2 + b.q2, a.q3 + b.q3}; }
inline Q4 normalize(Q4 q) {  // myQua.h:51-62, host branch
  double lenSq = q.q0 * q.q0 + q.q1 * q.q1 + q.q2 * q.q2 + q.q3 * q.q3;
  if (isZero(lenSq)) return q;
  double invLen = 1.0 / std::sqrt(lenSq);
  return Q4{q.q0 * invLen, q.q1 * invLen, q.q2 * invLen, q.q3 * invLen};
}
inline Q4 quaternionIntegration(Q4 q, V3 w, double dt) {  // myQua.h:77-93
  V3 v = w * (0.5 * dt);
  Q4 d;
  d.q0 = -q.q1 * v.x - q.q2 * v.y - q.q3 * v.z;
  d.q1 = q.q0 * v.x + q.q3 * v.y - q.q2 * v.z;
  d.q2 = -q.q3 * v.x + q.q0 * v.y + q.q1 * v.z;
  d.q3 = q.q2 * v.x - q.q1 * v.y + q.q0 * v.z;
  return normalize(q + d);
}
inline V3 rotate(Q4 q, V3 v) {  // rotateVectorByQuaternion, myQua.h:104-114
  V3 qv = v3(q.q1, q.q2, q.q3);
  V3 t = cross(qv, v) * 2.0;
  return v + (t * q.q0) + cross(qv, t);
}
inline V3 rotateInv(V3 v, Q4 q) {  // reverseRotateVectorByQuaternion, myQua.h:120-129
  V3 qv = v3(-q.q1, -q.q2, -q.q3);
  V3 t = cross(qv, v) * 2.0;
  return v + (t * q.q0) + cross(qv, t);
}

// v rotated by the rotation vector omega (Rodrigues), rotateVector, myVec.h:143-184 (host branch).
inline V3 rotateVector(V3 v, V3 omega) {
  const double theta2 = dot(omega, omega);
  const V3 oxv = cross(omega, v);
  const V3 oxoxv = cross(omega, oxv);
  if (isZero(theta2)) return v + oxv + 0.5 * oxoxv;
  const double invTheta = 1.0 / std::sqrt(theta2);
  const double theta = theta2 * invTheta;
  const double A = std::sin(theta) * invTheta;
  const double B = (1.0 - std::cos(theta)) * (invTheta * invTheta);
  return v + A * oxv + B * oxoxv;
}
// (rotation by omega) x q, normalised; rotateQuaternion, myQua.h:131-179 (host branch).
inline Q4 rotateQuaternion(Q4 q, V3 omega) {
  const double theta2 = omega.x * omega.x + omega.y * omega.y + omega.z * omega.z;
  if (isZero(theta2)) return q;
  const double theta = std::sqrt(theta2);
  const double invTheta = 1.0 / theta;
  const double ux = omega.x * invTheta, uy = omega.y * invTheta, uz = omega.z * invTheta;
  const double c = std::cos(0.5 * theta), sn = std::sin(0.5 * theta);
  const double r0 = c, r1 = sn * ux, r2 = sn * uy, r3 = sn * uz;
  Q4 o;
  o.q0 = r0 * q.q0 - r1 * q.q1 - r2 * q.q2 - r3 * q.q3;
  o.q1 = r0 * q.q1 + r1 * q.q0 + r2 * q.q3 - r3 * q.q2;
  o.q2 = r0 * q.q2 - r1 * q.q3 + r2 * q.q0 + r3 * q.q1;
  o.q3 = r0 * q.q3 + r1 * q.q2 - r2 * q.q1 + r3 * q.q0;
  return normalize(o);
}
inline V3 operator*(S6 m, V3 v) {  // myMat.h:80-87
  return v3(m.xx * v.x + m.xy * v.y + m.xz * v.z, m.xy * v.x + m.yy * v.y + m.yz * v.z,
            m.xz * v.x + m.yz * v.y + m.zz * v.z);
}
inline S6 rotateInvInertia(Q4 q, S6 I) {  // myMat.h:107-158
  double q0 = q.q0, q1 = q.q1, q2 = q.q2, q3 = q.q3;
  double q11 = 2 * q1 * q1, q22 = 2 * q2 * q2, q33 = 2 * q3 * q3;
  double a = 1.0 - q22 - q33, b = 2 * (q1 * q2 - q0 * q3), c = 2 * (q1 * q3 + q0 * q2);
  double d = 2 * (q1 * q2 + q0 * q3), e = 1.0 - q11 - q33, f = 2 * (q2 * q3 - q0 * q1);
  double g = 2 * (q1 * q3 - q0 * q2), h = 2 * (q2 * q3 + q0 * q1), i = 1.0 - q11 - q22;
  double r0x = I.xx * a + I.xy * b + I.xz * c, r0y = I.xy * a + I.yy * b + I.yz * c,
         r0z = I.xz * a + I.yz * b + I.zz * c;
  double r1x = I.xx * d + I.xy * e + I.xz * f, r1y = I.xy * d + I.yy * e + I.yz * f,
         r1z = I.xz * d + I.yz * e + I.zz * f;
  double r2x = I.xx * g + I.xy * h + I.xz * i, r2y = I.xy * g + I.yy * h + I.yz * i,
         r2z = I.xz * g + I.yz * h + I.zz * i;
  return S6{r0x * a + r0y * b + r0z * c, r1x * d + r1y * e + r1z * f, r2x * g + r2y * h + r2z * i,
            r0x * d + r0y * e + r0z * f, r0x * g + r0y * h + r0z * i, r1x * g + r1y * h + r1z * i};
}
inline S6 inverse(S6 A) {  // myMat.h:163-195
  const double det = A.xx * (A.yy * A.zz - A.yz * A.yz) + A.xy * (A.yz * A.xz - A.xy * A.zz) +
                     A.xz * (A.xy * A.yz - A.yy * A.xz);
  if (std::fabs(det) < 1e-30) return S6{0, 0, 0, 0, 0, 0};
  const double di = 1.0 / det;
  return S6{(A.yy * A.zz - A.yz * A.yz) * di,  (A.xx * A.zz - A.xz * A.xz) * di,
            (A.xx * A.yy - A.xy * A.xy) * di,  -(A.xy * A.zz - A.xz * A.yz) * di,
            (A.xy * A.yz - A.xz * A.yy) * di,  -(A.xx * A.yz - A.xz * A.xy) * di};
}

// Trilinear value, LSParticleContactDetectionKernel.cuh:4-32 (same expression order).
inline double interpValue(double x, double y, double z, const double* p) {
  const double wx0 = 1.0 - x, wx1 = x, wy0 = 1.0 - y, wy1 = y, wz0 = 1.0 - z, wz1 = z;
  return p[0] * wx0 * wy0 * wz0 + p[1] * wx1 * wy0 * wz0 + p[2] * wx0 * wy1 * wz0 +
         p[3] * wx1 * wy1 * wz0 + p[4] * wx0 * wy0 * wz1 + p[5] * wx1 * wy0 * wz1 +
         p[6] * wx0 * wy1 * wz1 + p[7] * wx1 * wy1 * wz1;
}
// Gradient, LSParticleContactDetectionKernel.cuh:34-77.  The reference declares the first
// parameter `const int x`, so the fractional x coordinate arrives truncated to 0; with the
// quirk switch on this function receives x = 0 exactly as the reference does.
inline V3 interpGradient(int xi, double y, double z, const double* p) {
  const double wx0 = 1.0 - xi, wx1 = xi, wy0 = 1.0 - y, wy1 = y, wz0 = 1.0 - z, wz1 = z;
  const double gx = (p[1] - p[0]) * wy0 * wz0 + (p[3] - p[2]) * wy1 * wz0 +
                    (p[5] - p[4]) * wy0 * wz1 + (p[7] - p[6]) * wy1 * wz1;
  const double gy = (p[2] - p[0]) * wx0 * wz0 + (p[3] - p[1]) * wx1 * wz0 +
                    (p[6] - p[4]) * wx0 * wz1 + (p[7] - p[5]) * wx1 * wz1;
  const double gz = (p[4] - p[0]) * wx0 * wy0 + (p[5] - p[1]) * wx1 * wy0 +
                    (p[6] - p[2]) * wx0 * wy1 + (p[7] - p[3]) * wx1 * wy1;
  return v3(gx, gy, gz);
}
// Same with a real-valued x (quirk switch off).
inline V3 interpGradientExact(double x, double y, double z, const double* p) {
  const double wx0 = 1.0 - x, wx1 = x, wy0 = 1.0 - y, wy1 = y, wz0 = 1.0 - z, wz1 = z;
  const double gx = (p[1] - p[0]) * wy0 * wz0 + (p[3] - p[2]) * wy1 * wz0 +
                    (p[5] - p[4]) * wy0 * wz1 + (p[7] - p[6]) * wy1 * wz1;
  const double gy = (p[2] - p[0]) * wx0 * wz0 + (p[3] - p[1]) * wx1 * wz0 +
                    (p[6] - p[4]) * wx0 * wz1 + (p[7] - p[5]) * wx1 * wz1;
  const double gz = (p[4] - p[0]) * wx0 * wy0 + (p[5] - p[1]) * wx1 * wy0 +
                    (p[6] - p[2]) * wx0 * wy1 + (p[7] - p[3]) * wx1 * wy1;
  return v3(gx, gy, gz);
}

// ---------------------------------------------------------------------------------------
// Scene
// ---------------------------------------------------------------------------------------
enum Quirk : int {
  QUIRK_INT_X_GRADIENT = 1,      // LSParticleContactDetectionKernel.cuh:34
  QUIRK_GHOST_Z_VS_MAXY = 2,     // mySpatialGrid.cu:236
  QUIRK_WRAP_AFTER_BUILD = 4,    // LSSolver.h:1071-1074 (contacts built before the wrap)
  QUIRK_ALL = 7,
};

struct Shape {
  int nx = 0, ny = 0, nz = 0;
  V3 origin{0, 0, 0};
  double spacing = 0;
  std::vector<float> sdf;  // x fastest
  std::vector<V3> nodes;   // local (unscaled) boundary nodes
};

struct Grid {  // spatialGrid::set, mySpatialGrid.h:212-240
  V3 mn{0, 0, 0}, mx{0, 0, 0}, inv{0, 0, 0};
  I3 size{2, 2, 2};
  void set(V3 minB, V3 maxB, double cell) {
    if (maxB.x <= minB.x) maxB.x = minB.x + cell;
    if (maxB.y <= minB.y) maxB.y = minB.y + cell;
    if (maxB.z <= minB.z) maxB.z = minB.z + cell;
    mn = minB;
    mx = maxB;
    size = I3{2, 2, 2};
    inv = v3(0, 0, 0);
    if (isZero(cell)) return;
    const V3 d = maxB - minB;
    size.x = d.x > 2. * cell ? int(d.x / cell) : 2;
    size.y = d.y > 2. * cell ? int(d.y / cell) : 2;
    size.z = d.z > 2. * cell ? int(d.z / cell) : 2;
    inv.x = (d.x > 0.0) ? (double(size.x) / d.x) : 0.0;
    inv.y = (d.y > 0.0) ? (double(size.y) / d.y) : 0.0;
    inv.z = (d.z > 0.0) ? (double(size.z) / d.z) : 0.0;
  }
  int ncell() const { return size.x * size.y * size.z; }
  I3 cellOf(V3 p) const {  // calculateGridPosition, mySpatialGrid.h:24-31
    return I3{int((p.x - mn.x) * inv.x), int((p.y - mn.y) * inv.y), int((p.z - mn.z) * inv.z)};
  }
  int key(V3 p) const {  // calculateHashKernel, mySpatialGrid.cu:4-26
    if (p.x < mn.x) return -1; else if (p.x >= mx.x) return -1;
    if (p.y < mn.y) return -1; else if (p.y >= mx.y) return -1;
    if (p.z < mn.z) return -1; else if (p.z >= mx.z) return -1;
    return linearIndex3D(cellOf(p), size);
  }
  // Clamped node cell and the +-1 search window, LSParticleContactDetectionKernel.cu:35-43.
  void window(V3 p, I3& c, I3& lo, I3& hi) const {
    c = cellOf(p);
    lo = I3{-1, -1, -1};
    hi = I3{1, 1, 1};
    if (c.x <= 0) { c.x = 0; lo.x = 0; }
    if (c.x >= size.x - 1) { c.x = size.x - 1; hi.x = 0; }
    if (c.y <= 0) { c.y = 0; lo.y = 0; }
    if (c.y >= size.y - 1) { c.y = size.y - 1; hi.y = 0; }
    if (c.z <= 0) { c.z = 0; lo.z = 0; }
    if (c.z >= size.z - 1) { c.z = size.z - 1; hi.z = 0; }
  }
  I3 decode(int key) const {
    I3 r;
    r.x = key % size.x;
    r.y = (key / size.x) % size.y;
    r.z = key / (size.x * size.y);
    return r;
  }
};

// Body set (particles or walls): SoA of the reference's per-body arrays (LSParticle.h:745-769).
struct Bodies {
  std::vector<int> shape;
  std::vector<double> scale, radius, invM, kn, ks, mu, e;
  std::vector<V3> pos, vel, angvel, force, torque;
  std::vector<Q4> q;
  std::vector<S6> invI;
  std::vector<int> nodeStart;  // prefix of boundary-node counts (particles only)
  size_t n() const { return pos.size(); }
};

struct Contact {
  int master;   // global boundary-node index
  int slave;    // partner index (particle, wall or ghost source particle)
  V3 point, normal, spring;
  double overlap, En, Es;
};

struct Bond {  // VBondedInteraction.h:74-195
  int i, j;
  int active;
  double B1, B2, B3, B4, radius, L0, sigmaT, cohesion, mu;
  V3 n1i, n2i, n3i, pi, n1j, n2j, n3j, pj;
  V3 point;
  double Un, Us, Ub, Ut, sigmaMax, tauMax;
};

enum ListId { LIST_PP = 0, LIST_PW = 1, LIST_GX = 2, LIST_GY = 3, LIST_GXY = 4, NUM_LISTS = 5 };

struct Ghost {  // PeriodicBoundary.h:7-170
  std::vector<V3> pos, vel;
  std::vector<Q4> q;
  std::vector<int> key;
};

struct Counters {
  int64_t pairs = 0;        // candidate (master particle, partner) pairs examined
  int64_t nodeTests = 0;    // nodes tested against a partner
  int64_t sdfQueries = 0;   // 8-corner SDF interpolations
  int64_t contacts = 0;
};

class Solver {
 public:
  std::vector<Shape> shapes;
  Bodies P, W;
  std::vector<Bond> bonds;
  V3 domMin{0, 0, 0}, domMax{1, 1, 1}, g{0, 0, 0};
  double dt = 1e-5;
  bool perX = false, perY = false;
  // quarter-sector rotational periodicity about the z axis through the domain's lower corner
  // (LSSolver::addPeriodicBoundarySector, PeriodicBoundary.h:313-420); its three image sets are
  // rotated by +90, 180 and -90 degrees and use the lists of the x, y and xy image sets
  bool perSector = false;
  int quirks = QUIRK_ALL;
  bool literal = false;  // true: node-centric 27-cell traversal exactly as the reference
  Grid gridP, gridW;
  std::vector<Contact> lists[NUM_LISTS];
  Ghost ghost[3];
  Counters counters;
  int64_t stepCount = 0;
  double time = 0.0;  // accumulated as LSSolver::compute does (time += timeStep, LSSolver.h:1013)
  // LSSolver::addExternalForceTorque(time) (LSSolver.h:508), called after all contact and bond
  // forces and before the second half-kick (:1009,1042,1079,1117)
  std::function<void(double)> external;

  // ---- setup ----
  int addShape(const float* sdf, int nx, int ny, int nz, V3 origin, double spacing, const double* nodes,
               int nNodes) {
    Shape s;
    s.nx = nx; s.ny = ny; s.nz = nz; s.origin = origin; s.spacing = spacing;
    s.sdf.assign(sdf, sdf + size_t(nx) * ny * nz);
    for (int i = 0; i < nNodes; ++i) s.nodes.push_back(v3(nodes[3 * i], nodes[3 * i + 1], nodes[3 * i + 2]));
    shapes.push_back(std::move(s));
    return int(shapes.size()) - 1;
  }
  // LSParticle::add (LSParticle.h:115-212): clamps and radius = max |node|.
  void addBody(Bodies& B, bool wall, int shape, double scale, V3 x, Q4 q, V3 v, V3 w, double invM, S6 invI,
               double kn, double ks, double mu, double e, double wallRadius) {
    const Shape& s = shapes[shape];
    double r = 0.;
    if (!wall) {
      int start = B.nodeStart.empty() ? 0 : B.nodeStart.back();
      if (B.nodeStart.empty()) B.nodeStart.push_back(0);
      for (const V3& n : s.nodes) r = std::max(length(n * scale), r);
      B.nodeStart.push_back(start + int(s.nodes.size()));
    } else {
      r = wallRadius;
    }
    B.shape.push_back(shape);
    B.scale.push_back(scale);
    B.radius.push_back(r);
    B.pos.push_back(x);
    B.vel.push_back(v);
    B.angvel.push_back(w);
    B.force.push_back(v3(0, 0, 0));
    B.torque.push_back(v3(0, 0, 0));
    B.q.push_back(normalize(q));
    B.invM.push_back(invM);
    B.invI.push_back(invI);
    B.kn.push_back(kn > 0. ? kn : 0.);
    B.ks.push_back(ks > 0. ? ks : 0.);
    B.mu.push_back(mu > 0. ? mu : 0.);
    double ee = e;
    if (ee <= 0.0 || ee > 1.0) ee = 1.0;
    B.e.push_back(ee);
  }
  int numNodes() const { return P.nodeStart.empty() ? 0 : P.nodeStart.back(); }

  // LSParticle::initialize (LSParticle.h:241-256): cell = 2 * max radius.
  void initialize() {
    double cellP = 0., cellW = 0.;
    for (double r : P.radius) cellP = std::max(cellP, r);
    for (double r : W.radius) cellW = std::max(cellW, r);
    gridP.set(domMin, domMax, P.n() ? cellP * 2.0 : 0.);
    gridW.set(domMin, domMax, W.n() ? cellW * 2.0 : 0.);
    for (auto& l : lists) l.clear();
    primePending = true;
  }

  // ---- per-body geometry of the scaled shape ----
  V3 nodeLocal(const Bodies& B, int body, int k) const {
    return shapes[B.shape[body]].nodes[k] * B.scale[body];
  }
  V3 gridOrigin(const Bodies& B, int body) const { return shapes[B.shape[body]].origin * B.scale[body]; }
  double invSpacing(const Bodies& B, int body) const {
    return 1. / (shapes[B.shape[body]].spacing * B.scale[body]);
  }
  double sdfAt(const Bodies& B, int body, int lin) const {
    return double(shapes[B.shape[body]].sdf[lin]) * B.scale[body];
  }

  // Node-vs-SDF test (LSParticleContactDetectionKernel.cu:62-115 / :355-409).  Returns true
  // and fills the contact when overlap >= 0.  `rel` is the node minus partner position.
  bool nodeVsSdf(const Bodies& B, int b, V3 xw, V3 xB, Q4 qB, Contact& c, Counters& cn) {
    const V3 rel = xw - xB;
    const V3 loc = rotateInv(rel, qB);
    const V3 o = gridOrigin(B, b);
    const double is = invSpacing(B, b);
    const Shape& s = shapes[B.shape[b]];
    const I3 n{s.nx, s.ny, s.nz};
    const double gx = (loc.x - o.x) * is, gy = (loc.y - o.y) * is, gz = (loc.z - o.z) * is;
    int i0 = (int)std::floor(gx), j0 = (int)std::floor(gy), k0 = (int)std::floor(gz);
    if (i0 < 0 || j0 < 0 || k0 < 0) return false;
    if (i0 >= n.x - 1 || j0 >= n.y - 1 || k0 >= n.z - 1) return false;
    const int i1 = i0 + 1, j1 = j0 + 1, k1 = k0 + 1;
    const double x = gx - double(i0), y = gy - double(j0), z = gz - double(k0);
    double p[8];
    p[0] = sdfAt(B, b, linearIndex3D(I3{i0, j0, k0}, n));
    p[1] = sdfAt(B, b, linearIndex3D(I3{i1, j0, k0}, n));
    p[2] = sdfAt(B, b, linearIndex3D(I3{i0, j1, k0}, n));
    p[3] = sdfAt(B, b, linearIndex3D(I3{i1, j1, k0}, n));
    p[4] = sdfAt(B, b, linearIndex3D(I3{i0, j0, k1}, n));
    p[5] = sdfAt(B, b, linearIndex3D(I3{i1, j0, k1}, n));
    p[6] = sdfAt(B, b, linearIndex3D(I3{i0, j1, k1}, n));
    p[7] = sdfAt(B, b, linearIndex3D(I3{i1, j1, k1}, n));
    cn.sdfQueries++;
    const double ov = -interpValue(x, y, z, p);
    if (!(ov >= 0.)) return false;
    V3 nc = (quirks & QUIRK_INT_X_GRADIENT) ? interpGradient(int(x), y, z, p) : interpGradientExact(x, y, z, p);
    nc = rotate(qB, nc);
    nc = normalize(nc);
    c.point = xw + 0.5 * ov * nc;
    c.normal = nc;
    c.overlap = ov;
    c.slave = b;
    return true;
  }

  // ---- phases ----
  void kick(double h) {  // particleIntegrationKernel.cu:3-24
    for (size_t i = 0; i < P.n(); ++i) {
      double invM = P.invM[i];
      if (isZero(invM)) continue;
      P.vel[i] += (P.force[i] * invM + g) * h;
      P.angvel[i] += (rotateInvInertia(P.q[i], P.invI[i]) * P.torque[i]) * h;
    }
  }
  static void drift(Bodies& B, double h) {  // particleIntegrationKernel.cu:26-40
    for (size_t i = 0; i < B.n(); ++i) {
      B.pos[i] += h * B.vel[i];
      B.q[i] = quaternionIntegration(B.q[i], B.angvel[i], h);
    }
  }

  // Sorted (key, index) list with cell start/end (buildHashStartEnd.cu:18-137).  Ties are
  // broken by index (stable); only traversal order depends on it, never the contact set.
  struct CellTable {
    std::vector<int> sortedIdx, start, end;
  };
  static CellTable buildCells(const std::vector<int>& keys, int ncell) {
    CellTable t;
    std::vector<int> idx(keys.size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = int(i);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return keys[a] < keys[b]; });
    t.sortedIdx = idx;
    t.start.assign(ncell, -1);
    t.end.assign(ncell, -1);
    for (size_t s = 0; s < idx.size(); ++s) {
      int h = keys[idx[s]];
      if (h < 0 || h >= ncell) continue;
      if (s == 0 || keys[idx[s - 1]] != h) t.start[h] = int(s);
      if (s == idx.size() - 1 || keys[idx[s + 1]] != h) t.end[h] = int(s + 1);
    }
    return t;
  }

  typedef std::unordered_map<uint64_t, V3> History;
  static uint64_t hkey(int node, int slave) { return (uint64_t(uint32_t(node)) << 32) | uint32_t(slave); }
  static History makeHistory(const std::vector<Contact>& l) {
    History h;
    h.reserve(l.size() * 2 + 1);
    for (const Contact& c : l) h.emplace(hkey(c.master, c.slave), c.spring);  // first match wins
    return h;
  }

  // Particle-particle contacts, node-centric literal traversal (LSParticleContactDetectionKernel.cu:124-297).
  void buildPP_literal(const std::vector<int>& keyP, const CellTable& ct, const History& hist,
                       std::vector<Contact>& out) {
    overMasters(out, [&](size_t A, std::vector<Contact>& out, Counters& cn) {
      for (int k = P.nodeStart[A]; k < P.nodeStart[A + 1]; ++k) {
        const V3 xw = rotate(P.q[A], nodeLocal(P, int(A), k - P.nodeStart[A])) + P.pos[A];
        I3 c, lo, hi;
        gridP.window(xw, c, lo, hi);
        for (int zz = lo.z; zz <= hi.z; zz++)
          for (int yy = lo.y; yy <= hi.y; yy++)
            for (int xx = lo.x; xx <= hi.x; xx++) {
              const int h = linearIndex3D(I3{c.x + xx, c.y + yy, c.z + zz}, gridP.size);
              if (ct.start[h] == -1) continue;
              for (int s = ct.start[h]; s < ct.end[h]; ++s) {
                const int B = ct.sortedIdx[s];
                if (int(A) >= B) continue;
                cn.nodeTests++;
                const V3 rel = xw - P.pos[B];
                const double rB = P.radius[B];
                if (lengthSquared(rel) > rB * rB) continue;
                Contact ct2;
                if (!nodeVsSdf(P, B, xw, P.pos[B], P.q[B], ct2, cn)) continue;
                finishContact(ct2, k, hist, cn);
                out.push_back(ct2);
              }
            }
      }
    });
    (void)keyP;
  }

  // Same contact set from bounding-sphere pairs: for A < B with |xA-xB| <= rA+rB, test all
  // nodes of A.  Equivalent because a node within rB of xB lies in a cell within +-1 of B's
  // cell whenever the cell width is >= 2*max radius (the reference's choice).
  void buildPP_pairs(const std::vector<int>& keyP, const History& hist, std::vector<Contact>& out) {
    const CellTable ct = buildCells(keyP, gridP.ncell());
    overMasters(out, [&](size_t A, std::vector<Contact>& out, Counters& cn) {
      std::vector<int> cand;
      // candidate partners: B > A, valid key, bounding spheres overlap (with slack).  B lies
      // within rA + rB <= 2 max radius <= one cell width of A, so the clamped 27-cell window
      // around A's centre holds every candidate.
      cand.clear();
      I3 c, lo, hi;
      gridP.window(P.pos[A], c, lo, hi);
      for (int zz = lo.z; zz <= hi.z; zz++)
        for (int yy = lo.y; yy <= hi.y; yy++)
          for (int xx = lo.x; xx <= hi.x; xx++) {
            const int h = linearIndex3D(I3{c.x + xx, c.y + yy, c.z + zz}, gridP.size);
            if (ct.start[h] == -1) continue;
            for (int s = ct.start[h]; s < ct.end[h]; ++s) {
              const int B = ct.sortedIdx[s];
              if (B <= int(A)) continue;
              const double rr = (P.radius[A] + P.radius[B]) * (1.0 + 1e-9) + 1e-12;
              if (lengthSquared(P.pos[A] - P.pos[B]) > rr * rr) continue;
              cand.push_back(B);
            }
          }
      std::sort(cand.begin(), cand.end());
      cn.pairs += int64_t(cand.size());
      for (int k = P.nodeStart[A]; k < P.nodeStart[A + 1]; ++k) {
        const V3 xw = rotate(P.q[A], nodeLocal(P, int(A), k - P.nodeStart[A])) + P.pos[A];
        for (int B : cand) {
          cn.nodeTests++;
          const V3 rel = xw - P.pos[B];
          const double rB = P.radius[B];
          if (lengthSquared(rel) > rB * rB) continue;
          Contact c;
          if (!nodeVsSdf(P, B, xw, P.pos[B], P.q[B], c, cn)) continue;
          finishContact(c, k, hist, cn);
          out.push_back(c);
        }
      }
    });
  }

  void finishContact(Contact& c, int node, const History& hist, Counters& cn) const {
    c.master = node;
    auto it = hist.find(hkey(node, c.slave));
    c.spring = (it == hist.end()) ? v3(0., 0., 0.) : it->second;
    c.En = 0.;
    c.Es = 0.;
    cn.contacts++;
  }

  // Runs fn(A, out, counters) for every master particle A and leaves in `out` the contacts in
  // increasing A (the serial order).  With threads > 1 the masters are cut into contiguous
  // chunks taken by OpenMP threads; each chunk fills its own list and counters, concatenated in
  // chunk order afterwards, so lists and counters are the same for any thread count.
  int threads = 1;
  template <class Fn>
  void overMasters(std::vector<Contact>& out, Fn fn) {
    out.clear();
    const size_t n = P.n();
    if (threads <= 1 || n < 2) {
      for (size_t A = 0; A < n; ++A) fn(A, out, counters);
      return;
    }
    const int nChunks = int(std::min<size_t>(n, size_t(threads) * 16));
    std::vector<std::vector<Contact>> part(nChunks);
    std::vector<Counters> cnt(nChunks);
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
    for (int ch = 0; ch < nChunks; ++ch) {
      const size_t lo = n * size_t(ch) / size_t(nChunks), hi = n * size_t(ch + 1) / size_t(nChunks);
      for (size_t A = lo; A < hi; ++A) fn(A, part[ch], cnt[ch]);
    }
    for (int ch = 0; ch < nChunks; ++ch) {
      out.insert(out.end(), part[ch].begin(), part[ch].end());
      counters.pairs += cnt[ch].pairs;
      counters.nodeTests += cnt[ch].nodeTests;
      counters.sdfQueries += cnt[ch].sdfQueries;
      counters.contacts += cnt[ch].contacts;
    }
  }

  // Fixed-partner contacts (walls, ghosts): LSParticleContactDetectionKernel.cu:418-590.  No
  // A<B filter and no radius cull: any node inside the partner's grid box is tested.
  void buildFixed(const Bodies& F, const std::vector<V3>& fpos, const std::vector<Q4>& fq,
                  const std::vector<int>& fkey, const Grid& grid, const History& hist,
                  std::vector<Contact>& out) {
    const CellTable ct = buildCells(fkey, grid.ncell());
    overMasters(out, [&](size_t A, std::vector<Contact>& out, Counters& cn) {
      for (int k = P.nodeStart[A]; k < P.nodeStart[A + 1]; ++k) {
        const V3 xw = rotate(P.q[A], nodeLocal(P, int(A), k - P.nodeStart[A])) + P.pos[A];
        I3 c, lo, hi;
        grid.window(xw, c, lo, hi);
        for (int zz = lo.z; zz <= hi.z; zz++)
          for (int yy = lo.y; yy <= hi.y; yy++)
            for (int xx = lo.x; xx <= hi.x; xx++) {
              const int h = linearIndex3D(I3{c.x + xx, c.y + yy, c.z + zz}, grid.size);
              if (ct.start[h] == -1) continue;
              for (int s = ct.start[h]; s < ct.end[h]; ++s) {
                const int B = ct.sortedIdx[s];
                cn.nodeTests++;
                Contact cc;
                if (!nodeVsSdf(F, B, xw, fpos[B], fq[B], cc, cn)) continue;
                finishContact(cc, k, hist, cn);
                out.push_back(cc);
              }
            }
      }
    });
  }

  // Largest distance from a body's centre to a corner of its (scaled) grid box: no node outside
  // that sphere can lie inside the grid.
  double gridBoxRadius(const Bodies& B, int b) const {
    const Shape& s = shapes[B.shape[b]];
    double r2 = 0.;
    for (int k = 0; k < 8; ++k) {
      const V3 c = s.origin + s.spacing * v3((k & 1) ? double(s.nx - 1) : 0., (k & 2) ? double(s.ny - 1) : 0.,
                                             (k & 4) ? double(s.nz - 1) : 0.);
      r2 = std::max(r2, lengthSquared(c));
    }
    return std::sqrt(r2) * B.scale[b];
  }

  // The same contact set as buildFixed, from partner candidates per master: a fixed partner B
  // can hold a node of A only if A's bounding sphere meets the sphere around B's grid box.  Each
  // node then keeps the literal traversal's conditions (B's key in the node's clamped +-1 window,
  // the node inside B's grid) and the same node-vs-SDF arithmetic, so contacts are identical;
  // only their order within the list differs.
  void buildFixed_pairs(const Bodies& F, const std::vector<V3>& fpos, const std::vector<Q4>& fq,
                        const std::vector<int>& fkey, const Grid& grid, const History& hist,
                        std::vector<Contact>& out) {
    std::vector<double> R(F.n(), 0.);
    double rA = 0., rF = 0.;
    for (size_t A = 0; A < P.n(); ++A) rA = std::max(rA, P.radius[A]);
    for (size_t B = 0; B < F.n(); ++B) {
      if (fkey[B] < 0) continue;
      R[B] = gridBoxRadius(F, int(B));
      rF = std::max(rF, R[B]);
    }
    // hash of the partners with a key, on cells of the largest reach
    const double h = (rA + rF) * (1.0 + 1e-9) + 1e-12;
    auto cellKey = [&](int64_t x, int64_t y, int64_t z) {
      return ((x + (1ll << 20)) << 42) | ((y + (1ll << 20)) << 21) | (z + (1ll << 20));
    };
    auto cellOf = [&](V3 p, int64_t& x, int64_t& y, int64_t& z) {
      x = int64_t(std::floor(p.x / h)); y = int64_t(std::floor(p.y / h)); z = int64_t(std::floor(p.z / h));
    };
    std::unordered_map<int64_t, std::vector<int>> cells;
    for (size_t B = 0; B < F.n(); ++B) {
      if (fkey[B] < 0) continue;
      int64_t x, y, z;
      cellOf(fpos[B], x, y, z);
      cells[cellKey(x, y, z)].push_back(int(B));
    }
    overMasters(out, [&](size_t A, std::vector<Contact>& out, Counters& cn) {
      std::vector<int> cand;
      int64_t x, y, z;
      cellOf(P.pos[A], x, y, z);
      for (int64_t dz = -1; dz <= 1; ++dz)
        for (int64_t dy = -1; dy <= 1; ++dy)
          for (int64_t dx = -1; dx <= 1; ++dx) {
            auto it = cells.find(cellKey(x + dx, y + dy, z + dz));
            if (it == cells.end()) continue;
            for (int B : it->second) {
              const double rr = (P.radius[A] + R[B]) * (1.0 + 1e-9) + 1e-12;
              if (lengthSquared(P.pos[A] - fpos[B]) <= rr * rr) cand.push_back(B);
            }
          }
      std::sort(cand.begin(), cand.end());
      cn.pairs += int64_t(cand.size());
      if (cand.empty()) return;
      for (int k = P.nodeStart[A]; k < P.nodeStart[A + 1]; ++k) {
        const V3 xw = rotate(P.q[A], nodeLocal(P, int(A), k - P.nodeStart[A])) + P.pos[A];
        I3 c, lo, hi;
        grid.window(xw, c, lo, hi);
        for (int B : cand) {
          const I3 kc = grid.decode(fkey[B]);
          if (kc.x < c.x + lo.x || kc.x > c.x + hi.x || kc.y < c.y + lo.y || kc.y > c.y + hi.y ||
              kc.z < c.z + lo.z || kc.z > c.z + hi.z)
            continue;
          cn.nodeTests++;
          Contact cc;
          if (!nodeVsSdf(F, B, xw, fpos[B], fq[B], cc, cn)) continue;
          finishContact(cc, k, hist, cn);
          out.push_back(cc);
        }
      }
    });
  }

  // ---- contact law (contactKernel.cuh:4-77) ----
  static V3 integrateSlidingSpring(V3 s0, V3 sv, V3 n, V3 Fn, double mu, double k, double c, double dt) {
    V3 spring = v3(0., 0., 0.);
    if (mu > 0. && k > 0.) {
      V3 s1 = s0 - dot(s0, n) * n;
      const double a1 = length(s1);
      if (!isZero(a1)) s1 *= length(s0) / a1;
      spring = s1 + sv * dt;
      V3 Fs = -k * spring - c * sv;
      const double aF = length(Fs);
      const double aN = length(Fn);
      if (!isZero(aF) && aF > mu * aN) {
        double ratio = mu * aN / aF;
        Fs *= ratio;
        spring = -(Fs + c * sv) / k;
      }
    }
    return spring;
  }
  static void linearContact(V3& F, V3& spring, V3 vc, V3 n, double delta, double dt, double kn, double ks,
                            double mu, double res, double effM) {
    if (delta > 0.) {
      const double logRes = std::log(res);
      const double diss = -logRes / double(sqrtf(float(logRes * logRes + kPi * kPi)));
      const double cn = 2. * diss * double(sqrtf(float(effM * kn)));
      const V3 vn = dot(vc, n) * n;
      const V3 Fn = kn * delta * n - cn * vn;
      const double cs = 2. * diss * double(sqrtf(float(effM * ks)));
      const V3 vs = vc - vn;
      spring = integrateSlidingSpring(spring, vs, n, Fn, mu, ks, cs, dt);
      const V3 Fs = -ks * spring - cs * vs;
      F = Fn + Fs;
    } else {
      spring = v3(0., 0., 0.);
    }
  }

  std::vector<V3> posForce;  // positions used by the force kernels (post-wrap)

  // contactKernel.cu:46-118
  void forcesPP(std::vector<Contact>& l) {
    for (Contact& c : l) {
      const int i = nodeOwner(c.master), j = c.slave;
      const V3 ri = P.pos[i], rj = P.pos[j];
      const V3 vi = P.vel[i], wi = P.angvel[i], vj = P.vel[j], wj = P.angvel[j];
      const V3 rc = c.point, n = c.normal;
      const V3 vc = vi + cross(wi, rc - ri) - (vj + cross(wj, rc - rj));
      double kn = 0., ks = 0., effM = 0.;
      if (P.kn[i] > 0. && P.kn[j] > 0.) kn = P.kn[i] * P.kn[j] / (P.kn[i] + P.kn[j]);
      if (P.ks[i] > 0. && P.ks[j] > 0.) ks = P.ks[i] * P.ks[j] / (P.ks[i] + P.ks[j]);
      if (P.invM[i] > 0. || P.invM[j] > 0.) effM = 1. / (P.invM[i] + P.invM[j]);
      const double res = std::fmin(P.e[i], P.e[j]), mu = std::fmin(P.mu[i], P.mu[j]);
      V3 F = v3(0., 0., 0.);
      V3 s = c.spring;
      linearContact(F, s, vc, n, c.overlap, dt, kn, ks, mu, res, effM);
      c.spring = s;
      c.En = 0.5 * c.overlap * c.overlap * kn;
      c.Es = 0.5 * lengthSquared(s) * ks;
      P.force[i] += F;
      P.torque[i] += cross(rc - ri, F);
      P.force[j] += -F;
      P.torque[j] += cross(rc - rj, -F);
    }
  }
  // contactKernel.cu:120-191
  void forcesPW(std::vector<Contact>& l) {
    for (Contact& c : l) {
      const int i = nodeOwner(c.master), j = c.slave;
      const V3 ri = P.pos[i], rj = W.pos[j];
      const V3 vc = P.vel[i] + cross(P.angvel[i], c.point - ri) - (W.vel[j] + cross(W.angvel[j], c.point - rj));
      double effM = 0.;
      if (P.invM[i] > 0.) effM = 1. / P.invM[i];
      const double mu = std::fmin(P.mu[i], W.mu[j]), res = std::fmin(P.e[i], W.e[j]);
      V3 F = v3(0., 0., 0.);
      V3 s = c.spring;
      linearContact(F, s, vc, c.normal, c.overlap, dt, P.kn[i], P.ks[i], mu, res, effM);
      c.spring = s;
      c.En = 0.5 * c.overlap * c.overlap * P.kn[i];
      c.Es = 0.5 * lengthSquared(s) * P.ks[i];
      P.force[i] += F;
      P.torque[i] += cross(c.point - ri, F);
    }
  }
  // contactKernel.cu:193-271 (partner pose and velocity from the ghost, angular velocity from
  // the real particle; reaction skipped when i == j)
  void forcesGhost(std::vector<Contact>& l, const Ghost& G) {
    for (Contact& c : l) {
      const int i = nodeOwner(c.master), j = c.slave;
      const V3 ri = P.pos[i], rj = G.pos[j];
      const V3 vi = P.vel[i], wi = P.angvel[i], vj = G.vel[j], wj = P.angvel[j];
      const V3 rc = c.point;
      const V3 vc = vi + cross(wi, rc - ri) - (vj + cross(wj, rc - rj));
      double kn = 0., ks = 0., effM = 0.;
      if (P.kn[i] > 0. && P.kn[j] > 0.) kn = P.kn[i] * P.kn[j] / (P.kn[i] + P.kn[j]);
      if (P.ks[i] > 0. && P.ks[j] > 0.) ks = P.ks[i] * P.ks[j] / (P.ks[i] + P.ks[j]);
      if (P.invM[i] > 0. || P.invM[j] > 0.) effM = 1. / (P.invM[i] + P.invM[j]);
      const double res = std::fmin(P.e[i], P.e[j]), mu = std::fmin(P.mu[i], P.mu[j]);
      V3 F = v3(0., 0., 0.);
      V3 s = c.spring;
      linearContact(F, s, vc, c.normal, c.overlap, dt, kn, ks, mu, res, effM);
      c.spring = s;
      c.En = 0.5 * c.overlap * c.overlap * kn;
      c.Es = 0.5 * lengthSquared(s) * ks;
      P.force[i] += F;
      P.torque[i] += cross(rc - ri, F);
      if (i != j) {
        P.force[j] += -F;
        P.torque[j] += cross(rc - rj, -F);
      }
    }
  }

  // ---- bonds (contactKernel.cuh:138-199, contactKernel.cu:273-351) ----
  static int vbond(V3& F, V3& Mij, V3& Mji, Bond& b, V3 rbi, V3 rbj, V3 n1i, V3 n2i, V3 n3i, V3 n1j, V3 n2j,
                   V3 n3j) {
    const V3 rbij = rbi - rbj;
    const double D = length(rbij);
    if (isZero(D)) return 1;
    const V3 nij = normalize(rbij);
    const V3 e = -nij;
    const double D0 = b.L0;
    b.Un = 0.5 * b.B1 * (D - D0) * (D - D0);
    b.Us = b.B2 * (0.5 * dot((n1j - n1i), e) - 0.25 * dot(n1i, n1j) + 0.75);
    b.Ub = (0.25 * b.B2 + b.B3 + 0.5 * b.B4) * (dot(n1i, n1j) + 1.);
    b.Ut = -0.5 * b.B4 * (dot(n1i, n1j) + dot(n2i, n2j) + dot(n3i, n3j) - 1);
    F = b.B1 * (D - D0) * e + b.B2 / (2 * D) * ((n1j - n1i) - dot((n1j - n1i), e) * e);
    const V3 MTB = b.B3 * cross(n1j, n1i) - 0.5 * b.B4 * (cross(n2j, n2i) + cross(n3j, n3i));
    Mij = -0.5 * b.B2 * cross(e, n1i) + MTB;
    Mji = 0.5 * b.B2 * cross(e, n1j) - MTB;
    const double Fn = dot(F, nij);
    const V3 Fs = F - Fn * nij;
    const double Tt = dot(MTB, nij);
    const V3 Tb = MTB - Tt * nij;
    const double A = b.radius * b.radius * kPi;
    const double I = A * b.radius * b.radius / 4.;
    const double J = 2 * I;
    b.sigmaMax = -Fn / A + length(Tb) / I * b.radius;
    b.tauMax = length(Fs) / A + std::fabs(Tt) / J * b.radius;
    if (b.sigmaT > 0. && b.sigmaMax > b.sigmaT) return 0;
    else if (b.cohesion > 0. && b.tauMax > b.cohesion - b.mu * b.sigmaMax) return 0;
    return 1;
  }
  void forcesBonds() {
    for (Bond& b : bonds) {
      if (b.active == 0) {
        b.Un = b.Us = b.Ub = b.Ut = b.sigmaMax = b.tauMax = 0.;
        continue;
      }
      const int i = b.i, j = b.j;
      const Q4 qi = P.q[i], qj = P.q[j];
      const V3 n1i = rotate(qi, b.n1i), n2i = rotate(qi, b.n2i), n3i = rotate(qi, b.n3i);
      const V3 n1j = rotate(qj, b.n1j), n2j = rotate(qj, b.n2j), n3j = rotate(qj, b.n3j);
      const V3 ri = P.pos[i], rj = P.pos[j];
      const V3 rbi = rotate(qi, b.pi) + ri, rbj = rotate(qj, b.pj) + rj;
      b.point = 0.5 * (rbi + rbj);
      V3 F = v3(0, 0, 0), Mij = v3(0, 0, 0), Mji = v3(0, 0, 0);
      b.active = vbond(F, Mij, Mji, b, rbi, rbj, n1i, n2i, n3i, n1j, n2j, n3j);
      P.force[i] += F;
      P.torque[i] += Mij + cross(rbi - ri, F);
      P.force[j] += -F;
      P.torque[j] += Mji + cross(rbj - rj, -F);
    }
  }
  // VBondedInteraction::add (VBondedInteraction.h:74-195).  Returns false on rejected input.
  bool addBond(int i, int j, V3 point, V3 normal, double r, double L, double E, double nu, double sigmaT,
               double coh, double mu) {
    if (i < 0 || j < 0 || i >= int(P.n()) || j >= int(P.n())) return false;
    if (isZero(lengthSquared(normal)) || r <= 0. || L <= 0. || E <= 0. || nu <= -1.0 || nu >= 0.5) return false;
    Bond b{};
    const double G = E / (2. * (1. + nu));
    const double A = kPi * r * r;
    const double I = A * r * r / 4.;
    const double J = 2. * I;
    const double Kn = E * A / L, Kt = G * J / L, Ks = 12. * E * I / (L * L * L), Kb = E * I / L;
    b.B1 = Kn; b.B2 = L * L * Ks; b.B4 = Kt; b.B3 = Kb - b.B2 / 4. - b.B4 / 2.;
    const V3 n = normal / length(normal);
    const V3 gpi = point + 0.5 * L * n, gpj = point - 0.5 * L * n;
    b.pi = rotateInv(gpi - P.pos[i], P.q[i]);
    b.pj = rotateInv(gpj - P.pos[j], P.q[j]);
    const V3 N1i = -n;
    const V3 ref = (std::fabs(N1i.x) < 0.9) ? v3(1., 0., 0.) : v3(0., 1., 0.);
    const V3 N2i = normalize(cross(ref, N1i));
    const V3 N3i = normalize(cross(N1i, N2i));
    const V3 N1j = -N1i, N2j = N2i, N3j = N3i;
    // the rotated frames are stored as they come out of the rotation (VBondedInteraction.h:161-166)
    b.n1i = rotateInv(N1i, P.q[i]);
    b.n2i = rotateInv(N2i, P.q[i]);
    b.n3i = rotateInv(N3i, P.q[i]);
    b.n1j = rotateInv(N1j, P.q[j]);
    b.n2j = rotateInv(N2j, P.q[j]);
    b.n3j = rotateInv(N3j, P.q[j]);
    b.i = i; b.j = j; b.active = 1; b.point = point; b.radius = r; b.L0 = L;
    b.sigmaT = sigmaT; b.cohesion = coh; b.mu = mu;
    bonds.push_back(b);
    return true;
  }

  int nodeOwner(int node) const {
    // nodeStart is a prefix; binary search for the owning particle
    auto it = std::upper_bound(P.nodeStart.begin(), P.nodeStart.end(), node);
    return int(it - P.nodeStart.begin()) - 1;
  }

  // ---- periodic XY (mySpatialGrid.cu:28-56, 87-152, 214-240; PeriodicBoundary.h:215-289) ----
  void wrapX() {
    const double Lx = gridP.mx.x - gridP.mn.x;
    for (V3& p : P.pos) {
      if (p.x < gridP.mn.x) p += v3(Lx, 0., 0.);
      else if (p.x >= gridP.mx.x) p -= v3(Lx, 0., 0.);
    }
  }
  void wrapY() {
    const double Ly = gridP.mx.y - gridP.mn.y;
    for (V3& p : P.pos) {
      if (p.y < gridP.mn.y) p += v3(0., Ly, 0.);
      else if (p.y >= gridP.mx.y) p -= v3(0., Ly, 0.);
    }
  }
  int ghostKey(V3 p) const {
    I3 c = gridP.cellOf(p);
    bool out = false;
    if (p.x < gridP.mn.x) { c.x = 0; out = true; }
    else if (p.x >= gridP.mx.x) { c.x = gridP.size.x - 1; out = true; }
    if (p.y < gridP.mn.y) { c.y = 0; out = true; }
    else if (p.y >= gridP.mx.y) { c.y = gridP.size.y - 1; out = true; }
    const double zmax = (quirks & QUIRK_GHOST_Z_VS_MAXY) ? gridP.mx.y : gridP.mx.z;
    if (p.z < gridP.mn.z) { c.z = 0; out = true; }
    else if (p.z >= zmax) { c.z = gridP.size.z - 1; out = true; }
    return out ? linearIndex3D(c, gridP.size) : -1;
  }
  void makeGhost(int set) {  // set 0: +x, 1: +y, 2: +x+y
    const V3 ox = v3(gridP.mx.x - gridP.mn.x, 0., 0.), oy = v3(0., gridP.mx.y - gridP.mn.y, 0.);
    Ghost& G = ghost[set];
    G.pos.resize(P.n());
    G.key.resize(P.n());
    for (size_t i = 0; i < P.n(); ++i) {
      const V3 p = P.pos[i];
      G.pos[i] = p;
      const I3 c = gridP.cellOf(p);
      if (set == 0 && c.x == 0) G.pos[i] += ox;
      if (set == 1 && c.y == 0) G.pos[i] += oy;
      if (set == 2 && c.x == 0 && c.y == 0) G.pos[i] += ox + oy;
      G.key[i] = ghostKey(G.pos[i]);
    }
    G.q = P.q;
    G.vel = P.vel;
  }

  // ---- quarter-sector periodicity (mySpatialGrid.cu:58-85, 154-212) ----
  // Bodies that left the sector across x = min.x (with y >= min.y) turn by -90 degrees about the
  // corner's z axis, across y = min.y (with x >= min.x) by +90: position, velocity and orientation
  // (the angular velocity is left as it is).
  void wrapSector() {
    const V3 mn = gridP.mn;
    const V3 rm = v3(0., 0., -0.5 * kPi), rp = v3(0., 0., 0.5 * kPi);
    for (size_t i = 0; i < P.n(); ++i) {
      V3& p = P.pos[i];
      if (p.x < mn.x && p.y >= mn.y) {
        P.vel[i] = rotateVector(P.vel[i], rm);
        p = mn + rotateVector(p - mn, rm);
        P.q[i] = rotateQuaternion(P.q[i], rm);
      }
      if (p.y < mn.y && p.x >= mn.x) {
        P.vel[i] = rotateVector(P.vel[i], rp);
        p = mn + rotateVector(p - mn, rp);
        P.q[i] = rotateQuaternion(P.q[i], rp);
      }
    }
  }
  // Image sets: 0 turned by +90 degrees for bodies in the first y cell, 1 by 180 for bodies in the
  // first x and y cells, 2 by -90 for bodies in the first x cell; other copies keep the body's pose.
  void makeSectorGhosts() {
    const V3 mn = gridP.mn;
    const V3 ang[3] = {v3(0., 0., 0.5 * kPi), v3(0., 0., kPi), v3(0., 0., -0.5 * kPi)};
    for (int set = 0; set < 3; ++set) {
      Ghost& G = ghost[set];
      G.pos = P.pos;
      G.vel = P.vel;
      G.q = P.q;
      G.key.resize(P.n());
      for (size_t i = 0; i < P.n(); ++i) {
        const I3 c = gridP.cellOf(P.pos[i]);
        const bool on = set == 0 ? c.y == 0 : (set == 1 ? (c.x == 0 && c.y == 0) : c.x == 0);
        if (on) {
          G.vel[i] = rotateVector(P.vel[i], ang[set]);
          G.pos[i] = mn + rotateVector(P.pos[i] - mn, ang[set]);
          G.q[i] = rotateQuaternion(P.q[i], ang[set]);
        }
        G.key[i] = ghostKey(G.pos[i]);
      }
    }
  }

  // ---- detection: keys, particle-particle and wall lists, periodic wrap, ghost lists ----
  // The part of a step between the drift and the force kernels (LSSolver.h:1003-1004,
  // 1071-1074): contact lists rebuilt at the current poses, each carrying its springs over
  // from the list it replaces.
  void detect() {
    std::vector<int> keyP(P.n()), keyW(W.n());
    for (size_t i = 0; i < P.n(); ++i) keyP[i] = gridP.key(P.pos[i]);
    for (size_t i = 0; i < W.n(); ++i) keyW[i] = gridW.key(W.pos[i]);

    const bool periodic = perX || perY;
    // Contact lists built from pre-wrap positions (the wrap follows the main build).
    if ((periodic || perSector) && !(quirks & QUIRK_WRAP_AFTER_BUILD)) {
      if (perX) wrapX();
      if (perY) wrapY();
      if (perSector) wrapSector();
      for (size_t i = 0; i < P.n(); ++i) keyP[i] = gridP.key(P.pos[i]);
    }
    {
      History h = makeHistory(lists[LIST_PP]);
      if (literal) buildPP_literal(keyP, buildCells(keyP, gridP.ncell()), h, lists[LIST_PP]);
      else buildPP_pairs(keyP, h, lists[LIST_PP]);
    }
    {
      History h = makeHistory(lists[LIST_PW]);
      if (literal) buildFixed(W, W.pos, W.q, keyW, gridW, h, lists[LIST_PW]);
      else buildFixed_pairs(W, W.pos, W.q, keyW, gridW, h, lists[LIST_PW]);
    }
    if (periodic) {
      if (quirks & QUIRK_WRAP_AFTER_BUILD) {
        if (perX) { wrapX(); makeGhost(0); }
        if (perY) { wrapY(); makeGhost(1); }
      } else {
        if (perX) makeGhost(0);
        if (perY) makeGhost(1);
      }
      if (perX && perY) makeGhost(2);
      for (int s = 0; s < 3; ++s) {
        const bool on = (s == 0 && perX) || (s == 1 && perY) || (s == 2 && perX && perY);
        if (!on) { lists[LIST_GX + s].clear(); continue; }
        History h = makeHistory(lists[LIST_GX + s]);
        if (literal) buildFixed(P, ghost[s].pos, ghost[s].q, ghost[s].key, gridP, h, lists[LIST_GX + s]);
        else buildFixed_pairs(P, ghost[s].pos, ghost[s].q, ghost[s].key, gridP, h, lists[LIST_GX + s]);
      }
    }
    if (perSector) {
      if (quirks & QUIRK_WRAP_AFTER_BUILD) wrapSector();
      makeSectorGhosts();
      for (int s = 0; s < 3; ++s) {
        History h = makeHistory(lists[LIST_GX + s]);
        if (literal) buildFixed(P, ghost[s].pos, ghost[s].q, ghost[s].key, gridP, h, lists[LIST_GX + s]);
        else buildFixed_pairs(P, ghost[s].pos, ghost[s].q, ghost[s].key, gridP, h, lists[LIST_GX + s]);
      }
    }
  }

  // ---- the pre-loop build of each LSSolver::compute* call (LSSolver.h:996-998, 1029-1031,
  // 1061-1066, 1097-1104): detection at the current poses with no integration and no forces.
  // Its effects: particles outside a periodic cell are wrapped before the first kick, and the
  // first step finds its springs in lists built at the phase's starting poses.
  bool primePending = false;  // set by initialize: the first step is preceded by it
  void prime() {
    detect();
    primePending = false;
  }

  // ---- one step: LSSolver::compute* loop body (LSSolver.h:994-1132) ----
  void step() {
    if (primePending) prime();
    kick(0.5 * dt);
    drift(P, dt);
    drift(W, dt);
    detect();
    const bool periodic = perX || perY;
    for (size_t i = 0; i < P.n(); ++i) { P.force[i] = v3(0, 0, 0); P.torque[i] = v3(0, 0, 0); }
    forcesPP(lists[LIST_PP]);
    forcesPW(lists[LIST_PW]);
    if (periodic) {
      if (perX) forcesGhost(lists[LIST_GX], ghost[0]);
      if (perY) forcesGhost(lists[LIST_GY], ghost[1]);
      if (perX && perY) forcesGhost(lists[LIST_GXY], ghost[2]);
    }
    if (perSector)
      for (int s = 0; s < 3; ++s) forcesGhost(lists[LIST_GX + s], ghost[s]);
    if (!bonds.empty()) forcesBonds();
    if (external) external(time);
    kick(0.5 * dt);
    stepCount++;
    time += dt;
  }
};

// Mass, centroid and inertia from an SDF grid with a smoothed Heaviside (eps = 1.5 cells),
// LSSolver.h:51-163.  Loops in the reference's x-outer, z-inner order.
struct MassProps {
  double mass, invMass;
  V3 centroid;
  S6 inertia, invInertia;
};
inline MassProps massFromSdf(const double* sdf, int nx, int ny, int nz, V3 origin, double h, double density) {
  auto H = [](double phi, double eps) -> double {
    if (eps <= 0.0) return (phi > 0.0) ? 0.0 : 1.0;
    if (phi < -eps) return 1.0;
    if (phi > eps) return 0.0;
    const double x = -phi / eps;
    return 0.5 * (1.0 + x + std::sin(kPi * x) / kPi);
  };
  const I3 n{nx, ny, nz};
  MassProps m{};
  double mass = 0.;
  const double mg = density * h * h * h;
  for (int x = 0; x < nx; x++)
    for (int y = 0; y < ny; y++)
      for (int z = 0; z < nz; z++) mass += H(sdf[linearIndex3D(I3{x, y, z}, n)] / h, 1.5);
  mass *= mg;
  m.mass = mass;
  m.invMass = 0.;
  m.centroid = v3(0., 0., 0.);
  m.inertia = S6{0, 0, 0, 0, 0, 0};
  m.invInertia = S6{0, 0, 0, 0, 0, 0};
  if (mass > 0.) {
    m.invMass = 1. / mass;
    V3 c = v3(0., 0., 0.);
    for (int x = 0; x < nx; x++)
      for (int y = 0; y < ny; y++)
        for (int z = 0; z < nz; z++) {
          const double Hv = H(sdf[linearIndex3D(I3{x, y, z}, n)] / h, 1.5);
          c.x += Hv * (origin.x + double(x) * h);
          c.y += Hv * (origin.y + double(y) * h);
          c.z += Hv * (origin.z + double(z) * h);
        }
    c *= mg / mass;
    S6 I{0, 0, 0, 0, 0, 0};
    for (int x = 0; x < nx; x++)
      for (int y = 0; y < ny; y++)
        for (int z = 0; z < nz; z++) {
          const double Hv = H(sdf[linearIndex3D(I3{x, y, z}, n)] / h, 1.5);
          V3 r = origin + h * v3(double(x), double(y), double(z)) - c;
          I.xx += Hv * (r.y * r.y + r.z * r.z) * mg;
          I.yy += Hv * (r.x * r.x + r.z * r.z) * mg;
          I.zz += Hv * (r.y * r.y + r.x * r.x) * mg;
          I.xy -= Hv * r.x * r.y * mg;
          I.xz -= Hv * r.x * r.z * mg;
          I.yz -= Hv * r.y * r.z * mg;
        }
    m.centroid = c;
    m.inertia = I;
    m.invInertia = inverse(I);
  }
  return m;
}

}  // namespace lsdem_oracle
