// Host value types and vector helpers of the LSSolver drop-in: the names apps use from the
// reference's myVec.h / myQua.h / myMat.h (quaternion {q0 scalar, q1..q3}, symMatrix
// {xx, yy, zz, xy, xz, yz}, pi(), dot, cross, length, normalize).  double3 / int3 and their
// arithmetic operators are HIP's vector types.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

struct quaternion {
  double q0, q1, q2, q3;
};
inline quaternion make_quaternion(double q0, double q1, double q2, double q3) { return quaternion{q0, q1, q2, q3}; }

struct symMatrix {
  double xx, yy, zz, xy, xz, yz;
};
inline symMatrix make_symMatrix(double xx, double yy, double zz, double xy, double xz, double yz) {
  return symMatrix{xx, yy, zz, xy, xz, yz};
}

inline constexpr double pi() { return 3.14159265358979323846; }

inline double dot(const double3& a, const double3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double3 cross(const double3& a, const double3& b) {
  return make_double3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline double lengthSquared(const double3& v) { return dot(v, v); }
inline double length(const double3& v) { return std::sqrt(dot(v, v)); }
inline double3 normalize(const double3& v) {
  const double l2 = dot(v, v);
  if (l2 == 0.0) return make_double3(0.0, 0.0, 0.0);
  const double s = 1.0 / std::sqrt(l2);
  return make_double3(v.x * s, v.y * s, v.z * s);
}

inline quaternion normalize(const quaternion& q) {
  const double l2 = q.q0 * q.q0 + q.q1 * q.q1 + q.q2 * q.q2 + q.q3 * q.q3;
  if (l2 == 0.0) return make_quaternion(1.0, 0.0, 0.0, 0.0);
  const double s = 1.0 / std::sqrt(l2);
  return make_quaternion(q.q0 * s, q.q1 * s, q.q2 * s, q.q3 * s);
}

// v' = q v q*, q = (q0; q1, q2, q3) unit
inline double3 rotateVectorByQuaternion(const quaternion& q, const double3& v) {
  const double3 u = make_double3(q.q1, q.q2, q.q3);
  const double3 t = cross(u, v);
  const double3 t2 = make_double3(2.0 * t.x, 2.0 * t.y, 2.0 * t.z);
  const double3 c = cross(u, t2);
  return make_double3(v.x + q.q0 * t2.x + c.x, v.y + q.q0 * t2.y + c.y, v.z + q.q0 * t2.z + c.z);
}

// Axis-angle rotation (axis need not be unit).
inline quaternion quaternionFromAxisAngle(const double3& axis, double angle) {
  const double3 a = normalize(axis);
  const double s = std::sin(0.5 * angle);
  return make_quaternion(std::cos(0.5 * angle), a.x * s, a.y * s, a.z * s);
}

inline symMatrix inverse(const symMatrix& m) {
  const double a = m.xx, b = m.yy, c = m.zz, d = m.xy, e = m.xz, f = m.yz;
  const double det = a * (b * c - f * f) - d * (d * c - f * e) + e * (d * f - b * e);
  if (det == 0.0) return make_symMatrix(0, 0, 0, 0, 0, 0);
  const double inv = 1.0 / det;
  return make_symMatrix((b * c - f * f) * inv, (a * c - e * e) * inv, (a * b - d * d) * inv, (e * f - d * c) * inv,
                        (d * f - e * b) * inv, (d * e - a * f) * inv);
}

// x-fastest linear index of a grid node (the SDF layout, ix + nx (iy + ny iz)).
inline int linearIndex3D(const int3 p, const int3 size) { return p.x + size.x * (p.y + size.y * p.z); }
