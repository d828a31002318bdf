// FP64 vector, quaternion and symmetric-tensor math for the LS-DEM kernels.
//
// Every expression keeps the evaluation order of the reference's math headers
// (FunDEM/kernel/CUDAKernelFunction/myUtility/myVec.h, myQua.h, myMat.h) because the contact
// set is compared bit for bit with a host restatement.  The library is compiled with
// -ffp-contract=off so that no multiply-add is fused behind our back.  The reference's device
// branch uses rsqrt() inside normalize(); this library uses 1/sqrt() (its host branch), which
// is correctly rounded and therefore identical on host and device.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#define LSD_HD __host__ __device__ __forceinline__

namespace lsd {

struct d3 { double x, y, z; };
struct q4 { double q0, q1, q2, q3; };
struct s6 { double xx, yy, zz, xy, xz, yz; };
struct i3 { int x, y, z; };

constexpr double kPi = 3.14159265358979323846;

LSD_HD d3 mk3(double x, double y, double z) { return d3{x, y, z}; }
LSD_HD bool isZero(double x) { return fabs(x) <= 1.0e-20; }
LSD_HD d3 operator+(d3 a, d3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
LSD_HD d3 operator-(d3 a, d3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
LSD_HD d3 operator*(d3 a, double s) { return mk3(a.x * s, a.y * s, a.z * s); }
LSD_HD d3 operator*(double s, d3 a) { return mk3(a.x * s, a.y * s, a.z * s); }
LSD_HD d3 operator/(d3 a, double s) { double inv = 1.0 / s; return mk3(a.x * inv, a.y * inv, a.z * inv); }
LSD_HD d3& operator+=(d3& a, d3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
LSD_HD d3& operator-=(d3& a, d3 b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
LSD_HD d3& operator*=(d3& a, double s) { a.x *= s; a.y *= s; a.z *= s; return a; }
LSD_HD d3 operator-(d3 a) { return mk3(-a.x, -a.y, -a.z); }
LSD_HD double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
LSD_HD d3 cross(d3 a, d3 b) {
  return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
LSD_HD double len2(d3 v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
LSD_HD double len(d3 v) { return sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
LSD_HD d3 normalize(d3 v) {
  double l2 = v.x * v.x + v.y * v.y + v.z * v.z;
  if (isZero(l2)) return mk3(0.0, 0.0, 0.0);
  double il = 1.0 / sqrt(l2);
  return mk3(v.x * il, v.y * il, v.z * il);
}
LSD_HD int lin3(i3 c, i3 n) { return c.z * n.y * n.x + c.y * n.x + c.x; }

LSD_HD q4 qadd(q4 a, q4 b) { return q4{a.q0 + b.q0, a.q1 + b.q1, a.q2 + b.q2, a.q3 + b.q3}; }
LSD_HD q4 qnormalize(q4 q) {
  double l2 = q.q0 * q.q0 + q.q1 * q.q1 + q.q2 * q.q2 + q.q3 * q.q3;
  if (isZero(l2)) return q;
  double il = 1.0 / sqrt(l2);
  return q4{q.q0 * il, q.q1 * il, q.q2 * il, q.q3 * il};
}
// Explicit Euler on q' = 0.5 (0, w) (x) q with w in the world frame, then normalise.
LSD_HD q4 qintegrate(q4 q, d3 w, double dt) {
  d3 v = w * (0.5 * dt);
  q4 d;
  d.q0 = -q.q1 * v.x - q.q2 * v.y - q.q3 * v.z;
  d.q1 = q.q0 * v.x + q.q3 * v.y - q.q2 * v.z;
  d.q2 = -q.q3 * v.x + q.q0 * v.y + q.q1 * v.z;
  d.q3 = q.q2 * v.x - q.q1 * v.y + q.q0 * v.z;
  return qnormalize(qadd(q, d));
}
LSD_HD d3 rot(q4 q, d3 v) {  // v' = v + q0 t + qv x t, t = 2 qv x v
  d3 qv = mk3(q.q1, q.q2, q.q3);
  d3 t = cross(qv, v) * 2.0;
  return v + (t * q.q0) + cross(qv, t);
}
LSD_HD d3 rotInv(d3 v, q4 q) {
  d3 qv = mk3(-q.q1, -q.q2, -q.q3);
  d3 t = cross(qv, v) * 2.0;
  return v + (t * q.q0) + cross(qv, t);
}
// v turned about the axis of w by the rotation whose coefficients A = sin|w| / |w| and
// B = (1 - cos|w|) / |w|^2 the host computed: v + A (w x v) + B (w x (w x v)), the expression of
// rotateVector (myVec.h:143-184).
LSD_HD d3 rodrigues(d3 v, d3 w, double A, double B) {
  const d3 a = cross(w, v);
  const d3 b = cross(w, a);
  return v + A * a + B * b;
}
// r (x) q, normalised: the Hamilton product of rotateQuaternion (myQua.h:131-179), r the rotation's
// unit quaternion computed on the host.
LSD_HD q4 qrotate(q4 r, q4 q) {
  q4 o;
  o.q0 = r.q0 * q.q0 - r.q1 * q.q1 - r.q2 * q.q2 - r.q3 * q.q3;
  o.q1 = r.q0 * q.q1 + r.q1 * q.q0 + r.q2 * q.q3 - r.q3 * q.q2;
  o.q2 = r.q0 * q.q2 - r.q1 * q.q3 + r.q2 * q.q0 + r.q3 * q.q1;
  o.q3 = r.q0 * q.q3 + r.q1 * q.q2 - r.q2 * q.q1 + r.q3 * q.q0;
  return qnormalize(o);
}
LSD_HD d3 mul(s6 m, d3 v) {
  return mk3(m.xx * v.x + m.xy * v.y + m.xz * v.z, m.xy * v.x + m.yy * v.y + m.yz * v.z,
             m.xz * v.x + m.yz * v.y + m.zz * v.z);
}
// R (local inverse inertia) R^T with R from q.
LSD_HD s6 rotInvInertia(q4 q, s6 I) {
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
  return s6{r0x * a + r0y * b + r0z * c, r1x * d + r1y * e + r1z * f, r2x * g + r2y * h + r2z * i,
            r0x * d + r0y * e + r0z * f, r0x * g + r0y * h + r0z * i, r1x * g + r1y * h + r1z * i};
}

// Trilinear interpolation of 8 corner values p[0..7] ordered (i,j,k) = 000,100,010,110,001,101,011,111.
LSD_HD double triValue(double x, double y, double z, const double* p) {
  const double wx0 = 1.0 - x, wx1 = x, wy0 = 1.0 - y, wy1 = y, wz0 = 1.0 - z, wz1 = z;
  return p[0] * wx0 * wy0 * wz0 + p[1] * wx1 * wy0 * wz0 + p[2] * wx0 * wy1 * wz0 +
         p[3] * wx1 * wy1 * wz0 + p[4] * wx0 * wy0 * wz1 + p[5] * wx1 * wy0 * wz1 +
         p[6] * wx0 * wy1 * wz1 + p[7] * wx1 * wy1 * wz1;
}
// Gradient in grid units.  xw is the x weight actually used: the reference passes its x
// fraction through an `int` parameter, i.e. 0, unless that behaviour is switched off.
LSD_HD d3 triGradient(double xw, double y, double z, const double* p) {
  const double wx0 = 1.0 - xw, wx1 = xw, wy0 = 1.0 - y, wy1 = y, wz0 = 1.0 - z, wz1 = z;
  const double gx = (p[1] - p[0]) * wy0 * wz0 + (p[3] - p[2]) * wy1 * wz0 +
                    (p[5] - p[4]) * wy0 * wz1 + (p[7] - p[6]) * wy1 * wz1;
  const double gy = (p[2] - p[0]) * wx0 * wz0 + (p[3] - p[1]) * wx1 * wz0 +
                    (p[6] - p[4]) * wx0 * wz1 + (p[7] - p[5]) * wx1 * wz1;
  const double gz = (p[4] - p[0]) * wx0 * wy0 + (p[5] - p[1]) * wx1 * wy0 +
                    (p[6] - p[2]) * wx0 * wy1 + (p[7] - p[3]) * wx1 * wy1;
  return mk3(gx, gy, gz);
}

// Tangential spring update and Coulomb cap; linear spring-dashpot with restitution-derived
// damping evaluated through float square roots (as the reference does).
LSD_HD d3 slidingSpring(d3 s0, d3 sv, d3 n, d3 Fn, double mu, double k, double c, double dt) {
  d3 spring = mk3(0., 0., 0.);
  if (mu > 0. && k > 0.) {
    d3 s1 = s0 - dot(s0, n) * n;
    const double a1 = len(s1);
    if (!isZero(a1)) s1 *= len(s0) / a1;
    spring = s1 + sv * dt;
    d3 Fs = -k * spring - c * sv;
    const double aF = len(Fs);
    const double aN = len(Fn);
    if (!isZero(aF) && aF > mu * aN) {
      double ratio = mu * aN / aF;
      Fs *= ratio;
      spring = -(Fs + c * sv) / k;
    }
  }
  return spring;
}
// Force and updated tangential spring of a contact (returned by value: reference outputs kept the
// pair in scratch memory on the GPU).
struct ContactLaw {
  d3 F, spring;
};
LSD_HD ContactLaw linearContact(d3 spring, d3 vc, d3 n, double delta, double dt, double kn, double ks, double mu,
                                double res, double effM) {
  d3 F = mk3(0., 0., 0.);
  if (delta > 0.) {
    const double logRes = log(res);
    const double diss = -logRes / double(sqrtf(float(logRes * logRes + kPi * kPi)));
    const double cn = 2. * diss * double(sqrtf(float(effM * kn)));
    const d3 vn = dot(vc, n) * n;
    const d3 Fn = kn * delta * n - cn * vn;
    const double cs = 2. * diss * double(sqrtf(float(effM * ks)));
    const d3 vs = vc - vn;
    spring = slidingSpring(spring, vs, n, Fn, mu, ks, cs, dt);
    const d3 Fs = -ks * spring - cs * vs;
    F = Fn + Fs;
  } else {
    spring = mk3(0., 0., 0.);
  }
  return ContactLaw{F, spring};
}

}  // namespace lsd
