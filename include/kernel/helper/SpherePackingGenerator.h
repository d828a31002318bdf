// Sphere packings that FunDEM apps place particles at (FunDEM/kernel/helper/
// SpherePackingGenerator.h:15-555), written anew for this library's LSSolver drop-in.
//
// The random packings are random sequential addition: radii drawn uniform in [rMin, rMax] with
// std::mt19937(seed), sorted largest first, then each sphere tries up to maxAttemptsPerSphere
// uniform positions inside the container and keeps the first one that overlaps no placed sphere.
// The generator is consumed in the reference's order: all radii first, then per attempt the
// box's x, y, z, or the cylinder's radial u, angle v and axial w.  (The reference draws a box
// attempt's three coordinates inside one call's argument list, whose evaluation order C++ leaves
// to the compiler; here they are drawn x, y, z.)
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <functional>
#include <random>
#include <stdexcept>
#include <unordered_map>
#include <vector>

#include "../lsdem_types.h"

namespace SpherePacking {

struct Pack {
  std::vector<double3> centers;
  std::vector<double> radii;
};

namespace detail {

// Spheres already placed, bucketed on cubic cells of edge 2 rMax anchored at `origin`: a sphere
// can only overlap spheres whose centres lie in its own or the 26 neighbouring cells.
class PlacedSpheres {
 public:
  PlacedSpheres(Pack& pack, double3 origin, double cell) : pack_(pack), origin_(origin), cell_(cell) {}

  bool overlaps(const double3& c, double r) const {
    int x, y, z;
    cellOf(c, x, y, z);
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          auto it = cells_.find(key(x + dx, y + dy, z + dz));
          if (it == cells_.end()) continue;
          for (int id : it->second) {
            const double3 d = c - pack_.centers[size_t(id)];
            const double rr = r + pack_.radii[size_t(id)];
            if (dot(d, d) < rr * rr) return true;
          }
        }
    return false;
  }

  void add(const double3& c, double r) {
    int x, y, z;
    cellOf(c, x, y, z);
    cells_[key(x, y, z)].push_back(int(pack_.centers.size()));
    pack_.centers.push_back(c);
    pack_.radii.push_back(r);
  }

 private:
  void cellOf(const double3& p, int& x, int& y, int& z) const {
    x = int(std::floor((p.x - origin_.x) / cell_));
    y = int(std::floor((p.y - origin_.y) / cell_));
    z = int(std::floor((p.z - origin_.z) / cell_));
  }
  static int64_t key(int x, int y, int z) {
    const int64_t o = int64_t(1) << 20;
    return ((int64_t(x) + o) << 42) ^ ((int64_t(y) + o) << 21) ^ (int64_t(z) + o);
  }
  Pack& pack_;
  double3 origin_;
  double cell_;
  std::unordered_map<int64_t, std::vector<int>> cells_;
};

inline std::vector<double> radiiLargestFirst(std::mt19937& rng, int n, double rMin, double rMax) {
  std::uniform_real_distribution<double> R(rMin, rMax);
  std::vector<double> r(size_t(std::max(n, 0)));
  for (double& x : r) x = R(rng);
  std::sort(r.begin(), r.end(), std::greater<double>());
  return r;
}

inline bool insideBox(const double3& c, double r, const double3& lo, const double3& hi) {
  return c.x - r >= lo.x && c.x + r <= hi.x && c.y - r >= lo.y && c.y + r <= hi.y && c.z - r >= lo.z &&
         c.z + r <= hi.z;
}

}  // namespace detail

// Up to nSpheres non-overlapping spheres inside the box, largest first.
inline Pack buildNonOverlappingInBox_LargeFirst(const double3 boxMin, const double3 boxMax, const int nSpheres,
                                                const double rMin, const double rMax, const uint32_t seed = 12345u,
                                                const int maxAttemptsPerSphere = 2000) {
  Pack out;
  if (nSpheres <= 0 || rMin <= 0.0 || rMax < rMin) return out;
  const double3 lo = make_double3(std::min(boxMin.x, boxMax.x), std::min(boxMin.y, boxMax.y), std::min(boxMin.z, boxMax.z));
  const double3 hi = make_double3(std::max(boxMin.x, boxMax.x), std::max(boxMin.y, boxMax.y), std::max(boxMin.z, boxMax.z));
  const double3 ext = hi - lo;
  if (ext.x < 2.0 * rMin || ext.y < 2.0 * rMin || ext.z < 2.0 * rMin) return out;
  std::mt19937 rng(seed);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  const std::vector<double> radii = detail::radiiLargestFirst(rng, nSpheres, rMin, rMax);
  detail::PlacedSpheres placed(out, lo, 2.0 * rMax);
  out.centers.reserve(radii.size());
  out.radii.reserve(radii.size());
  for (double r : radii) {
    if (ext.x < 2.0 * r || ext.y < 2.0 * r || ext.z < 2.0 * r) continue;
    for (int a = 0; a < maxAttemptsPerSphere; ++a) {
      const double x = (lo.x + r) + U(rng) * ((hi.x - r) - (lo.x + r));
      const double y = (lo.y + r) + U(rng) * ((hi.y - r) - (lo.y + r));
      const double z = (lo.z + r) + U(rng) * ((hi.z - r) - (lo.z + r));
      const double3 c = make_double3(x, y, z);
      if (placed.overlaps(c, r)) continue;
      placed.add(c, r);
      break;
    }
  }
  return out;
}

// Up to nSpheres non-overlapping spheres inside the cylinder from A to B, largest first.
inline Pack buildNonOverlappingInCylinder_LargeFirst(const double3 cylinderPointA, const double3 cylinderPointB,
                                                     const double cylinderRadius, const int nSpheres,
                                                     const double rMin, const double rMax,
                                                     const uint32_t seed = 12345u,
                                                     const int maxAttemptsPerSphere = 2000) {
  Pack out;
  if (nSpheres <= 0 || rMin <= 0.0 || rMax < rMin || cylinderRadius <= 0.0) return out;
  const double3 axis = cylinderPointB - cylinderPointA;
  const double L = length(axis);
  if (L <= 0.0 || L < 2.0 * rMin || cylinderRadius < rMin) return out;
  const double3 dir = axis / L;
  const double3 ref = std::fabs(dir.x) < 0.9 ? make_double3(1.0, 0.0, 0.0) : make_double3(0.0, 1.0, 0.0);
  const double3 u = normalize(cross(dir, ref));
  const double3 w = normalize(cross(dir, u));
  std::mt19937 rng(seed);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  const std::vector<double> radii = detail::radiiLargestFirst(rng, nSpheres, rMin, rMax);
  const double3 boxLo = make_double3(std::min(cylinderPointA.x, cylinderPointB.x) - cylinderRadius,
                                     std::min(cylinderPointA.y, cylinderPointB.y) - cylinderRadius,
                                     std::min(cylinderPointA.z, cylinderPointB.z) - cylinderRadius);
  detail::PlacedSpheres placed(out, boxLo, 2.0 * rMax);
  auto fits = [&](const double3& c, double r) {
    const double3 rel = c - cylinderPointA;
    const double t = dot(rel, dir);
    if (t < r || t > L - r) return false;
    const double3 radial = c - (cylinderPointA + t * dir);
    const double room = cylinderRadius - r;
    return room >= 0.0 && dot(radial, radial) <= room * room;
  };
  out.centers.reserve(radii.size());
  out.radii.reserve(radii.size());
  for (double r : radii) {
    if (L < 2.0 * r || cylinderRadius < r) continue;
    for (int a = 0; a < maxAttemptsPerSphere; ++a) {
      // uniform over the disc of radius R - r (radius sqrt(u) scaled), uniform along [r, L - r]
      const double ru = U(rng), av = U(rng);
      const double rad = (cylinderRadius - r) * std::sqrt(ru), ang = 2.0 * pi() * av;
      const double t = r + U(rng) * ((L - r) - r);
      const double3 c = cylinderPointA + t * dir + (rad * std::cos(ang)) * u + (rad * std::sin(ang)) * w;
      if (!fits(c, r) || placed.overlaps(c, r)) continue;
      placed.add(c, r);
      break;
    }
  }
  return out;
}

// Simple cubic lattice of spheres of radius r, pitch 2r, starting r inside the box's lower corner;
// only sites with the sphere inside the box (to 1e-12).
inline Pack buildRegularInBox(const double3 boxMin, const double3 boxMax, const double sphereRadius) {
  if (sphereRadius <= 0.0) throw std::runtime_error("sphereRadius must be positive.");
  if (boxMax.x <= boxMin.x || boxMax.y <= boxMin.y || boxMax.z <= boxMin.z)
    throw std::runtime_error("Invalid box: boxMax must be larger than boxMin.");
  Pack p;
  const double r = sphereRadius, eps = 1e-12, step = 2.0 * r;
  const double3 lo = boxMin + make_double3(r, r, r), hi = boxMax - make_double3(r, r, r);
  if (lo.x > hi.x || lo.y > hi.y || lo.z > hi.z) return p;
  for (double z = lo.z; z <= hi.z + eps; z += step)
    for (double y = lo.y; y <= hi.y + eps; y += step)
      for (double x = lo.x; x <= hi.x + eps; x += step) {
        p.centers.push_back(make_double3(x, y, z));
        p.radii.push_back(r);
      }
  return p;
}

// Staggered close packing: rows 2r apart in x, sqrt(3) r in y, layers sqrt(8/3) r in z; odd rows
// shifted r in x, odd layers r in x and sqrt(3) r / 3 in y; only spheres inside the box.
inline Pack buildHexInBox(const double3 boxMin, const double3 boxMax, const double sphereRadius) {
  if (sphereRadius <= 0.0) throw std::runtime_error("sphereRadius must be positive.");
  if (boxMax.x <= boxMin.x || boxMax.y <= boxMin.y || boxMax.z <= boxMin.z)
    throw std::runtime_error("Invalid box: boxMax must be larger than boxMin.");
  Pack p;
  const double r = sphereRadius, eps = 1e-12;
  const double3 lo = boxMin + make_double3(r, r, r), hi = boxMax - make_double3(r, r, r);
  if (lo.x > hi.x || lo.y > hi.y || lo.z > hi.z) return p;
  const double dx = 2.0 * r, dy = std::sqrt(3.0) * r, dz = std::sqrt(8.0 / 3.0) * r;
  int layer = 0;
  for (double z = lo.z; z <= hi.z + eps; z += dz, ++layer) {
    const bool oddLayer = layer % 2 == 1;
    int row = 0;
    for (double y = lo.y + (oddLayer ? std::sqrt(3.0) * r / 3.0 : 0.0); y <= hi.y + eps; y += dy, ++row) {
      double x0 = lo.x;
      if (row % 2 == 1) x0 += r;
      if (oddLayer) x0 += r;
      for (double x = x0; x <= hi.x + eps; x += dx) {
        const double3 c = make_double3(x, y, z);
        if (detail::insideBox(c, r, boxMin, boxMax)) {
          p.centers.push_back(c);
          p.radii.push_back(r);
        }
      }
    }
  }
  return p;
}

}  // namespace SpherePacking
