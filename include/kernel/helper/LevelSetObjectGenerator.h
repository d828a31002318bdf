// Level-set shape generators of the LSSolver drop-in: the classes FunDEM apps use to build a
// particle's or a wall's signed-distance grid and its surface mesh
// (FunDEM/kernel/helper/LevelSetObjectGenerator.h:56-2136), written anew for this library.
//
// Same public interface and results as the reference:
//  - a grid node holds value / |grad value| of the shape's implicit function, on a grid padded by
//    `padding` cells and snapped to multiples of the spacing (:122-224); negative inside unless the
//    shape is "positive inside" (walls);
//  - the surface mesh's vertices are what apps pass as boundary nodes, so the meshes are built in
//    the reference's vertex order (icosahedron subdivisions with a midpoint cache, box corners,
//    cylinder rings);
//  - TriangleMeshParticle's distance is the exact distance to the mesh (closest point through a
//    bounding-volume hierarchy), its sign from the winding number (sum of solid angles).
// Host code only; the grids go to LSSolver::addLSParticle / addWall.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <fstream>
#include <limits>
#include <random>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../lsdem_types.h"

namespace LevelSetObject {

namespace detail {
inline double3 vmin(const double3& a, const double3& b) {
  return make_double3(std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z));
}
inline double3 vmax(const double3& a, const double3& b) {
  return make_double3(std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z));
}
inline double comp(const double3& v, int axis) { return axis == 0 ? v.x : (axis == 1 ? v.y : v.z); }
// |v|^e with |0|^e = 0 for e > 0
inline double apow(double v, double e) { return std::pow(std::fabs(v), e); }
inline double sgn1(double v) { return v >= 0.0 ? 1.0 : -1.0; }

// Closest point of triangle (a, b, c) to p: the Voronoi-region walk of Ericson, "Real-Time
// Collision Detection" (2005), section 5.1.5.
inline double3 closestOnTriangle(const double3& p, const double3& a, const double3& b, const double3& c) {
  const double3 ab = b - a, ac = c - a, ap = p - a;
  const double d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0 && d2 <= 0.0) return a;
  const double3 bp = p - b;
  const double d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0 && d4 <= d3) return b;
  const double vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) return a + (d1 / (d1 - d3)) * ab;
  const double3 cp = p - c;
  const double d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0 && d5 <= d6) return c;
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) return a + (d2 / (d2 - d6)) * ac;
  const double va = d3 * d6 - d5 * d4;
  if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) return b + ((d4 - d3) / ((d4 - d3) + (d5 - d6))) * (c - b);
  const double den = 1.0 / (va + vb + vc);
  return a + (vb * den) * ab + (vc * den) * ac;
}

// Signed solid angle of triangle (a, b, c) seen from p (Van Oosterom and Strackee, IEEE Trans.
// Biomed. Eng. 30 (1983) 125-126); 0 when p sits on a vertex.
inline double solidAngle(const double3& p, const double3& a, const double3& b, const double3& c) {
  const double3 r0 = a - p, r1 = b - p, r2 = c - p;
  const double l0 = length(r0), l1 = length(r1), l2 = length(r2);
  if (l0 < 1e-30 || l1 < 1e-30 || l2 < 1e-30) return 0.0;
  const double num = dot(r0, cross(r1, r2));
  const double den = l0 * l1 * l2 + dot(r0, r1) * l2 + dot(r1, r2) * l0 + dot(r2, r0) * l1;
  return 2.0 * std::atan2(num, den);
}

// Icosahedron on the unit sphere's golden-rectangle corners (0, +-1, +-phi) and its cyclic
// permutations, in the reference's vertex and face order.
inline void icosahedron(std::vector<double3>& v, std::vector<int3>& f) {
  const double t = 0.5 * (1.0 + std::sqrt(5.0));
  v = {make_double3(-1, t, 0), make_double3(1, t, 0),   make_double3(-1, -t, 0), make_double3(1, -t, 0),
       make_double3(0, -1, t), make_double3(0, 1, t),   make_double3(0, -1, -t), make_double3(0, 1, -t),
       make_double3(t, 0, -1), make_double3(t, 0, 1),   make_double3(-t, 0, -1), make_double3(-t, 0, 1)};
  f = {make_int3(0, 11, 5), make_int3(0, 5, 1),  make_int3(0, 1, 7),   make_int3(0, 7, 10), make_int3(0, 10, 11),
       make_int3(1, 5, 9),  make_int3(5, 11, 4), make_int3(11, 10, 2), make_int3(10, 7, 6), make_int3(7, 1, 8),
       make_int3(3, 9, 4),  make_int3(3, 4, 2),  make_int3(3, 2, 6),   make_int3(3, 6, 8),  make_int3(3, 8, 9),
       make_int3(4, 9, 5),  make_int3(2, 4, 11), make_int3(6, 2, 10),  make_int3(8, 6, 7),  make_int3(9, 8, 1)};
}

// One midpoint subdivision: every face becomes four; the midpoint of an edge is created once (at
// the first face that meets the edge, edges taken in the order (0,1), (1,2), (2,0)) and mapped by
// `place` onto the surface.
template <class Place>
void subdivide(std::vector<double3>& v, std::vector<int3>& f, Place place) {
  std::unordered_map<uint64_t, int> mid;
  mid.reserve(f.size() * 2);
  auto midpoint = [&](int a, int b) {
    const uint64_t key = (uint64_t(uint32_t(std::min(a, b))) << 32) | uint32_t(std::max(a, b));
    auto it = mid.find(key);
    if (it != mid.end()) return it->second;
    const int id = int(v.size());
    const double3 m = 0.5 * (v[size_t(std::min(a, b))] + v[size_t(std::max(a, b))]);
    v.push_back(place(m));
    mid.emplace(key, id);
    return id;
  };
  std::vector<int3> out;
  out.reserve(f.size() * 4);
  for (const int3& t : f) {
    const int m01 = midpoint(t.x, t.y), m12 = midpoint(t.y, t.z), m20 = midpoint(t.z, t.x);
    out.push_back(make_int3(t.x, m01, m20));
    out.push_back(make_int3(t.y, m12, m01));
    out.push_back(make_int3(t.z, m20, m12));
    out.push_back(make_int3(m01, m12, m20));
  }
  f.swap(out);
}

// Faces of a star-shaped mesh around the origin turned to face outward.
inline void faceOutward(const std::vector<double3>& v, std::vector<int3>& f) {
  for (int3& t : f) {
    const double3 a = v[size_t(t.x)], b = v[size_t(t.y)], c = v[size_t(t.z)];
    if (dot(cross(b - a, c - a), (a + b + c) / 3.0) < 0.0) std::swap(t.y, t.z);
  }
}
}  // namespace detail

// Signed-distance grid: x fastest, value < 0 inside unless the object is positive inside.
class GridInfo {
 public:
  double3 gridOrigin{0.0, 0.0, 0.0};
  int3 gridNodeSize{0, 0, 0};
  double gridNodeSpacing{0.0};
  std::vector<double> gridNodeSignedDistance;
};

class Base {
 protected:
  GridInfo grid_;

 public:
  virtual ~Base() = default;

  void clearGrid() { grid_ = GridInfo{}; }
  bool isGridBuilt() const { return !grid_.gridNodeSignedDistance.empty(); }

  virtual bool isValid() const = 0;
  virtual void evaluateImplicitFunctionValueAndGradient(double& value, double3& gradient,
                                                        const double3& point) const = 0;
  virtual double3 boundingBoxMin() const = 0;
  virtual double3 boundingBoxMax() const = 0;

  // Grid of spacing h over the bounding box padded by max(1, padding) cells, its corners snapped
  // outward to multiples of h; each node holds value / |gradient| (or +-h/2 where the gradient
  // vanishes).
  void buildGrid(const double h, const int padding = 2) {
    clearGrid();
    if (!isValid() || !(h > 0.0)) return;
    const double pad = double(std::max(1, padding)) * h;
    double3 lo = boundingBoxMin() - make_double3(pad, pad, pad);
    double3 hi = boundingBoxMax() + make_double3(pad, pad, pad);
    lo = make_double3(h * std::floor(lo.x / h), h * std::floor(lo.y / h), h * std::floor(lo.z / h));
    hi = make_double3(h * std::ceil(hi.x / h), h * std::ceil(hi.y / h), h * std::ceil(hi.z / h));
    const int nx = int(std::floor((hi.x - lo.x) / h + 0.5)) + 1;
    const int ny = int(std::floor((hi.y - lo.y) / h + 0.5)) + 1;
    const int nz = int(std::floor((hi.z - lo.z) / h + 0.5)) + 1;
    if (nx <= 0 || ny <= 0 || nz <= 0) return;
    grid_.gridNodeSpacing = h;
    grid_.gridOrigin = lo;
    grid_.gridNodeSize = make_int3(nx, ny, nz);
    grid_.gridNodeSignedDistance.assign(size_t(nx) * size_t(ny) * size_t(nz), 0.0);
    for (int k = 0; k < nz; ++k) {
      const double z = lo.z + double(k) * h;
      for (int j = 0; j < ny; ++j) {
        const double y = lo.y + double(j) * h;
        for (int i = 0; i < nx; ++i) {
          const double x = lo.x + double(i) * h;
          double f = 0.0;
          double3 g = make_double3(0.0, 0.0, 0.0);
          evaluateImplicitFunctionValueAndGradient(f, g, make_double3(x, y, z));
          const double gn = length(g);
          grid_.gridNodeSignedDistance[size_t(i) + size_t(nx) * (size_t(j) + size_t(ny) * size_t(k))] =
              gn > 1e-12 ? f / gn : (f <= 0.0 ? -0.5 : 0.5) * h;
        }
      }
    }
  }

  // Spacing = the bounding box's largest edge / nodesPerDiameter.
  void buildGridByResolution(const int nodesPerDiameter = 50, const int padding = 2) {
    if (!isValid() || nodesPerDiameter <= 0) return;
    const double3 e = boundingBoxMax() - boundingBoxMin();
    const double d = std::max(e.x, std::max(e.y, e.z));
    if (!(d > 0.0)) return;
    buildGrid(d / double(nodesPerDiameter), padding);
  }

  // The grid as VTK vertex cells with point data signedDistance and gridIJK, in
  // <prefix>LSGrid.vtu (ASCII).
  void outputGridVTU(const std::string& prefix) const {
    if (!isGridBuilt()) return;
    const int nx = grid_.gridNodeSize.x, ny = grid_.gridNodeSize.y, nz = grid_.gridNodeSize.z;
    const size_t n = size_t(nx) * size_t(ny) * size_t(nz);
    if (grid_.gridNodeSignedDistance.size() != n) return;
    const std::string name = prefix + "LSGrid.vtu";
    std::ofstream o(name);
    if (!o) throw std::runtime_error("Cannot open " + name);
    const double h = grid_.gridNodeSpacing;
    o << "<?xml version=\"1.0\"?>\n"
      << "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
      << "  <UnstructuredGrid>\n    <Piece NumberOfPoints=\"" << n << "\" NumberOfCells=\"" << n << "\">\n"
      << "      <Points>\n        <DataArray type=\"Float32\" NumberOfComponents=\"3\" format=\"ascii\">\n";
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
          o << "          " << float(grid_.gridOrigin.x + i * h) << ' ' << float(grid_.gridOrigin.y + j * h) << ' '
            << float(grid_.gridOrigin.z + k * h) << '\n';
    o << "        </DataArray>\n      </Points>\n      <Cells>\n"
      << "        <DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\">\n";
    for (size_t c = 0; c < n; ++c) o << "          " << c << '\n';
    o << "        </DataArray>\n        <DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n";
    for (size_t c = 0; c < n; ++c) o << "          " << (c + 1) << '\n';
    o << "        </DataArray>\n        <DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n";
    for (size_t c = 0; c < n; ++c) o << "          1\n";
    o << "        </DataArray>\n      </Cells>\n      <PointData Scalars=\"signedDistance\">\n"
      << "        <DataArray type=\"Float32\" Name=\"signedDistance\" format=\"ascii\">\n";
    for (size_t c = 0; c < n; ++c) o << "          " << float(grid_.gridNodeSignedDistance[c]) << '\n';
    o << "        </DataArray>\n"
      << "        <DataArray type=\"Int32\" Name=\"gridIJK\" NumberOfComponents=\"3\" format=\"ascii\">\n";
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) o << "          " << i << ' ' << j << ' ' << k << '\n';
    o << "        </DataArray>\n      </PointData>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n";
  }

  const GridInfo& gridInfo() const { return grid_; }
  GridInfo& gridInfo() { return grid_; }
};

// A closed triangle mesh: exact distance to the mesh, negative inside (winding number > 1/2).
class TriangleMeshParticle : public Base {
 protected:
  std::vector<double3> vertexPosition_;
  std::vector<int3> triangleVertexIndex_;
  bool levelSetPositiveInside_{false};
  double3 boxMin_{0.0, 0.0, 0.0}, boxMax_{0.0, 0.0, 0.0};

  // Bounding-volume hierarchy over the triangles: nodes split at the median centroid along the
  // longest axis of their centroid bounds; leaves hold at most kLeaf triangles.
  struct Node {
    double3 lo, hi;
    int left = -1, right = -1;  // children (inner node)
    int first = 0, count = 0;   // triangle range in tri_ (leaf when count > 0)
  };
  static constexpr int kLeaf = 4;
  std::vector<Node> nodes_;
  std::vector<int> tri_;

  void refreshBounds() {
    if (vertexPosition_.empty()) {
      boxMin_ = boxMax_ = make_double3(0.0, 0.0, 0.0);
      return;
    }
    boxMin_ = boxMax_ = vertexPosition_[0];
    for (const double3& p : vertexPosition_) {
      boxMin_ = detail::vmin(boxMin_, p);
      boxMax_ = detail::vmax(boxMax_, p);
    }
  }

  void buildHierarchy() {
    nodes_.clear();
    tri_.clear();
    const size_t nt = triangleVertexIndex_.size();
    if (nt == 0 || !TriangleMeshParticle::isValid()) return;
    std::vector<double3> cen(nt);
    for (size_t t = 0; t < nt; ++t) {
      const int3 f = triangleVertexIndex_[t];
      cen[t] = (vertexPosition_[size_t(f.x)] + vertexPosition_[size_t(f.y)] + vertexPosition_[size_t(f.z)]) / 3.0;
    }
    tri_.resize(nt);
    for (size_t t = 0; t < nt; ++t) tri_[t] = int(t);
    nodes_.reserve(2 * nt / kLeaf + 2);
    // explicit stack of (node, first, count)
    std::vector<std::array<int, 3>> work;
    nodes_.push_back(Node{});
    work.push_back({0, 0, int(nt)});
    while (!work.empty()) {
      const auto w = work.back();
      work.pop_back();
      const int id = w[0], first = w[1], count = w[2];
      double3 lo = make_double3(1e300, 1e300, 1e300), hi = -lo, clo = lo, chi = hi;
      for (int k = first; k < first + count; ++k) {
        const int3 f = triangleVertexIndex_[size_t(tri_[size_t(k)])];
        for (int c : {f.x, f.y, f.z}) {
          lo = detail::vmin(lo, vertexPosition_[size_t(c)]);
          hi = detail::vmax(hi, vertexPosition_[size_t(c)]);
        }
        clo = detail::vmin(clo, cen[size_t(tri_[size_t(k)])]);
        chi = detail::vmax(chi, cen[size_t(tri_[size_t(k)])]);
      }
      nodes_[size_t(id)].lo = lo;
      nodes_[size_t(id)].hi = hi;
      const double3 ext = chi - clo;
      const int axis = (ext.x >= ext.y && ext.x >= ext.z) ? 0 : (ext.y >= ext.z ? 1 : 2);
      if (count <= kLeaf || detail::comp(ext, axis) <= 0.0) {
        nodes_[size_t(id)].first = first;
        nodes_[size_t(id)].count = count;
        continue;
      }
      const int half = count / 2;
      std::nth_element(tri_.begin() + first, tri_.begin() + first + half, tri_.begin() + first + count,
                       [&](int a, int b) { return detail::comp(cen[size_t(a)], axis) < detail::comp(cen[size_t(b)], axis); });
      const int l = int(nodes_.size());
      nodes_.push_back(Node{});
      nodes_.push_back(Node{});
      nodes_[size_t(id)].left = l;
      nodes_[size_t(id)].right = l + 1;
      work.push_back({l, first, half});
      work.push_back({l + 1, first + half, count - half});
    }
  }

  static double boxDistance2(const Node& n, const double3& p) {
    const double dx = std::max(std::max(n.lo.x - p.x, 0.0), p.x - n.hi.x);
    const double dy = std::max(std::max(n.lo.y - p.y, 0.0), p.y - n.hi.y);
    const double dz = std::max(std::max(n.lo.z - p.z, 0.0), p.z - n.hi.z);
    return dx * dx + dy * dy + dz * dz;
  }

  // Closest point of the mesh to p (nearer child first, subtrees farther than the best pruned).
  double3 closestPoint(const double3& p) const {
    double best = std::numeric_limits<double>::max();
    double3 bestPt = p;
    if (nodes_.empty()) return bestPt;
    std::vector<int> stack;
    stack.reserve(64);
    stack.push_back(0);
    while (!stack.empty()) {
      const Node& n = nodes_[size_t(stack.back())];
      stack.pop_back();
      if (boxDistance2(n, p) > best) continue;
      if (n.count > 0) {
        for (int k = n.first; k < n.first + n.count; ++k) {
          const int3 f = triangleVertexIndex_[size_t(tri_[size_t(k)])];
          const double3 c = detail::closestOnTriangle(p, vertexPosition_[size_t(f.x)], vertexPosition_[size_t(f.y)],
                                                      vertexPosition_[size_t(f.z)]);
          const double d2 = lengthSquared(p - c);
          if (d2 < best) {
            best = d2;
            bestPt = c;
          }
        }
        continue;
      }
      const double dl = boxDistance2(nodes_[size_t(n.left)], p), dr = boxDistance2(nodes_[size_t(n.right)], p);
      if (dl <= dr) {
        stack.push_back(n.right);
        stack.push_back(n.left);
      } else {
        stack.push_back(n.left);
        stack.push_back(n.right);
      }
    }
    return bestPt;
  }

  // Winding number of the closed mesh around p above one half.
  bool inside(const double3& p) const {
    double w = 0.0;
    for (const int3& f : triangleVertexIndex_)
      w += detail::solidAngle(p, vertexPosition_[size_t(f.x)], vertexPosition_[size_t(f.y)], vertexPosition_[size_t(f.z)]);
    return std::fabs(w / (4.0 * pi())) > 0.5;
  }

  void setMeshInternal(const std::vector<double3>& v, const std::vector<int3>& f) {
    vertexPosition_ = v;
    triangleVertexIndex_ = f;
    refreshBounds();
    buildHierarchy();
    clearGrid();
  }

  static void flipIfPositiveInside(bool positiveInside, double& value, double3& gradient) {
    if (positiveInside) {
      value = -value;
      gradient = -gradient;
    }
  }

 public:
  TriangleMeshParticle() = default;
  TriangleMeshParticle(const std::vector<double3>& v, const std::vector<int3>& f) { setMeshInternal(v, f); }

  void setMesh(const std::vector<double3>& v, const std::vector<int3>& f) { setMeshInternal(v, f); }
  void setLevelSetPositiveInside(bool on) { levelSetPositiveInside_ = on; }
  void clearMesh() {
    vertexPosition_.clear();
    triangleVertexIndex_.clear();
    nodes_.clear();
    tri_.clear();
    refreshBounds();
    clearGrid();
  }

  bool isValid() const override {
    if (vertexPosition_.empty() || triangleVertexIndex_.empty()) return false;
    const int n = int(vertexPosition_.size());
    for (const int3& t : triangleVertexIndex_) {
      if (t.x < 0 || t.y < 0 || t.z < 0 || t.x >= n || t.y >= n || t.z >= n) return false;
      if (t.x == t.y || t.y == t.z || t.z == t.x) return false;
    }
    return true;
  }

  void evaluateImplicitFunctionValueAndGradient(double& value, double3& gradient, const double3& p) const override {
    if (!isValid()) {
      value = 1.0;
      gradient = make_double3(0.0, 0.0, 0.0);
      return;
    }
    const double3 c = closestPoint(p);
    const double d = length(p - c);
    const bool in = inside(p);
    value = in ? -d : d;
    gradient = d > 1e-30 ? (in ? -1.0 : 1.0) * ((p - c) / d) : make_double3(0.0, 0.0, 0.0);
    flipIfPositiveInside(levelSetPositiveInside_, value, gradient);
  }

  double3 boundingBoxMin() const override { return boxMin_; }
  double3 boundingBoxMax() const override { return boxMax_; }

  const std::vector<double3>& vertexPosition() const { return vertexPosition_; }
  std::vector<double3>& vertexPosition() { return vertexPosition_; }
  const std::vector<int3>& triangleVertexIndex() const { return triangleVertexIndex_; }
  std::vector<int3>& triangleVertexIndex() { return triangleVertexIndex_; }
  const double3& meshBoundingBoxMin() const { return boxMin_; }
  const double3& meshBoundingBoxMax() const { return boxMax_; }
};

// Superellipsoid (|x/rx|^(2/ee) + |y/ry|^(2/ee))^(ee/en) + |z/rz|^(2/en) = 1; mesh: an icosahedron
// subdivided `subdivisionLevel` times (0..8) with its vertices projected radially onto the surface.
class Superellipsoid : public TriangleMeshParticle {
 protected:
  double rx_{1.0}, ry_{1.0}, rz_{1.0}, ee_{1.0}, en_{1.0};
  int level_{4};

  bool paramsValid() const { return rx_ > 0.0 && ry_ > 0.0 && rz_ > 0.0 && ee_ > 0.0 && en_ > 0.0 && level_ >= 0; }

  // The surface point in direction d: d / |d| scaled by F(d / |d|)^(-en / 2).
  double3 radial(const double3& d) const {
    const double l2 = lengthSquared(d);
    if (l2 <= 1e-30) return make_double3(0.0, 0.0, 0.0);
    const double3 u = d / std::sqrt(l2);
    const double s = detail::apow(detail::apow(u.x / rx_, 2.0 / ee_) + detail::apow(u.y / ry_, 2.0 / ee_), ee_ / en_) +
                     detail::apow(u.z / rz_, 2.0 / en_);
    if (s <= 1e-30) return make_double3(0.0, 0.0, 0.0);
    return u * std::pow(s, -0.5 * en_);
  }

  void rebuildMesh() {
    clearMesh();
    if (!paramsValid()) return;
    std::vector<double3> v;
    std::vector<int3> f;
    detail::icosahedron(v, f);
    for (double3& p : v) p = normalize(p);
    detail::faceOutward(v, f);
    for (int l = 0; l < level_; ++l) {
      detail::subdivide(v, f, [](const double3& m) { return normalize(m); });
      detail::faceOutward(v, f);
    }
    for (double3& p : v) p = radial(p);
    detail::faceOutward(v, f);
    setMeshInternal(v, f);
  }

 public:
  Superellipsoid() = default;
  Superellipsoid(double rx, double ry, double rz, double ee, double en, int subdivisionLevel = 4)
      : rx_(rx), ry_(ry), rz_(rz), ee_(ee), en_(en), level_(std::max(0, std::min(subdivisionLevel, 8))) {
    rebuildMesh();
  }

  void setParams(double rx, double ry, double rz, double ee, double en, int subdivisionLevel = 4) {
    rx_ = rx;
    ry_ = ry;
    rz_ = rz;
    ee_ = ee;
    en_ = en;
    level_ = std::max(0, std::min(subdivisionLevel, 8));
    rebuildMesh();
  }
  void setSubdivisionLevel(int subdivisionLevel) {
    level_ = std::max(0, std::min(subdivisionLevel, 8));
    rebuildMesh();
  }

  bool isValid() const override { return paramsValid() && TriangleMeshParticle::isValid(); }

  void evaluateImplicitFunctionValueAndGradient(double& value, double3& g, const double3& p) const override {
    if (!paramsValid()) {
      value = 1.0;
      g = make_double3(0.0, 0.0, 0.0);
      return;
    }
    // (reciprocal radii multiplied in, and the products grouped, as the reference does, so that
    // grids come out the same to the last bit)
    const double ix = 1.0 / rx_, iy = 1.0 / ry_, iz = 1.0 / rz_;
    const double ax = std::fabs(p.x * ix), ay = std::fabs(p.y * iy), az = std::fabs(p.z * iz);
    const double pxy = 2.0 / ee_, pz = 2.0 / en_, q = ee_ / en_;
    const double sxy = detail::apow(ax, pxy) + detail::apow(ay, pxy);
    value = detail::apow(sxy, q) + detail::apow(az, pz) - 1.0;
    const double tiny = 1e-30;
    const double outer = q * detail::apow(std::max(sxy, tiny), q - 1.0);
    g = make_double3(outer * (pxy * detail::apow(std::max(ax, tiny), pxy - 1.0) * detail::sgn1(p.x) * ix),
                     outer * (pxy * detail::apow(std::max(ay, tiny), pxy - 1.0) * detail::sgn1(p.y) * iy),
                     pz * detail::apow(std::max(az, tiny), pz - 1.0) * detail::sgn1(p.z) * iz);
    flipIfPositiveInside(levelSetPositiveInside_, value, g);
  }

  double3 boundingBoxMin() const override { return make_double3(-rx_, -ry_, -rz_); }
  double3 boundingBoxMax() const override { return make_double3(rx_, ry_, rz_); }

  double rx() const { return rx_; }
  double ry() const { return ry_; }
  double rz() const { return rz_; }
  double ee() const { return ee_; }
  double en() const { return en_; }
};

// Sphere |p| = radius; mesh: the icosahedron subdivided until it has at least
// targetSurfacePointCount vertices (10 4^L + 2), midpoints pushed out to the sphere.
class Sphere : public TriangleMeshParticle {
 protected:
  double radius_{0.0};
  size_t target_{0};

  void rebuildMesh() {
    if (radius_ <= 0.0 || target_ < 4) {
      clearMesh();
      return;
    }
    int level = 0;
    if (target_ > 12) {
      level = 12;
      for (int l = 0; l <= 12; ++l)
        if (size_t(10.0 * std::pow(4.0, l) + 2.0) >= target_) {
          level = l;
          break;
        }
    }
    const double r = radius_;
    auto onSphere = [r](const double3& p) {
      const double l = length(p);
      return l < 1e-30 ? make_double3(0.0, 0.0, 0.0) : (r / l) * p;
    };
    std::vector<double3> v;
    std::vector<int3> f;
    detail::icosahedron(v, f);
    for (double3& p : v) p = onSphere(p);
    for (int l = 0; l < level; ++l) detail::subdivide(v, f, onSphere);
    setMeshInternal(v, f);
  }

 public:
  Sphere() = default;
  Sphere(double radius, size_t targetSurfacePointCount) : radius_(radius), target_(targetSurfacePointCount) {
    rebuildMesh();
  }

  void setParams(double radius, size_t targetSurfacePointCount) {
    radius_ = radius;
    target_ = targetSurfacePointCount;
    rebuildMesh();
  }
  void clearSphere() {
    radius_ = 0.0;
    target_ = 0;
    levelSetPositiveInside_ = false;
    clearMesh();
  }

  bool isValid() const override { return radius_ > 0.0 && target_ >= 4 && TriangleMeshParticle::isValid(); }

  void evaluateImplicitFunctionValueAndGradient(double& value, double3& g, const double3& p) const override {
    if (radius_ <= 0.0) {
      value = 1.0;
      g = make_double3(0.0, 0.0, 0.0);
      return;
    }
    const double d = length(p);
    value = d - radius_;
    // (at the centre the gradient is left as passed in)
    if (d > 1e-30) g = p / d;
    flipIfPositiveInside(levelSetPositiveInside_, value, g);
  }

  double3 boundingBoxMin() const override { return make_double3(-radius_, -radius_, -radius_); }
  double3 boundingBoxMax() const override { return make_double3(radius_, radius_, radius_); }
  double radius() const { return radius_; }
};

// Box of edges lx, ly, lz centred at the origin, positive inside (a container wall).
class BoxWall : public TriangleMeshParticle {
 protected:
  double lx_{1.0}, ly_{1.0}, lz_{1.0};

  void rebuildMesh() {
    std::vector<double3> v;
    std::vector<int3> f;
    if (lx_ > 0.0 && ly_ > 0.0 && lz_ > 0.0) {
      const double x = 0.5 * lx_, y = 0.5 * ly_, z = 0.5 * lz_;
      v = {make_double3(-x, -y, -z), make_double3(x, -y, -z), make_double3(x, y, -z), make_double3(-x, y, -z),
           make_double3(-x, -y, z),  make_double3(x, -y, z),  make_double3(x, y, z),  make_double3(-x, y, z)};
      f = {make_int3(0, 2, 1), make_int3(0, 3, 2), make_int3(4, 5, 6), make_int3(4, 6, 7),
           make_int3(0, 7, 3), make_int3(0, 4, 7), make_int3(1, 2, 6), make_int3(1, 6, 5),
           make_int3(0, 1, 5), make_int3(0, 5, 4), make_int3(3, 7, 6), make_int3(3, 6, 2)};
    }
    setMeshInternal(v, f);
  }

 public:
  BoxWall() { levelSetPositiveInside_ = true; }
  BoxWall(double lx, double ly, double lz) : lx_(lx), ly_(ly), lz_(lz) {
    levelSetPositiveInside_ = true;
    rebuildMesh();
  }

  void setParams(double lx, double ly, double lz) {
    lx_ = lx;
    ly_ = ly;
    lz_ = lz;
    rebuildMesh();
    clearGrid();
  }

  bool isValid() const override { return lx_ > 0.0 && ly_ > 0.0 && lz_ > 0.0; }

  // Exact box distance: |max(q, 0)| + min(max_i q_i, 0), q = |p| - half edges; the gradient is the
  // outward direction of the nearest face (inside) or of the nearest surface point (outside).
  void evaluateImplicitFunctionValueAndGradient(double& value, double3& g, const double3& p) const override {
    if (!isValid()) {
      value = 1.0;
      g = make_double3(0.0, 0.0, 0.0);
      return;
    }
    const double3 h = make_double3(0.5 * lx_, 0.5 * ly_, 0.5 * lz_);
    const double3 q = make_double3(std::fabs(p.x) - h.x, std::fabs(p.y) - h.y, std::fabs(p.z) - h.z);
    const double3 qo = make_double3(std::max(q.x, 0.0), std::max(q.y, 0.0), std::max(q.z, 0.0));
    const double out = length(qo);
    value = out + std::min(std::max(q.x, std::max(q.y, q.z)), 0.0);
    if (out > 1e-30) {
      g = qo / out;
      if (p.x < 0.0 && q.x > 0.0) g.x = -g.x;
      if (p.y < 0.0 && q.y > 0.0) g.y = -g.y;
      if (p.z < 0.0 && q.z > 0.0) g.z = -g.z;
    } else {
      const double fx = h.x - std::fabs(p.x), fy = h.y - std::fabs(p.y), fz = h.z - std::fabs(p.z);
      if (fx <= fy && fx <= fz) g = make_double3(p.x >= 0.0 ? 1.0 : -1.0, 0.0, 0.0);
      else if (fy <= fx && fy <= fz) g = make_double3(0.0, p.y >= 0.0 ? 1.0 : -1.0, 0.0);
      else g = make_double3(0.0, 0.0, p.z >= 0.0 ? 1.0 : -1.0);
    }
    flipIfPositiveInside(levelSetPositiveInside_, value, g);
  }

  double3 boundingBoxMin() const override { return make_double3(-0.5 * lx_, -0.5 * ly_, -0.5 * lz_); }
  double3 boundingBoxMax() const override { return make_double3(0.5 * lx_, 0.5 * ly_, 0.5 * lz_); }
  double lx() const { return lx_; }
  double ly() const { return ly_; }
  double lz() const { return lz_; }
};

// Closed cylinder between pointA and pointB with the given radius, positive inside (a drum).
// Mesh: rings of numSegmentsCircumference vertices at numSegmentsAxial + 1 stations, then the two
// cap centres.  Implicit function, as the reference's (LevelSetObjectGenerator.h:2047-2111): inside,
// the nearer of mantle and caps; in the caps' slab but outside the mantle, the radial distance;
// outside both, the distance to the rim; the gradient of a cap region is the axis vector B - A
// itself (so the grid holds the axial distance over |B - A| there).
class CylinderWall : public TriangleMeshParticle {
 protected:
  double3 a_{0.0, 0.0, -0.5}, b_{0.0, 0.0, 0.5};
  double radius_{1.0};
  int nCirc_{256}, nAxial_{1};

  void rebuildMesh() {
    std::vector<double3> v;
    std::vector<int3> f;
    if (!isValid()) {
      setMeshInternal(v, f);
      return;
    }
    const double3 axis = b_ - a_;
    const double3 dir = axis / length(axis);
    const double3 ref = std::fabs(dir.x) < 0.9 ? make_double3(1.0, 0.0, 0.0) : make_double3(0.0, 1.0, 0.0);
    const double3 u = normalize(cross(dir, ref));
    const double3 w = normalize(cross(dir, u));
    for (int s = 0; s <= nAxial_; ++s) {
      const double3 c = a_ + (double(s) / double(nAxial_)) * axis;
      for (int k = 0; k < nCirc_; ++k) {
        const double th = 2.0 * pi() * double(k) / double(nCirc_);
        v.push_back(c + radius_ * (std::cos(th) * u + std::sin(th) * w));
      }
    }
    auto at = [&](int s, int k) { return s * nCirc_ + ((k % nCirc_) + nCirc_) % nCirc_; };
    for (int s = 0; s < nAxial_; ++s)
      for (int k = 0; k < nCirc_; ++k) {
        f.push_back(make_int3(at(s, k), at(s + 1, k), at(s + 1, k + 1)));
        f.push_back(make_int3(at(s, k), at(s + 1, k + 1), at(s, k + 1)));
      }
    const int ca = int(v.size());
    v.push_back(a_);
    const int cb = int(v.size());
    v.push_back(b_);
    for (int k = 0; k < nCirc_; ++k) f.push_back(make_int3(ca, at(0, k + 1), at(0, k)));
    for (int k = 0; k < nCirc_; ++k) f.push_back(make_int3(cb, at(nAxial_, k), at(nAxial_, k + 1)));
    setMeshInternal(v, f);
  }

 public:
  CylinderWall() { levelSetPositiveInside_ = true; }
  CylinderWall(const double3& pointA, const double3& pointB, double radius, int numSegmentsCircumference = 256,
               int numSegmentsAxial = 1)
      : a_(pointA), b_(pointB), radius_(radius), nCirc_(std::max(3, numSegmentsCircumference)),
        nAxial_(std::max(1, numSegmentsAxial)) {
    levelSetPositiveInside_ = true;
    rebuildMesh();
  }

  void setParams(const double3& pointA, const double3& pointB, double radius, int numSegmentsCircumference = 256,
                 int numSegmentsAxial = 1) {
    a_ = pointA;
    b_ = pointB;
    radius_ = radius;
    nCirc_ = std::max(3, numSegmentsCircumference);
    nAxial_ = std::max(1, numSegmentsAxial);
    rebuildMesh();
    clearGrid();
  }
  void setMeshResolution(int numSegmentsCircumference, int numSegmentsAxial) {
    nCirc_ = std::max(3, numSegmentsCircumference);
    nAxial_ = std::max(1, numSegmentsAxial);
    rebuildMesh();
    clearGrid();
  }

  bool isValid() const override { return radius_ > 0.0 && length(b_ - a_) > 1e-30; }

  void evaluateImplicitFunctionValueAndGradient(double& value, double3& g, const double3& p) const override {
    if (!isValid()) {
      value = 1.0;
      g = make_double3(0.0, 0.0, 0.0);
      return;
    }
    const double3 axis = b_ - a_;
    const double L = length(axis);
    const double3 dir = axis / L;
    const double3 rel = p - a_;
    const double t = dot(rel, dir);
    const double3 onAxis = a_ + std::min(std::max(t, 0.0), L) * dir;
    const bool radialIn = length(cross(rel, dir)) <= radius_;
    const bool axialIn = t >= 0.0 && t <= L;
    if (radialIn && axialIn) {
      const double3 rv = p - onAxis;
      const double dr = length(rv) - radius_;
      const double da = t < 0.5 * L ? -t : t - L;
      if (std::fabs(dr) <= std::fabs(da)) {
        value = dr;
        g = normalize(rv);
      } else {
        value = da;
        g = t < 0.5 * L ? -axis : axis;
      }
    } else if (axialIn) {  // beside the mantle
      const double3 rv = p - onAxis;
      value = length(rv) - radius_;
      g = normalize(rv);
    } else if (!radialIn) {  // beyond a cap and outside the mantle: the rim
      const double3 rim = onAxis + normalize(rel - t * dir) * radius_;
      value = length(p - rim);
      g = normalize(p - rim);
    } else {  // beyond a cap, within the radius
      value = t < 0.0 ? -t : t - L;
      g = t < 0.0 ? -axis : axis;
    }
    flipIfPositiveInside(levelSetPositiveInside_, value, g);
  }

  double3 boundingBoxMin() const override {
    return make_double3(std::min(a_.x, b_.x) - radius_, std::min(a_.y, b_.y) - radius_, std::min(a_.z, b_.z) - radius_);
  }
  double3 boundingBoxMax() const override {
    return make_double3(std::max(a_.x, b_.x) + radius_, std::max(a_.y, b_.y) + radius_, std::max(a_.z, b_.z) + radius_);
  }
  const double3& pointA() const { return a_; }
  const double3& pointB() const { return b_; }
  double radius() const { return radius_; }
};

}  // namespace LevelSetObject
