// Host checks of the drop-in's shape, packing and OBJ helpers (include/kernel/helper/*.h), built
// and run by tests/test_helpers.py.  Every check pins a result against an independent
// formulation: closed-form distances, brute force over all triangles, counts by formula.
#include <cstdio>
#include <cstdlib>
#include <fstream>

#include "kernel/helper/LevelSetObjectGenerator.h"
#include "kernel/helper/OBJLoader.h"
#include "kernel/helper/SpherePackingGenerator.h"

static int failures = 0;
#define CHECK(cond, ...)                    \
  do {                                      \
    if (!(cond)) {                          \
      std::printf("FAIL %s: ", #cond);      \
      std::printf(__VA_ARGS__);             \
      std::printf("\n");                    \
      ++failures;                           \
    }                                       \
  } while (0)

static double gridAt(const LevelSetObject::GridInfo& g, int i, int j, int k) {
  return g.gridNodeSignedDistance[size_t(i) + size_t(g.gridNodeSize.x) * (size_t(j) + size_t(g.gridNodeSize.y) * k)];
}

// brute-force distance to a triangle mesh: closest point over every triangle, sampled barycentric
// minimum refined by the exact edge/vertex/face cases through the helper's own routine would not
// be independent, so the check uses a dense barycentric sampling bound instead
static double bruteDistance(const std::vector<double3>& v, const std::vector<int3>& f, const double3& p, int samples) {
  double best = 1e300;
  for (const int3& t : f)
    for (int a = 0; a <= samples; ++a)
      for (int b = 0; a + b <= samples; ++b) {
        const double u = double(a) / samples, w = double(b) / samples;
        const double3 q = (1.0 - u - w) * v[t.x] + u * v[t.y] + w * v[t.z];
        best = std::min(best, length(p - q));
      }
  return best;
}

int main(int argc, char** argv) {
  // Sphere: the grid holds |p| - r exactly; 10 4^L + 2 vertices on the sphere
  {
    LevelSetObject::Sphere s;
    s.setParams(0.5, 600);
    CHECK(s.vertexPosition().size() == 642, "vertices %zu", s.vertexPosition().size());
    for (const double3& p : s.vertexPosition()) CHECK(std::fabs(length(p) - 0.5) < 1e-14, "vertex radius");
    s.buildGridByResolution(20);
    const auto& g = s.gridInfo();
    CHECK(std::fabs(g.gridNodeSpacing - 0.05) < 1e-15, "spacing %g", g.gridNodeSpacing);
    double worst = 0.0;
    for (int k = 0; k < g.gridNodeSize.z; ++k)
      for (int j = 0; j < g.gridNodeSize.y; ++j)
        for (int i = 0; i < g.gridNodeSize.x; ++i) {
          const double3 p = g.gridOrigin + g.gridNodeSpacing * make_double3(i, j, k);
          if (length(p) < 1e-12) continue;
          worst = std::max(worst, std::fabs(gridAt(g, i, j, k) - (length(p) - 0.5)));
        }
    CHECK(worst < 1e-14, "sphere grid err %g", worst);
    // padding: 2 cells beyond the bounding box on every side, snapped to the spacing
    CHECK(std::fabs(g.gridOrigin.x + 0.6) < 1e-12 && g.gridNodeSize.x == 25, "origin %g size %d", g.gridOrigin.x,
          g.gridNodeSize.x);
  }
  // Superellipsoid with ee = en = 1 is the ellipsoid: vertices on the surface; the grid's sign
  // and zero set match the closed form; 2562 vertices at level 4
  {
    LevelSetObject::Superellipsoid e;
    e.setParams(0.5, 0.4, 0.3, 1.0, 1.0);
    CHECK(e.vertexPosition().size() == 2562, "vertices %zu", e.vertexPosition().size());
    for (const double3& p : e.vertexPosition()) {
      const double f = p.x * p.x / 0.25 + p.y * p.y / 0.16 + p.z * p.z / 0.09 - 1.0;
      CHECK(std::fabs(f) < 1e-12, "vertex off the surface %g", f);
    }
    // outward faces: every face normal points away from the centre
    for (const int3& t : e.triangleVertexIndex()) {
      const double3 a = e.vertexPosition()[t.x], b = e.vertexPosition()[t.y], c = e.vertexPosition()[t.z];
      CHECK(dot(cross(b - a, c - a), a + b + c) > 0.0, "inward face");
    }
    e.buildGridByResolution(40);
    const auto& g = e.gridInfo();
    int bad = 0;
    for (int k = 0; k < g.gridNodeSize.z; ++k)
      for (int j = 0; j < g.gridNodeSize.y; ++j)
        for (int i = 0; i < g.gridNodeSize.x; ++i) {
          const double3 p = g.gridOrigin + g.gridNodeSpacing * make_double3(i, j, k);
          const double f = p.x * p.x / 0.25 + p.y * p.y / 0.16 + p.z * p.z / 0.09 - 1.0;
          const double grad = 2.0 * std::sqrt(p.x * p.x / 0.0625 + p.y * p.y / 0.0256 + p.z * p.z / 0.0081);
          if (grad < 1e-9) continue;
          if (std::fabs(gridAt(g, i, j, k) - f / grad) > 1e-12 * std::max(1.0, std::fabs(f / grad))) ++bad;
        }
    CHECK(bad == 0, "%d ellipsoid grid nodes differ from f/|grad f|", bad);
  }
  // Triangle mesh: BVH distance equals brute force, sign from the winding number
  {
    LevelSetObject::Sphere s;
    s.setParams(0.4, 162);
    LevelSetObject::TriangleMeshParticle m;
    m.setMesh(s.vertexPosition(), s.triangleVertexIndex());
    const double3 pts[] = {make_double3(0.0, 0.0, 0.0), make_double3(0.1, -0.2, 0.05), make_double3(0.7, 0.1, -0.3),
                           make_double3(-0.39, 0.0, 0.0), make_double3(0.2, 0.5, 0.6)};
    for (const double3& p : pts) {
      double f;
      double3 gr;
      m.evaluateImplicitFunctionValueAndGradient(f, gr, p);
      const double d = bruteDistance(m.vertexPosition(), m.triangleVertexIndex(), p, 60);
      const bool in = length(p) < 0.38;  // inside the inscribed sphere of the icosphere
      const bool out = length(p) > 0.41;
      if (in) CHECK(f < 0.0, "point inside has value %g", f);
      if (out) CHECK(f > 0.0, "point outside has value %g", f);
      CHECK(std::fabs(std::fabs(f) - d) < 2e-3 && std::fabs(f) <= d + 1e-12, "|f| %g vs brute %g", std::fabs(f), d);
      CHECK(std::fabs(length(gr) - 1.0) < 1e-12, "gradient not unit");
    }
    m.setLevelSetPositiveInside(true);
    double f;
    double3 gr;
    m.evaluateImplicitFunctionValueAndGradient(f, gr, make_double3(0.0, 0.0, 0.0));
    CHECK(f > 0.0, "positive-inside flips the sign");
  }
  // Box wall: positive inside, distance to the nearest face
  {
    LevelSetObject::BoxWall b;
    b.setParams(2.0, 4.0, 6.0);
    CHECK(b.vertexPosition().size() == 8 && b.triangleVertexIndex().size() == 12, "box mesh");
    double f;
    double3 gr;
    b.evaluateImplicitFunctionValueAndGradient(f, gr, make_double3(0.2, 0.0, 0.0));
    CHECK(std::fabs(f - 0.8) < 1e-15, "inside value %g", f);
    b.evaluateImplicitFunctionValueAndGradient(f, gr, make_double3(2.0, 3.0, 0.0));
    CHECK(std::fabs(f + std::sqrt(2.0)) < 1e-15, "outside corner value %g", f);
  }
  // Cylinder wall (axis x from -1 to 1, radius 0.5): positive inside, radial distance near the
  // mantle; rings of 16 vertices plus two cap centres
  {
    LevelSetObject::CylinderWall c;
    c.setParams(make_double3(-1.0, 0.0, 0.0), make_double3(1.0, 0.0, 0.0), 0.5, 16, 2);
    CHECK(c.vertexPosition().size() == 3 * 16 + 2, "cylinder vertices %zu", c.vertexPosition().size());
    double f;
    double3 gr;
    c.evaluateImplicitFunctionValueAndGradient(f, gr, make_double3(0.0, 0.3, 0.0));
    CHECK(std::fabs(f - 0.2) < 1e-15, "mantle value %g", f);
    c.evaluateImplicitFunctionValueAndGradient(f, gr, make_double3(0.0, 0.8, 0.0));
    CHECK(std::fabs(f + 0.3) < 1e-15, "outside value %g", f);
    c.buildGridByResolution(30);
    CHECK(c.isGridBuilt(), "cylinder grid");
  }
  // Packings: inside the container, no overlaps, largest first
  {
    auto p = SpherePacking::buildNonOverlappingInBox_LargeFirst(make_double3(0, 0, 0), make_double3(1, 1, 2), 400, 0.05,
                                                                0.1, 7u);
    CHECK(p.centers.size() > 100, "box pack has %zu spheres", p.centers.size());
    for (size_t i = 0; i < p.centers.size(); ++i) {
      const double3 c = p.centers[i];
      const double r = p.radii[i];
      CHECK(c.x - r >= 0 && c.x + r <= 1 && c.y - r >= 0 && c.y + r <= 1 && c.z - r >= 0 && c.z + r <= 2, "outside");
      if (i) CHECK(p.radii[i] <= p.radii[i - 1], "not largest first");
      for (size_t j = 0; j < i; ++j) CHECK(length(c - p.centers[j]) >= r + p.radii[j], "overlap %zu %zu", i, j);
    }
    auto q = SpherePacking::buildNonOverlappingInCylinder_LargeFirst(make_double3(-1, 0, 0), make_double3(1, 0, 0), 0.6,
                                                                     300, 0.05, 0.1, 3u);
    CHECK(q.centers.size() > 50, "cylinder pack has %zu spheres", q.centers.size());
    for (size_t i = 0; i < q.centers.size(); ++i) {
      const double3 c = q.centers[i];
      const double r = q.radii[i];
      CHECK(c.x - r >= -1 - 1e-12 && c.x + r <= 1 + 1e-12 && std::sqrt(c.y * c.y + c.z * c.z) + r <= 0.6 + 1e-12,
            "outside the cylinder");
      for (size_t j = 0; j < i; ++j) CHECK(length(c - q.centers[j]) >= r + q.radii[j], "overlap");
    }
    auto reg = SpherePacking::buildRegularInBox(make_double3(0, 0, 0), make_double3(1, 1, 3), 0.1);
    CHECK(reg.centers.size() == 5 * 5 * 15, "regular pack %zu", reg.centers.size());
    auto hex = SpherePacking::buildHexInBox(make_double3(0, 0, 0), make_double3(1, 1, 1), 0.1);
    for (size_t i = 0; i < hex.centers.size(); ++i)
      for (size_t j = 0; j < i; ++j) CHECK(length(hex.centers[i] - hex.centers[j]) >= 0.2 - 1e-12, "hex overlap");
    CHECK(hex.centers.size() > 80, "hex pack %zu", hex.centers.size());
  }
  // OBJ: quads split into fans, negative and slash-form indices
  {
    const char* path = argc > 1 ? argv[1] : "check_helpers.obj";
    {
      std::ofstream o(path);
      o << "# test\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvn 0 0 1\nf 1//1 2//1 3//1 4//1\nf -4 -2 -1\n";
    }
    std::vector<double3> v;
    std::vector<int3> f;
    const bool ok = OBJLoader::loadOBJMesh(path, v, f, 0, nullptr, false);
    CHECK(ok && v.size() == 4 && f.size() == 3, "obj %zu %zu", v.size(), f.size());
    if (f.size() == 3) {
      CHECK(f[0].x == 0 && f[0].y == 1 && f[0].z == 2 && f[1].x == 0 && f[1].y == 2 && f[1].z == 3, "fan");
      CHECK(f[2].x == 0 && f[2].y == 2 && f[2].z == 3, "negative indices");
    }
    CHECK(!OBJLoader::loadOBJMesh("/nonexistent/file.obj", v, f, 0, nullptr, false), "missing file");
  }
  std::printf(failures ? "FAILED %d\n" : "ALL OK\n", failures);
  return failures ? 1 : 0;
}
