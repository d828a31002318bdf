// A level-set sphere dropped from 1 m onto a plane wall with restitution 1, written against the
// LSSolver drop-in exactly as a FunDEM app is (compare FunDEM/apps/test/LSSphereBouncing.cpp).
// Usage: ls_sphere_bouncing [output-dir-prefix]
#include "kernel/LSSolver.h"

int main(const int argc, char** argv) {
  LSSolver solver_(argc > 1 ? argv[1] : "LSSphereBouncing");

  const double radius = 0.01, density = 2000., height = 1.;
  const double3 planeMin = make_double3(-0.5, -0.5, 0.), planeMax = make_double3(0.5, 0.5, 0.);
  const double collideTime = 0.005;
  const double mass = (4. / 3.) * pi() * radius * radius * radius * density;
  const double I = 0.4 * mass * radius * radius;
  const double kn = 0.5 * mass * pi() * pi() / (collideTime * collideTime);

  // sphere SDF on a 13^3 grid, nodes on latitude/longitude rings with a triangle mesh
  const double h = radius / 5.;
  const int3 n = make_int3(13, 13, 13);
  const double3 origin = make_double3(-0.5 * h * (n.x - 1), -0.5 * h * (n.y - 1), -0.5 * h * (n.z - 1));
  std::vector<double> sdf(size_t(n.x) * n.y * n.z);
  for (int z = 0; z < n.z; ++z)
    for (int y = 0; y < n.y; ++y)
      for (int x = 0; x < n.x; ++x)
        sdf[linearIndex3D(make_int3(x, y, z), n)] = length(origin + make_double3(x, y, z) * h) - radius;
  std::vector<double3> nodes;
  std::vector<int3> tris;
  const int seg = 6;
  for (int i = 0; i < seg; ++i)
    for (int j = 0; j < seg; ++j) {
      const double th = i * pi() / (seg - 1), ph = j * 2. * pi() / seg;
      nodes.push_back(make_double3(radius * std::sin(th) * std::cos(ph), radius * std::sin(th) * std::sin(ph),
                                   radius * std::cos(th)));
    }
  for (int i = 0; i + 1 < seg; ++i)
    for (int j = 0; j < seg; ++j) {
      tris.push_back(make_int3(i * seg + j, (i + 1) * seg + j, (i + 1) * seg + (j + 1) % seg));
      tris.push_back(make_int3(i * seg + j, (i + 1) * seg + (j + 1) % seg, i * seg + (j + 1) % seg));
    }
  solver_.addLSParticle(nodes, sdf, origin, n, h, make_double3(0., 0., 1.), make_double3(0., 0., 0.),
                        make_double3(0., 0., 0.), make_quaternion(1., 0., 0., 0.), mass,
                        make_symMatrix(I, I, I, 0., 0., 0.), kn, 0., 0., 1., tris);

  // plane z = 0 as a 2^3 grid holding phi = z
  const double ph = std::max((planeMax - planeMin).x, (planeMax - planeMin).y);
  const int3 pn = make_int3(2, 2, 2);
  const double3 po = 0.5 * (planeMin + planeMax) - 0.5 * make_double3(ph, ph, ph);
  std::vector<double> psdf(8);
  for (int z = 0; z < 2; ++z)
    for (int y = 0; y < 2; ++y)
      for (int x = 0; x < 2; ++x) psdf[linearIndex3D(make_int3(x, y, z), pn)] = po.z + z * ph;
  const std::vector<double3> verts = {make_double3(planeMin.x, planeMin.y, 0.), make_double3(planeMax.x, planeMin.y, 0.),
                                      make_double3(planeMax.x, planeMax.y, 0.), make_double3(planeMin.x, planeMax.y, 0.)};
  const std::vector<int3> ptris = {make_int3(0, 1, 2), make_int3(0, 2, 3)};
  solver_.addWall(verts, ptris, psdf, po, pn, ph, 0.5 * (planeMin + planeMax), make_quaternion(1., 0., 0., 0.), 0.);

  solver_.solve(planeMin - make_double3(0., 0., height), planeMax + make_double3(0., 0., height),
                make_double3(0., 0., -9.81), collideTime / 50., 2., 200, argc, argv);
  return 0;
}
