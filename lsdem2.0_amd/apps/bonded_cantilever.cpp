// A cantilever of bonded level-set spheres under a tip load, written against the LSSolver drop-in
// and its helper headers as a FunDEM app is (the scenario of the reference's bonded-beam tutorial,
// FunDEM/apps/tutorial/tutorial3-bondedLSParticleBeam.cpp, with a small load): 11 steel spheres of
// radius 0.2 m in a row, neighbours joined by parallel bonds of the same radius, the first sphere
// held in place; a transverse tip force ramps up over 0.05 s and global damping settles the beam.
// Prints the tip deflection next to Euler-Bernoulli's F L^3 / (3 E I), and what the interaction
// getters report.  Usage: bonded_cantilever [output-dir-prefix]
#include <cstdio>

#include "globalDamping.h"
#include "kernel/LSSolver.h"
#include "kernel/helper/LevelSetObjectGenerator.h"

class Cantilever : public LSSolver {
 public:
  explicit Cantilever(const std::string& dir) : LSSolver(dir) {}
  size_t tip = 0;
  double load = 0.0, ramp = 0.05, damping = 0.1;

  void addExternalForceTorque(const double time) override {
    const double s = time < ramp ? time / ramp : 1.0;
    launchAddGlobalConstantForceTorque(getLSParticle().force(), getLSParticle().torque(),
                                       make_double3(0.0, 0.0, -load * s), make_double3(0.0, 0.0, 0.0), tip, 0);
    launchAddGlobalDampingForceTorque(getLSParticle().force(), getLSParticle().torque(), getLSParticle().velocity(),
                                      getLSParticle().angularVelocity(), damping, getLSParticle().num_device(),
                                      getLSParticle().gridDim(), getLSParticle().blockDim(), 0);
  }
  LSContactListView& contacts() { return getLSParticleInteraction(); }
  LSBondListView& bonds() { return getLSParticleBondedInteraction(); }
  // particle centres, copied from the device view the hook saw last
  std::vector<double3> positions() {
    std::vector<double3> x(getLSParticle().num_device());
    if (!x.empty() && hipMemcpy(x.data(), getLSParticle().position(), x.size() * sizeof(double3),
                                hipMemcpyDeviceToHost) != hipSuccess)
      x.clear();
    return x;
  }
};

int main(const int argc, char** argv) {
  const int n = 11;
  const double a = 0.2, density = 7800.0, E = 200.0e9, nu = 0.3, F = 5.0e4;
  Cantilever beam(argc > 1 ? argv[1] : "BondedCantilever");
  beam.tip = n - 1;
  beam.load = F;
  const double mass = 4.0 / 3.0 * pi() * a * a * a * density;
  const double I = 0.4 * mass * a * a;
  LevelSetObject::Sphere sphere;
  sphere.setParams(a, 600);  // 642 surface nodes
  sphere.buildGridByResolution(20);
  for (int i = 0; i < n; ++i) {
    const double3 x = make_double3(2.0 * a * i, 0.0, 0.0);
    beam.addLSParticle(sphere.vertexPosition(), sphere.gridInfo().gridNodeSignedDistance, sphere.gridInfo().gridOrigin,
                       sphere.gridInfo().gridNodeSize, sphere.gridInfo().gridNodeSpacing, x, make_double3(0, 0, 0),
                       make_double3(0, 0, 0), make_quaternion(1, 0, 0, 0), i > 0 ? mass : 0.0,
                       make_symMatrix(I, I, I, 0, 0, 0), 0.0, 0.0, 0.0, 1.0, sphere.triangleVertexIndex());
    if (i > 0)
      beam.addBondedInteraction(i - 1, i, x - make_double3(a, 0, 0), make_double3(-1, 0, 0), a, 2.0 * a, E, nu);
  }
  beam.solve(make_double3(-2 * a, -4 * a, -4 * a), make_double3(2 * a * n, 4 * a, 4 * a), make_double3(0, 0, 0), 1.0e-5,
             1.0, 2, argc, argv);

  const std::vector<double3> x = beam.positions();
  if (int(x.size()) != n) {
    std::printf("no particle state\n");
    return 1;
  }
  const double L = 2.0 * a * (n - 1);
  const double EI = E * pi() * a * a * a * a / 4.0;
  const double expected = F * L * L * L / (3.0 * EI);
  std::printf("tip deflection %.9e m, Euler-Bernoulli %.9e m, ratio %.6f\n", -x[n - 1].z, expected,
              -x[n - 1].z / expected);
  std::printf("bonds %zu active %d, particle contacts %zu\n", beam.bonds().numPair(),
              [&] {
                int k = 0;
                for (int f : beam.bonds().activatedFlagHostRef()) k += f;
                return k;
              }(),
              beam.contacts().numPair_device());
  return 0;
}
