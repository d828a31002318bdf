// LSSolver: source-compatible drop-in for FunDEM's header-only solver class
// (FunDEM/kernel/LSSolver.h:11-520), running on liblsdem.so (include/lsdem.h) on one AMD GPU.
//
// Apps keep `#include "kernel/LSSolver.h"`, derive from LSSolver, override
// addExternalForceTorque(time) and call the same setup methods and solve().  Differences:
//  - the stream argument of the constructor is accepted and ignored: the library runs on its
//    own stream, and hooks receive it through getLSParticle().stream();
//  - identical level-set grids are stored once (deduplicated by content);
//  - addPeriodicBoundarySector() cannot be combined with x/y periodicity (the library rejects it);
//  - a later solve() continues from the state of the previous one; if bodies or bonds were added
//    in between, or the domain/time step/gravity/periodicity changed, the library context is
//    rebuilt from the current state and the tangential contact history starts empty;
// Output goes to <dir>_phase<k>/{LSParticle,LSParticleInteraction,Wall,LSParticle-WallInteraction}
// as in the reference (LSSolver.h:598-660), plus one .pvd collection per series.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <filesystem>
#include <functional>
#include <iomanip>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../lsdem.h"
#include "lsdem_types.h"
#include "lsdem_vtu.h"

// Device views of a body set, as the reference's LSParticle exposes them to apps
// (LSParticle.h:661-670).  Valid inside addExternalForceTorque.
class LSParticleDeviceView {
 public:
  double3* force() const { return reinterpret_cast<double3*>(v_.force); }
  double3* torque() const { return reinterpret_cast<double3*>(v_.torque); }
  double3* velocity() const { return reinterpret_cast<double3*>(v_.velocity); }
  double3* angularVelocity() const { return reinterpret_cast<double3*>(v_.angular_velocity); }
  const double3* position() const { return reinterpret_cast<const double3*>(v_.position); }
  const quaternion* orientation() const { return reinterpret_cast<const quaternion*>(v_.quaternion); }
  size_t num_device() const { return size_t(v_.num); }
  size_t gridDim() const { return size_t(v_.grid_dim); }
  size_t blockDim() const { return size_t(v_.block_dim); }
  hipStream_t stream() const { return static_cast<hipStream_t>(v_.stream); }
  void set(const lsdem_device_view& v) { v_ = v; }

 private:
  lsdem_device_view v_{};
};

// The walls' device arrays, as getWall() returns them (LSSolver.h:512; the LSParticle accessors of
// LSParticle.h:660-671).  Walls are kinematic here: no force or torque arrays.
class LSWallDeviceView {
 public:
  const double3* position() const { return reinterpret_cast<const double3*>(v_.position); }
  const quaternion* orientation() const { return reinterpret_cast<const quaternion*>(v_.quaternion); }
  double3* velocity() const { return reinterpret_cast<double3*>(v_.velocity); }
  double3* angularVelocity() const { return reinterpret_cast<double3*>(v_.angular_velocity); }
  size_t num_device() const { return size_t(v_.num); }
  size_t gridDim() const { return size_t(v_.grid_dim); }
  size_t blockDim() const { return size_t(v_.block_dim); }
  hipStream_t stream() const { return static_cast<hipStream_t>(v_.stream); }
  void set(const lsdem_wall_device_view& v) { v_ = v; }

 private:
  lsdem_wall_device_view v_{};
};

// Host copies of one contact list of the last step, with the names of the reference's
// SolidInteraction host references (SolidInteraction.h:335-374); each call downloads the list,
// sorted by (master node, slave).  Empty before the first solve().
class LSContactListView {
 public:
  size_t numPair_device() const { return size_t(count()); }
  std::vector<double3> contactPointHostRef() const { return vec3(kPoint); }
  std::vector<double3> contactNormalHostRef() const { return vec3(kNormal); }
  std::vector<double3> slidingSpringHostRef() const { return vec3(kSpring); }
  std::vector<double> contactOverlapHostRef() const { return vec1(kOverlap); }
  std::vector<double> normalElasticEnergyHostRef() const { return vec1(kEn); }
  std::vector<double> slidingElasticEnergyHostRef() const { return vec1(kEs); }
  std::vector<int> masterIDHostRef() const { return ids(true); }
  std::vector<int> slaveIDHostRef() const { return ids(false); }
  void bind(lsdem_ctx* const* ctx, int32_t list) {
    ctx_ = ctx;
    list_ = list;
  }

 private:
  enum { kPoint, kNormal, kSpring, kOverlap, kEn, kEs };
  lsdem_ctx* const* ctx_ = nullptr;
  int32_t list_ = LSDEM_LIST_PP;
  int64_t count() const {
    int64_t n = 0;
    if (ctx_ && *ctx_)
      lsdem_get_contacts(*ctx_, list_, 0, &n, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    return n;
  }
  std::vector<double> raw(int what, int width) const {
    int64_t n = count();
    std::vector<double> d(static_cast<size_t>(n) * width);
    if (n == 0) return d;
    double* a[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    a[what] = d.data();
    lsdem_get_contacts(*ctx_, list_, n, &n, nullptr, nullptr, a[kPoint], a[kNormal], a[kOverlap], a[kSpring], a[kEn],
                       a[kEs]);
    return d;
  }
  std::vector<double3> vec3(int what) const {
    const std::vector<double> d = raw(what, 3);
    std::vector<double3> v(d.size() / 3);
    for (size_t k = 0; k < v.size(); ++k) v[k] = make_double3(d[3 * k], d[3 * k + 1], d[3 * k + 2]);
    return v;
  }
  std::vector<double> vec1(int what) const { return raw(what, 1); }
  std::vector<int> ids(bool master) const {
    int64_t n = count();
    std::vector<int32_t> m(static_cast<size_t>(n)), sl(static_cast<size_t>(n));
    if (n) lsdem_get_contacts(*ctx_, list_, n, &n, m.data(), sl.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    const std::vector<int32_t>& r = master ? m : sl;
    return std::vector<int>(r.begin(), r.end());
  }
};

// Host copies of the bond state, with the names of the reference's VBondedInteraction
// (VBondedInteraction.h:489-520): activated flag, centre point, energies and peak stresses.
class LSBondListView {
 public:
  size_t numPair() const { return n_; }
  size_t numPair_device() const { return n_; }
  std::vector<int> activatedFlagHostRef() const {
    std::vector<int32_t> a(n_);
    if (n_) lsdem_get_bonds(*ctx_, a.data(), nullptr, nullptr, nullptr, nullptr);
    return std::vector<int>(a.begin(), a.end());
  }
  std::vector<double3> centerPointHostRef() const {
    std::vector<double> d(3 * n_);
    if (n_) lsdem_get_bonds(*ctx_, nullptr, d.data(), nullptr, nullptr, nullptr);
    std::vector<double3> v(n_);
    for (size_t k = 0; k < n_; ++k) v[k] = make_double3(d[3 * k], d[3 * k + 1], d[3 * k + 2]);
    return v;
  }
  // Un, Us, Ub, Ut per bond
  std::vector<double> elasticEnergyHostRef(int which) const {
    std::vector<double> e(4 * n_), out(n_);
    if (n_) lsdem_get_bonds(*ctx_, nullptr, nullptr, e.data(), nullptr, nullptr);
    for (size_t k = 0; k < n_; ++k) out[k] = e[4 * k + size_t(which & 3)];
    return out;
  }
  std::vector<double> maxNormalStressHostRef() const { return stress(true); }
  std::vector<double> maxShearStressHostRef() const { return stress(false); }
  void bind(lsdem_ctx* const* ctx, size_t n) {
    ctx_ = ctx;
    n_ = (ctx && *ctx) ? n : 0;
  }

 private:
  lsdem_ctx* const* ctx_ = nullptr;
  size_t n_ = 0;
  std::vector<double> stress(bool normal) const {
    std::vector<double> s(n_), t(n_);
    if (n_) lsdem_get_bonds(*ctx_, nullptr, nullptr, nullptr, s.data(), t.data());
    return normal ? s : t;
  }
};

class LSSolver {
 public:
  LSSolver(const std::string dir = "Problem", hipStream_t stream = 0, const size_t maxGPUThread = 256,
           const int device = 0)
      : dir_(dir), maxGPUThread_(maxGPUThread), device_(device) {
    (void)stream;
  }

  virtual ~LSSolver() {
    if (ctx_) lsdem_destroy(ctx_);
  }

  LSSolver(const LSSolver&) = delete;
  LSSolver& operator=(const LSSolver&) = delete;

  // Free particle; mass, centroid and inertia from the SDF (LSSolver.h:51-163).
  void addLSParticle(const std::vector<double3>& boundaryNodeLocalPosition,
                     const std::vector<double>& gridNodeSignedDistance, const double3 gridNodeLocalOrigin,
                     const int3 gridNodeSize, const double gridNodeSpacing, const double3 position,
                     const double3 velocity, const double3 angularVelocity, const quaternion orientation,
                     const double normalStiffness, const double shearStiffness, const double frictionCoefficient,
                     const double restitutionCoefficient, const double density,
                     const std::vector<int3>& boundaryNodeConnectivity = {}) {
    if (!validGrid("LSParticle", gridNodeSignedDistance, gridNodeSize, gridNodeSpacing)) return;
    double mass = 0, c[3] = {0, 0, 0}, inertia[6], invI[6];
    const double o[3] = {gridNodeLocalOrigin.x, gridNodeLocalOrigin.y, gridNodeLocalOrigin.z};
    lsdem_mass_properties(gridNodeSignedDistance.data(), gridNodeSize.x, gridNodeSize.y, gridNodeSize.z, o,
                          gridNodeSpacing, density, &mass, c, inertia, invI);
    const double3 cen = make_double3(c[0], c[1], c[2]);
    const quaternion q = normalize(orientation);
    std::vector<double3> nodes;
    nodes.reserve(boundaryNodeLocalPosition.size());
    for (const auto& p : boundaryNodeLocalPosition) nodes.push_back(p - cen);
    const symMatrix Ii = mass > 0 ? make_symMatrix(invI[0], invI[1], invI[2], invI[3], invI[4], invI[5])
                                  : make_symMatrix(0, 0, 0, 0, 0, 0);
    pushParticle(nodes, gridNodeSignedDistance, gridNodeLocalOrigin - cen, gridNodeSize, gridNodeSpacing,
                 position + rotateVectorByQuaternion(q, cen), velocity, angularVelocity, q,
                 mass > 0 ? 1.0 / mass : 0.0, Ii, normalStiffness, shearStiffness, frictionCoefficient,
                 restitutionCoefficient, boundaryNodeConnectivity);
  }

  // Particle with given mass and inertia (LSSolver.h:190-223); mass 0 fixes it in place.
  void addLSParticle(const std::vector<double3>& boundaryNodeLocalPosition,
                     const std::vector<double>& gridNodeSignedDistance, const double3 gridNodeLocalOrigin,
                     const int3 gridNodeSize, const double gridNodeSpacing, const double3 position,
                     const double3 velocity, const double3 angularVelocity, const quaternion orientation,
                     const double mass, const symMatrix inertiaTensor, const double normalStiffness,
                     const double shearStiffness, const double frictionCoefficient,
                     const double restitutionCoefficient, const std::vector<int3>& boundaryNodeConnectivity = {}) {
    if (!validGrid("LSParticle", gridNodeSignedDistance, gridNodeSize, gridNodeSpacing)) return;
    pushParticle(boundaryNodeLocalPosition, gridNodeSignedDistance, gridNodeLocalOrigin, gridNodeSize,
                 gridNodeSpacing, position, velocity, angularVelocity, normalize(orientation),
                 mass > 0. ? 1. / mass : 0., mass > 0. ? inverse(inertiaTensor) : make_symMatrix(0, 0, 0, 0, 0, 0),
                 normalStiffness, shearStiffness, frictionCoefficient, restitutionCoefficient,
                 boundaryNodeConnectivity);
  }

  // Immovable level-set wall with a mesh for output (LSSolver.h:245-272).
  void addWall(const std::vector<double3>& objVertexLocalPosition, const std::vector<int3>& objTriangleVertexID,
               const std::vector<double>& gridNodeSignedDistance, const double3 gridNodeLocalOrigin,
               const int3 gridNodeSize, const double gridNodeSpacing, const double3 position,
               const quaternion orientation, const double frictionCoefficient,
               const double restitutionCoefficient = 1.) {
    if (!validGrid("LSParticle", gridNodeSignedDistance, gridNodeSize, gridNodeSpacing)) return;
    Wall w;
    w.shape = shapeId(gridNodeSignedDistance, gridNodeLocalOrigin, gridNodeSize, gridNodeSpacing, {});
    double r = 0;
    for (const auto& p : objVertexLocalPosition) r = std::max(r, length(p));
    w.radius = r;
    w.pos = position;
    w.q = normalize(orientation);
    w.mu = frictionCoefficient > 0 ? frictionCoefficient : 0.0;
    w.e = (restitutionCoefficient <= 0.0 || restitutionCoefficient > 1.0) ? 1.0 : restitutionCoefficient;
    w.verts = objVertexLocalPosition;
    w.tris = objTriangleVertexID;
    walls_.push_back(w);
    dirty_ = true;
  }

  void moveLSParticle(const size_t index, const double3 offSet) {
    if (index >= parts_.size()) return;
    syncDown();
    parts_[index].pos = parts_[index].pos + offSet;
    pushState();
  }

  void moveWall(const size_t index, const double3 offset) {
    if (index >= walls_.size()) return;
    syncDown();
    walls_[index].pos = walls_[index].pos + offset;
    pushState();
  }

  void setFixedVelocityToWall(const size_t index, const double3 velocity) {
    if (index >= walls_.size()) return;
    syncDown();
    walls_[index].vel = velocity;
    pushState();
  }

  void setFixedAngularVelocityToWall(const size_t index, const double3 angularVelocity) {
    if (index >= walls_.size()) return;
    syncDown();
    walls_[index].angvel = angularVelocity;
    pushState();
  }

  // Parallel bond between two particles at their current poses (LSSolver.h:344-372).
  void addBondedInteraction(const int masterObjectID, const int slaveObjectID, const double3 bondPoint,
                            const double3 bondNormal, const double radius, const double initialLength,
                            const double YoungsModulus, const double poissonRatio, const double tensileStrength = 0.,
                            const double cohesion = 0., const double frictionCoefficient = 0.) {
    if (masterObjectID >= int(parts_.size()) || slaveObjectID >= int(parts_.size())) return;
    if (masterObjectID < 0 || slaveObjectID < 0 || masterObjectID == slaveObjectID) {
      std::cerr << "[VBondedInteraction] Invalid object IDs: " << masterObjectID << ", " << slaveObjectID << "."
                << std::endl;
      return;
    }
    if (radius <= 0.0 || initialLength <= 0.0 || YoungsModulus <= 0.0 || length(bondNormal) == 0.0) {
      std::cerr << "[VBondedInteraction] Invalid bond parameters." << std::endl;
      return;
    }
    syncDown();
    bonds_.push_back(Bond{masterObjectID, slaveObjectID, bondPoint, bondNormal, radius, initialLength,
                          YoungsModulus, poissonRatio, tensileStrength, cohesion, frictionCoefficient});
    dirty_ = true;
  }

  // One bond per particle-particle contact of the last step (LSSolver.h:385-417).
  void createBondsFromContacts(const double radius, const double initialLength, const double YoungsModulus,
                               const double poissonRatio, const double tensileStrength = 0.,
                               const double cohesion = 0., const double frictionCoefficient = 0.) {
    if (!ctx_) return;
    int64_t n = 0;
    lsdem_get_contacts(ctx_, LSDEM_LIST_PP, 0, &n, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                       nullptr);
    std::vector<int32_t> master(n), slave(n);
    std::vector<double> point(3 * n), normal(3 * n);
    lsdem_get_contacts(ctx_, LSDEM_LIST_PP, n, &n, master.data(), slave.data(), point.data(), normal.data(), nullptr,
                       nullptr, nullptr, nullptr);
    for (int64_t k = 0; k < n; ++k) {
      const int i = nodeOwner(master[k]);
      addBondedInteraction(i, slave[k], make_double3(point[3 * k], point[3 * k + 1], point[3 * k + 2]),
                           make_double3(normal[3 * k], normal[3 * k + 1], normal[3 * k + 2]), radius, initialLength,
                           YoungsModulus, poissonRatio, tensileStrength, cohesion, frictionCoefficient);
    }
  }

  void addPeriodicBoundaryXD() { periodicX_ = true; }
  void addPeriodicBoundaryYD() { periodicY_ = true; }
  // Quarter-sector rotational periodicity about the z axis through the domain's lower corner
  // (LSSolver.h:443-451).
  void addPeriodicBoundarySector() { periodicSector_ = true; }
  void removePeriodicBoundary() {
    periodicX_ = false;
    periodicY_ = false;
    periodicSector_ = false;
  }

  // Runs floor(T/dt) + 2 steps and writes numFrame + 2 frames (LSSolver.h:478-505, 994-1024).
  void solve(const double3 minDomain, const double3 maxDomain, const double3 gravity, const double timeStep,
             const double maximumTime, const size_t numFrame, const int argc, char** argv) {
    if (!checkSolverInput(minDomain, maxDomain, timeStep, maximumTime)) return;
    if (hipSetDevice(device_) != hipSuccess) {
      std::cout << "hipSetDevice( " << device_ << " ) failed! Is an AMD GPU visible?" << std::endl;
      std::exit(1);
    }
    phase_ += 1;
    updateDir(argc, argv);
    prepareDirs();
    std::cout << "[Solver] Uploading..." << std::endl;
    if (!upload(minDomain, maxDomain, gravity, timeStep)) return;
    std::cout << "[Solver] Upload Completed" << std::endl;

    const size_t numStep = size_t(maximumTime / timeStep) + 1;
    size_t frameInterval = numStep;
    if (numFrame > 0) frameInterval = numStep / numFrame;
    if (frameInterval < 1) frameInterval = 1;

    lsdem_set_time(ctx_, 0.0);
    size_t iStep = 0, iFrame = 0;
    // the pre-loop contact build and periodic wrap of LSSolver::compute* (LSSolver.h:996-998,
    // 1061-1066); frame 0 shows its contacts
    if (lsdem_begin_phase(ctx_) != LSDEM_OK) {
      std::cerr << "[Solver] " << lsdem_last_error() << std::endl;
      return;
    }
    output(iFrame, iStep, 0.0);
    while (iStep <= numStep) {
      // steps until the next frame (or the end of the loop)
      size_t k = frameInterval - (iStep % frameInterval);
      k = std::min(k, numStep + 1 - iStep);
      if (lsdem_step(ctx_, int32_t(k)) != LSDEM_OK) {
        std::cerr << "[Solver] " << lsdem_last_error() << std::endl;
        return;
      }
      iStep += k;
      if (iStep % frameInterval == 0) output(++iFrame, iStep, currentTime());
    }
    output(++iFrame, iStep, currentTime());
    std::cout << "[Solver] Computation Completed" << std::endl;
  }

 protected:
  virtual void addExternalForceTorque(const double time) { (void)time; }

  LSParticleDeviceView& getLSParticle() { return view_; }
  // (valid once solve() has set the simulation up; empty before)
  LSWallDeviceView& getWall() {
    lsdem_wall_device_view v{};
    if (ctx_ && lsdem_get_wall_device_view(ctx_, &v) == LSDEM_OK) wallView_.set(v);
    return wallView_;
  }
  // Contact lists and bonds of the last step, as host copies (the reference returns its
  // SolidInteraction / VBondedInteraction objects, LSSolver.h:510-518).
  LSContactListView& getLSParticleInteraction() {
    ppView_.bind(&ctx_, LSDEM_LIST_PP);
    return ppView_;
  }
  LSContactListView& getLSParticleWallInteraction() {
    pwView_.bind(&ctx_, LSDEM_LIST_PW);
    return pwView_;
  }
  LSBondListView& getLSParticleBondedInteraction() {
    bondView_.bind(&ctx_, dirty_ ? 0 : bonds_.size());
    return bondView_;
  }

 private:
  struct Particle {
    int shape;
    double3 pos, vel, angvel;
    quaternion q;
    double invM;
    symMatrix invI;
    double kn, ks, mu, e;
    int nodeBase;  // index of the first boundary node (global node numbering)
  };
  struct Wall {
    int shape;
    double radius;
    double3 pos, vel = make_double3(0, 0, 0), angvel = make_double3(0, 0, 0);
    quaternion q;
    double mu, e;
    std::vector<double3> verts;
    std::vector<int3> tris;
  };
  struct Bond {
    int i, j;
    double3 point, normal;
    double r, L, E, nu, st, c, mu;
  };
  struct Shape {
    std::vector<float> sdf;
    double3 origin;
    int3 size;
    double h;
    std::vector<double3> nodes;
  };

  std::string dir_;
  size_t maxGPUThread_;
  int device_;
  int phase_ = 0;
  bool periodicX_ = false, periodicY_ = false, periodicSector_ = false;
  std::vector<Shape> shapes_;
  std::map<std::string, int> shapeKey_;
  std::vector<Particle> parts_;
  std::vector<Wall> walls_;
  std::vector<Bond> bonds_;
  std::vector<int3> meshTris_;  // boundary-node triangles, global node numbering
  int nNodes_ = 0;
  lsdem_ctx* ctx_ = nullptr;
  bool dirty_ = true;
  double3 cfgMin_{}, cfgMax_{}, cfgG_{};
  double cfgDt_ = 0;
  bool cfgPX_ = false, cfgPY_ = false, cfgPS_ = false;
  LSParticleDeviceView view_;
  LSWallDeviceView wallView_;
  LSContactListView ppView_, pwView_;
  LSBondListView bondView_;
  std::map<std::string, std::vector<std::pair<double, std::string>>> pvd_;

  static bool validGrid(const char* who, const std::vector<double>& sdf, int3 n, double h) {
    if (n.x < 2 || n.y < 2 || n.z < 2) {
      std::cerr << "[" << who << "] Invalid level-set host grid size: (" << n.x << ", " << n.y << ", " << n.z << ")."
                << std::endl;
      return false;
    }
    if (h <= 0.0) {
      std::cerr << "[" << who << "] Invalid level-set host grid spacing: " << h << "." << std::endl;
      return false;
    }
    const size_t want = size_t(n.x) * size_t(n.y) * size_t(n.z);
    if (want != sdf.size()) {
      std::cerr << "[" << who << "] Inconsistent level-set host data size. Expected " << want << ", got "
                << sdf.size() << "." << std::endl;
      return false;
    }
    return true;
  }

  int shapeId(const std::vector<double>& sdf, double3 origin, int3 n, double h, const std::vector<double3>& nodes) {
    Shape s;
    s.sdf.assign(sdf.begin(), sdf.end());
    s.origin = origin;
    s.size = n;
    s.h = h;
    s.nodes = nodes;
    std::string key;
    auto add = [&](const void* p, size_t b) { key.append(static_cast<const char*>(p), b); };
    add(s.sdf.data(), s.sdf.size() * sizeof(float));
    add(&origin, sizeof origin);
    add(&n, sizeof n);
    add(&h, sizeof h);
    if (!nodes.empty()) add(nodes.data(), nodes.size() * sizeof(double3));
    const auto it = shapeKey_.find(key);
    if (it != shapeKey_.end()) return it->second;
    shapes_.push_back(std::move(s));
    const int id = int(shapes_.size()) - 1;
    shapeKey_.emplace(std::move(key), id);
    return id;
  }

  void pushParticle(const std::vector<double3>& nodes, const std::vector<double>& sdf, double3 origin, int3 n,
                    double h, double3 x, double3 v, double3 w, quaternion q, double invM, symMatrix invI, double kn,
                    double ks, double mu, double e, const std::vector<int3>& tris) {
    syncDown();
    Particle p;
    p.shape = shapeId(sdf, origin, n, h, nodes);
    p.pos = x;
    p.vel = v;
    p.angvel = w;
    p.q = q;
    p.invM = invM;
    p.invI = invI;
    p.kn = kn;
    p.ks = ks;
    p.mu = mu;
    p.e = e;
    p.nodeBase = nNodes_;
    for (const auto& t : tris) meshTris_.push_back(make_int3(t.x + nNodes_, t.y + nNodes_, t.z + nNodes_));
    nNodes_ += int(nodes.size());
    parts_.push_back(p);
    dirty_ = true;
  }

  int nodeOwner(int node) const {
    int lo = 0, hi = int(parts_.size()) - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) / 2;
      if (parts_[mid].nodeBase <= node) lo = mid;
      else hi = mid - 1;
    }
    return lo;
  }

  double currentTime() const {
    double t = 0;
    if (ctx_) lsdem_get_time(ctx_, &t);
    return t;
  }

  static void hookTrampoline(void* user, double time, const lsdem_device_view* v) {
    LSSolver* s = static_cast<LSSolver*>(user);
    s->view_.set(*v);
    s->addExternalForceTorque(time);
  }

  // Host copies of the particle and wall state from the library (before edits between phases).
  void syncDown() {
    if (!ctx_) return;
    const size_t n = parts_.size(), m = walls_.size();
    std::vector<double> x(3 * n), q(4 * n), v(3 * n), w(3 * n);
    if (n && lsdem_get_particles(ctx_, x.data(), q.data(), v.data(), w.data(), nullptr, nullptr) == LSDEM_OK) {
      for (size_t i = 0; i < n; ++i) {
        parts_[i].pos = make_double3(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
        parts_[i].q = make_quaternion(q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]);
        parts_[i].vel = make_double3(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
        parts_[i].angvel = make_double3(w[3 * i], w[3 * i + 1], w[3 * i + 2]);
      }
    }
    std::vector<double> wx(3 * m), wq(4 * m), wv(3 * m), ww(3 * m);
    if (m && lsdem_get_walls(ctx_, wx.data(), wq.data(), wv.data(), ww.data()) == LSDEM_OK) {
      for (size_t i = 0; i < m; ++i) {
        walls_[i].pos = make_double3(wx[3 * i], wx[3 * i + 1], wx[3 * i + 2]);
        walls_[i].q = make_quaternion(wq[4 * i], wq[4 * i + 1], wq[4 * i + 2], wq[4 * i + 3]);
      }
    }
  }

  // Push edited host state back into a live context (pose, velocity; history kept).
  void pushState() {
    if (!ctx_ || dirty_) return;
    const size_t n = parts_.size(), m = walls_.size();
    std::vector<double> x, q, v, w;
    for (const auto& p : parts_) {
      x.insert(x.end(), {p.pos.x, p.pos.y, p.pos.z});
      q.insert(q.end(), {p.q.q0, p.q.q1, p.q.q2, p.q.q3});
      v.insert(v.end(), {p.vel.x, p.vel.y, p.vel.z});
      w.insert(w.end(), {p.angvel.x, p.angvel.y, p.angvel.z});
    }
    if (n) lsdem_set_particles(ctx_, x.data(), q.data(), v.data(), w.data(), nullptr, nullptr);
    std::vector<double> wx, wq, wv, ww;
    for (const auto& s : walls_) {
      wx.insert(wx.end(), {s.pos.x, s.pos.y, s.pos.z});
      wq.insert(wq.end(), {s.q.q0, s.q.q1, s.q.q2, s.q.q3});
      wv.insert(wv.end(), {s.vel.x, s.vel.y, s.vel.z});
      ww.insert(ww.end(), {s.angvel.x, s.angvel.y, s.angvel.z});
    }
    if (m) lsdem_set_walls(ctx_, wx.data(), wq.data(), wv.data(), ww.data());
  }

  bool sameConfig(double3 mn, double3 mx, double3 g, double dt) const {
    auto eq = [](double3 a, double3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; };
    return eq(mn, cfgMin_) && eq(mx, cfgMax_) && eq(g, cfgG_) && dt == cfgDt_ && periodicX_ == cfgPX_ &&
           periodicY_ == cfgPY_ && periodicSector_ == cfgPS_;
  }

  bool upload(double3 mn, double3 mx, double3 g, double dt) {
    if (ctx_ && !dirty_ && sameConfig(mn, mx, g, dt)) return true;
    syncDown();
    if (ctx_) {
      lsdem_destroy(ctx_);
      ctx_ = nullptr;
    }
    lsdem_config cfg;
    lsdem_config_default(&cfg);
    cfg.domain_min[0] = mn.x; cfg.domain_min[1] = mn.y; cfg.domain_min[2] = mn.z;
    cfg.domain_max[0] = mx.x; cfg.domain_max[1] = mx.y; cfg.domain_max[2] = mx.z;
    cfg.gravity[0] = g.x; cfg.gravity[1] = g.y; cfg.gravity[2] = g.z;
    cfg.dt = dt;
    cfg.periodic_x = periodicX_;
    cfg.periodic_y = periodicY_;
    cfg.periodic_sector = periodicSector_;
    cfg.block_size = int32_t(maxGPUThread_);
    cfg.device = device_;
    auto check = [&](int rc) {
      if (rc != LSDEM_OK) std::cerr << "[Solver] " << lsdem_last_error() << std::endl;
      return rc == LSDEM_OK;
    };
    if (!check(lsdem_create(&cfg, &ctx_))) return false;
    for (const auto& s : shapes_) {
      const double o[3] = {s.origin.x, s.origin.y, s.origin.z};
      int32_t id = 0;
      if (!check(lsdem_add_shape(ctx_, s.sdf.data(), s.size.x, s.size.y, s.size.z, o, s.h,
                                 s.nodes.empty() ? nullptr : reinterpret_cast<const double*>(s.nodes.data()),
                                 int32_t(s.nodes.size()), &id)))
        return false;
    }
    const int32_t n = int32_t(parts_.size());
    if (n) {
      std::vector<int32_t> shape(n);
      std::vector<double> scale(n, 1.0), x, q, v, w, im(n), ii, kn(n), ks(n), mu(n), e(n);
      for (int32_t i = 0; i < n; ++i) {
        const Particle& p = parts_[i];
        shape[i] = p.shape;
        x.insert(x.end(), {p.pos.x, p.pos.y, p.pos.z});
        q.insert(q.end(), {p.q.q0, p.q.q1, p.q.q2, p.q.q3});
        v.insert(v.end(), {p.vel.x, p.vel.y, p.vel.z});
        w.insert(w.end(), {p.angvel.x, p.angvel.y, p.angvel.z});
        im[i] = p.invM;
        ii.insert(ii.end(), {p.invI.xx, p.invI.yy, p.invI.zz, p.invI.xy, p.invI.xz, p.invI.yz});
        kn[i] = p.kn; ks[i] = p.ks; mu[i] = p.mu; e[i] = p.e;
      }
      if (!check(lsdem_add_particles(ctx_, n, shape.data(), scale.data(), x.data(), q.data(), v.data(), w.data(),
                                     im.data(), ii.data(), kn.data(), ks.data(), mu.data(), e.data())))
        return false;
    }
    const int32_t m = int32_t(walls_.size());
    if (m) {
      std::vector<int32_t> shape(m);
      std::vector<double> r(m), x, q, v, w, mu(m), e(m);
      for (int32_t i = 0; i < m; ++i) {
        const Wall& s = walls_[i];
        shape[i] = s.shape;
        r[i] = s.radius;
        x.insert(x.end(), {s.pos.x, s.pos.y, s.pos.z});
        q.insert(q.end(), {s.q.q0, s.q.q1, s.q.q2, s.q.q3});
        v.insert(v.end(), {s.vel.x, s.vel.y, s.vel.z});
        w.insert(w.end(), {s.angvel.x, s.angvel.y, s.angvel.z});
        mu[i] = s.mu; e[i] = s.e;
      }
      if (!check(lsdem_add_walls(ctx_, m, shape.data(), r.data(), x.data(), q.data(), v.data(), w.data(), mu.data(),
                                 e.data())))
        return false;
    }
    for (const Bond& b : bonds_) {
      const double pt[3] = {b.point.x, b.point.y, b.point.z}, nm[3] = {b.normal.x, b.normal.y, b.normal.z};
      if (lsdem_add_bond(ctx_, b.i, b.j, pt, nm, b.r, b.L, b.E, b.nu, b.st, b.c, b.mu) != LSDEM_OK)
        std::cerr << "[VBondedInteraction] " << lsdem_last_error() << std::endl;
    }
    if (!check(lsdem_initialize(ctx_))) return false;
    lsdem_set_external_force(ctx_, &LSSolver::hookTrampoline, this);
    lsdem_device_view v;
    if (lsdem_get_device_view(ctx_, &v) == LSDEM_OK) view_.set(v);
    cfgMin_ = mn; cfgMax_ = mx; cfgG_ = g; cfgDt_ = dt; cfgPX_ = periodicX_; cfgPY_ = periodicY_; cfgPS_ = periodicSector_;
    dirty_ = false;
    return true;
  }

  bool checkSolverInput(double3 mn, double3 mx, double dt, double T) {
    const double3 d = mx - mn;
    if (d.x <= 0.0 || d.y <= 0.0 || d.z <= 0.0) {
      std::cerr << "[Solver] Invalid simulation domain size: (" << d.x << ", " << d.y << ", " << d.z << ")."
                << std::endl;
      return false;
    }
    if (dt <= 0.0) {
      std::cerr << "[Solver] Invalid timeStep: " << dt << "." << std::endl;
      return false;
    }
    if (T < 0.0) {
      std::cerr << "[Solver] Invalid maximumTime: " << T << "." << std::endl;
      return false;
    }
    return true;
  }

  // Relative output directories resolve against the nearest "build" ancestor of the executable,
  // else its folder (LSSolver.h:535-562).
  void updateDir(int argc, char** argv) {
    namespace fs = std::filesystem;
    const fs::path p(dir_);
    if (p.is_absolute()) return;
    fs::path base = fs::current_path();
    if (argc > 0 && argv && argv[0]) {
      fs::path exe = fs::weakly_canonical(fs::absolute(fs::path(argv[0])));
      base = exe.parent_path();
      for (fs::path d = base; !d.empty(); d = d.parent_path()) {
        if (d.filename() == "build") {
          base = d;
          break;
        }
        if (d.parent_path() == d) break;
      }
    }
    dir_ = (base / p).string();
  }

  std::string phaseDir() const { return dir_ + "_phase" + std::to_string(phase_); }

  void prepareDirs() {
    namespace fs = std::filesystem;
    for (const char* sub : {"LSParticle", "LSParticleInteraction", "Wall", "LSParticle-WallInteraction"}) {
      const fs::path d = fs::path(phaseDir()) / sub;
      std::error_code ec;
      fs::create_directories(d, ec);
      for (auto& f : fs::directory_iterator(d, ec))
        if (f.path().extension() == ".vtu" || f.path().extension() == ".pvd") fs::remove(f.path(), ec);
    }
    pvd_.clear();
  }

  static std::string frameName(const std::string& stem, size_t iFrame) {
    std::ostringstream s;
    s << stem << "_" << std::setw(4) << std::setfill('0') << iFrame << ".vtu";
    return s.str();
  }

  void emit(const std::string& sub, const std::string& stem, size_t iFrame, double time, size_t iStep,
            const std::function<void(const std::string&)>& writer) {
    const std::string name = frameName(stem, iFrame);
    const std::string dir = phaseDir() + "/" + sub;
    writer(dir + "/" + name);
    auto& frames = pvd_[sub + "/" + stem];
    frames.emplace_back(time, name);
    lsdem_vtu::writePvd(dir + "/" + stem + ".pvd", frames);
    (void)iStep;
  }

  void output(size_t iFrame, size_t iStep, double time) {
    std::cout << "[Solver] Phase " << phase_ << ": ------ Frame " << iFrame << " at Time " << time << " ------ "
              << std::endl;
    syncDown();
    std::cout << "[Solver] Outputting... " << std::endl;
    using lsdem_vtu::array;
    const size_t n = parts_.size();
    if (n) {
      // boundary nodes with the particles' meshes
      std::vector<float> pts, vel;
      std::vector<int32_t> pid;
      for (size_t i = 0; i < n; ++i) {
        const Particle& p = parts_[i];
        for (const auto& nd : shapes_[p.shape].nodes) {
          const double3 xw = p.pos + rotateVectorByQuaternion(p.q, nd);
          const double3 vw = p.vel + cross(p.angvel, xw - p.pos);
          pts.insert(pts.end(), {float(xw.x), float(xw.y), float(xw.z)});
          vel.insert(vel.end(), {float(vw.x), float(vw.y), float(vw.z)});
          pid.push_back(int32_t(i));
        }
      }
      std::vector<int32_t> conn, offs;
      std::vector<uint8_t> types;
      for (size_t t = 0; t < meshTris_.size(); ++t) {
        conn.insert(conn.end(), {meshTris_[t].x, meshTris_[t].y, meshTris_[t].z});
        offs.push_back(int32_t(3 * (t + 1)));
        types.push_back(5);
      }
      emit("LSParticle", "LSObjectMesh", iFrame, time, iStep, [&](const std::string& f) {
        lsdem_vtu::write(f, time, iStep, pts, &conn, &offs, &types,
                         {array("particleID", "Int32", 0, pid), array("velocity", "Float32", 3, vel)}, "velocity");
      });
      // particle centres
      std::vector<float> c, v, w, o, r;
      for (const auto& p : parts_) {
        c.insert(c.end(), {float(p.pos.x), float(p.pos.y), float(p.pos.z)});
        v.insert(v.end(), {float(p.vel.x), float(p.vel.y), float(p.vel.z)});
        w.insert(w.end(), {float(p.angvel.x), float(p.angvel.y), float(p.angvel.z)});
        o.insert(o.end(), {float(p.q.q0), float(p.q.q1), float(p.q.q2), float(p.q.q3)});
        double rr = 0;
        for (const auto& nd : shapes_[p.shape].nodes) rr = std::max(rr, length(nd));
        r.push_back(float(rr));
      }
      emit("LSParticle", "Particle", iFrame, time, iStep, [&](const std::string& f) {
        lsdem_vtu::write(f, time, iStep, c, nullptr, nullptr, nullptr,
                         {array("velocity", "Float32", 3, v), array("angularVelocity", "Float32", 3, w),
                          array("orientation", "Float32", 4, o), array("radius", "Float32", 1, r)},
                         "velocity");
      });
    }
    writeContacts(LSDEM_LIST_PP, "LSParticleInteraction", iFrame, time, iStep);
    writeContacts(LSDEM_LIST_PW, "LSParticle-WallInteraction", iFrame, time, iStep);
    writeBonds(iFrame, time, iStep);
    writeWalls(iFrame, time, iStep);
    std::cout << "[Solver] Output Completed" << std::endl;
  }

  void writeContacts(int list, const std::string& sub, size_t iFrame, double time, size_t iStep) {
    int64_t n = 0;
    if (ctx_)
      lsdem_get_contacts(ctx_, list, 0, &n, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    std::vector<double> pt(3 * n), nm(3 * n), sp(3 * n), en(n), es(n);
    if (n)
      lsdem_get_contacts(ctx_, list, n, &n, nullptr, nullptr, pt.data(), nm.data(), nullptr, sp.data(), en.data(),
                         es.data());
    std::vector<float> p(3 * n), nrm(3 * n), sd(3 * n), fe(n), fs(n);
    for (int64_t k = 0; k < n; ++k) {
      const double3 s = normalize(make_double3(sp[3 * k], sp[3 * k + 1], sp[3 * k + 2]));
      for (int a = 0; a < 3; ++a) {
        p[3 * k + a] = float(pt[3 * k + a]);
        nrm[3 * k + a] = float(nm[3 * k + a]);
      }
      sd[3 * k] = float(s.x); sd[3 * k + 1] = float(s.y); sd[3 * k + 2] = float(s.z);
      fe[k] = float(en[k]);
      fs[k] = float(es[k]);
    }
    std::vector<int32_t> conn, offs;
    std::vector<uint8_t> types;
    lsdem_vtu::vertexCells(size_t(n), conn, offs, types);
    using lsdem_vtu::array;
    emit(sub, "LSInteraction", iFrame, time, iStep, [&](const std::string& f) {
      lsdem_vtu::write(f, time, iStep, p, &conn, &offs, &types,
                       {array("contactNormal", "Float32", 3, nrm), array("slidingDirection", "Float32", 3, sd),
                        array("normalElasticEnergy", "Float32", 0, fe),
                        array("slidingElasticEnergy", "Float32", 0, fs)});
    });
  }

  void writeBonds(size_t iFrame, double time, size_t iStep) {
    const size_t n = bonds_.size();
    if (!n) return;
    std::vector<int32_t> act(n, 1);
    std::vector<double> pt(3 * n), en(4 * n), sm(n), tm(n);
    for (size_t k = 0; k < n; ++k) {
      pt[3 * k] = bonds_[k].point.x; pt[3 * k + 1] = bonds_[k].point.y; pt[3 * k + 2] = bonds_[k].point.z;
    }
    if (ctx_) lsdem_get_bonds(ctx_, act.data(), pt.data(), en.data(), sm.data(), tm.data());
    std::vector<float> p(3 * n), un(n), us(n), ub(n), ut(n), sn(n), ss(n);
    std::vector<int32_t> mid(n), sid(n);
    for (size_t k = 0; k < n; ++k) {
      for (int a = 0; a < 3; ++a) p[3 * k + a] = float(pt[3 * k + a]);
      mid[k] = bonds_[k].i;
      sid[k] = bonds_[k].j;
      un[k] = float(en[4 * k]); us[k] = float(en[4 * k + 1]); ub[k] = float(en[4 * k + 2]); ut[k] = float(en[4 * k + 3]);
      sn[k] = float(sm[k]);
      ss[k] = float(tm[k]);
    }
    std::vector<int32_t> conn, offs;
    std::vector<uint8_t> types;
    lsdem_vtu::vertexCells(n, conn, offs, types);
    using lsdem_vtu::array;
    emit("LSParticleInteraction", "VBondedInteraction", iFrame, time, iStep, [&](const std::string& f) {
      lsdem_vtu::write(f, time, iStep, p, &conn, &offs, &types,
                       {array("masterObjectID", "Int32", 0, mid), array("slaveObjectID", "Int32", 0, sid),
                        array("activated", "Int32", 0, act), array("normalElasticEnergy", "Float32", 0, un),
                        array("shearElasticEnergy", "Float32", 0, us), array("bendingElasticEnergy", "Float32", 0, ub),
                        array("torsionElasticEnergy", "Float32", 0, ut), array("maxNormalStress", "Float32", 0, sn),
                        array("maxShearStress", "Float32", 0, ss)},
                       "", "activated");
    });
  }

  void writeWalls(size_t iFrame, double time, size_t iStep) {
    if (walls_.empty()) return;
    std::vector<float> pts, vel;
    std::vector<int32_t> pid, conn, offs;
    std::vector<uint8_t> types;
    int base = 0;
    for (size_t i = 0; i < walls_.size(); ++i) {
      const Wall& s = walls_[i];
      for (const auto& vtx : s.verts) {
        const double3 xw = s.pos + rotateVectorByQuaternion(s.q, vtx);
        const double3 vw = s.vel + cross(s.angvel, xw - s.pos);
        pts.insert(pts.end(), {float(xw.x), float(xw.y), float(xw.z)});
        vel.insert(vel.end(), {float(vw.x), float(vw.y), float(vw.z)});
        pid.push_back(int32_t(i));
      }
      for (const auto& t : s.tris) {
        conn.insert(conn.end(), {t.x + base, t.y + base, t.z + base});
        offs.push_back(int32_t(conn.size()));
        types.push_back(5);
      }
      base += int(s.verts.size());
    }
    using lsdem_vtu::array;
    emit("Wall", "LSObjectMesh", iFrame, time, iStep, [&](const std::string& f) {
      lsdem_vtu::write(f, time, iStep, pts, &conn, &offs, &types,
                       {array("particleID", "Int32", 0, pid), array("velocity", "Float32", 3, vel)}, "velocity");
    });
  }
};
