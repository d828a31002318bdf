// C API of the oracle (test infrastructure only).  Mirrors include/lsdem.h with the prefix
// lsdem_oracle_ so that tests drive the GPU library and the oracle with the same calls.
#include <cstdio>
#include <string>

#include "../include/lsdem.h"
#include "lsdem_oracle.hpp"

using namespace lsdem_oracle;

namespace {
thread_local std::string g_err;
int fail(int code, const char* msg) {
  g_err = msg;
  return code;
}
struct Ctx {
  Solver s;
  bool initialized = false;
};
V3 rd3(const double* p, size_t i) { return v3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
void wr3(double* p, size_t i, V3 v) {
  p[3 * i] = v.x;
  p[3 * i + 1] = v.y;
  p[3 * i + 2] = v.z;
}
}  // namespace

extern "C" {

const char* lsdem_oracle_last_error(void) { return g_err.c_str(); }

int lsdem_oracle_create(const lsdem_config* cfg, void** out) {
  if (!cfg || !out) return fail(LSDEM_ERR_ARG, "null argument");
  Ctx* c = new Ctx();
  c->s.domMin = v3(cfg->domain_min[0], cfg->domain_min[1], cfg->domain_min[2]);
  c->s.domMax = v3(cfg->domain_max[0], cfg->domain_max[1], cfg->domain_max[2]);
  c->s.g = v3(cfg->gravity[0], cfg->gravity[1], cfg->gravity[2]);
  c->s.dt = cfg->dt;
  c->s.perX = cfg->periodic_x != 0;
  c->s.perY = cfg->periodic_y != 0;
  c->s.perSector = cfg->periodic_sector != 0;
  if (c->s.perSector && (c->s.perX || c->s.perY)) {
    delete c;
    return fail(LSDEM_ERR_ARG, "sector periodicity with x/y periodicity");
  }
  c->s.quirks = cfg->quirks;
  *out = c;
  return LSDEM_OK;
}

void lsdem_oracle_destroy(void* ctx) { delete static_cast<Ctx*>(ctx); }

int lsdem_oracle_set_literal(void* ctx, int literal) {
  static_cast<Ctx*>(ctx)->s.literal = literal != 0;
  return LSDEM_OK;
}

// Quarter-sector rotational periodicity about the z axis through the domain's lower corner
// (LSSolver::addPeriodicBoundarySector); not combined with x/y periodicity.
int lsdem_oracle_set_sector(void* ctx, int on) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (on && (c->s.perX || c->s.perY)) return fail(LSDEM_ERR_ARG, "sector periodicity with x/y periodicity");
  c->s.perSector = on != 0;
  return LSDEM_OK;
}

// OpenMP threads of the detection loops (1: serial).  Lists and counters do not depend on it.
int lsdem_oracle_set_threads(void* ctx, int threads) {
  if (threads < 1) return fail(LSDEM_ERR_ARG, "threads < 1");
  static_cast<Ctx*>(ctx)->s.threads = threads;
  return LSDEM_OK;
}

int lsdem_oracle_add_shape(void* ctx, const float* sdf, int32_t nx, int32_t ny, int32_t nz,
                           const double origin[3], double spacing, const double* nodes, int32_t n_nodes,
                           int32_t* shape_id) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (nx < 2 || ny < 2 || nz < 2) return fail(LSDEM_ERR_ARG, "grid size < 2");
  if (spacing <= 0.0) return fail(LSDEM_ERR_ARG, "spacing <= 0");
  int id = c->s.addShape(sdf, nx, ny, nz, v3(origin[0], origin[1], origin[2]), spacing, nodes, n_nodes);
  if (shape_id) *shape_id = id;
  return LSDEM_OK;
}

int lsdem_oracle_add_particles(void* ctx, int32_t n, const int32_t* shape, const double* scale,
                               const double* pos, const double* quat, const double* vel, const double* angvel,
                               const double* inv_mass, const double* inv_inertia, const double* kn,
                               const double* ks, const double* mu, const double* restitution) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (c->initialized) return fail(LSDEM_ERR_STATE, "already initialized");
  for (int i = 0; i < n; ++i) {
    if (shape[i] < 0 || shape[i] >= int(c->s.shapes.size())) return fail(LSDEM_ERR_ARG, "bad shape id");
    const double* I = inv_inertia + 6 * i;
    c->s.addBody(c->s.P, false, shape[i], scale ? scale[i] : 1.0, rd3(pos, i),
                 Q4{quat[4 * i], quat[4 * i + 1], quat[4 * i + 2], quat[4 * i + 3]}, rd3(vel, i), rd3(angvel, i),
                 inv_mass[i], S6{I[0], I[1], I[2], I[3], I[4], I[5]}, kn[i], ks[i], mu[i], restitution[i], 0.);
  }
  return LSDEM_OK;
}

int lsdem_oracle_add_walls(void* ctx, int32_t n, const int32_t* shape, const double* bound_radius,
                           const double* pos, const double* quat, const double* vel, const double* angvel,
                           const double* mu, const double* restitution) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (c->initialized) return fail(LSDEM_ERR_STATE, "already initialized");
  for (int i = 0; i < n; ++i) {
    if (shape[i] < 0 || shape[i] >= int(c->s.shapes.size())) return fail(LSDEM_ERR_ARG, "bad shape id");
    c->s.addBody(c->s.W, true, shape[i], 1.0, rd3(pos, i),
                 Q4{quat[4 * i], quat[4 * i + 1], quat[4 * i + 2], quat[4 * i + 3]}, rd3(vel, i), rd3(angvel, i),
                 0., S6{0, 0, 0, 0, 0, 0}, 0., 0., mu[i], restitution[i], bound_radius[i]);
  }
  return LSDEM_OK;
}

int lsdem_oracle_add_bond(void* ctx, int32_t i, int32_t j, const double point[3], const double normal[3],
                          double radius, double length, double youngs, double poisson, double tensile,
                          double cohesion, double friction) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c->s.addBond(i, j, v3(point[0], point[1], point[2]), v3(normal[0], normal[1], normal[2]), radius, length,
                    youngs, poisson, tensile, cohesion, friction))
    return fail(LSDEM_ERR_ARG, "bond rejected");
  return LSDEM_OK;
}

int lsdem_oracle_initialize(void* ctx) {
  Ctx* c = static_cast<Ctx*>(ctx);
  c->s.initialize();
  c->initialized = true;
  return LSDEM_OK;
}

int lsdem_oracle_step(void* ctx, int32_t n) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c->initialized) return fail(LSDEM_ERR_STATE, "not initialized");
  for (int i = 0; i < n; ++i) {
    c->s.counters = Counters();
    c->s.step();
  }
  return LSDEM_OK;
}

// The pre-loop build of LSSolver::compute* now (lsdem_begin_phase).
int lsdem_oracle_begin_phase(void* ctx) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c->initialized) return fail(LSDEM_ERR_STATE, "not initialized");
  c->s.counters = Counters();
  c->s.prime();
  return LSDEM_OK;
}

// Ghost images of set g (0: +x, 1: +y, 2: +x+y) of the last detection: positions (n x 3) and
// keys (-1: the image does not exist), PeriodicBoundary.h:215-289.  n must be the particle count.
int lsdem_oracle_ghosts(void* ctx, int32_t set, int32_t n, double* pos, int32_t* key) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (set < 0 || set > 2) return fail(LSDEM_ERR_ARG, "bad ghost set");
  const Ghost& G = c->s.ghost[set];
  if (n != int(c->s.P.n()) || int(G.pos.size()) != n) return fail(LSDEM_ERR_STATE, "no ghost images of that size");
  for (int i = 0; i < n; ++i) {
    if (pos) wr3(pos, i, G.pos[i]);
    if (key) key[i] = G.key[i];
  }
  return LSDEM_OK;
}

int lsdem_oracle_get_particles(void* ctx, double* pos, double* quat, double* vel, double* angvel, double* force,
                               double* torque) {
  Ctx* c = static_cast<Ctx*>(ctx);
  const Bodies& P = c->s.P;
  for (size_t i = 0; i < P.n(); ++i) {
    if (pos) wr3(pos, i, P.pos[i]);
    if (vel) wr3(vel, i, P.vel[i]);
    if (angvel) wr3(angvel, i, P.angvel[i]);
    if (force) wr3(force, i, P.force[i]);
    if (torque) wr3(torque, i, P.torque[i]);
    if (quat) {
      quat[4 * i] = P.q[i].q0;
      quat[4 * i + 1] = P.q[i].q1;
      quat[4 * i + 2] = P.q[i].q2;
      quat[4 * i + 3] = P.q[i].q3;
    }
  }
  return LSDEM_OK;
}

int lsdem_oracle_set_particles(void* ctx, const double* pos, const double* quat, const double* vel,
                               const double* angvel, const double* force, const double* torque) {
  Ctx* c = static_cast<Ctx*>(ctx);
  Bodies& P = c->s.P;
  for (size_t i = 0; i < P.n(); ++i) {
    if (pos) P.pos[i] = rd3(pos, i);
    if (vel) P.vel[i] = rd3(vel, i);
    if (angvel) P.angvel[i] = rd3(angvel, i);
    if (force) P.force[i] = rd3(force, i);
    if (torque) P.torque[i] = rd3(torque, i);
    if (quat) P.q[i] = Q4{quat[4 * i], quat[4 * i + 1], quat[4 * i + 2], quat[4 * i + 3]};
  }
  return LSDEM_OK;
}

int lsdem_oracle_get_walls(void* ctx, double* pos, double* quat, double* vel, double* angvel) {
  Ctx* c = static_cast<Ctx*>(ctx);
  const Bodies& W = c->s.W;
  for (size_t i = 0; i < W.n(); ++i) {
    if (pos) wr3(pos, i, W.pos[i]);
    if (vel) wr3(vel, i, W.vel[i]);
    if (angvel) wr3(angvel, i, W.angvel[i]);
    if (quat) {
      quat[4 * i] = W.q[i].q0;
      quat[4 * i + 1] = W.q[i].q1;
      quat[4 * i + 2] = W.q[i].q2;
      quat[4 * i + 3] = W.q[i].q3;
    }
  }
  return LSDEM_OK;
}

int lsdem_oracle_set_walls(void* ctx, const double* pos, const double* quat, const double* vel,
                           const double* angvel) {
  Ctx* c = static_cast<Ctx*>(ctx);
  Bodies& W = c->s.W;
  for (size_t i = 0; i < W.n(); ++i) {
    if (pos) W.pos[i] = rd3(pos, i);
    if (vel) W.vel[i] = rd3(vel, i);
    if (angvel) W.angvel[i] = rd3(angvel, i);
    if (quat) W.q[i] = Q4{quat[4 * i], quat[4 * i + 1], quat[4 * i + 2], quat[4 * i + 3]};
  }
  return LSDEM_OK;
}

int lsdem_oracle_get_contacts(void* ctx, int32_t list, int64_t capacity, int64_t* count, int32_t* master,
                              int32_t* slave, double* point, double* normal, double* overlap, double* spring,
                              double* energy_n, double* energy_s) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (list < 0 || list >= NUM_LISTS) return fail(LSDEM_ERR_ARG, "bad list id");
  std::vector<Contact> l = c->s.lists[list];
  std::sort(l.begin(), l.end(), [](const Contact& a, const Contact& b) {
    return a.master != b.master ? a.master < b.master : a.slave < b.slave;
  });
  if (count) *count = int64_t(l.size());
  const int64_t m = std::min<int64_t>(capacity, int64_t(l.size()));
  for (int64_t k = 0; k < m; ++k) {
    const Contact& x = l[size_t(k)];
    if (master) master[k] = x.master;
    if (slave) slave[k] = x.slave;
    if (point) wr3(point, size_t(k), x.point);
    if (normal) wr3(normal, size_t(k), x.normal);
    if (overlap) overlap[k] = x.overlap;
    if (spring) wr3(spring, size_t(k), x.spring);
    if (energy_n) energy_n[k] = x.En;
    if (energy_s) energy_s[k] = x.Es;
  }
  return LSDEM_OK;
}

int lsdem_oracle_get_bonds(void* ctx, int32_t* active, double* point, double* energies4, double* sigma_max,
                           double* tau_max) {
  Ctx* c = static_cast<Ctx*>(ctx);
  for (size_t k = 0; k < c->s.bonds.size(); ++k) {
    const Bond& b = c->s.bonds[k];
    if (active) active[k] = b.active;
    if (point) wr3(point, k, b.point);
    if (energies4) {
      energies4[4 * k] = b.Un;
      energies4[4 * k + 1] = b.Us;
      energies4[4 * k + 2] = b.Ub;
      energies4[4 * k + 3] = b.Ut;
    }
    if (sigma_max) sigma_max[k] = b.sigmaMax;
    if (tau_max) tau_max[k] = b.tauMax;
  }
  return LSDEM_OK;
}

int lsdem_oracle_get_counters(void* ctx, lsdem_counters* out) {
  Ctx* c = static_cast<Ctx*>(ctx);
  std::memset(out, 0, sizeof(*out));
  out->steps = c->s.stepCount;
  out->pairs = c->s.counters.pairs;
  out->node_tests = c->s.counters.nodeTests;
  out->sdf_queries = c->s.counters.sdfQueries;
  out->contacts = c->s.counters.contacts;
  return LSDEM_OK;
}

// Reference grid geometry chosen at initialize (spatialGrid::set): min[3], max[3], inv[3], size[3].
int lsdem_oracle_grid(void* ctx, int which, double* mn, double* mx, double* inv, int32_t* size) {
  Ctx* c = static_cast<Ctx*>(ctx);
  const Grid& g = which == 0 ? c->s.gridP : c->s.gridW;
  mn[0] = g.mn.x; mn[1] = g.mn.y; mn[2] = g.mn.z;
  mx[0] = g.mx.x; mx[1] = g.mx.y; mx[2] = g.mx.z;
  inv[0] = g.inv.x; inv[1] = g.inv.y; inv[2] = g.inv.z;
  size[0] = g.size.x; size[1] = g.size.y; size[2] = g.size.z;
  return LSDEM_OK;
}

// Keys + stable sort + cell start/end of n points on a grid (buildHashStartEnd.cu:69-137).
int lsdem_oracle_hash_cells(const double* mn, const double* mx, double cell, int32_t n, const double* pos,
                            int32_t* keys, int32_t* sorted_idx, int32_t ncell_cap, int32_t* start, int32_t* end,
                            int32_t* ncell) {
  Grid g;
  g.set(v3(mn[0], mn[1], mn[2]), v3(mx[0], mx[1], mx[2]), cell);
  std::vector<int> k(n);
  for (int i = 0; i < n; ++i) k[i] = g.key(rd3(pos, i));
  Solver::CellTable t = Solver::buildCells(k, g.ncell());
  *ncell = g.ncell();
  if (g.ncell() > ncell_cap) return fail(LSDEM_ERR_ARG, "ncell_cap too small");
  for (int i = 0; i < n; ++i) {
    keys[i] = k[i];
    sorted_idx[i] = t.sortedIdx[i];
  }
  for (int h = 0; h < g.ncell(); ++h) {
    start[h] = t.start[h];
    end[h] = t.end[h];
  }
  return LSDEM_OK;
}

int lsdem_oracle_mass_properties(const double* sdf, int32_t nx, int32_t ny, int32_t nz, const double origin[3],
                                 double spacing, double density, double* mass, double centroid[3], double* inertia6,
                                 double* inv_inertia6) {
  MassProps m = massFromSdf(sdf, nx, ny, nz, v3(origin[0], origin[1], origin[2]), spacing, density);
  *mass = m.mass;
  centroid[0] = m.centroid.x; centroid[1] = m.centroid.y; centroid[2] = m.centroid.z;
  if (inertia6) {
    const S6& I = m.inertia;
    double v[6] = {I.xx, I.yy, I.zz, I.xy, I.xz, I.yz};
    std::memcpy(inertia6, v, sizeof v);
  }
  if (inv_inertia6) {
    const S6& I = m.invInertia;
    double v[6] = {I.xx, I.yy, I.zz, I.xy, I.xz, I.yz};
    std::memcpy(inv_inertia6, v, sizeof v);
  }
  return LSDEM_OK;
}

// Single-function probes used to pin the restated math against closed forms.
int lsdem_oracle_linear_contact(const double vc[3], const double n[3], double delta, double dt, double kn,
                                double ks, double mu, double res, double effM, double spring_io[3],
                                double force_out[3]) {
  V3 F = v3(0, 0, 0);
  V3 s = v3(spring_io[0], spring_io[1], spring_io[2]);
  Solver::linearContact(F, s, v3(vc[0], vc[1], vc[2]), v3(n[0], n[1], n[2]), delta, dt, kn, ks, mu, res, effM);
  spring_io[0] = s.x; spring_io[1] = s.y; spring_io[2] = s.z;
  force_out[0] = F.x; force_out[1] = F.y; force_out[2] = F.z;
  return LSDEM_OK;
}

int lsdem_oracle_interp(const double corners[8], double x, double y, double z, int int_x_quirk, double* value,
                        double grad[3]) {
  *value = interpValue(x, y, z, corners);
  V3 gr = int_x_quirk ? interpGradient(int(x), y, z, corners) : interpGradientExact(x, y, z, corners);
  grad[0] = gr.x; grad[1] = gr.y; grad[2] = gr.z;
  return LSDEM_OK;
}

int lsdem_oracle_set_external_force(void* ctx, lsdem_external_force_fn fn, void* user) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!fn) {
    c->s.external = nullptr;
    return LSDEM_OK;
  }
  c->s.external = [c, fn, user](double time) {
    Bodies& P = c->s.P;
    lsdem_device_view v;
    v.force = reinterpret_cast<double*>(P.force.data());
    v.torque = reinterpret_cast<double*>(P.torque.data());
    v.velocity = reinterpret_cast<double*>(P.vel.data());
    v.angular_velocity = reinterpret_cast<double*>(P.angvel.data());
    v.position = reinterpret_cast<const double*>(P.pos.data());
    v.quaternion = reinterpret_cast<const double*>(P.q.data());
    v.num = int64_t(P.n());
    v.block_dim = 256;
    v.grid_dim = (v.num + 255) / 256;
    v.stream = nullptr;
    fn(user, time, &v);
  };
  return LSDEM_OK;
}

int lsdem_oracle_get_time(void* ctx, double* time) {
  *time = static_cast<Ctx*>(ctx)->s.time;
  return LSDEM_OK;
}

int lsdem_oracle_set_time(void* ctx, double time) {
  static_cast<Ctx*>(ctx)->s.time = time;
  return LSDEM_OK;
}

}  // extern "C"
