/*
 * lsdem.h -- C ABI of the MI355X level-set DEM timestep library (liblsdem.so).
 *
 * One context = one simulation on one GPU and one HIP stream.  The context owns all device
 * memory; every host array passed in is copied.  Calls are not re-entrant per context and
 * the library starts no threads.  Every function returns LSDEM_OK (0) or an LSDEM_ERR_*
 * code; lsdem_last_error() gives the message of the last failure on the calling thread.
 *
 * Units and layouts: SI units; double3 values are 3 consecutive doubles (x, y, z);
 * quaternions are 4 doubles (q0 = scalar, q1, q2, q3); symmetric 3x3 tensors are 6 doubles
 * (xx, yy, zz, xy, xz, yz); SDF grids are float with x fastest (index ix + nx*(iy + ny*iz)).
 *
 * Each entry point names the interface of the reference (FunDEM, kaiqideng/LSDEM2.0) that it
 * replaces; paths are relative to the reference checkout.
 */
#ifndef LSDEM_H
#define LSDEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LSDEM_OK 0
#define LSDEM_ERR_ARG 1      /* invalid argument */
#define LSDEM_ERR_STATE 2    /* call not allowed in the current context state */
#define LSDEM_ERR_HIP 3      /* HIP runtime error */
#define LSDEM_ERR_CAPACITY 4 /* pair or contact buffer overflow; raise the capacities */
#define LSDEM_ERR_NOMEM 5

/* Reference behaviours kept by default (all on).  Each bit can be switched off. */
#define LSDEM_QUIRK_INT_X_GRADIENT 1  /* LSParticleContactDetectionKernel.cuh:34 `const int x` */
#define LSDEM_QUIRK_GHOST_Z_VS_MAXY 2 /* mySpatialGrid.cu:236 z tested against maxBound.y */
#define LSDEM_QUIRK_WRAP_AFTER_BUILD 4 /* LSSolver.h:1071-1074 wrap after the contact build */
#define LSDEM_QUIRK_ALL 7

/* Contact lists (SolidInteraction instances of the reference). */
#define LSDEM_LIST_PP 0  /* particle-particle, LSSolver.h:1136 LSParticleInteraction_ */
#define LSDEM_LIST_PW 1  /* particle-wall, LSSolver.h:1137 LSParticleWallInteraction_ */
#define LSDEM_LIST_GX 2  /* periodic ghost +x, PeriodicBoundary.h:306 */
#define LSDEM_LIST_GY 3  /* periodic ghost +y, PeriodicBoundary.h:307 */
#define LSDEM_LIST_GXY 4 /* periodic ghost +x+y, PeriodicBoundary.h:308 */

typedef struct lsdem_ctx lsdem_ctx;

typedef struct lsdem_config {
  double domain_min[3];   /* solve(minDomain, ...), LSSolver.h:478 */
  double domain_max[3];
  double gravity[3];
  double dt;
  int32_t periodic_x;     /* addPeriodicBoundaryXD, LSSolver.h:426 */
  int32_t periodic_y;     /* addPeriodicBoundaryYD, LSSolver.h:437 */
  int32_t quirks;         /* LSDEM_QUIRK_* bitmask */
  int32_t block_size;     /* threads per block of per-body kernels (maxGPUThread), 0 = 256 */
  int64_t pair_capacity;  /* candidate pairs per step, 0 = automatic */
  int64_t contact_capacity; /* contacts per step, 0 = automatic */
  int32_t device;         /* HIP device ordinal */
  int32_t no_sdf_cull;    /* 1: cull node clusters by bounding spheres / grid boxes only (no SDF
                             bound); for measuring what the SDF and surrogate culls save */
  double skin;            /* candidate pairs are kept for reuse within this distance and rebuilt
                             when a body has moved half of it (< 0: 8 % of the largest bounding
                             radius; 0: rebuilt every step).  Contacts do not depend on it. */
  int32_t no_certificates; /* 1: evaluate every candidate pair every step (no separation
                              certificates); contacts do not depend on it */
  int32_t broad_phase_launches; /* 1: run the broad phase as separate launches gated on the
                              device instead of one launch with grid barriers; results do not
                              depend on it */
  int32_t broad_phase_blocks; /* grid of the single-launch broad phase, blocks of 1024 threads (0:
                              one per 512 particles or per 16,384 search cells, at least 8); capped
                              at the CU count */
  int32_t narrow_batch;   /* candidate pairs per wavefront batch of the narrow phase: 1, 2, 4 or 8
                             (0: chosen each step from the number of pairs it evaluates) */
  int32_t no_fused_steps; /* 1: launch each step's kick/drift and its closing sums separately
                             instead of one launch per step boundary; results do not depend on it */
  int32_t no_recovery;    /* 1: a call that runs out of pair or contact room (or whose broad-phase
                             grid barrier gives up) returns the error with the state undefined;
                             0: the call is restored to its start, the buffers grown (or the broad
                             phase switched to separate launches) and the call run again, with
                             results identical to a call that had the room from the start */
  int32_t barrier_spin;   /* polls of a broad-phase grid barrier before it gives up (0: ~1 s; -1:
                             none, which makes the barriers give up: a test of the recovery) */
  int32_t broad_phase_coop; /* 1: launch the single-launch broad phase with hipLaunchCooperativeKernel
                             (the runtime checks that its blocks are resident together); 0: a plain
                             launch sized by the occupancy query, the bounded barriers and the
                             recovery covering the case where they are not */
  int32_t periodic_sector; /* 1: quarter-sector rotational periodicity about the z axis through
                             domain_min (LSSolver::addPeriodicBoundarySector, PeriodicBoundary.h:313-420):
                             a body that leaves across x = min (y = min) turns by -90 (+90) degrees
                             about that axis; images turned by +90, 180 and -90 degrees fill the lists
                             LSDEM_LIST_GX, _GY and _GXY.  Not combined with periodic_x / periodic_y. */
  int32_t narrow_blocks;  /* workgroups of the candidate kernel, 256 threads each (0: one per 2 bodies up to
                             8 per CU); a tuning knob, results do not depend on it */
} lsdem_config;

typedef struct lsdem_counters {
  int64_t steps;          /* steps taken */
  int64_t pairs;          /* candidate (master particle, partner) pairs, last step */
  int64_t node_tests;     /* nodes tested against a partner after cluster culling, last step */
  int64_t sdf_queries;    /* 8-corner SDF interpolations, last step */
  int64_t contacts;       /* contacts written (all lists), last step */
  int64_t pair_capacity;
  int64_t contact_capacity;
  int64_t recoveries;     /* calls run again after growing a buffer or switching the broad phase */
  int64_t reserved[8];
} lsdem_counters;

/* Fill *cfg with defaults (quirks all on, gravity 0, dt 1e-5, automatic capacities). */
void lsdem_config_default(lsdem_config* cfg);

/* Create a context on cfg->device.  Replaces LSSolver::LSSolver + activateGPUDevice
 * (LSSolver.h:14, :521-533); a missing device is an error return, not exit(1). */
int lsdem_create(const lsdem_config* cfg, lsdem_ctx** out);
void lsdem_destroy(lsdem_ctx* ctx);
const char* lsdem_last_error(void);

/* Register a level-set shape: SDF grid plus boundary nodes in the shape's local frame.
 * Bodies reference shapes by id with a per-body scale s (geometry s*x, SDF values s*phi), so
 * identical grids are stored once.  Replaces the per-particle SDF copies of
 * LSParticle::add (LSParticle.h:115-212). */
int lsdem_add_shape(lsdem_ctx* ctx, const float* sdf, int32_t nx, int32_t ny, int32_t nz,
                    const double origin[3], double spacing, const double* nodes, int32_t n_nodes,
                    int32_t* shape_id);

/* Append n particles.  Same clamps as LSParticle::add (LSParticle.h:188-202): negative kn,
 * ks, mu become 0, restitution outside (0,1] becomes 1, q is normalised, bounding radius
 * = max |s * node|.  inv_inertia is the local-frame inverse inertia tensor. */
int lsdem_add_particles(lsdem_ctx* ctx, int32_t n, const int32_t* shape, const double* scale,
                        const double* pos, const double* quat, const double* vel,
                        const double* angvel, const double* inv_mass, const double* inv_inertia,
                        const double* kn, const double* ks, const double* mu,
                        const double* restitution);

/* Append n walls (immovable level-set bodies with prescribed v, w).  bound_radius is the
 * largest |vertex| of the wall mesh, which sets the wall cell size as in LSSolver::addWall
 * (LSSolver.h:245-272) + LSParticle::initialize (LSParticle.h:253). */
int lsdem_add_walls(lsdem_ctx* ctx, int32_t n, const int32_t* shape, const double* bound_radius,
                    const double* pos, const double* quat, const double* vel, const double* angvel,
                    const double* mu, const double* restitution);

/* Add a parallel ("V") bond between particles i and j.  Replaces VBondedInteraction::add
 * (VBondedInteraction.h:74-195) via LSSolver::addBondedInteraction (LSSolver.h:344-372).
 * Uses the particles' current pose.  Returns LSDEM_ERR_ARG for the inputs the reference
 * rejects (zero normal, r <= 0, L <= 0, E <= 0, nu outside (-1, 0.5), ids out of range). */
int lsdem_add_bond(lsdem_ctx* ctx, int32_t i, int32_t j, const double point[3],
                   const double normal[3], double radius, double length, double youngs,
                   double poisson, double tensile_strength, double cohesion, double friction);

/* Upload and size device buffers.  Replaces LSSolver::upload (LSSolver.h:612-623). */
int lsdem_initialize(lsdem_ctx* ctx);

/* Advance n_steps timesteps on the context's stream; returns after the work is complete.
 * One step is the loop body of LSSolver::compute* (LSSolver.h:1000-1020 / :1104-1128):
 * half-kick, drift, broad phase, narrow phase with tangential history, contact and bond
 * forces, per-body reduction, half-kick.  Replaces the launchParticle*Integration,
 * launchUpdateSpatialGridHashStartEnd, launchBuildLevelSetBoundaryNode*Interactions1st/2nd,
 * buildPrefixSum, launchAdd*ForceTorque and launchCalculateGhost* launchers
 * (particleIntegrationKernel.cu:42,72; mySpatialGrid.h:33; LSParticleContactDetectionKernel.cuh:79-215;
 * buildHashStartEnd.cu:69,139; contactKernel.cuh:201-333). */
int lsdem_step(lsdem_ctx* ctx, int32_t n_steps);
/* The pre-loop build that opens each LSSolver::compute* call (LSSolver.h:996-998, 1029-1031,
 * 1061-1066, 1097-1104): contact lists at the current poses with springs carried over, then the
 * periodic wrap and the ghost images; no integration, no forces, no step counted.  It runs by
 * itself before the first lsdem_step after lsdem_initialize; the drop-in's solve() calls it at
 * the start of every phase.  Contacts are readable after it (the reference's frame 0). */
int lsdem_begin_phase(lsdem_ctx* ctx);
/* The same n_steps (even, >= 2) replayed from a HIP graph captured on the first call with that
 * count: one launch per call instead of one per kernel, for simulations small enough to be
 * launch-bound.  Results are identical to lsdem_step.  Not with an external-force hook, profiling
 * or the surrogate (LSDEM_ERR_STATE); a pending rebuild request runs the call eagerly.  No
 * counterpart in the reference. */
int lsdem_step_graph(lsdem_ctx* ctx, int32_t n_steps);

/* One step split into its phases, for tests and tools that inspect what a phase produced.  Called
 * in this order, the six make one step with the same results as lsdem_step(ctx, 1); a call out of
 * order fails with LSDEM_ERR_STATE and so do lsdem_step, lsdem_step_graph and lsdem_begin_phase
 * while a step driven this way is unfinished.  Each call waits for its launches and reports the
 * device's error flags (a capacity overflow abandons the step: LSDEM_ERR_CAPACITY, state undefined,
 * buffers not grown).  A step driven this way rebuilds its cell table and candidate pairs (a step
 * of lsdem_step reuses them within the skin; contacts and forces are the same either way).
 *   integrate: first half-kick, drift, periodic wrap, ghost images, cell keys
 *              (particleIntegrationKernel.cu:42; mySpatialGrid.cu:58-240; the pre-loop build first
 *              when one is pending, LSSolver.h:1061-1066);
 *   hash:      search cell of every body, counting sort by cell, cell start table
 *              (launchUpdateSpatialGridHashStartEnd, buildHashStartEnd.cu:69-139);
 *   pairs:     candidate pairs (bodies, walls, images) and per-body reaction lists
 *              (LSParticleContactDetectionKernel.cuh:79-215, count pass);
 *   narrow:    pair check, node-vs-SDF tests, contact law with tangential history
 *              (LSParticleContactDetectionKernel.cu:58-115; contactKernel.cuh:4-77);
 *   forces:    parallel-bond forces (contactKernel.cuh:138-199, contactKernel.cu:343-350);
 *   reduce:    per-body sums, second half-kick, external-force hook (contactKernel.cu:10-44;
 *              particleIntegrationKernel.cu:72); ends the step. */
int lsdem_phase_integrate(lsdem_ctx* ctx);
int lsdem_phase_hash(lsdem_ctx* ctx);
int lsdem_phase_pairs(lsdem_ctx* ctx);
int lsdem_phase_narrow(lsdem_ctx* ctx);
int lsdem_phase_forces(lsdem_ctx* ctx);
int lsdem_phase_reduce(lsdem_ctx* ctx);

/* The search grid and cell table of the last broad phase (the reference's hash / start / end
 * arrays, mySpatialGrid.h:33, buildHashStartEnd.cu:69-139, laid out as this library's grid):
 * origin and edge of the cubic cells, cells per axis; cell_of[i] the cell of body i's wrapped centre,
 * clamped to the grid; sorted the bodies by (cell, index); cell_start[c] .. cell_start[c + 1] the
 * range of cell c in sorted (ncell + 1 entries).  NULL arrays are skipped; cell_start needs
 * start_capacity >= ncell + 1 (else LSDEM_ERR_CAPACITY). */
int lsdem_get_cell_table(lsdem_ctx* ctx, double origin[3], double* cell_edge, int32_t dims[3],
                         int32_t* cell_of, int32_t* sorted, int64_t start_capacity, int32_t* cell_start);
/* Candidate pairs of the last broad phase, grouped by master body in ascending order: master,
 * partner (a body, a wall, or the source body of a periodic image) and type (0 body, 1 wall,
 * 2/3/4 the +x/+y/+xy images or, with sector periodicity, the +90/180/-90 degree images).  Writes at
 * most capacity records; *count receives the total.  Counterpart of the reference's per-node
 * candidate counts (LSParticleContactDetectionKernel.cuh:79-140), kept per body pair here. */
int lsdem_get_pairs(lsdem_ctx* ctx, int64_t capacity, int64_t* count, int32_t* master, int32_t* partner,
                    int32_t* type);

/* Copy particle state to / from the host; NULL pointers are skipped.  force/torque are the
 * totals of the last step.  Replaces LSParticle::finalize / copyHostToDevice
 * (LSParticle.h:699-743). */
int lsdem_get_particles(lsdem_ctx* ctx, double* pos, double* quat, double* vel, double* angvel,
                        double* force, double* torque);
int lsdem_set_particles(lsdem_ctx* ctx, const double* pos, const double* quat, const double* vel,
                        const double* angvel, const double* force, const double* torque);
int lsdem_get_walls(lsdem_ctx* ctx, double* pos, double* quat, double* vel, double* angvel);
int lsdem_set_walls(lsdem_ctx* ctx, const double* pos, const double* quat, const double* vel,
                    const double* angvel);

/* Contacts of one list from the last step, sorted by (master node, slave).  Writes at most
 * capacity records; *count receives the total.  NULL arrays are skipped.  Replaces the
 * SolidInteraction device arrays and their host copies (SolidInteraction.h:319-389). */
int lsdem_get_contacts(lsdem_ctx* ctx, int32_t list, int64_t capacity, int64_t* count,
                       int32_t* master, int32_t* slave, double* point, double* normal,
                       double* overlap, double* spring, double* energy_n, double* energy_s);

/* Bond state (VBondedInteraction.h device arrays): activated flag, centre point, Un, Us, Ub,
 * Ut, max normal stress, max shear stress.  NULL arrays are skipped. */
int lsdem_get_bonds(lsdem_ctx* ctx, int32_t* active, double* point, double* energies4,
                    double* sigma_max, double* tau_max);

int lsdem_get_counters(lsdem_ctx* ctx, lsdem_counters* out);

/* Per-phase timing of lsdem_step with HIP events recorded on the context's stream.  Phases:
 * 0 integrate (half-kick + drift + walls), 1 broad-phase sort, 2 candidate pairs,
 * 3 reaction lists, 4 pair check (liveness, separation certificates, map records of the
 * evaluated pairs), 5 node culls + exact test + contact law + history table (and the surrogate's
 * masks and audits), 6 bonds, 7 reduction + second half-kick.  With the single-launch broad phase
 * (the default) phases 1-3 are one launch, reported under 1.  When consecutive
 * steps of one call share a launch at the step boundary (the default), the next step's kick and
 * drift are reported under 7 and phase 0 holds only a call's first step.  Work counters are summed
 * over the profiled steps.
 * No counterpart in the reference (it has no timers). */
#define LSDEM_NUM_PHASES 8
typedef struct lsdem_profile {
  int64_t steps;               /* profiled steps */
  double ms[LSDEM_NUM_PHASES]; /* summed GPU time per phase */
  int64_t pairs, node_tests, sdf_queries, contacts; /* sums over the profiled steps */
  int64_t cluster_tests;       /* node clusters passed to the node prefilter, summed */
  int64_t master_nodes;        /* boundary nodes of the masters of the evaluated pairs, summed */
  int64_t evaluated_pairs;     /* pairs the narrow phase evaluated (the rest held a separation
                                  certificate or had no partner), summed */
  int64_t rebuilds;            /* steps that rebuilt the candidate pairs */
  int64_t sampled_steps;       /* steps whose phases were timed (ms[] sums over these) */
  int64_t reserved[3];
} lsdem_profile;
/* Start recording; at most max_steps steps are timed (events are preallocated).  `steps` and the
 * work counters cover every step until lsdem_profile_end. */
int lsdem_profile_begin(lsdem_ctx* ctx, int32_t max_steps);
/* The same, timing only every `every`-th step (the first one after this call, then every
 * every-th), at most max_samples of them: the event records of a timed step cost GPU time of their
 * own, which sampling spreads thin.  No counterpart in the reference. */
int lsdem_profile_begin_sampled(lsdem_ctx* ctx, int32_t max_samples, int32_t every);
/* Stop recording, wait for the work and return the totals. */
int lsdem_profile_end(lsdem_ctx* ctx, lsdem_profile* out);

/* Device arrays of the particles, as the reference's apps reach them through
 * getLSParticle().force() / torque() / velocity() / angularVelocity() and num_device(),
 * gridDim(), blockDim() (LSParticle.h:661-670): N x 3 doubles each (AoS, the layout of double3),
 * on the context's device, in the caller's particle order. */
typedef struct lsdem_device_view {
  double* force;
  double* torque;
  double* velocity;
  double* angular_velocity;
  const double* position;
  const double* quaternion;      /* N x 4, scalar first */
  int64_t num;                   /* num_device() */
  int64_t grid_dim, block_dim;   /* one thread per particle at the configured block size */
  void* stream;                  /* hipStream_t the step runs on */
} lsdem_device_view;

/* Device arrays of the walls, as apps reach them through getWall() (LSSolver.h:512) and its
 * LSParticle accessors position() / orientation() / velocity() / angularVelocity() and
 * num_device(), gridDim(), blockDim() (LSParticle.h:660-671): W x 3 doubles (W x 4, scalar first,
 * for the orientation), on the context's device, in the caller's wall order.  Walls are kinematic:
 * they carry no force or torque.  A velocity or angular velocity written here (on the stream, from a
 * hook) moves the wall from the next drift on. */
typedef struct lsdem_wall_device_view {
  const double* position;
  const double* quaternion;
  double* velocity;
  double* angular_velocity;
  int64_t num;
  int64_t grid_dim, block_dim;
  void* stream;
} lsdem_wall_device_view;
int lsdem_get_wall_device_view(lsdem_ctx* ctx, lsdem_wall_device_view* out);

/* External-force hook: replaces the virtual LSSolver::addExternalForceTorque(double time)
 * (LSSolver.h:508; called at :1009,1042,1079,1117 after all contact and bond forces and before
 * the second half-kick).  `time` is the simulation time at the start of the step, accumulated as
 * the reference does (time += dt per step).  The hook enqueues its work on view->stream or the
 * null stream; it must not synchronise the device.  fn = NULL removes the hook. */
typedef void (*lsdem_external_force_fn)(void* user, double time, const lsdem_device_view* view);
int lsdem_set_external_force(lsdem_ctx* ctx, lsdem_external_force_fn fn, void* user);
/* The same views outside a step (after lsdem_initialize). */
int lsdem_get_device_view(lsdem_ctx* ctx, lsdem_device_view* out);
/* Simulation time (sum of dt over the steps taken); set_time restarts the clock, as each
 * LSSolver::solve phase does (compute(..., time = 0.), LSSolver.h:994-996). */
int lsdem_get_time(lsdem_ctx* ctx, double* time);
int lsdem_set_time(lsdem_ctx* ctx, double time);

/* ---- Learned culling surrogate (no counterpart in the reference) ----
 * A Fourier-feature MLP per shape family (every shape used by a particle): 16 sin/cos frequencies
 * of the normalised local point (96 features) + a 16-d learned shape embedding -> 4 x 64 SiLU ->
 * 1 distance.  BF16 weights, FP32 accumulation (MFMA).  When enabled, node clusters whose
 * surrogate distance exceeds their Lipschitz reach plus the band (band_factor x the maximum
 * validation error) are culled before any node is transformed; everything else takes the exact
 * grid path.  Every audit_every steps the step runs the exact path and checks that each contact
 * lies in a cluster the surrogate keeps (missed_contacts must stay 0); pairs skipped under a
 * separation certificate of the exact culls alone hold no contact and are not re-evaluated. */
typedef struct lsdem_surrogate_params {
  int32_t iterations;      /* optimizer steps (3000) */
  int32_t batch;           /* samples per step, a multiple of 64 (4096) */
  int64_t samples;         /* training samples drawn from the grids (2,000,000) */
  int64_t validation;      /* validation samples for the error bound (262,144) */
  double lr, lr_final;     /* Adam step size, cosine decay (2e-3 -> 1e-4) */
  double clamp;            /* |target| clamp in units of the family length scale (0.6) */
  double band_factor;      /* band = band_factor x max validation error (2) */
  int32_t audit_every;     /* exact audit period in steps, 0 = never (100) */
  int32_t pad;
  uint64_t seed;
  int64_t reserved[8];
} lsdem_surrogate_params;
typedef struct lsdem_surrogate_report {
  int32_t trained, enabled, n_shapes, n_params;
  int64_t iterations;
  double final_loss;       /* training MSE of the last step (normalised units) */
  double max_error;        /* max |surrogate - clamped SDF| on the validation set (unscaled SDF units) */
  double mean_error;
  double band;             /* cull band (unscaled SDF units) */
  double train_ms;         /* GPU time of the optimizer steps */
  double length_scale;     /* family length scale (1 / input normalisation) */
  int64_t audits, audited_contacts, missed_contacts;
  int64_t reserved[8];
} lsdem_surrogate_report;
void lsdem_surrogate_params_default(lsdem_surrogate_params* p);
/* Train on the GPU (after lsdem_initialize); the network is then available to the step. */
int lsdem_surrogate_train(lsdem_ctx* ctx, const lsdem_surrogate_params* p);
int lsdem_surrogate_enable(lsdem_ctx* ctx, int32_t on);
int lsdem_surrogate_set_band(lsdem_ctx* ctx, double band, int32_t audit_every);
int lsdem_surrogate_get_report(lsdem_ctx* ctx, lsdem_surrogate_report* r);
/* FP32 parameters: W1 (112x64, row = input), b1, W2, b2, W3, b3, W4, b4 (64x64, 64), w5 (64),
 * b5, then the embeddings (n_shapes x 16, family slots in shape-id order). */
int lsdem_surrogate_get_weights(lsdem_ctx* ctx, float* w, int64_t count);
int lsdem_surrogate_set_weights(lsdem_ctx* ctx, const float* w, int64_t count);
/* Surrogate SDF (unscaled units) at unscaled local points of the given shapes. */
int lsdem_surrogate_eval(lsdem_ctx* ctx, int64_t n, const float* points, const int32_t* shape_ids, float* out);
/* Mean-squared-error loss against normalised targets and its gradient (same layout as the
 * weights), computed by the training kernels; n a multiple of 64. */
int lsdem_surrogate_gradient(lsdem_ctx* ctx, int64_t n, const float* points, const int32_t* shape_ids,
                             const float* targets, float* grad, double* loss);
/* lsdem_surrogate_gradient without any device allocation, for callers that cannot allocate
 * (code under HIP graph capture, tight training or serving loops).  No counterpart in the
 * reference.  points (n x 3), shape_ids (n), targets (n) and the outputs grad (same layout as the
 * weights) and loss (one double, may be NULL) are DEVICE pointers on the context's device; the
 * results equal lsdem_surrogate_gradient's bit for bit.  Intermediate data (family slots, per-block
 * partial gradients) live in a fixed 16 MiB array of device memory that the library holds on each
 * GPU until the process ends: n must be a positive multiple of 64 and at most what fits
 * (256 B + 4 n B + (n / 64) (19,778 + 16 x shapes) x 4 B; about 13,000 samples with 10 shapes),
 * else LSDEM_ERR_CAPACITY.  All calls on one GPU share that array: calls that overlap in time on
 * the same GPU (from several threads or contexts) race and must be serialised by the caller.  The
 * call synchronises the context's stream (to report shape ids outside the family as
 * LSDEM_ERR_ARG). */
int lsdem_surrogate_gradient_noalloc(lsdem_ctx* ctx, int64_t n, const float* points, const int32_t* shape_ids,
                                     const float* targets, float* grad, double* loss);

/* Mass, centroid (local frame), inertia and inverse inertia from an SDF grid with a smoothed
 * Heaviside of half-width 1.5 cells; host computation used by the first addLSParticle
 * overload (LSSolver.h:51-163).  inertia6/inv_inertia6 may be NULL. */
int lsdem_mass_properties(const double* sdf, int32_t nx, int32_t ny, int32_t nz,
                          const double origin[3], double spacing, double density, double* mass,
                          double centroid[3], double* inertia6, double* inv_inertia6);

#ifdef __cplusplus
}
#endif

#endif /* LSDEM_H */
