/* aemh.h -- C ABI of the MI355X (gfx950) MLS-MPM step for AE-ENGINE.
 *
 * Drop-in for the CPU hot path the engine runs every tick:
 *   ae_particle::mpm_solver::mpm_step_parallel_with_grid
 *       (ae_particle/src/mpm_solver.rs:1063-1084)
 *   ae_particle::mpm_solver::mpm_step_velocity_only_parallel
 *       (ae_particle/src/mpm_solver.rs:1337-1351)
 * following the FFI pattern of ae_cpp_kernel (ae_cpp_kernel/src/lib.rs:21-32,
 * cpp/simd_kernels.cpp:20): extern "C", plain pointers + sizes, caller-owned
 * host buffers borrowed only for the duration of a call.  Unlike the
 * template's `void` returns every entry point returns an aemh_status and
 * never aborts; aemh_last_error() gives the message.
 *
 * Particle data crosses the boundary in the reference's MpssBuffer layout
 * (ae_particle/src/mpss.rs:104-157): pos/vel as [n][3] f32, strain (F) and C
 * as [n][9] f32 row-major, jacobian/mass as [n] f32, material_idx as [n] u16.
 * State stays resident in HBM between calls; downloads return the original
 * slot order (the particle_indices semantics of the reference).
 *
 * Threading: one host thread per context at a time (the reference takes
 * &mut MpssBuffer).  All calls are synchronous on return.
 */
#ifndef AEMH_H_
#define AEMH_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct aemh_ctx aemh_ctx;

typedef enum {
  AEMH_OK = 0,
  AEMH_ERR_ARG = 1,        /* bad argument / size / null pointer          */
  AEMH_ERR_HIP = 2,        /* HIP runtime error (message has the name)    */
  AEMH_ERR_OOM = 3,        /* device allocation failed                    */
  AEMH_ERR_STATE = 4,      /* call order (e.g. step before upload)        */
  AEMH_ERR_OVERFLOW = 5,   /* Q32.32 accumulation could have wrapped      */
  AEMH_ERR_NO_DEVICE = 6   /* no gfx950 device / HIP not usable           */
} aemh_status;

/* Constitutive model ids: the Rust enum order of ConstitutiveModel
 * (ae_particle/src/constitutive.rs:11-27); AEMH_MODEL_SNOW is an addition
 * (Stomakhin 2013 clamp + J_p hardening, no reference counterpart). */
enum {
  AEMH_MODEL_NEO_HOOKEAN = 0,
  AEMH_MODEL_DRUCKER_PRAGER = 1,
  AEMH_MODEL_VON_MISES = 2,
  AEMH_MODEL_NEWTONIAN_FLUID = 3,
  AEMH_MODEL_COROTATED = 4,
  AEMH_MODEL_FIXED_COROTATED = 5,
  AEMH_MODEL_RIGID = 6,
  AEMH_MODEL_SNOW = 7
};

/* P2G accumulation: FP32 atomics, or Q32.32 int64 (FixedPoint format of
 * ae_physics/src/fixed_point.rs:4-38) -- the deterministic mode -- or FP32
 * sums in the reference serial substep's order (mpm_solver.rs:831-903, the
 * path the reference takes below 50,000 indices, :1077-1081): each node's
 * contributions added in index-list order, the reference's per-node G2P
 * order, and ln(J) from the host C library's logf (as Rust's f32::ln) over
 * [0.1, 10] (a 223 MB table built on first use).  Bit-exact with a CPU run of
 * the reference's serial FP32 arithmetic for the reference's material models;
 * the snow model's hardening exp (an addition) is the portable one. */
enum { AEMH_ACCUM_F32 = 0, AEMH_ACCUM_Q32_32 = 1, AEMH_ACCUM_F32_SERIAL = 2 };
/* G2P x/z clamp: reference serial path [0,(n-1)dx] (mpm_solver.rs:1002-1015),
 * parallel path [o, o+(n-1)dx] (mpm_solver.rs:1287-1290), or AUTO = the
 * reference's own switch at 50,000 indices (mpm_solver.rs:1077). */
enum { AEMH_CLAMP_AUTO = 0, AEMH_CLAMP_SERIAL = 1, AEMH_CLAMP_WINDOW = 2 };
/* Particle storage of C and F in HBM (compute is always FP32). */
enum { AEMH_STORE_F32 = 0, AEMH_STORE_BF16 = 1, AEMH_STORE_F16 = 2 };

/* MpmConfig (ae_particle/src/mpm_solver.rs:44-70) plus the options of this
 * implementation.  aemh_config_default() gives the reference defaults:
 * substeps 1, gravity (0,-9.81,0), damping 0.995, no wall margin. */
typedef struct {
  uint32_t substeps;
  float gravity[3];
  float damping;
  uint32_t bc_margin;      /* 0: floor rule only (mpm_solver.rs:131-136);
                              3: plus the 3-cell wall margin of
                              ae_compute/src/mpm_compute.rs:343-389     */
  float bc_restitution;
  uint32_t accum;          /* AEMH_ACCUM_*                               */
  uint32_t clamp_mode;     /* AEMH_CLAMP_*                               */
  uint32_t storage;        /* AEMH_STORE_*                               */
  uint32_t resort_interval;/* 0: automatic; k >= 1: G2P re-sorts the
                              working set by block every k-th substep and
                              updates it in place in between (results do
                              not depend on it in the Q32.32 and
                              reference-order modes).  No reference
                              counterpart (a tuning knob)               */
  uint32_t reserved[3];
} aemh_config;

/* One independent scene of a batched step: a contiguous range of the active
 * particle set, its own grid (the engine's near/mid tier grids with their
 * moving-window origins, ae_engine/src/managers/simulation.rs:541-607). */
typedef struct {
  uint32_t first;          /* first position in the active index list     */
  uint32_t count;          /* particles of this scene                      */
  uint32_t nx, ny, nz;     /* grid nodes                                   */
  float dx;
  float origin[3];
  float dt;                /* substep dt of this scene                     */
  uint32_t velocity_only;  /* 1: mpm_substep_velocity_only semantics       */
  uint32_t clamp_mode;     /* AEMH_CLAMP_*                                 */
  uint32_t reserved;
} aemh_scene;

void aemh_config_default(aemh_config* cfg);
const char* aemh_version(void);

/* Lifecycle ------------------------------------------------------------ */
aemh_status aemh_create(aemh_ctx** out, int device);
aemh_status aemh_destroy(aemh_ctx* ctx);
const char* aemh_last_error(const aemh_ctx* ctx);

/* Configuration -------------------------------------------------------- */
aemh_status aemh_set_config(aemh_ctx* ctx, const aemh_config* cfg);
/* Grid (ae_particle/src/mpm_solver.rs:73-103, origin :86). */
aemh_status aemh_set_grid(aemh_ctx* ctx, uint32_t nx, uint32_t ny, uint32_t nz,
                          float dx, const float origin[3]);
/* material_table / model_table (mpm_solver.rs:1067-1068): params is n x 7 f32
 * in MaterialParams field order (constitutive.rs:31-39), models n bytes.
 * ext (may be NULL) is n x 3 f32 (theta_c, theta_s, xi) for AEMH_MODEL_SNOW. */
aemh_status aemh_set_materials(aemh_ctx* ctx, uint32_t n, const float* params7,
                               const uint8_t* models, const float* ext3);

/* Particle state (MpssBuffer arrays, mpss.rs:104-157) ------------------- */
aemh_status aemh_upload_particles(aemh_ctx* ctx, uint32_t n, const float* pos3,
                                  const float* vel3, const float* C9,
                                  const float* F9, const float* J,
                                  const float* mass, const uint16_t* mat);
/* Plastic volume ratio J_p of the snow model (n floats; default 1). */
aemh_status aemh_upload_plastic_j(aemh_ctx* ctx, const float* Jp);
aemh_status aemh_download_particles(aemh_ctx* ctx, float* pos3, float* vel3,
                                    float* C9, float* F9, float* J);
aemh_status aemh_download_plastic_j(aemh_ctx* ctx, float* Jp);
/* The active set: particle_indices of mpm_step_parallel_with_grid. */
aemh_status aemh_set_indices(aemh_ctx* ctx, const uint32_t* idx, uint32_t n_idx);

/* Resident MpssBuffer (ae_particle/src/mpss.rs:104-297, 431-445) ------- */
/* Slot value of a spawn that found the buffer full (spawn() -> None). */
#define AEMH_NO_SLOT 0xFFFFFFFFu
/* MpssBuffer::new(capacity) (mpss.rs:161-187), resident in HBM: every slot
 * inactive with the new() defaults (strain and C zero, jacobian and mass 1,
 * lifetime f32::MAX, age 0, material 0).  Replaces any uploaded particles.
 * While a buffer is resident, aemh_step and aemh_mpm_step_parallel_with_grid
 * end with update_lifetimes(total_dt) as the reference steps do
 * (mpm_solver.rs:489, :508, :1083), and aemh_step without an index list
 * advances the active slots below count, the serial substep's set
 * (mpm_solver.rs:187-193). */
aemh_status aemh_mpss_reserve(aemh_ctx* ctx, uint32_t capacity);
/* spawn() n times (mpss.rs:189-203): free-list slots last-in first-out, then
 * fresh slots at count; each is reset (reset_slot, mpss.rs:215-232: pos and
 * vel 0, strain identity, jacobian 1, mass 1, material 0, lifetime
 * f32::MAX, age 0; C keeps its contents) and made active.  slots_out[n]
 * gets the slot or AEMH_NO_SLOT; *n_spawned (may be NULL) the number made. */
aemh_status aemh_mpss_spawn(aemh_ctx* ctx, uint32_t n, uint32_t* slots_out,
                            uint32_t* n_spawned);
/* kill() for each slot in order (mpss.rs:205-212); inactive or out-of-range
 * slots are ignored as the reference ignores them. */
aemh_status aemh_mpss_kill(aemh_ctx* ctx, const uint32_t* slots, uint32_t n);
/* compact() (mpss.rs:283-297): the active slots move to the front in slot
 * order with every field; the vacated slots hold what the reference's swap
 * sequence leaves there; the free list is cleared. */
aemh_status aemh_mpss_compact(aemh_ctx* ctx);
/* update_lifetimes(dt) (mpss.rs:431-445): active slots below count age by
 * dt; those with age >= lifetime are killed in ascending slot order.  The
 * expired slots (ascending) go to expired_out, at most max_out of them
 * (NULL: not returned); *n_expired gets their total. */
aemh_status aemh_mpss_update_lifetimes(aemh_ctx* ctx, float dt, uint32_t* expired_out,
                                       uint32_t max_out, uint32_t* n_expired);
/* Writes / reads the fields of n given slots (rows in slot-list order; NULL
 * fields are skipped): the pub field access of MpssBuffer (mpss.rs:104-157),
 * plus J_p of the snow model. */
aemh_status aemh_mpss_write(aemh_ctx* ctx, const uint32_t* slots, uint32_t n,
                            const float* pos3, const float* vel3, const float* C9,
                            const float* F9, const float* J, const float* Jp,
                            const float* mass, const uint16_t* mat,
                            const float* lifetime, const float* age);
aemh_status aemh_mpss_read(aemh_ctx* ctx, const uint32_t* slots, uint32_t n,
                           float* pos3, float* vel3, float* C9, float* F9, float* J,
                           float* Jp, float* mass, uint16_t* mat, float* lifetime,
                           float* age);
/* len() / capacity and the active flags (capacity bytes; NULL: skipped). */
aemh_status aemh_mpss_state(aemh_ctx* ctx, uint32_t* count, uint32_t* capacity,
                            uint8_t* active);

/* LOD tier orchestrator (ae_engine/src/managers/simulation.rs:480-700) -- */
/* PhysicsLod (simulation.rs:27-48) plus the tier grids and the mid cadence. */
typedef struct {
  float near_field_distance;          /* 50 m                                  */
  float mid_field_distance;           /* 200 m                                 */
  uint32_t far_field_update_interval; /* 60 ticks (0 acts as 1, :594)        */
  uint32_t max_near_particles;        /* 10,000                               */
  uint32_t mid_interval;              /* MID_INTERVAL = 12 ticks (:586)       */
  uint32_t near_n;                    /* next_pow2(ceil(2 near / 1.6)) = 64  */
  float near_dx;                      /* 1.6 m (:219-227)                     */
  uint32_t mid_n;                     /* 16 (:234)                            */
  float mid_dx;                       /* 2 mid / 16 = 25 m (:235)             */
  uint32_t reserved[3];
} aemh_lod_config;

/* What one tick did: LodStats (cached list sizes, :533-536), which tiers
 * ran, the moving-window origins (:540-556), and the slots the tick's
 * update_lifetimes calls killed. */
typedef struct {
  uint32_t near, mid, far;
  uint32_t reclassified;
  uint32_t near_ran, mid_ran, far_ran;
  uint32_t expired;
  float near_origin[3];
  float mid_origin[3];
} aemh_tier_stats;

void aemh_lod_config_default(aemh_lod_config* lod);
/* Sets the tier configuration and starts the tick counters, the cached
 * player position (origin) and the cached lists over, as a new
 * SimulationManager does (simulation.rs:311-317).  Without this call the
 * first tick uses aemh_lod_config_default. */
aemh_status aemh_tier_set_config(aemh_ctx* ctx, const aemh_lod_config* lod);
/* The MPSS part of SimulationManager::update_fields_and_particles
 * (simulation.rs:474-697) on the resident buffer: reclassification when the
 * player moved > 5 m, the near list is empty or every 8th tick (near / mid /
 * far by distance, near culled to the max_near closest with the rest moved
 * to mid); moving-window origins; the near full step
 * (mpm_step_parallel_with_grid, which ages lifetimes) and, every
 * mid_interval ticks, the mid velocity-only step with dt x mid_interval,
 * both scenes in one batched launch; every far_field_update_interval ticks
 * the far gravity update (velocity and height; the coarse temperature field
 * is not part of this library); then update_lifetimes(dt).  Material and
 * config come from aemh_set_materials / aemh_set_config.  stats may be NULL. */
aemh_status aemh_tier_tick(aemh_ctx* ctx, const float player[3], float dt,
                           aemh_tier_stats* stats);
/* Cached lists (NULL: skipped; each array needs room for its count) and
 * their sizes counts3[3] (NULL: skipped). */
aemh_status aemh_tier_lists(aemh_ctx* ctx, uint32_t* near, uint32_t* mid, uint32_t* far,
                            uint32_t* counts3);
/* Render readback (ae_engine/src/lib.rs:1891-1936): of the first `take`
 * entries of cached list `tier` (0 near, 1 mid, 2 far), those still live
 * (slot < count and active), in list order: positions pos3[n][3] and slots
 * (either may be NULL); *n_out gets n. */
aemh_status aemh_tier_render(aemh_ctx* ctx, uint32_t tier, uint32_t take, float* pos3,
                             uint32_t* slots, uint32_t* n_out);

/* Steps ----------------------------------------------------------------- */
/* mpm_step_parallel_with_grid (mpm_solver.rs:1063-1084): substeps of
 * total_dt/substeps over the active set on the context grid. */
aemh_status aemh_step(aemh_ctx* ctx, float total_dt);
/* Same, with the index list given per call as the reference does; the
 * active set is re-gathered only when the list changes. */
aemh_status aemh_mpm_step_parallel_with_grid(aemh_ctx* ctx, const uint32_t* idx,
                                             uint32_t n_idx, float total_dt);
/* mpm_step_velocity_only_parallel (mpm_solver.rs:1337-1351). */
aemh_status aemh_step_velocity_only(aemh_ctx* ctx, float total_dt);
/* aemh_step without waiting for the GPU: no counterpart in the reference
 * (whose step returns when done).  For callers that advance several steps
 * between host reads (substep loops, benchmarks): the host enqueues the
 * next step while the GPU runs the last.  Every later call on the context
 * is ordered after it; downloads wait for it.  aemh_sync waits and returns
 * the deterministic mode's range check (AEMH_ERR_OVERFLOW) of the steps
 * since the last synchronous call; between synchronous calls the re-sort
 * decision uses only the fixed cadence.  With a resident MpssBuffer (whose
 * step ages lifetimes) it is aemh_step. */
aemh_status aemh_step_async(aemh_ctx* ctx, float total_dt);
aemh_status aemh_sync(aemh_ctx* ctx);

/* Batched scenes: every scene advances in one launch per phase, the whole
 * substep captured in a hipGraph and replayed n_substeps times. */
aemh_status aemh_set_scenes(aemh_ctx* ctx, const aemh_scene* scenes, uint32_t n);
aemh_status aemh_batch_step(aemh_ctx* ctx, uint32_t n_substeps);
/* Times candidate launch shapes on the current particles and scenes and
 * keeps the fastest of each: P2G LDS tiles per workgroup (deterministic
 * mode: 2/3/4, a budget below 4 must beat 4 by 3 % per tile; FP32 mode: 5)
 * together with the tile flush (one HBM atomic per node value, or partial
 * sums stored per tile and added by the grid update) and the lane
 * assignment: lanes keep their sorted particles, or are re-assigned by rank
 * within their base cell so one LDS atomic's lanes hit distinct nodes
 * (layers; own lanes must win by 5 % in the deterministic mode and when most
 * particles are in velocity-only scenes); then grid-update workgroups (1024..16384) and the G2P workgroup
 * (64/128/256/512, sizes other than 256 must win by 3 %).  The particle state
 * is not advanced.  Called once by the first batch step; plain steps use the
 * defaults unless this is called.  No reference counterpart. */
aemh_status aemh_autotune(aemh_ctx* ctx);
/* The launch shape in use and, after aemh_autotune, the winning times in ms
 * (P2G + grid update, grid update, G2P; zeros before tuning).  NULL outputs
 * are skipped.  p2g_partials carries three P2G flags: bit 0 the partial-sum
 * tile flush, bit 1 the pair reduction (FP32 mode: lanes holding two
 * particles of one base cell add their node values across the pair before
 * the LDS atomics; set only through aemh_set_launch_shape), bit 2 the
 * layers. */
aemh_status aemh_get_launch_shape(aemh_ctx* ctx, uint32_t* p2g_tiles, uint32_t* p2g_partials,
                                  uint32_t* g2p_wg, uint32_t* upd_wgs, float* tuned_ms3);
/* Sets the launch shape directly (results do not depend on it, bit for bit
 * in the deterministic mode; p2g_tiles 2-5, default 3; p2g_partials 0-7 as
 * above, default 0). */
aemh_status aemh_set_launch_shape(aemh_ctx* ctx, uint32_t p2g_tiles, uint32_t p2g_partials,
                                  uint32_t g2p_wg, uint32_t upd_wgs);

/* Grid readback for parity checks (reference layout, x-fastest index
 * i + j nx + k nx ny, mpm_solver.rs:157-159) of scene `scene`, as left by
 * the last substep: after P2G (mass, momentum, force) and after the update
 * (velocity).  NULL outputs are skipped.  The FP32 mode (AEMH_ACCUM_F32)
 * accumulates the force into the momentum as its impulse: momentum3 then
 * holds p + f dt and force3 is zero.  The Q32.32 and reference-order modes
 * return m, p and f apart. */
aemh_status aemh_download_grid(aemh_ctx* ctx, uint32_t scene, float* mass,
                               float* momentum3, float* force3, float* vel3);
/* Q32.32 raw sums (7 per node: m, p0..2, f0..2) of the last substep. */
aemh_status aemh_download_grid_q32(aemh_ctx* ctx, uint32_t scene, int64_t* q7);

/* FP64 reductions over the active set: total mass, momentum[3], kinetic
 * energy (ae_physics/src/mpm.rs:752-771 diagnostics, in FP64). out[5]. */
aemh_status aemh_conservation_f64(aemh_ctx* ctx, double* out5);
/* The same FP64 sums, bit for bit, with no device allocation: no counterpart
 * in the reference.  For callers that cannot allocate device memory between
 * steps (tight stepping loops that take these diagnostics every few steps,
 * harnesses that forbid hipMalloc/hipFree).  The reduction's partial sums
 * live in a fixed static device buffer of 5 + 5 x 512 doubles (20.5 KB per
 * device, held until the process ends), enough for any particle count.
 * Needs a current working set (after a step or aemh_set_indices; otherwise
 * AEMH_ERR_STATE, since gathering one allocates).  All calls on one device
 * share the buffer: calls that overlap on one device race. */
aemh_status aemh_conservation_f64_noalloc(aemh_ctx* ctx, double* out5);

/* Snow-model SVDs that took the FP64 path (cond(F) > 1e4) since the
 * particles were uploaded.  No reference counterpart (a diagnostic). */
aemh_status aemh_svd_fp64_count(aemh_ctx* ctx, uint64_t* count);

/* Work of the last substep: active 4^3 node blocks (grid nodes touched =
 * 64 x blocks) and particles advanced. */
aemh_status aemh_last_substep_info(aemh_ctx* ctx, uint32_t* active_blocks,
                                   uint32_t* particles);

/* Per-kernel timing: while enabled, substeps launch eagerly on the context
 * stream (instead of the graph replay) and each section's HIP event pair is
 * recorded by its kernels' own dispatches (hipExtLaunchKernel start and stop
 * events); aemh_profile_get gives the average ms per section.  Sections of
 * a substep: "count", "scan", "scatter" (with the active-block list and the
 * accumulator clear), "p2g", "grid_update", "g2p"; tier ticks add
 * "lod_classify", "lod_select", "near_cull", "far_update", "lifetimes".
 * No reference counterpart. */
aemh_status aemh_profile_enable(aemh_ctx* ctx, int on);
aemh_status aemh_profile_get(aemh_ctx* ctx, const char* kernel, double* avg_ms,
                             uint64_t* count);
aemh_status aemh_profile_reset(aemh_ctx* ctx);

/* Deterministic scene generators (xoshiro256** seeded by splitmix64), host
 * only, for the benchmark workloads.  Fill caller arrays of n particles. */
void aemh_scene_uniform_box(uint64_t seed, uint32_t n, const float lo[3],
                            const float hi[3], float* pos3);

/* ------------------------------------------------------------------------
 * All-integer MLS-MPM solver: ae_physics::mpm::MpmSimulation
 * (ae_physics/src/mpm.rs:116-792) on FixedPoint Q32.32
 * (ae_physics/src/fixed_point.rs).  Every FixedPoint crosses the boundary
 * as its raw int64 (value = raw / 2^32).  Results are bit-identical to the
 * reference's serial loops: particle order, ids, fracture appends and
 * events included.  The grid scatter adds integers in parallel, which is
 * exact while no saturating prefix sum of the reference's serial order can
 * clip: the device proves that from bounds on the sums of |contribution|
 * (whole substep, then per cell) and otherwise rebuilds that substep's grid
 * with the serial loop (counted in aemh_fixed_state).
 * ---------------------------------------------------------------------- */
typedef struct aemh_fixed aemh_fixed;

/* MpmMaterialModel (mpm.rs:5-13); model_param carries the variant's field */
typedef enum {
  AEMH_FX_ELASTIC = 0,
  AEMH_FX_ELASTOPLASTIC = 1,
  AEMH_FX_GRANULAR = 2,  /* model_param: friction_angle (degrees)   */
  AEMH_FX_BRITTLE = 3,   /* model_param: fracture_strain           */
  AEMH_FX_FLUID = 4,     /* model_param: viscosity                 */
  AEMH_FX_SNOW = 5       /* model_param: hardening                 */
} aemh_fixed_model;

/* MpmConfig (mpm.rs:15-31) */
typedef struct {
  uint32_t grid_resolution[3];
  int64_t cell_size;
  int64_t youngs_modulus;
  int64_t poissons_ratio;
  int64_t yield_stress;
  int64_t hardening;
  int64_t density;
  int64_t gravity[3];
  int64_t dt;
  uint32_t substeps;
  uint32_t material_model; /* aemh_fixed_model */
  int64_t model_param;
  uint32_t enable_fracture;
  int64_t fracture_strain;
  uint64_t particle_count;
} aemh_fixed_config;

/* The non-particle fields of MpmSimulation (mpm.rs:116-126) and the config
 * fields the material models change while stepping (:464-470). */
typedef struct {
  uint64_t particles;            /* particles.len()                        */
  uint64_t step_count;
  double time;
  uint64_t next_particle_id;
  uint64_t fracture_events;      /* fracture_events.len()                  */
  int64_t hardening;
  int64_t fracture_strain;
  uint32_t enable_fracture;
  uint32_t material_model;
  int64_t model_param;
  uint64_t checked_substeps;     /* substeps whose scatter was bounded per cell */
  uint64_t serial_substeps;      /* substeps whose grid was rebuilt serially    */
  /* largest whole-substep sum of ((|c| >> 24) + 1) over nonzero scatter
   * contributions c; the atomic scatter alone is proven exact below 2^39
   * (UINT64_MAX once a substep went past it) */
  uint64_t max_scatter_bound;
} aemh_fixed_state;

/* FractureEvent (mpm.rs:132-138) */
typedef struct {
  double time;
  float position[3];
  float strain_magnitude;
  uint32_t num_fragments;
  uint32_t pad;
} aemh_fracture_event;

/* active_particles, fractured_count, total_mass, average_temperature,
 * kinetic_energy, total_strain_energy (mpm.rs:744-778) */
typedef struct {
  uint64_t active_particles;
  uint64_t fractured_count;
  int64_t total_mass;
  int64_t average_temperature;
  int64_t kinetic_energy;
  int64_t total_strain_energy;
} aemh_fixed_stats;

/* MpmParticleSnapshot (mpm.rs:94-103) */
typedef struct {
  float position[3];
  float velocity[3];
  float stress_magnitude;
  float plastic_strain;
  float temperature;
  float color[4];
  uint32_t fractured;
} aemh_fixed_snapshot;

/* MpmConfig::default (mpm.rs:33-52). */
void aemh_fixed_config_default(aemh_fixed_config* cfg);
/* MpmSimulation::new (mpm.rs:159-177): the grid, and config.particle_count
 * particles placed by init_particles (:179-206). */
aemh_status aemh_fixed_create(aemh_fixed** out, int device, const aemh_fixed_config* cfg);
void aemh_fixed_destroy(aemh_fixed* sim);
const char* aemh_fixed_last_error(const aemh_fixed* sim);
/* add_particle (mpm.rs:212-233): pos3, vel3 raw, mass raw, color4; the id
 * is written to *id_out (may be NULL). */
aemh_status aemh_fixed_add_particle(aemh_fixed* sim, const int64_t* pos3, const int64_t* vel3,
                                    int64_t mass, const float* color4, uint64_t* id_out);
/* n add_particle calls in order, in one upload: pos3/vel3 [n][3], mass [n],
 * color4 [n][4] (NULL: the default color).  *first_id_out gets the first
 * id (ids are consecutive).  No reference counterpart beyond the loop. */
aemh_status aemh_fixed_add_particles(aemh_fixed* sim, uint64_t n, const int64_t* pos3,
                                     const int64_t* vel3, const int64_t* mass,
                                     const float* color4, uint64_t* first_id_out);
/* remove_particle (mpm.rs:235-242): *removed = 1 if the id was found. */
aemh_status aemh_fixed_remove_particle(aemh_fixed* sim, uint64_t id, int* removed);
/* step (mpm.rs:244-251), n_steps times. */
aemh_status aemh_fixed_step(aemh_fixed* sim, uint32_t n_steps);
/* set_material_model (mpm.rs:780-782). */
aemh_status aemh_fixed_set_material_model(aemh_fixed* sim, uint32_t model, int64_t model_param);
/* Interval, in substeps, at which the cell-ordered P2G re-sorts its order
 * from the live particle records (default 1: every substep; between
 * re-sorts it reuses the last order, which stays exact but loses runs of
 * equal cells as particles move).  Results do not depend on it, bit for
 * bit.  Accepts 1..65536.  No reference counterpart (a tuning knob). */
aemh_status aemh_fixed_set_sort_interval(aemh_fixed* sim, uint32_t substeps);
/* enable_fracture / disable_fracture (mpm.rs:784-791). */
aemh_status aemh_fixed_enable_fracture(aemh_fixed* sim, int64_t threshold);
aemh_status aemh_fixed_disable_fracture(aemh_fixed* sim);
aemh_status aemh_fixed_get_state(aemh_fixed* sim, aemh_fixed_state* out);
/* The public `particles` vector (mpm.rs:119), records [first, first+count):
 * q29 [count][29] raw int64 = position3, velocity3, deformation_gradient9
 * (column-major: x_axis, y_axis, z_axis), stress9 (column-major), mass,
 * volume, plastic_strain, temperature, accumulated_strain; flags3 [count][3]
 * = active, fractured, num_connections; ids9 [count][9] = parent_id,
 * connected_to[8]; color4 [count][4].  NULL outputs are skipped. */
aemh_status aemh_fixed_read_particles(aemh_fixed* sim, uint64_t first, uint64_t count, int64_t* q29,
                                      uint8_t* flags3, uint64_t* ids9, float* color4);
/* Overwrites records [first, first+count) (first+count <= particles), as a
 * caller assigning the public fields does.  All four arrays required. */
aemh_status aemh_fixed_write_particles(aemh_fixed* sim, uint64_t first, uint64_t count,
                                       const int64_t* q29, const uint8_t* flags3,
                                       const uint64_t* ids9, const float* color4);
/* fracture_events (mpm.rs:125) [first, first+count). */
aemh_status aemh_fixed_fracture_events(aemh_fixed* sim, uint64_t first, uint64_t count,
                                       aemh_fracture_event* out);
aemh_status aemh_fixed_get_stats(aemh_fixed* sim, aemh_fixed_stats* out);
/* snapshots (mpm.rs:718-742): active particles in order, at most cap;
 * *n_out = the number of active particles. */
aemh_status aemh_fixed_snapshots(aemh_fixed* sim, aemh_fixed_snapshot* out, uint64_t cap,
                                 uint64_t* n_out);

#ifdef __cplusplus
}
#endif
#endif /* AEMH_H_ */
