// refcpu.h -- CPU restatement of AE-ENGINE's MLS-MPM hot path (the oracle).
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
// calls this code; only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may.  It restates, in plain C++ compiled with
// -ffp-contract=off (Rust never fuses multiply-adds), the reference files:
//   ae_particle/src/mpm_solver.rs      (FP32 solver, grid, stencil, guards)
//   ae_particle/src/constitutive.rs    (material models, polar decomposition)
//   ae_physics/src/fixed_point.rs      (Q32.32 FixedPoint cast/add rules)
//   ae_compute/src/mpm_compute.rs      (3-cell wall margin, row "grid_op")
// Each function cites the file:line it follows.  Functions that have no
// counterpart in the reference (snow plasticity, SVD, portable ln/exp) say so.
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

// Constitutive model ids follow the Rust enum's declaration order
// (ae_particle/src/constitutive.rs:11-27); SNOW is an addition.
enum {
  REF_NEO_HOOKEAN = 0,
  REF_DRUCKER_PRAGER = 1,
  REF_VON_MISES = 2,
  REF_NEWTONIAN_FLUID = 3,
  REF_COROTATED = 4,
  REF_FIXED_COROTATED = 5,
  REF_RIGID = 6,
  REF_SNOW = 7
};

enum { REF_ACCUM_F32 = 0, REF_ACCUM_Q32 = 1 };
// Position clamp used by G2P:
//   SERIAL  -> x,z clamped to [0, (n-1)dx]         (mpm_solver.rs:1002-1015)
//   WINDOW  -> x,z clamped to [o, o + (n-1)dx]     (mpm_solver.rs:1287-1290)
//   AUTO    -> SERIAL below 50,000 indices, else WINDOW (mpm_solver.rs:1077)
enum { REF_CLAMP_AUTO = 0, REF_CLAMP_SERIAL = 1, REF_CLAMP_WINDOW = 2 };

typedef struct {
  uint32_t nx, ny, nz;
  float dx;
  float origin[3];
  uint32_t substeps;
  float gravity[3];
  float damping;
  uint32_t bc_margin;     // 0: floor rule only; 3: ae_compute 3-cell walls
  float bc_restitution;
  uint32_t accum;         // REF_ACCUM_*
  uint32_t clamp_mode;    // REF_CLAMP_*
  uint32_t det_math;      // 0: libm logf/expf (as the reference calls them);
                          // 1: portable ln and exp (deterministic mode);
                          // 2: libm logf, portable exp (the reference-order
                          //    FP32 mode: exp only appears in the snow model,
                          //    which has no reference counterpart)
  uint32_t threads;       // <=1: serial; >1: fold/reduce P2G + parallel G2P
} RefConfig;

// Materials: params is n x 7 floats in MaterialParams field order
// (young_modulus, poisson_ratio, yield_stress, hardening, density,
// friction_angle, cohesion; constitutive.rs:31-39); models is n bytes;
// ext is n x 3 floats (theta_c, theta_s, xi) used only by REF_SNOW.
typedef struct {
  uint32_t n;
  const float* params;
  const uint8_t* models;
  const float* ext;
} RefMaterials;

// ---- scalar pieces (exported for known-answer tests) --------------------
float ref_b_spline(float x);
float ref_b_spline_deriv(float x);
int64_t ref_fp_from_f32(float v);
float ref_fp_to_f32(int64_t raw);
int64_t ref_fp_add(int64_t a, int64_t b);
int64_t ref_fp_mul(int64_t a, int64_t b);
int64_t ref_fp_div(int64_t a, int64_t b);
int64_t ref_fp_sqrt(int64_t a);
float ref_det_logf(float x);
float ref_det_expf(float x);
// returns plastic flag; stress and new_F are row-major 3x3
int ref_compute_stress(int model, const float* params7, const float* ext3,
                       const float* F, float J, float Jp, int det_math,
                       float* stress, float* new_F);
// One-sided Jacobi SVD (FP32 with FP64 fallback when cond > 1e4):
// F = U diag(sig) V^T, U and V rotations.  Returns 1 if FP64 was used.
int ref_svd3(const float* F, float* U, float* sig, float* V);
int ref_svd3_d(const double* F, double* U, double* sig, double* V);
// Snow clamp of G2P (no reference counterpart): F and Jp updated in place.
void ref_snow_clamp(const float* ext3, float* F, float* Jp);

// ---- one full step (substeps inside) -----------------------------------
// Arrays are the reference's MpssBuffer layout: pos/vel [n][3], F/C [n][9]
// row-major, J/mass/Jp [n], mat [n].  idx lists the particles to advance
// (mpm_step_parallel_with_grid's particle_indices).  When grid_* outputs are
// non-null the grid of the LAST substep is returned: after P2G (q32 raw
// sums, or float m/momentum/force) and after the grid update (velocity).
int ref_step_f32(const RefConfig* cfg, const RefMaterials* mats,
                 uint32_t n, float* pos, float* vel, float* F, float* J,
                 float* C, const float* mass, const uint16_t* mat, float* Jp,
                 const uint32_t* idx, uint32_t n_idx, float total_dt,
                 float* grid_m, float* grid_p, float* grid_f, float* grid_v,
                 int64_t* grid_q);
int ref_step_f64(const RefConfig* cfg, const RefMaterials* mats,
                 uint32_t n, double* pos, double* vel, double* F, double* J,
                 double* C, const double* mass, const uint16_t* mat,
                 double* Jp, const uint32_t* idx, uint32_t n_idx,
                 double total_dt);
// mpm_step_velocity_only_parallel (mpm_solver.rs:1337-1572)
int ref_step_velocity_only(const RefConfig* cfg, uint32_t n, float* pos,
                           float* vel, const float* mass, const uint32_t* idx,
                           uint32_t n_idx, float total_dt);

#ifdef __cplusplus
}
#endif
