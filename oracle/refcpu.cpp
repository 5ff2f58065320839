// refcpu.cpp -- CPU restatement of AE-ENGINE's MLS-MPM hot path.
//
// TEST INFRASTRUCTURE ONLY (the oracle): used by tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg as the checker and the timed CPU baseline.
// The shipped library never links or calls it.
//
// Built with -ffp-contract=off so every float expression is evaluated in the
// order the Rust source writes it, with no fused multiply-add, as rustc does.
// Expressions below keep the reference's parenthesisation on purpose: the
// left-to-right evaluation order IS the specification being restated.
//
// Parity status per function (see the pinning tests in tests/test_oracle.py):
//   b_spline / b_spline_deriv ........ pinned (reference known answers)
//   FixedPoint cast / add / mul / div / sqrt .. pinned (reference tests,
//                                       exact __int128 closed forms)
//   compute_stress (all 7 models) .... pinned (reference known answers +
//                                       closed forms)
//   P2G / grid update / G2P .......... pinned (closed forms: free fall,
//                                       partition deficit, linear-field APIC,
//                                       mass/momentum conservation; the 1-D
//                                       force derivative and its update term,
//                                       grad v and F' = (I + dt grad v) F on
//                                       one particle, tests/test_oracle_pins.py)
//   snow clamp + SVD (no reference) .. pinned (reconstruction/orthogonality,
//                                       known singular values)
//   det_logf / det_expf (no reference) pinned against libm double log/exp
//   multi-step FP32 trajectories ..... parity unpinned beyond the above: no
//                                       Rust toolchain exists to run the
//                                       reference itself.
#include "refcpu.h"

// Planted mistakes, compiled in only by tests/test_oracle_pins.py (one
// -DORACLE_MUT_* flag per build of a throw-away copy) to show that the
// closed-form pins fail on them; the oracle proper is built without any.
//   ORACLE_MUT_TEXTBOOK_P2G   P2G force with the tensor-product gradient
//                             N'(x)N(y)N(z) instead of the 1-D N'(x) (Q3)
//   ORACLE_MUT_NO_FORCE_TERM  grid update without f/m (mpm_solver.rs:127-129)
//   ORACLE_MUT_TEXTBOOK_G2P   G2P grad v = -sum g (x) grad w instead of
//                             -sum w g (x) N' (Q4)
//   ORACLE_MUT_TRANSPOSE_GV   grad v stored transposed
//   ORACLE_MUT_F_RIGHT        F' = F (I + dt grad v) instead of (I + dt grad v) F

#include <math.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

namespace {

// ------------------------------------------------------------------------
// Quadratic B-spline and derivative: ae_particle/src/mpm_solver.rs:442-464
// ------------------------------------------------------------------------
template <class R>
inline R b_spline(R x) {
  R ax = fabs(x);
  if (ax < R(0.5)) {
    return R(0.75) - ax * ax;
  } else if (ax < R(1.5)) {
    R t = R(1.5) - ax;
    return R(0.5) * t * t;
  }
  return R(0);
}

template <class R>
inline R signum(R x) {  // Rust f32::signum: +1 for +0.0, -1 for -0.0, NaN->NaN
  if (x != x) return x;
  return signbit(x) ? R(-1) : R(1);
}

template <class R>
inline R b_spline_deriv(R x) {
  R ax = fabs(x);
  if (ax < R(0.5)) {
    return R(-2.0) * x;
  } else if (ax < R(1.5)) {
    return -(R(1.5) - ax) * signum(x);
  }
  return R(0);
}

// Rust `f32 as isize` / `f64 as i64`: truncate toward zero, saturate, NaN->0.
template <class R>
inline int64_t sat_cast_i64(R v) {
  if (v != v) return 0;
  if (v >= R(9223372036854775807.0)) return INT64_MAX;
  if (v <= R(-9223372036854775808.0)) return INT64_MIN;
  return (int64_t)v;
}

// Rust f32::max / f32::min semantics for one NaN operand (returns the other).
template <class R>
inline R rmax(R a, R b) { return fmax(a, b); }
template <class R>
inline R rmin(R a, R b) { return fmin(a, b); }
// Rust f32::clamp: NaN passes through.
template <class R>
inline R rclamp(R x, R lo, R hi) {
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}

// ------------------------------------------------------------------------
// FixedPoint Q32.32: ae_physics/src/fixed_point.rs
// ------------------------------------------------------------------------
// from_f32: (value as f64 * 2^32) as i64 -- truncating, saturating, NaN->0
// (fixed_point.rs:22-25).  The product is exact (24-bit mantissa scaled by 2^32).
inline int64_t fp_from_f32(float v) { return sat_cast_i64((double)v * 4294967296.0); }
// to_f32: raw as f32 / 2^32 (fixed_point.rs:36-38).
inline float fp_to_f32(int64_t raw) { return (float)raw / 4294967296.0f; }
// Add saturates (fixed_point.rs:187-189 via Add, :292-296).
inline int64_t fp_add(int64_t a, int64_t b) {
  int64_t r;
  if (__builtin_add_overflow(a, b, &r)) return a > 0 ? INT64_MAX : INT64_MIN;
  return r;
}

// ------------------------------------------------------------------------
// Portable transcendental functions for the deterministic (Q32.32) mode.
// No counterpart in the reference (it calls libm's logf, constitutive.rs:384).
// ln: x = m 2^e, m in [sqrt(1/2), sqrt(2)), ln m = 2 atanh(s), s=(m-1)/(m+1),
// the same identity as FixedPoint::ln (fixed_point.rs:254-271), evaluated in
// double to 12 odd terms and rounded once to float.
// exp: k = round(x/ln2), r = x - k ln2, Taylor to r^14, scaled by 2^k.
// Only IEEE +,-,*,/ in double with a fixed order: identical bits on any
// IEEE-754 machine compiled without contraction.
// ------------------------------------------------------------------------
const double kLn2Hi = 6.93147180369123816490e-01;  // 0x3FE62E42FEE00000
const double kLn2Lo = 1.90821492927058770002e-10;
const double kInvLn2 = 1.44269504088896338700e+00;

double det_log_d(double x) {
  if (x != x) return x;
  if (x <= 0.0) return x == 0.0 ? -INFINITY : NAN;
  if (x == INFINITY) return x;
  uint64_t b;
  memcpy(&b, &x, 8);
  int e = (int)((b >> 52) & 0x7ff) - 1023;
  uint64_t mb = (b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
  double m;
  memcpy(&m, &mb, 8);
  if (m > 1.4142135623730951) {
    m = m * 0.5;
    e += 1;
  }
  double s = (m - 1.0) / (m + 1.0);
  double s2 = s * s;
  double p = 1.0 / 23.0;
  p = p * s2 + 1.0 / 21.0;
  p = p * s2 + 1.0 / 19.0;
  p = p * s2 + 1.0 / 17.0;
  p = p * s2 + 1.0 / 15.0;
  p = p * s2 + 1.0 / 13.0;
  p = p * s2 + 1.0 / 11.0;
  p = p * s2 + 1.0 / 9.0;
  p = p * s2 + 1.0 / 7.0;
  p = p * s2 + 1.0 / 5.0;
  p = p * s2 + 1.0 / 3.0;
  p = p * s2 + 1.0;
  double r = 2.0 * s * p;
  double ed = (double)e;
  return ed * kLn2Hi + (ed * kLn2Lo + r);
}

double det_exp_d(double x) {
  if (x != x) return x;
  if (x > 709.0) return INFINITY;
  if (x < -745.0) return 0.0;
  double kf = floor(x * kInvLn2 + 0.5);
  double r = (x - kf * kLn2Hi) - kf * kLn2Lo;
  double p = 1.0 / 87178291200.0;            // 1/14!
  p = p * r + 1.0 / 6227020800.0;            // 1/13!
  p = p * r + 1.0 / 479001600.0;             // 1/12!
  p = p * r + 1.0 / 39916800.0;
  p = p * r + 1.0 / 3628800.0;
  p = p * r + 1.0 / 362880.0;
  p = p * r + 1.0 / 40320.0;
  p = p * r + 1.0 / 5040.0;
  p = p * r + 1.0 / 720.0;
  p = p * r + 1.0 / 120.0;
  p = p * r + 1.0 / 24.0;
  p = p * r + 1.0 / 6.0;
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  int k = (int)kf;
  // 2^k as two factors so subnormal results stay exact-scaled.
  int k1 = k / 2, k2 = k - k1;
  uint64_t b1 = (uint64_t)(k1 + 1023) << 52, b2 = (uint64_t)(k2 + 1023) << 52;
  double s1, s2;
  memcpy(&s1, &b1, 8);
  memcpy(&s2, &b2, 8);
  return (p * s1) * s2;
}

float det_logf(float x) { return (float)det_log_d((double)x); }
float det_expf(float x) {
  if (x > 89.0f) return INFINITY;
  if (x < -104.0f) return 0.0f;
  return (float)det_exp_d((double)x);
}

template <class R>
inline R r_ln(R x, int det);
template <>
inline float r_ln<float>(float x, int det) { return det == 1 ? det_logf(x) : logf(x); }
template <>
inline double r_ln<double>(double x, int) { return log(x); }
template <class R>
inline R r_exp(R x, int det);
template <>
inline float r_exp<float>(float x, int det) { return det ? det_expf(x) : expf(x); }
template <>
inline double r_exp<double>(double x, int) { return exp(x); }

// ------------------------------------------------------------------------
// 3x3 helpers (row-major): constitutive.rs:254-281
// ------------------------------------------------------------------------
template <class R>
inline void identity3(R* m) {
  for (int i = 0; i < 9; i++) m[i] = (i % 4 == 0) ? R(1) : R(0);
}
template <class R>
inline void mat_mul(const R* a, const R* b, R* o) {
  R t[9];
  t[0] = a[0] * b[0] + a[1] * b[3] + a[2] * b[6];
  t[1] = a[0] * b[1] + a[1] * b[4] + a[2] * b[7];
  t[2] = a[0] * b[2] + a[1] * b[5] + a[2] * b[8];
  t[3] = a[3] * b[0] + a[4] * b[3] + a[5] * b[6];
  t[4] = a[3] * b[1] + a[4] * b[4] + a[5] * b[7];
  t[5] = a[3] * b[2] + a[4] * b[5] + a[5] * b[8];
  t[6] = a[6] * b[0] + a[7] * b[3] + a[8] * b[6];
  t[7] = a[6] * b[1] + a[7] * b[4] + a[8] * b[7];
  t[8] = a[6] * b[2] + a[7] * b[5] + a[8] * b[8];
  memcpy(o, t, sizeof(t));
}
template <class R>
inline void transpose3(const R* m, R* o) {
  R t[9] = {m[0], m[3], m[6], m[1], m[4], m[7], m[2], m[5], m[8]};
  memcpy(o, t, sizeof(t));
}
// Cofactor determinant in the order G2P writes it (mpm_solver.rs:1038-1040).
template <class R>
inline R det3(const R* f) {
  return f[0] * (f[4] * f[8] - f[5] * f[7]) - f[1] * (f[3] * f[8] - f[5] * f[6]) +
         f[2] * (f[3] * f[7] - f[4] * f[6]);
}

// `inv_transpose` (constitutive.rs:306-324).  NB: as written it returns the
// inverse (adjugate / det), not its transpose; restated literally.
template <class R>
inline void inv_transpose(const R* m, R* o) {
  R det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
          m[2] * (m[3] * m[7] - m[4] * m[6]);
  if (fabs(det) < R(1e-15)) {
    identity3(o);
    return;
  }
  R inv_det = R(1.0) / det;
  o[0] = (m[4] * m[8] - m[5] * m[7]) * inv_det;
  o[1] = -(m[1] * m[8] - m[2] * m[7]) * inv_det;
  o[2] = (m[1] * m[5] - m[2] * m[4]) * inv_det;
  o[3] = -(m[3] * m[8] - m[5] * m[6]) * inv_det;
  o[4] = (m[0] * m[8] - m[2] * m[6]) * inv_det;
  o[5] = -(m[0] * m[5] - m[2] * m[3]) * inv_det;
  o[6] = (m[3] * m[7] - m[4] * m[6]) * inv_det;
  o[7] = -(m[0] * m[7] - m[1] * m[6]) * inv_det;
  o[8] = (m[0] * m[4] - m[1] * m[3]) * inv_det;
}

// polar_decomp (constitutive.rs:284-304): 8 averaging iterations.
template <class R>
inline void polar_decomp(const R* F, R* r, R* s) {
  memcpy(r, F, 9 * sizeof(R));
  for (int it = 0; it < 8; it++) {
    R rit[9], nr[9];
    inv_transpose(r, rit);
    R diff = 0;
    for (int i = 0; i < 9; i++) {
      nr[i] = R(0.5) * (r[i] + rit[i]);
      diff += fabs(nr[i] - r[i]);
    }
    memcpy(r, nr, sizeof(nr));
    if (diff < R(1e-10)) break;
  }
  R rt[9];
  transpose3(r, rt);
  mat_mul(rt, F, s);
}

// ------------------------------------------------------------------------
// Material parameters: constitutive.rs:55-82 (nu clamped to [-0.999, 0.4999])
// ------------------------------------------------------------------------
template <class R>
struct Mat {
  R E, nu, yield_stress, hardening, density, friction_angle, cohesion;
  R theta_c, theta_s, xi;
  int model;
  R clamped_nu() const { return rclamp(nu, R(-0.999), R(0.4999)); }
  R lambda() const {
    R n = clamped_nu();
    return E * n / ((R(1.0) + n) * (R(1.0) - R(2.0) * n));
  }
  R mu() const {
    R n = clamped_nu();
    return E / (R(2.0) * (R(1.0) + n));
  }
  R bulk() const {
    R n = clamped_nu();
    return E / (R(3.0) * (R(1.0) - R(2.0) * n));
  }
};

template <class R>
Mat<R> load_mat(const RefMaterials* mats, uint32_t idx) {
  // material_table.get(idx).unwrap_or(&material_table[0]) and
  // model_table.get(idx).unwrap_or(&NeoHookean) (mpm_solver.rs:840-841).
  uint32_t pi = idx < mats->n ? idx : 0;
  const float* p = mats->params + 7 * (size_t)pi;
  Mat<R> m;
  m.E = p[0];
  m.nu = p[1];
  m.yield_stress = p[2];
  m.hardening = p[3];
  m.density = p[4];
  m.friction_angle = p[5];
  m.cohesion = p[6];
  m.model = idx < mats->n ? mats->models[idx] : REF_NEO_HOOKEAN;
  if (mats->ext) {
    m.theta_c = mats->ext[3 * (size_t)pi + 0];
    m.theta_s = mats->ext[3 * (size_t)pi + 1];
    m.xi = mats->ext[3 * (size_t)pi + 2];
  } else {
    m.theta_c = R(2.5e-2);
    m.theta_s = R(7.5e-3);
    m.xi = R(10);
  }
  return m;
}

// ------------------------------------------------------------------------
// Constitutive models: constitutive.rs:370-617.  Each returns the plastic flag.
// ------------------------------------------------------------------------
template <class R>
int neo_hookean_mu_lambda(const R* F, R J, R mu, R lambda, int det, R* stress) {
  // constitutive.rs:375-393
  R ft[9], b[9];
  transpose3(F, ft);
  mat_mul(F, ft, b);
  R p_term = (J > R(0)) ? lambda * r_ln(J, det) / J : R(0);
  R scale = mu / rmax(J, R(1e-10));
  for (int i = 0; i < 9; i++) {
    R dev = (i % 4 == 0) ? b[i] - R(1.0) + p_term / scale : b[i];
    stress[i] = scale * dev;
  }
  return 0;
}

template <class R>
int compute_stress(const Mat<R>& m, const R* F, R J, R Jp, int det, R* stress, R* new_F) {
  memcpy(new_F, F, 9 * sizeof(R));
  switch (m.model) {
    case REF_RIGID:  // constitutive.rs:370-373 (identity stress, as written)
      identity3(stress);
      return 0;
    case REF_NEO_HOOKEAN:
      return neo_hookean_mu_lambda(F, J, m.mu(), m.lambda(), det, stress);
    case REF_COROTATED: {  // constitutive.rs:395-415
      R mu = m.mu(), lambda = m.lambda();
      R r[9], s[9];
      polar_decomp(F, r, s);
      for (int i = 0; i < 9; i++) stress[i] = R(2.0) * mu * (F[i] - r[i]);
      for (int i = 0; i < 9; i += 4) stress[i] += lambda * (J - R(1.0));
      return 0;
    }
    case REF_FIXED_COROTATED: {  // constitutive.rs:417-441
      R mu = m.mu(), lambda = m.lambda();
      R r[9], s[9], rt[9], diff[9], drt[9];
      polar_decomp(F, r, s);
      transpose3(r, rt);
      for (int i = 0; i < 9; i++) diff[i] = F[i] - r[i];
      mat_mul(diff, rt, drt);
      for (int i = 0; i < 9; i++) stress[i] = R(2.0) * mu * drt[i];
      for (int i = 0; i < 9; i += 4) stress[i] += lambda * (J - R(1.0)) * J;
      return 0;
    }
    case REF_DRUCKER_PRAGER:
    case REF_VON_MISES: {  // constitutive.rs:443-601
      R r[9], s[9];
      polar_decomp(F, r, s);
      R mu = m.mu(), lambda = m.lambda();
      R eps[9] = {s[0] - R(1.0), s[1], s[2], s[3], s[4] - R(1.0), s[5], s[6], s[7], s[8] - R(1.0)};
      R trace_e = eps[0] + eps[4] + eps[8];
      R tr3 = trace_e / R(3.0);
      R ed[9];
      memcpy(ed, eps, sizeof(ed));
      ed[0] -= tr3;
      ed[4] -= tr3;
      ed[8] -= tr3;
      R sd[9];
      for (int i = 0; i < 9; i++) sd[i] = R(2.0) * mu * ed[i];
      R s_norm = sqrt(R(0.5) * (sd[0] * sd[0] + sd[4] * sd[4] + sd[8] * sd[8] +
                                R(2.0) * (sd[1] * sd[1] + sd[2] * sd[2] + sd[3] * sd[3] +
                                          sd[5] * sd[5] + sd[6] * sd[6] + sd[7] * sd[7])));
      int plastic = 0;
      R scale = R(1);
      if (m.model == REF_DRUCKER_PRAGER) {
        R pressure = -lambda * trace_e;
        // f32::to_radians: self * (PI / 180.0) with the constant in f32
        const R rads_per_deg = R(3.14159265358979323846) / R(180.0);
        R phi = m.friction_angle * rads_per_deg;
        R sphi = sin(phi), cphi = cos(phi);
        R sqrt3 = sqrt(R(3.0));
        R alpha = (R(2.0) * sphi) / (sqrt3 * (R(3.0) - sphi));
        R k = (R(6.0) * m.cohesion * cphi) / (sqrt3 * (R(3.0) - sphi));
        R yield_val = s_norm + alpha * pressure - k;
        if (yield_val > R(0)) {
          scale = rmax(k - alpha * pressure, R(0)) / rmax(s_norm, R(1e-10));
          plastic = 1;
        }
        if (plastic) {
          R ns[9];
          for (int i = 0; i < 9; i++) {
            R id = (i % 4 == 0) ? R(1.0) : R(0.0);
            ns[i] = id + scale * (s[i] - id);
          }
          for (int i = 0; i < 9; i += 4) ns[i] += (R(1.0) - scale) * tr3;
          mat_mul(r, ns, new_F);
          for (int i = 0; i < 9; i++) sd[i] *= scale;
        }
        for (int i = 0; i < 9; i++) stress[i] = sd[i];
        for (int i = 0; i < 9; i += 4) stress[i] -= pressure;
        return plastic;
      } else {
        R yield_threshold = m.yield_stress / sqrt(R(3.0));
        if (s_norm > yield_threshold) {
          scale = yield_threshold / s_norm;
          R ns[9];
          for (int i = 0; i < 9; i++) {
            R id = (i % 4 == 0) ? R(1.0) : R(0.0);
            ns[i] = id + scale * (s[i] - id);
          }
          for (int i = 0; i < 9; i += 4) ns[i] += (R(1.0) - scale) * tr3;
          mat_mul(r, ns, new_F);
          for (int i = 0; i < 9; i++) sd[i] *= scale;
          plastic = 1;
        }
        R pressure = -lambda * trace_e;
        for (int i = 0; i < 9; i++) stress[i] = sd[i];
        for (int i = 0; i < 9; i += 4) stress[i] -= pressure;
        return plastic;
      }
    }
    case REF_NEWTONIAN_FLUID: {  // constitutive.rs:603-617
      R k = m.bulk();
      R sd = k * (J - R(1.0));
      for (int i = 0; i < 9; i++) stress[i] = R(0);
      stress[0] = sd;
      stress[4] = sd;
      stress[8] = sd;
      return 0;
    }
    case REF_SNOW: {
      // No reference counterpart (ae_particle has no snow model; the
      // ae_physics snow clamp, mpm.rs:565-639, has no J_p hardening).
      // Stomakhin et al. 2013 hardening h = exp(xi (1 - J_p)), clamped to
      // [0.1, 5], applied to mu and lambda of the Neo-Hookean stress that the
      // FP32 solver uses (stored J, constitutive.rs:375-393).
      R h = r_exp(m.xi * (R(1.0) - Jp), det);
      h = rmin(rmax(h, R(0.1)), R(5.0));
      return neo_hookean_mu_lambda(F, J, m.mu() * h, m.lambda() * h, det, stress);
    }
    default:
      return neo_hookean_mu_lambda(F, J, m.mu(), m.lambda(), det, stress);
  }
}

// ------------------------------------------------------------------------
// One-sided (Hestenes) Jacobi SVD, at most 6 sweeps of (0,1),(0,2),(1,2).
// A rotation is skipped when gamma^2 <= eps^2 alpha beta (eps = 2^-24 for
// float, 2^-53 for double); a sweep with no rotation ends the loop, which
// gives the same values as any further sweeps.
// No reference counterpart: the reference's snow SVD is fixed-point Jacobi on
// F^T F (ae_physics/src/mpm.rs:794-951), which squares cond(F).  This form
// orthogonalises the columns of F directly; U and V are rotations and a
// reflection is carried by negating the smallest singular value.
// ------------------------------------------------------------------------
template <class R>
constexpr R jacobi_tol();
template <>
constexpr float jacobi_tol<float>() { return 3.5527136788005009e-15f; }  // 2^-48
template <>
constexpr double jacobi_tol<double>() { return 1.2325951644078310e-32; }  // 2^-106

template <class R>
void svd3_jacobi(const R* F, R* U, R* sig, R* V) {
  R a[9];
  memcpy(a, F, sizeof(a));
  identity3(V);
  const int P[3] = {0, 0, 1}, Q[3] = {1, 2, 2};
  for (int sweep = 0; sweep < 6; sweep++) {
    bool rotated = false;
    for (int pq = 0; pq < 3; pq++) {
      int p = P[pq], q = Q[pq];
      R alpha = a[p] * a[p] + a[3 + p] * a[3 + p] + a[6 + p] * a[6 + p];
      R beta = a[q] * a[q] + a[3 + q] * a[3 + q] + a[6 + q] * a[6 + q];
      R gamma = a[p] * a[q] + a[3 + p] * a[3 + q] + a[6 + p] * a[6 + q];
      if (gamma * gamma <= jacobi_tol<R>() * (alpha * beta)) continue;
      rotated = true;
      R zeta = (beta - alpha) / (R(2.0) * gamma);
      R t = R(1.0) / (fabs(zeta) + sqrt(R(1.0) + zeta * zeta));
      if (zeta < R(0)) t = -t;
      R c = R(1.0) / sqrt(R(1.0) + t * t);
      R s = c * t;
      for (int r = 0; r < 3; r++) {
        R ap = a[3 * r + p], aq = a[3 * r + q];
        a[3 * r + p] = c * ap - s * aq;
        a[3 * r + q] = s * ap + c * aq;
        R vp = V[3 * r + p], vq = V[3 * r + q];
        V[3 * r + p] = c * vp - s * vq;
        V[3 * r + q] = s * vp + c * vq;
      }
    }
    if (!rotated) break;
  }
  for (int c = 0; c < 3; c++) {
    R n2 = a[c] * a[c] + a[3 + c] * a[3 + c] + a[6 + c] * a[6 + c];
    sig[c] = sqrt(n2);
  }
  // Columns of U; a vanishing column is rebuilt from the other two.
  int zero_col = -1;
  for (int c = 0; c < 3; c++) {
    if (sig[c] > R(1e-30)) {
      R inv = R(1.0) / sig[c];
      for (int r = 0; r < 3; r++) U[3 * r + c] = a[3 * r + c] * inv;
    } else {
      zero_col = c;
      for (int r = 0; r < 3; r++) U[3 * r + c] = R(0);
    }
  }
  if (zero_col >= 0) {
    int c1 = (zero_col + 1) % 3, c2 = (zero_col + 2) % 3;
    R x = U[3 * 1 + c1] * U[3 * 2 + c2] - U[3 * 2 + c1] * U[3 * 1 + c2];
    R y = U[3 * 2 + c1] * U[3 * 0 + c2] - U[3 * 0 + c1] * U[3 * 2 + c2];
    R z = U[3 * 0 + c1] * U[3 * 1 + c2] - U[3 * 1 + c1] * U[3 * 0 + c2];
    R n = sqrt(x * x + y * y + z * z);
    if (n > R(1e-30)) {
      U[zero_col] = x / n;
      U[3 + zero_col] = y / n;
      U[6 + zero_col] = z / n;
    } else {
      identity3(U);
    }
  }
  // Make U a rotation: a reflection moves to the smallest singular value.
  R du = det3(U);
  if (du < R(0)) {
    int k = 0;
    if (sig[1] < sig[k]) k = 1;
    if (sig[2] < sig[k]) k = 2;
    sig[k] = -sig[k];
    for (int r = 0; r < 3; r++) U[3 * r + k] = -U[3 * r + k];
  }
}

// FP32 SVD with FP64 fallback when the FP32 condition estimate exceeds 1e4.
int svd3_f32(const float* F, float* U, float* sig, float* V) {
  svd3_jacobi<float>(F, U, sig, V);
  float smax = fmaxf(fabsf(sig[0]), fmaxf(fabsf(sig[1]), fabsf(sig[2])));
  float smin = fminf(fabsf(sig[0]), fminf(fabsf(sig[1]), fabsf(sig[2])));
  if (!(smin * 1e4f >= smax)) {
    double Fd[9], Ud[9], sd[3], Vd[9];
    for (int i = 0; i < 9; i++) Fd[i] = F[i];
    svd3_jacobi<double>(Fd, Ud, sd, Vd);
    for (int i = 0; i < 9; i++) {
      U[i] = (float)Ud[i];
      V[i] = (float)Vd[i];
    }
    for (int i = 0; i < 3; i++) sig[i] = (float)sd[i];
    return 1;
  }
  return 0;
}
int svd3_r(const float* F, float* U, float* sig, float* V) { return svd3_f32(F, U, sig, V); }
int svd3_r(const double* F, double* U, double* sig, double* V) {
  svd3_jacobi<double>(F, U, sig, V);
  return 1;
}

// Snow plasticity clamp applied in G2P after F' = (I + dt grad v) F.
// Stomakhin et al. 2013: sigma clamped to [1-theta_c, 1+theta_s];
// J_p *= sigma/sigma_hat per axis; F <- U diag(sigma_hat) V^T.  F and J_p stay
// unchanged when no singular value is clamped.  Two tests on C = F^T F
// (whose eigenvalues are sigma^2) decide that without an SVD: its Gershgorin
// discs lying inside [lo^2, hi^2], or else C - lo^2 I and hi^2 I - C both
// positive definite by Sylvester's criterion (every leading minor > 0), which
// holds exactly when every eigenvalue lies strictly inside (lo^2, hi^2).
template <class R>
bool sym3_strictly_inside(R c00, R c11, R c22, R c01, R c02, R c12, R lo2, R hi2) {
  // det [[m0, x, y], [x, m1, z], [y, z, m2]] = m0 (m1 m2 - z^2) - x (x m2 - z y) + y (x z - m1 y)
  R a0 = c00 - lo2, a1 = c11 - lo2, a2 = c22 - lo2;
  R b0 = hi2 - c00, b1 = hi2 - c11, b2 = hi2 - c22;
  R da2 = a0 * a1 - c01 * c01;
  R db2 = b0 * b1 - c01 * c01;
  R da3 = a0 * (a1 * a2 - c12 * c12) - c01 * (c01 * a2 - c12 * c02) + c02 * (c01 * c12 - a1 * c02);
  // hi^2 I - C has off-diagonals -c01, -c02, -c12
  R db3 = b0 * (b1 * b2 - c12 * c12) - c01 * (c01 * b2 + c12 * c02) - c02 * (c01 * c12 + b1 * c02);
  return a0 > R(0) && da2 > R(0) && da3 > R(0) && b0 > R(0) && db2 > R(0) && db3 > R(0);
}

template <class R>
void snow_clamp(const Mat<R>& m, R* F, R* Jp) {
  R lo = R(1.0) - m.theta_c, hi = R(1.0) + m.theta_s;
  {
    R c00 = F[0] * F[0] + F[3] * F[3] + F[6] * F[6];
    R c11 = F[1] * F[1] + F[4] * F[4] + F[7] * F[7];
    R c22 = F[2] * F[2] + F[5] * F[5] + F[8] * F[8];
    R s01 = F[0] * F[1] + F[3] * F[4] + F[6] * F[7];
    R s02 = F[0] * F[2] + F[3] * F[5] + F[6] * F[8];
    R s12 = F[1] * F[2] + F[4] * F[5] + F[7] * F[8];
    R c01 = fabs(s01), c02 = fabs(s02), c12 = fabs(s12);
    R lo2 = lo * lo, hi2 = hi * hi;
    R r0 = c01 + c02, r1 = c01 + c12, r2 = c02 + c12;
    if (c00 - r0 >= lo2 && c00 + r0 <= hi2 && c11 - r1 >= lo2 && c11 + r1 <= hi2 &&
        c22 - r2 >= lo2 && c22 + r2 <= hi2)
      return;
    if (sym3_strictly_inside(c00, c11, c22, s01, s02, s12, lo2, hi2)) return;
  }
  R U[9], s[3], V[9];
  svd3_r(F, U, s, V);
  R sh[3];
  bool clamped = false;
  for (int d = 0; d < 3; d++) {
    sh[d] = rmin(rmax(s[d], lo), hi);
    clamped = clamped || sh[d] != s[d];
  }
  if (!clamped) return;
  R jp = *Jp;
  for (int d = 0; d < 3; d++) jp = jp * s[d] / sh[d];
  *Jp = jp;
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++)
      F[3 * r + c] = U[3 * r + 0] * sh[0] * V[3 * c + 0] + U[3 * r + 1] * sh[1] * V[3 * c + 1] +
                     U[3 * r + 2] * sh[2] * V[3 * c + 2];
}

// ------------------------------------------------------------------------
// Grid (mpm_solver.rs:73-176): SoA, x-fastest index i + j nx + k nx ny.
// ------------------------------------------------------------------------
template <class R>
struct Grid {
  int64_t nx, ny, nz;
  R dx, inv_dx;
  R o[3];
  std::vector<R> m, v, f;     // v holds momentum after P2G, velocity after update
  std::vector<int64_t> q;     // Q32.32 mode: 7 raw sums per node
  size_t nodes() const { return (size_t)(nx * ny * nz); }
};

// Stencil set-up shared by P2G and G2P (mpm_solver.rs:853-861): base =
// floor((x-o)/dx - 0.5) (saturating cast), fx taken with the unclamped base,
// then the base is clamped at 0.
template <class R>
struct Stencil {
  uint64_t g[3];
  R fx[3];
};
template <class R>
inline Stencil<R> make_stencil(const Grid<R>& G, const R* pos) {
  Stencil<R> s;
  for (int a = 0; a < 3; a++) {
    R t = (pos[a] - G.o[a]) * G.inv_dx - R(0.5);
    int64_t gi = sat_cast_i64(floor(t));
    s.fx[a] = (pos[a] - G.o[a]) * G.inv_dx - R(0.5) - (R)gi;
    s.g[a] = gi > 0 ? (uint64_t)gi : 0;
  }
  return s;
}

// ---- P2G (mpm_solver.rs:831-903 serial; :1108-1193 per-thread fold) ------
template <class R>
void p2g_particle(Grid<R>& G, R* gm, R* gv, R* gf, int64_t* gq, int q32,
                  const Mat<R>& mt, const R* pos, const R* vel, R mass, const R* F, R J,
                  const R* c, R Jp, int det) {
  R stress[9], nf[9];
  compute_stress(mt, F, J, Jp, det, stress, nf);  // new_F discarded (Q11)
  R vol = mass / mt.density;
  R sv[9];
  for (int k = 0; k < 9; k++) sv[k] = vol * stress[k];
  Stencil<R> s = make_stencil(G, pos);
  const int64_t nxny = G.nx * G.ny;
  for (int dk = 0; dk < 3; dk++)
    for (int dj = 0; dj < 3; dj++)
      for (int di = 0; di < 3; di++) {
        R wx = b_spline((R)di - s.fx[0]);
        R wy = b_spline((R)dj - s.fx[1]);
        R wz = b_spline((R)dk - s.fx[2]);
        R w = wx * wy * wz;
        if (w < R(1e-10)) continue;
        uint64_t gx = s.g[0] + di, gy = s.g[1] + dj, gz = s.g[2] + dk;
        if (gx >= (uint64_t)G.nx || gy >= (uint64_t)G.ny || gz >= (uint64_t)G.nz) continue;
        size_t gid = (size_t)(gx + gy * G.nx + gz * nxny);
        R dxg = ((R)gx + R(0.5)) * G.dx + G.o[0] - pos[0];
        R dyg = ((R)gy + R(0.5)) * G.dx + G.o[1] - pos[1];
        R dzg = ((R)gz + R(0.5)) * G.dx + G.o[2] - pos[2];
        R vx = vel[0] + c[0] * dxg + c[1] * dyg + c[2] * dzg;
        R vy = vel[1] + c[3] * dxg + c[4] * dyg + c[5] * dzg;
        R vz = vel[2] + c[6] * dxg + c[7] * dyg + c[8] * dzg;
        R dwx = b_spline_deriv((R)di - s.fx[0]) * G.inv_dx;
        R dwy = b_spline_deriv((R)dj - s.fx[1]) * G.inv_dx;
        R dwz = b_spline_deriv((R)dk - s.fx[2]) * G.inv_dx;
#ifdef ORACLE_MUT_TEXTBOOK_P2G
        dwx *= wy * wz;
        dwy *= wx * wz;
        dwz *= wx * wy;
#endif
        R cm = w * mass;
        R cp0 = w * mass * vx, cp1 = w * mass * vy, cp2 = w * mass * vz;
        R cf0 = sv[0] * dwx + sv[1] * dwy + sv[2] * dwz;
        R cf1 = sv[3] * dwx + sv[4] * dwy + sv[5] * dwz;
        R cf2 = sv[6] * dwx + sv[7] * dwy + sv[8] * dwz;
        if (q32) {
          int64_t* q = gq + 7 * gid;
          q[0] = fp_add(q[0], fp_from_f32((float)cm));
          q[1] = fp_add(q[1], fp_from_f32((float)cp0));
          q[2] = fp_add(q[2], fp_from_f32((float)cp1));
          q[3] = fp_add(q[3], fp_from_f32((float)cp2));
          q[4] = fp_add(q[4], fp_from_f32((float)cf0));
          q[5] = fp_add(q[5], fp_from_f32((float)cf1));
          q[6] = fp_add(q[6], fp_from_f32((float)cf2));
        } else {
          gm[gid] += cm;
          gv[3 * gid + 0] += cp0;
          gv[3 * gid + 1] += cp1;
          gv[3 * gid + 2] += cp2;
          gf[3 * gid + 0] += cf0;
          gf[3 * gid + 1] += cf1;
          gf[3 * gid + 2] += cf2;
        }
      }
}

// ---- Grid update: Grid::update_parallel (mpm_solver.rs:112-139), then the
// optional wall margin of ae_compute/src/mpm_compute.rs:343-389 --------------
template <class R>
void grid_update_range(Grid<R>& G, const RefConfig* cfg, R dt, size_t n0, size_t n1) {
  const int64_t nx = G.nx, ny = G.ny, nz = G.nz;
  const int64_t margin = cfg->bc_margin;
  const R rest = cfg->bc_restitution;
  const R g0 = cfg->gravity[0], g1 = cfg->gravity[1], g2 = cfg->gravity[2];
  for (size_t gid = n0; gid < n1; gid++) {
    R m = G.m[gid];
    if (!(m > R(1e-12))) continue;
    R* v = &G.v[3 * gid];
    const R* f = &G.f[3 * gid];
    v[0] /= m;
    v[1] /= m;
    v[2] /= m;
#ifdef ORACLE_MUT_NO_FORCE_TERM
    v[0] += g0 * dt;
    v[1] += g1 * dt;
    v[2] += g2 * dt;
#else
    v[0] += (g0 + f[0] / m) * dt;
    v[1] += (g1 + f[1] / m) * dt;
    v[2] += (g2 + f[2] / m) * dt;
#endif
    int64_t j = ((int64_t)gid % (nx * ny)) / nx;
    if (j == 0) {
      v[1] = rmax(v[1], R(0.0));
      v[0] *= R(0.9);
      v[2] *= R(0.9);
    }
    if (margin > 0) {
      int64_t i = (int64_t)gid % nx, k = (int64_t)gid / (nx * ny);
      if (i < margin && v[0] < R(0)) v[0] = -v[0] * rest;
      if (i >= nx - margin && v[0] > R(0)) v[0] = -v[0] * rest;
      if (j < margin && v[1] < R(0)) v[1] = -v[1] * rest;
      if (j >= ny - margin && v[1] > R(0)) v[1] = -v[1] * rest;
      if (k < margin && v[2] < R(0)) v[2] = -v[2] * rest;
      if (k >= nz - margin && v[2] > R(0)) v[2] = -v[2] * rest;
    }
  }
}

// ---- G2P (mpm_solver.rs:914-1053 serial; :1208-1325 parallel) -------------
template <class R>
void g2p_particle(const Grid<R>& G, const RefConfig* cfg, int window_clamp, R dt,
                  const Mat<R>& mt, R* pos, R* vel, R* F, R* J, R* c, R* Jp) {
  Stencil<R> s = make_stencil(G, pos);
  const int64_t nxny = G.nx * G.ny;
  R nv[3] = {0, 0, 0}, nc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, gv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int dk = 0; dk < 3; dk++)
    for (int dj = 0; dj < 3; dj++)
      for (int di = 0; di < 3; di++) {
        R wx = b_spline((R)di - s.fx[0]);
        R wy = b_spline((R)dj - s.fx[1]);
        R wz = b_spline((R)dk - s.fx[2]);
        R w = wx * wy * wz;
        if (w < R(1e-10)) continue;
        uint64_t gx = s.g[0] + di, gy = s.g[1] + dj, gz = s.g[2] + dk;
        if (gx >= (uint64_t)G.nx || gy >= (uint64_t)G.ny || gz >= (uint64_t)G.nz) continue;
        size_t gid = (size_t)(gx + gy * G.nx + gz * nxny);
        if (G.m[gid] < R(1e-12)) continue;
        const R* g = &G.v[3 * gid];
        nv[0] += w * g[0];
        nv[1] += w * g[1];
        nv[2] += w * g[2];
        R dxg = ((R)gx + R(0.5)) * G.dx + G.o[0] - pos[0];
        R dyg = ((R)gy + R(0.5)) * G.dx + G.o[1] - pos[1];
        R dzg = ((R)gz + R(0.5)) * G.dx + G.o[2] - pos[2];
        R scale = R(4.0) * G.inv_dx * G.inv_dx * w;
        nc[0] += scale * g[0] * dxg;
        nc[1] += scale * g[0] * dyg;
        nc[2] += scale * g[0] * dzg;
        nc[3] += scale * g[1] * dxg;
        nc[4] += scale * g[1] * dyg;
        nc[5] += scale * g[1] * dzg;
        nc[6] += scale * g[2] * dxg;
        nc[7] += scale * g[2] * dyg;
        nc[8] += scale * g[2] * dzg;
        R dwx = b_spline_deriv((R)di - s.fx[0]) * G.inv_dx;
        R dwy = b_spline_deriv((R)dj - s.fx[1]) * G.inv_dx;
        R dwz = b_spline_deriv((R)dk - s.fx[2]) * G.inv_dx;
#if defined(ORACLE_MUT_TEXTBOOK_G2P)
        for (int r = 0; r < 3; r++) {
          gv[3 * r + 0] -= g[r] * (dwx * wy * wz);
          gv[3 * r + 1] -= g[r] * (dwy * wx * wz);
          gv[3 * r + 2] -= g[r] * (dwz * wx * wy);
        }
#elif defined(ORACLE_MUT_TRANSPOSE_GV)
        for (int r = 0; r < 3; r++) {
          gv[r + 0] -= w * g[r] * dwx;
          gv[r + 3] -= w * g[r] * dwy;
          gv[r + 6] -= w * g[r] * dwz;
        }
#else
        gv[0] -= w * g[0] * dwx;
        gv[1] -= w * g[0] * dwy;
        gv[2] -= w * g[0] * dwz;
        gv[3] -= w * g[1] * dwx;
        gv[4] -= w * g[1] * dwy;
        gv[5] -= w * g[1] * dwz;
        gv[6] -= w * g[2] * dwx;
        gv[7] -= w * g[2] * dwy;
        gv[8] -= w * g[2] * dwz;
#endif
      }
  // Q7: non-finite velocity zeroes v and C (grad v is kept), then +-50 clamp.
  for (int a = 0; a < 3; a++) {
    if (!isfinite(nv[a])) {
      nv[0] = nv[1] = nv[2] = 0;
      for (int k = 0; k < 9; k++) nc[k] = 0;
      break;
    }
  }
  const R max_v = R(50.0);
  for (int a = 0; a < 3; a++) nv[a] = rclamp(nv[a], -max_v, max_v);
  // Q8: x += v dt with the undamped v; guards and clamps.
  R np[3] = {pos[0] + nv[0] * dt, pos[1] + nv[1] * dt, pos[2] + nv[2] * dt};
  for (int a = 0; a < 3; a++) {
    if (!isfinite(np[a])) {
      np[0] = R(6.0);
      np[1] = R(5.0);
      np[2] = R(6.0);
      nv[0] = nv[1] = nv[2] = 0;
      break;
    }
  }
  if (np[1] < R(0.0)) {
    np[1] = R(0.0);
    nv[1] = rmax(nv[1], R(0.0)) * R(0.5);
  }
  const R gmx = (R)(G.nx - 1) * G.dx;
  const R gmz = (R)(G.nz - 1) * G.dx;
  if (!window_clamp) {
    if (np[0] < R(0.0)) {
      np[0] = R(0.0);
      nv[0] = rmax(nv[0], R(0.0)) * R(0.5);
    } else if (np[0] > gmx) {
      np[0] = gmx;
      nv[0] = rmin(nv[0], R(0.0)) * R(0.5);
    }
    if (np[2] < R(0.0)) {
      np[2] = R(0.0);
      nv[2] = rmax(nv[2], R(0.0)) * R(0.5);
    } else if (np[2] > gmz) {
      np[2] = gmz;
      nv[2] = rmin(nv[2], R(0.0)) * R(0.5);
    }
  } else {
    if (np[0] < G.o[0]) {
      np[0] = G.o[0];
      nv[0] = rmax(nv[0], R(0.0)) * R(0.5);
    } else if (np[0] > gmx + G.o[0]) {
      np[0] = gmx + G.o[0];
      nv[0] = rmin(nv[0], R(0.0)) * R(0.5);
    }
    if (np[2] < G.o[2]) {
      np[2] = G.o[2];
      nv[2] = rmax(nv[2], R(0.0)) * R(0.5);
    } else if (np[2] > gmz + G.o[2]) {
      np[2] = gmz + G.o[2];
      nv[2] = rmin(nv[2], R(0.0)) * R(0.5);
    }
  }
  // Q9: F' = (I + dt grad v) F; validity; J = clamp(det F', 0.1, 10).
  R nf[9];
  for (int r = 0; r < 3; r++)
    for (int cc = 0; cc < 3; cc++) {
      R sum = R(0.0);
      for (int k = 0; k < 3; k++) {
        R kron = (r == k) ? R(1.0) : R(0.0);
        R term = kron + gv[3 * r + k] * dt;
#ifdef ORACLE_MUT_F_RIGHT
        sum += F[3 * r + k] * ((k == cc ? R(1.0) : R(0.0)) + gv[3 * k + cc] * dt);
        (void)term;
#else
        sum += term * F[3 * k + cc];
#endif
      }
      nf[3 * r + cc] = sum;
    }
  int f_valid = 1;
  for (int k = 0; k < 9; k++)
    if (!isfinite(nf[k]) || fabs(nf[k]) > R(1e3)) {
      f_valid = 0;
      break;
    }
  R fj;
  if (f_valid) {
    if (mt.model == REF_SNOW) snow_clamp(mt, nf, Jp);
    fj = rclamp(det3(nf), R(0.1), R(10.0));
  } else {
    identity3(nf);
    fj = R(1.0);
  }
  // Q10: store v * damping.
  vel[0] = nv[0] * (R)cfg->damping;
  vel[1] = nv[1] * (R)cfg->damping;
  vel[2] = nv[2] * (R)cfg->damping;
  for (int k = 0; k < 9; k++) c[k] = nc[k];
  pos[0] = np[0];
  pos[1] = np[1];
  pos[2] = np[2];
  memcpy(F, nf, sizeof(nf));
  *J = fj;
}

template <class Fn>
void parallel_for(uint32_t threads, size_t n, Fn fn) {
  if (threads <= 1 || n < 2) {
    fn(0, (size_t)0, n);
    return;
  }
  std::vector<std::thread> ts;
  size_t chunk = (n + threads - 1) / threads;
  for (uint32_t t = 0; t < threads; t++) {
    size_t a = t * chunk, b = std::min(n, a + chunk);
    if (a >= b) break;
    ts.emplace_back([=]() { fn(t, a, b); });
  }
  for (auto& th : ts) th.join();
}

template <class R>
Grid<R> make_grid(const RefConfig* cfg) {
  Grid<R> G;
  G.nx = cfg->nx;
  G.ny = cfg->ny;
  G.nz = cfg->nz;
  // Rust keeps dx and inv_dx = 1/dx as f32 (mpm_solver.rs:90-102).
  G.dx = (R)cfg->dx;
  G.inv_dx = R(1.0) / G.dx;
  for (int a = 0; a < 3; a++) G.o[a] = (R)cfg->origin[a];
  size_t n = G.nodes();
  G.m.assign(n, 0);
  G.v.assign(3 * n, 0);
  G.f.assign(3 * n, 0);
  if (cfg->accum == REF_ACCUM_Q32) G.q.assign(7 * n, 0);
  return G;
}

template <class R>
int step_impl(const RefConfig* cfg, const RefMaterials* mats, uint32_t n, R* pos, R* vel, R* F,
              R* J, R* C, const R* mass, const uint16_t* mat, R* Jp, const uint32_t* idx,
              uint32_t n_idx, R total_dt, float* out_m, float* out_p, float* out_f,
              float* out_v, int64_t* out_q) {
  if (n_idx == 0) return 0;  // mpm_solver.rs:1072-1074
  if (!cfg || !mats || mats->n == 0 || cfg->nx == 0 || cfg->ny == 0 || cfg->nz == 0 ||
      cfg->substeps == 0)
    return -1;
  for (uint32_t t = 0; t < n_idx; t++)
    if (idx[t] >= n) return -2;
  const int q32 = cfg->accum == REF_ACCUM_Q32;
  const int det = cfg->det_math <= 2 ? (int)cfg->det_math : 1;
  int window = cfg->clamp_mode == REF_CLAMP_WINDOW ||
               (cfg->clamp_mode == REF_CLAMP_AUTO && n_idx >= 50000);
  // sub_dt = total_dt / substeps as f32 (mpm_solver.rs:1075)
  const R dt = total_dt / (R)cfg->substeps;
  Grid<R> G = make_grid<R>(cfg);
  const size_t nn = G.nodes();
  std::vector<Mat<R>> mtab(mats->n);
  for (uint32_t i = 0; i < mats->n; i++) mtab[i] = load_mat<R>(mats, i);
  Mat<R> fallback = load_mat<R>(mats, 0xffffffffu);
  auto mat_of = [&](uint32_t p) -> const Mat<R>& {
    uint32_t mi = mat ? mat[p] : 0;
    return mi < mats->n ? mtab[mi] : fallback;
  };
  const uint32_t T = cfg->threads > 1 ? cfg->threads : 1;
  std::vector<R> tm, tv, tf;
  std::vector<int64_t> tq;
  for (uint32_t sub = 0; sub < cfg->substeps; sub++) {
    // grid.reset() (mpm_solver.rs:105-109)
    std::fill(G.m.begin(), G.m.end(), R(0));
    std::fill(G.v.begin(), G.v.end(), R(0));
    std::fill(G.f.begin(), G.f.end(), R(0));
    if (q32) std::fill(G.q.begin(), G.q.end(), 0);
    if (T == 1) {
      for (uint32_t t = 0; t < n_idx; t++) {
        uint32_t p = idx[t];
        R jp = Jp ? Jp[p] : R(1);
        p2g_particle(G, G.m.data(), G.v.data(), G.f.data(), q32 ? G.q.data() : nullptr, q32,
                     mat_of(p), pos + 3 * (size_t)p, vel + 3 * (size_t)p, mass[p],
                     F + 9 * (size_t)p, J[p], C + 9 * (size_t)p, jp, det);
      }
    } else {
      // fold: one private grid per worker; reduce: merge (mpm_solver.rs:1108-1197)
      if (q32) {
        tq.assign((size_t)T * 7 * nn, 0);
      } else {
        tm.assign((size_t)T * nn, 0);
        tv.assign((size_t)T * 3 * nn, 0);
        tf.assign((size_t)T * 3 * nn, 0);
      }
      parallel_for(T, n_idx, [&](uint32_t t, size_t a, size_t b) {
        R* lm = q32 ? nullptr : tm.data() + (size_t)t * nn;
        R* lv = q32 ? nullptr : tv.data() + (size_t)t * 3 * nn;
        R* lf = q32 ? nullptr : tf.data() + (size_t)t * 3 * nn;
        int64_t* lq = q32 ? tq.data() + (size_t)t * 7 * nn : nullptr;
        for (size_t k = a; k < b; k++) {
          uint32_t p = idx[k];
          R jp = Jp ? Jp[p] : R(1);
          p2g_particle(G, lm, lv, lf, lq, q32, mat_of(p), pos + 3 * (size_t)p,
                       vel + 3 * (size_t)p, mass[p], F + 9 * (size_t)p, J[p],
                       C + 9 * (size_t)p, jp, det);
        }
      });
      parallel_for(T, nn, [&](uint32_t, size_t a, size_t b) {
        for (uint32_t t = 0; t < T; t++) {
          for (size_t g = a; g < b; g++) {
            if (q32) {
              for (int k = 0; k < 7; k++)
                G.q[7 * g + k] = fp_add(G.q[7 * g + k], tq[(size_t)t * 7 * nn + 7 * g + k]);
            } else {
              G.m[g] += tm[(size_t)t * nn + g];
              for (int k = 0; k < 3; k++) {
                G.v[3 * g + k] += tv[(size_t)t * 3 * nn + 3 * g + k];
                G.f[3 * g + k] += tf[(size_t)t * 3 * nn + 3 * g + k];
              }
            }
          }
        }
      });
    }
    if (q32) {
      // Q32.32 sums become the f32 grid the update reads (FixedPoint::to_f32)
      for (size_t g = 0; g < nn; g++) {
        const int64_t* q = &G.q[7 * g];
        G.m[g] = (R)fp_to_f32(q[0]);
        for (int k = 0; k < 3; k++) {
          G.v[3 * g + k] = (R)fp_to_f32(q[1 + k]);
          G.f[3 * g + k] = (R)fp_to_f32(q[4 + k]);
        }
      }
    }
    const bool last = sub + 1 == cfg->substeps;
    if (last) {
      if (out_q && q32) memcpy(out_q, G.q.data(), sizeof(int64_t) * 7 * nn);
      for (size_t g = 0; g < nn; g++) {
        if (out_m) out_m[g] = (float)G.m[g];
        for (int k = 0; k < 3; k++) {
          if (out_p) out_p[3 * g + k] = (float)G.v[3 * g + k];
          if (out_f) out_f[3 * g + k] = (float)G.f[3 * g + k];
        }
      }
    }
    parallel_for(T, nn, [&](uint32_t, size_t a, size_t b) { grid_update_range(G, cfg, dt, a, b); });
    if (last && out_v)
      for (size_t g = 0; g < 3 * nn; g++) out_v[g] = (float)G.v[g];
    parallel_for(T, n_idx, [&](uint32_t, size_t a, size_t b) {
      for (size_t k = a; k < b; k++) {
        uint32_t p = idx[k];
        R jp = Jp ? Jp[p] : R(1);
        g2p_particle(G, cfg, window, dt, mat_of(p), pos + 3 * (size_t)p, vel + 3 * (size_t)p,
                     F + 9 * (size_t)p, J + p, C + 9 * (size_t)p, &jp);
        if (Jp) Jp[p] = jp;
      }
    });
  }
  return 0;
}

}  // namespace

extern "C" {

float ref_b_spline(float x) { return b_spline<float>(x); }
float ref_b_spline_deriv(float x) { return b_spline_deriv<float>(x); }
int64_t ref_fp_from_f32(float v) { return fp_from_f32(v); }
float ref_fp_to_f32(int64_t raw) { return fp_to_f32(raw); }
int64_t ref_fp_add(int64_t a, int64_t b) { return fp_add(a, b); }
// FixedPoint * : i128 product >> 32 (truncating division), wrapping cast
// (fixed_point.rs:68-71, 195-203).
int64_t ref_fp_mul(int64_t a, int64_t b) {
  __int128 p = (__int128)a * (__int128)b;
  return (int64_t)(p / ((__int128)1 << 32));
}
// FixedPoint / : (i128(a) << 32) / b, truncating; /0 saturates by sign
// (fixed_point.rs:73-79, 331-348).
int64_t ref_fp_div(int64_t a, int64_t b) {
  if (b == 0) return a == 0 ? 0 : (a > 0 ? INT64_MAX : INT64_MIN);
  __int128 p = (__int128)a * ((__int128)1 << 32);
  return (int64_t)(p / b);
}
// FixedPoint::sqrt: floor(sqrt(raw * 2^32)) (fixed_point.rs:81-88, 856-867).
int64_t ref_fp_sqrt(int64_t a) {
  if (a <= 0) return 0;
  unsigned __int128 r = (unsigned __int128)a << 32;
  unsigned __int128 x = (unsigned __int128)sqrtl((long double)r);
  while (x * x > r) x--;
  while ((x + 1) * (x + 1) <= r) x++;
  return (int64_t)x;
}
float ref_det_logf(float x) { return det_logf(x); }
float ref_det_expf(float x) { return det_expf(x); }

int ref_compute_stress(int model, const float* params7, const float* ext3, const float* F,
                       float J, float Jp, int det_math, float* stress, float* new_F) {
  uint8_t mdl = (uint8_t)model;
  RefMaterials mats = {1, params7, &mdl, ext3};
  Mat<float> m = load_mat<float>(&mats, 0);
  return compute_stress(m, F, J, Jp, det_math, stress, new_F);
}

int ref_svd3(const float* F, float* U, float* sig, float* V) { return svd3_f32(F, U, sig, V); }
int ref_svd3_d(const double* F, double* U, double* sig, double* V) {
  svd3_jacobi<double>(F, U, sig, V);
  return 1;
}

void ref_snow_clamp(const float* ext3, float* F, float* Jp) {
  Mat<float> m{};
  m.theta_c = ext3[0];
  m.theta_s = ext3[1];
  m.xi = ext3[2];
  m.model = REF_SNOW;
  snow_clamp(m, F, Jp);
}

int ref_step_f32(const RefConfig* cfg, const RefMaterials* mats, uint32_t n, float* pos,
                 float* vel, float* F, float* J, float* C, const float* mass,
                 const uint16_t* mat, float* Jp, const uint32_t* idx, uint32_t n_idx,
                 float total_dt, float* grid_m, float* grid_p, float* grid_f, float* grid_v,
                 int64_t* grid_q) {
  return step_impl<float>(cfg, mats, n, pos, vel, F, J, C, mass, mat, Jp, idx, n_idx, total_dt,
                          grid_m, grid_p, grid_f, grid_v, grid_q);
}

int ref_step_f64(const RefConfig* cfg, const RefMaterials* mats, uint32_t n, double* pos,
                 double* vel, double* F, double* J, double* C, const double* mass,
                 const uint16_t* mat, double* Jp, const uint32_t* idx, uint32_t n_idx,
                 double total_dt) {
  RefConfig c = *cfg;
  c.accum = REF_ACCUM_F32;
  return step_impl<double>(&c, mats, n, pos, vel, F, J, C, mass, mat, Jp, idx, n_idx, total_dt,
                           nullptr, nullptr, nullptr, nullptr, nullptr);
}

// Velocity-only tier: mpm_substep_velocity_only (mpm_solver.rs:1353-1572).
// Separable weights (w = wx * (wy * wz)), per-axis skips, NaN position keeps
// the old position, x/z lower clamp at 0 and upper clamp at o + (n-1)dx.
int ref_step_velocity_only(const RefConfig* cfg, uint32_t n, float* pos, float* vel,
                           const float* mass, const uint32_t* idx, uint32_t n_idx,
                           float total_dt) {
  if (n_idx == 0) return 0;
  if (!cfg || cfg->nx == 0 || cfg->ny == 0 || cfg->nz == 0 || cfg->substeps == 0) return -1;
  for (uint32_t t = 0; t < n_idx; t++)
    if (idx[t] >= n) return -2;
  const int q32 = cfg->accum == REF_ACCUM_Q32;
  const float dt = total_dt / (float)cfg->substeps;
  Grid<float> G = make_grid<float>(cfg);
  const size_t nn = G.nodes();
  const int64_t nx = G.nx, ny = G.ny, nz = G.nz, nxny = nx * ny;
  for (uint32_t sub = 0; sub < cfg->substeps; sub++) {
    std::fill(G.m.begin(), G.m.end(), 0.f);
    std::fill(G.v.begin(), G.v.end(), 0.f);
    std::fill(G.f.begin(), G.f.end(), 0.f);
    if (q32) std::fill(G.q.begin(), G.q.end(), 0);
    for (uint32_t t = 0; t < n_idx; t++) {
      uint32_t i = idx[t];
      const float* p = pos + 3 * (size_t)i;
      const float* v = vel + 3 * (size_t)i;
      float m = mass[i];
      Stencil<float> s = make_stencil(G, p);
      float wx[3], wy[3], wz[3];
      for (int a = 0; a < 3; a++) {
        wx[a] = b_spline((float)a - s.fx[0]);
        wy[a] = b_spline((float)a - s.fx[1]);
        wz[a] = b_spline((float)a - s.fx[2]);
      }
      for (int dk = 0; dk < 3; dk++) {
        if (wz[dk] < 1e-10f) continue;
        uint64_t gz = s.g[2] + dk;
        if (gz >= (uint64_t)nz) continue;
        for (int dj = 0; dj < 3; dj++) {
          if (wy[dj] < 1e-10f) continue;
          uint64_t gy = s.g[1] + dj;
          if (gy >= (uint64_t)ny) continue;
          float wyz = wy[dj] * wz[dk];
          for (int di = 0; di < 3; di++) {
            float w = wx[di] * wyz;
            if (w < 1e-10f) continue;
            uint64_t gx = s.g[0] + di;
            if (gx >= (uint64_t)nx) continue;
            size_t gid = (size_t)(gx + gy * nx + gz * nxny);
            float c0 = w * m, c1 = w * m * v[0], c2 = w * m * v[1], c3 = w * m * v[2];
            if (q32) {
              int64_t* q = &G.q[7 * gid];
              q[0] = fp_add(q[0], fp_from_f32(c0));
              q[1] = fp_add(q[1], fp_from_f32(c1));
              q[2] = fp_add(q[2], fp_from_f32(c2));
              q[3] = fp_add(q[3], fp_from_f32(c3));
            } else {
              G.m[gid] += c0;
              G.v[3 * gid + 0] += c1;
              G.v[3 * gid + 1] += c2;
              G.v[3 * gid + 2] += c3;
            }
          }
        }
      }
    }
    if (q32)
      for (size_t g = 0; g < nn; g++) {
        G.m[g] = fp_to_f32(G.q[7 * g]);
        for (int k = 0; k < 3; k++) G.v[3 * g + k] = fp_to_f32(G.q[7 * g + 1 + k]);
      }
    grid_update_range(G, cfg, dt, 0, nn);
    const float gmx = (float)(nx - 1) * G.dx, gmz = (float)(nz - 1) * G.dx;
    for (uint32_t t = 0; t < n_idx; t++) {
      uint32_t i = idx[t];
      float* p = pos + 3 * (size_t)i;
      float* v = vel + 3 * (size_t)i;
      Stencil<float> s = make_stencil(G, p);
      float wx[3], wy[3], wz[3];
      for (int a = 0; a < 3; a++) {
        wx[a] = b_spline((float)a - s.fx[0]);
        wy[a] = b_spline((float)a - s.fx[1]);
        wz[a] = b_spline((float)a - s.fx[2]);
      }
      float nv[3] = {0, 0, 0};
      for (int dk = 0; dk < 3; dk++) {
        if (wz[dk] < 1e-10f) continue;
        uint64_t gz = s.g[2] + dk;
        if (gz >= (uint64_t)nz) continue;
        for (int dj = 0; dj < 3; dj++) {
          if (wy[dj] < 1e-10f) continue;
          uint64_t gy = s.g[1] + dj;
          if (gy >= (uint64_t)ny) continue;
          float wyz = wy[dj] * wz[dk];
          for (int di = 0; di < 3; di++) {
            float w = wx[di] * wyz;
            if (w < 1e-10f) continue;
            uint64_t gx = s.g[0] + di;
            if (gx >= (uint64_t)nx) continue;
            size_t gid = (size_t)(gx + gy * nx + gz * nxny);
            if (G.m[gid] < 1e-12f) continue;
            nv[0] += w * G.v[3 * gid + 0];
            nv[1] += w * G.v[3 * gid + 1];
            nv[2] += w * G.v[3 * gid + 2];
          }
        }
      }
      for (int a = 0; a < 3; a++)
        if (!isfinite(nv[a])) {
          nv[0] = nv[1] = nv[2] = 0;
          break;
        }
      for (int a = 0; a < 3; a++) nv[a] = rclamp(nv[a], -50.0f, 50.0f);
      float np[3] = {p[0] + nv[0] * dt, p[1] + nv[1] * dt, p[2] + nv[2] * dt};
      for (int a = 0; a < 3; a++)
        if (!isfinite(np[a])) {
          np[0] = p[0];
          np[1] = p[1];
          np[2] = p[2];
          nv[0] = nv[1] = nv[2] = 0;
          break;
        }
      if (np[1] < 0.0f) {
        np[1] = 0.0f;
        nv[1] = fmaxf(nv[1], 0.0f) * 0.5f;
      }
      if (np[0] < 0.0f) {
        np[0] = 0.0f;
        nv[0] = fmaxf(nv[0], 0.0f) * 0.5f;
      } else if (np[0] > gmx + G.o[0]) {
        np[0] = gmx + G.o[0];
        nv[0] = fminf(nv[0], 0.0f) * 0.5f;
      }
      if (np[2] < 0.0f) {
        np[2] = 0.0f;
        nv[2] = fmaxf(nv[2], 0.0f) * 0.5f;
      } else if (np[2] > gmz + G.o[2]) {
        np[2] = gmz + G.o[2];
        nv[2] = fminf(nv[2], 0.0f) * 0.5f;
      }
      v[0] = nv[0] * cfg->damping;
      v[1] = nv[1] * cfg->damping;
      v[2] = nv[2] * cfg->damping;
      p[0] = np[0];
      p[1] = np[1];
      p[2] = np[2];
    }
  }
  return 0;
}

}  // extern "C"
