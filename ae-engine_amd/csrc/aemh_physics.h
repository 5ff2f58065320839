// aemh_physics.h -- per-particle constitutive update on the GPU.
//
// Included by k_fast.hip (AEMH_DET=0) and k_det.hip (AEMH_DET=1).  In the
// deterministic build the expressions keep the evaluation order of the
// reference (ae_particle/src/constitutive.rs) and transcendental functions
// come from the portable det_logf/det_expf, so results match a CPU build of
// the same arithmetic bit for bit.
#pragma once
#include "aemh_dev.h"

#ifndef AEMH_DET
#error "AEMH_DET must be defined to 0 or 1"
#endif

namespace aemh {
namespace {

#if AEMH_DET
__device__ __forceinline__ float m_ln(float x) { return det_logf(x); }
__device__ __forceinline__ float m_exp(float x) { return det_expf(x); }
#else
__device__ __forceinline__ float m_ln(float x) { return logf(x); }
__device__ __forceinline__ float m_exp(float x) { return expf(x); }
#endif

// The natural logarithm the Neo-Hookean stress takes of J: this build's
// m_ln, unless a caller passes another (the reference-order FP32 mode passes
// a table of the host C library's logf).
struct ModeLn {
  __device__ __forceinline__ float operator()(float x) const { return m_ln(x); }
};

// logf of the host C library (the function the reference's f32::ln calls,
// constitutive.rs:384), tabulated by the host for every float in the range
// G2P clamps J to, [0.1, 10] (mpm_solver.rs:1040-1042); other arguments (an
// uploaded J outside it) take the portable, correctly rounded det_logf
// (kLnTab* in aemh_common.h).
struct TableLn {
  const float* tab;
  __device__ __forceinline__ float operator()(float x) const {
    const uint32_t b = __float_as_uint(x) - kLnTabLo;  // positive floats order as their bits
    return b < kLnTabSize ? tab[b] : det_logf(x);
  }
};

template <typename T>
__device__ __forceinline__ void identity3(T* m) {
#pragma unroll
  for (int i = 0; i < 9; i++) m[i] = (i % 4 == 0) ? T(1) : T(0);
}
// constitutive.rs:264-276
template <typename T>
__device__ __forceinline__ void mat_mul(const T* a, const T* b, T* o) {
  T t[9];
  t[0] = a[0] * b[0] + a[1] * b[3] + a[2] * b[6];
  t[1] = a[0] * b[1] + a[1] * b[4] + a[2] * b[7];
  t[2] = a[0] * b[2] + a[1] * b[5] + a[2] * b[8];
  t[3] = a[3] * b[0] + a[4] * b[3] + a[5] * b[6];
  t[4] = a[3] * b[1] + a[4] * b[4] + a[5] * b[7];
  t[5] = a[3] * b[2] + a[4] * b[5] + a[5] * b[8];
  t[6] = a[6] * b[0] + a[7] * b[3] + a[8] * b[6];
  t[7] = a[6] * b[1] + a[7] * b[4] + a[8] * b[7];
  t[8] = a[6] * b[2] + a[7] * b[5] + a[8] * b[8];
#pragma unroll
  for (int i = 0; i < 9; i++) o[i] = t[i];
}
template <typename T>
__device__ __forceinline__ void transpose3(const T* m, T* o) {
  T t[9] = {m[0], m[3], m[6], m[1], m[4], m[7], m[2], m[5], m[8]};
#pragma unroll
  for (int i = 0; i < 9; i++) o[i] = t[i];
}

// `inv_transpose` of constitutive.rs:306-324 (returns adjugate/det as written).
__device__ __forceinline__ void inv_transpose(const float* m, float* o) {
  float det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
              m[2] * (m[3] * m[7] - m[4] * m[6]);
  if (fabsf(det) < 1e-15f) {
    identity3(o);
    return;
  }
  float inv_det = 1.0f / det;
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

// polar_decomp, constitutive.rs:284-304.
__device__ __noinline__ void polar_decomp(const float* F, float* r, float* s) {
#pragma unroll
  for (int i = 0; i < 9; i++) r[i] = F[i];
  for (int it = 0; it < 8; it++) {
    float rit[9], nr[9];
    inv_transpose(r, rit);
    float diff = 0.0f;
#pragma unroll
    for (int i = 0; i < 9; i++) {
      nr[i] = 0.5f * (r[i] + rit[i]);
      diff += fabsf(nr[i] - r[i]);
    }
#pragma unroll
    for (int i = 0; i < 9; i++) r[i] = nr[i];
    if (diff < 1e-10f) break;
  }
  float rt[9];
  transpose3(r, rt);
  mat_mul(rt, F, s);
}

// neo_hookean, constitutive.rs:375-393, with mu/lambda as arguments.
template <class Ln = ModeLn>
__device__ __forceinline__ void neo_hookean(const float* F, float J, float mu, float lambda,
                                            float* st, Ln ln = Ln()) {
  float ft[9], b[9];
  transpose3(F, ft);
  mat_mul(F, ft, b);
  float p_term = (J > 0.0f) ? lambda * ln(J) / J : 0.0f;
  float scale = mu / fmaxf(J, 1e-10f);
#pragma unroll
  for (int i = 0; i < 9; i++) {
    float dev = (i % 4 == 0) ? b[i] - 1.0f + p_term / scale : b[i];
    st[i] = scale * dev;
  }
}

// 3x3 matrix by value: the out-of-line models take and return it in
// registers, so callers need no stack copy of F or the stress.
struct Mat3 {
  float m[9];
};

// Drucker-Prager / von Mises, constitutive.rs:443-601 (stress only; the
// corrected F is discarded by P2G, mpm_solver.rs:843-844).
__device__ __noinline__ Mat3 plastic_stress(const MatDev M, const Mat3 Fm) {
  const float* F = Fm.m;
  Mat3 out;
  float* st = out.m;
  float r[9], s[9];
  polar_decomp(F, r, s);
  const float mu = M.mu, lambda = M.lambda;
  float eps[9] = {s[0] - 1.0f, s[1], s[2], s[3], s[4] - 1.0f, s[5], s[6], s[7], s[8] - 1.0f};
  float trace_e = eps[0] + eps[4] + eps[8];
  float tr3 = trace_e / 3.0f;
  eps[0] -= tr3;
  eps[4] -= tr3;
  eps[8] -= tr3;
  float sd[9];
#pragma unroll
  for (int i = 0; i < 9; i++) sd[i] = 2.0f * mu * eps[i];
  float s_norm = sqrtf(0.5f * (sd[0] * sd[0] + sd[4] * sd[4] + sd[8] * sd[8] +
                               2.0f * (sd[1] * sd[1] + sd[2] * sd[2] + sd[3] * sd[3] +
                                       sd[5] * sd[5] + sd[6] * sd[6] + sd[7] * sd[7])));
  float pressure = -lambda * trace_e;
  if (M.model == kModelDruckerPrager) {
    float yield_val = s_norm + M.dp_alpha * pressure - M.dp_k;
    if (yield_val > 0.0f) {
      float scale = fmaxf(M.dp_k - M.dp_alpha * pressure, 0.0f) / fmaxf(s_norm, 1e-10f);
#pragma unroll
      for (int i = 0; i < 9; i++) sd[i] *= scale;
    }
  } else {
    float yield_threshold = M.yield_stress / sqrtf(3.0f);
    if (s_norm > yield_threshold) {
      float scale = yield_threshold / s_norm;
#pragma unroll
      for (int i = 0; i < 9; i++) sd[i] *= scale;
    }
  }
#pragma unroll
  for (int i = 0; i < 9; i++) st[i] = sd[i];
  st[0] -= pressure;
  st[4] -= pressure;
  st[8] -= pressure;
  return out;
}

__device__ __noinline__ Mat3 corotated_stress(const MatDev M, const Mat3 Fm, float J) {
  const float* F = Fm.m;
  Mat3 out;
  float* st = out.m;
  float r[9], s[9];
  polar_decomp(F, r, s);
  if (M.model == kModelCorotated) {  // constitutive.rs:395-415
#pragma unroll
    for (int i = 0; i < 9; i++) st[i] = 2.0f * M.mu * (F[i] - r[i]);
    st[0] += M.lambda * (J - 1.0f);
    st[4] += M.lambda * (J - 1.0f);
    st[8] += M.lambda * (J - 1.0f);
  } else {  // fixed_corotated, constitutive.rs:417-441
    float rt[9], diff[9], drt[9];
    transpose3(r, rt);
#pragma unroll
    for (int i = 0; i < 9; i++) diff[i] = F[i] - r[i];
    mat_mul(diff, rt, drt);
#pragma unroll
    for (int i = 0; i < 9; i++) st[i] = 2.0f * M.mu * drt[i];
    st[0] += M.lambda * (J - 1.0f) * J;
    st[4] += M.lambda * (J - 1.0f) * J;
    st[8] += M.lambda * (J - 1.0f) * J;
  }
  return out;
}

// compute_stress dispatch, constitutive.rs:234-249, plus the snow model
// (Stomakhin 2013 hardening h = exp(xi (1 - J_p)) clamped to [0.1, 5] on
// mu and lambda of the Neo-Hookean stress; no reference counterpart).
template <class Ln = ModeLn>
__device__ __forceinline__ void compute_stress(const MatDev& M, const float* F, float J, float Jp,
                                               float* st, Ln ln = Ln()) {
  switch (M.model) {
    case kModelRigid:
      identity3(st);
      break;
    case kModelNewtonianFluid: {
      float sd = M.bulk * (J - 1.0f);
#pragma unroll
      for (int i = 0; i < 9; i++) st[i] = 0.0f;
      st[0] = sd;
      st[4] = sd;
      st[8] = sd;
      break;
    }
    case kModelCorotated:
    case kModelFixedCorotated: {
      Mat3 Fm;
#pragma unroll
      for (int i = 0; i < 9; i++) Fm.m[i] = F[i];
      const Mat3 o = corotated_stress(M, Fm, J);
#pragma unroll
      for (int i = 0; i < 9; i++) st[i] = o.m[i];
      break;
    }
    case kModelDruckerPrager:
    case kModelVonMises: {
      Mat3 Fm;
#pragma unroll
      for (int i = 0; i < 9; i++) Fm.m[i] = F[i];
      const Mat3 o = plastic_stress(M, Fm);
#pragma unroll
      for (int i = 0; i < 9; i++) st[i] = o.m[i];
      break;
    }
    case kModelSnow: {
      float h = m_exp(M.xi * (1.0f - Jp));
      h = fminf(fmaxf(h, 0.1f), 5.0f);
      neo_hookean(F, J, M.mu * h, M.lambda * h, st, ln);
      break;
    }
    default:
      neo_hookean(F, J, M.mu, M.lambda, st, ln);
      break;
  }
}

// One-sided Jacobi SVD (Hestenes): at most 6 sweeps over (0,1),(0,2),(1,2);
// U and V rotations, a reflection carried by the smallest singular value.
// A rotation is skipped when gamma^2 <= eps^2 alpha beta (columns orthogonal
// to working precision, eps = 2^-24 / 2^-53); a sweep that skips all three
// leaves every value unchanged, so stopping there equals running on.  Same
// operation sequence as the oracle's svd3_jacobi.
template <typename T>
struct JacobiTol;
template <>
struct JacobiTol<float> {
  static constexpr float v = 3.5527136788005009e-15f;  // 2^-48
};
template <>
struct JacobiTol<double> {
  static constexpr double v = 1.2325951644078310e-32;  // 2^-106
};

template <typename T>
__device__ __forceinline__ void svd3_jacobi_body(const T* F, T* U, T* sig, T* V) {
  T a[9];
#pragma unroll
  for (int i = 0; i < 9; i++) a[i] = F[i];
  identity3(V);
  for (int sweep = 0; sweep < 6; sweep++) {
    bool rotated = false;
#pragma unroll
    for (int pq = 0; pq < 3; pq++) {
      const int p = pq == 2 ? 1 : 0, q = pq == 0 ? 1 : 2;
      T alpha = a[p] * a[p] + a[3 + p] * a[3 + p] + a[6 + p] * a[6 + p];
      T beta = a[q] * a[q] + a[3 + q] * a[3 + q] + a[6 + q] * a[6 + q];
      T gamma = a[p] * a[q] + a[3 + p] * a[3 + q] + a[6 + p] * a[6 + q];
      if (gamma * gamma <= JacobiTol<T>::v * (alpha * beta)) continue;
      rotated = true;
      T zeta = (beta - alpha) / (T(2.0) * gamma);
      T t = T(1.0) / (fabs(zeta) + sqrt(T(1.0) + zeta * zeta));
      if (zeta < T(0)) t = -t;
      T c = T(1.0) / sqrt(T(1.0) + t * t);
      T s = c * t;
#pragma unroll
      for (int r = 0; r < 3; r++) {
        T ap = a[3 * r + p], aq = a[3 * r + q];
        a[3 * r + p] = c * ap - s * aq;
        a[3 * r + q] = s * ap + c * aq;
        T vp = V[3 * r + p], vq = V[3 * r + q];
        V[3 * r + p] = c * vp - s * vq;
        V[3 * r + q] = s * vp + c * vq;
      }
    }
    if (!rotated) break;
  }
#pragma unroll
  for (int c = 0; c < 3; c++) {
    T n2 = a[c] * a[c] + a[3 + c] * a[3 + c] + a[6 + c] * a[6 + c];
    sig[c] = sqrt(n2);
  }
  int zero_col = -1;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    if (sig[c] > T(1e-30)) {
      T inv = T(1.0) / sig[c];
      for (int r = 0; r < 3; r++) U[3 * r + c] = a[3 * r + c] * inv;
    } else {
      zero_col = c;
      for (int r = 0; r < 3; r++) U[3 * r + c] = T(0);
    }
  }
  if (zero_col >= 0) {
    int c1 = (zero_col + 1) % 3, c2 = (zero_col + 2) % 3;
    T x = U[3 * 1 + c1] * U[3 * 2 + c2] - U[3 * 2 + c1] * U[3 * 1 + c2];
    T y = U[3 * 2 + c1] * U[3 * 0 + c2] - U[3 * 0 + c1] * U[3 * 2 + c2];
    T z = U[3 * 0 + c1] * U[3 * 1 + c2] - U[3 * 1 + c1] * U[3 * 0 + c2];
    T n = sqrt(x * x + y * y + z * z);
    if (n > T(1e-30)) {
      U[zero_col] = x / n;
      U[3 + zero_col] = y / n;
      U[6 + zero_col] = z / n;
    } else {
      identity3(U);
    }
  }
  T du = det3(U);
  if (du < T(0)) {
    int k = 0;
    if (sig[1] < sig[k]) k = 1;
    if (sig[2] < sig[k]) k = 2;
    sig[k] = -sig[k];
    for (int r = 0; r < 3; r++) U[3 * r + k] = -U[3 * r + k];
  }
}

// The FP64 solve is rare (cond > 1e4): kept out of line so its registers do
// not weigh on the caller.
__device__ __noinline__ void svd3_jacobi_f64(const double* F, double* U, double* sig, double* V) {
  svd3_jacobi_body<double>(F, U, sig, V);
}

// FP32 SVD with the FP64 fallback for cond(F) > 1e4 (counted in *fp64_count
// when given).
__device__ __forceinline__ void svd3(const float* F, float* U, float* sig, float* V,
                                     uint32_t* fp64_count = nullptr) {
  svd3_jacobi_body<float>(F, U, sig, V);
  float smax = fmaxf(fabsf(sig[0]), fmaxf(fabsf(sig[1]), fabsf(sig[2])));
  float smin = fminf(fabsf(sig[0]), fminf(fabsf(sig[1]), fabsf(sig[2])));
  if (!(smin * 1e4f >= smax)) {
    if (fp64_count) atomicAdd(fp64_count, 1u);
    double Fd[9], Ud[9], sd[3], Vd[9];
    for (int i = 0; i < 9; i++) Fd[i] = F[i];
    svd3_jacobi_f64(Fd, Ud, sd, Vd);
    for (int i = 0; i < 9; i++) {
      U[i] = (float)Ud[i];
      V[i] = (float)Vd[i];
    }
    for (int i = 0; i < 3; i++) sig[i] = (float)sd[i];
  }
}

// Snow plasticity after the F update (Stomakhin et al. 2013): singular values
// outside [1 - theta_c, 1 + theta_s] are clamped, J_p *= sigma / sigma_hat,
// F = U diag(sigma_hat) V^T.  F and J_p are left unchanged when nothing is
// clamped, which two tests on C = F^T F (eigenvalues sigma^2) decide without
// an SVD: its Gershgorin discs lying inside [lo^2, hi^2], or else C - lo^2 I
// and hi^2 I - C both positive definite by Sylvester's criterion (every
// leading minor > 0: every eigenvalue strictly inside).  The same float
// operations as the oracle's.  No reference counterpart in the FP32 solver.
__device__ __forceinline__ bool sym3_strictly_inside(float c00, float c11, float c22, float c01,
                                                     float c02, float c12, float lo2, float hi2) {
  const float a0 = c00 - lo2, a1 = c11 - lo2, a2 = c22 - lo2;
  const float b0 = hi2 - c00, b1 = hi2 - c11, b2 = hi2 - c22;
  const float da2 = a0 * a1 - c01 * c01;
  const float db2 = b0 * b1 - c01 * c01;
  const float da3 = a0 * (a1 * a2 - c12 * c12) - c01 * (c01 * a2 - c12 * c02) + c02 * (c01 * c12 - a1 * c02);
  // hi^2 I - C has off-diagonals -c01, -c02, -c12
  const float db3 = b0 * (b1 * b2 - c12 * c12) - c01 * (c01 * b2 + c12 * c02) - c02 * (c01 * c12 + b1 * c02);
  return a0 > 0.0f && da2 > 0.0f && da3 > 0.0f && b0 > 0.0f && db2 > 0.0f && db3 > 0.0f;
}

// C = F^T F (diagonal c, signed off-diagonals s) and whether every singular
// value of F lies inside [lo, hi] by the two tests above.
struct SnowGram {
  float c00, c11, c22, s01, s02, s12;
};

__device__ __forceinline__ bool snow_inside(const float* F, float lo, float hi, SnowGram& g) {
  g.c00 = F[0] * F[0] + F[3] * F[3] + F[6] * F[6];
  g.c11 = F[1] * F[1] + F[4] * F[4] + F[7] * F[7];
  g.c22 = F[2] * F[2] + F[5] * F[5] + F[8] * F[8];
  g.s01 = F[0] * F[1] + F[3] * F[4] + F[6] * F[7];
  g.s02 = F[0] * F[2] + F[3] * F[5] + F[6] * F[8];
  g.s12 = F[1] * F[2] + F[4] * F[5] + F[7] * F[8];
  const float c01 = fabsf(g.s01), c02 = fabsf(g.s02), c12 = fabsf(g.s12);
  const float lo2 = lo * lo, hi2 = hi * hi;
  const float r0 = c01 + c02, r1 = c01 + c12, r2 = c02 + c12;
  if (g.c00 - r0 >= lo2 && g.c00 + r0 <= hi2 && g.c11 - r1 >= lo2 && g.c11 + r1 <= hi2 &&
      g.c22 - r2 >= lo2 && g.c22 + r2 <= hi2)
    return true;
  return sym3_strictly_inside(g.c00, g.c11, g.c22, g.s01, g.s02, g.s12, lo2, hi2);
}

__device__ __forceinline__ void snow_clamp_svd(const MatDev& M, float* F, float* Jp, uint32_t* fp64_count);

__device__ __forceinline__ void snow_clamp(const MatDev& M, float* F, float* Jp,
                                           uint32_t* fp64_count = nullptr) {
  SnowGram g;
  if (snow_inside(F, 1.0f - M.theta_c, 1.0f + M.theta_s, g)) return;
  snow_clamp_svd(M, F, Jp, fp64_count);
}

__device__ __forceinline__ void snow_clamp_svd(const MatDev& M, float* F, float* Jp, uint32_t* fp64_count) {
  const float lo = 1.0f - M.theta_c, hi = 1.0f + M.theta_s;
  float U[9], s[3], V[9];
  svd3(F, U, s, V, fp64_count);
  float sh[3];
  bool clamped = false;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    sh[d] = fminf(fmaxf(s[d], lo), hi);
    clamped = clamped || sh[d] != s[d];
  }
  if (!clamped) return;
  float jp = *Jp;
#pragma unroll
  for (int d = 0; d < 3; d++) jp = jp * s[d] / sh[d];
  *Jp = jp;
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++)
      F[3 * r + c] = U[3 * r + 0] * sh[0] * V[3 * c + 0] + U[3 * r + 1] * sh[1] * V[3 * c + 1] +
                     U[3 * r + 2] * sh[2] * V[3 * c + 2];
}

#if !AEMH_DET
// One cyclic-Jacobi rotation of the symmetric 3x3 A in the (p, q) plane
// (r the third index), zeroing a_pq; the rotation P (P_pp = P_qq = c,
// P_pq = s, P_qp = -s) is accumulated into the eigenvector columns of V.
// t = tan of the angle, the smaller root of t^2 + 2 theta t - 1 = 0 with
// theta = (a_qq - a_pp) / (2 a_pq) (Jacobi's method as in Golub and Van
// Loan, 8.5); a_pq = 0 gives t = 0 (theta = +-inf or 0/0).
template <int P, int Q>
__device__ __forceinline__ void sym3_rotate(float& app, float& aqq, float& apq, float& arp, float& arq,
                                            float* V) {
  const float th = (aqq - app) * __builtin_amdgcn_rcpf(2.0f * apq);
  float t = __builtin_amdgcn_rcpf(fabsf(th) + __builtin_amdgcn_sqrtf(fmaf(th, th, 1.0f)));
  t = apq == 0.0f ? 0.0f : copysignf(t, th);
  const float c = __builtin_amdgcn_rsqf(fmaf(t, t, 1.0f)), s = t * c;
  app = fmaf(-t, apq, app);
  aqq = fmaf(t, apq, aqq);
  apq = 0.0f;
  const float rp = arp, rq = arq;
  arp = c * rp - s * rq;
  arq = s * rp + c * rq;
#pragma unroll
  for (int r = 0; r < 3; r++) {
    const float vp = V[3 * r + P], vq = V[3 * r + Q];
    V[3 * r + P] = c * vp - s * vq;
    V[3 * r + Q] = s * vp + c * vq;
  }
}

// FP32-mode snow clamp: the same function as snow_clamp, through the
// eigen-decomposition of C = F^T F = V diag(lambda) V^T instead of the SVD.
// With sigma_i = sqrt(lambda_i) and U = F V diag(1 / sigma),
//   U diag(sigma_hat) V^T = F (I + V diag(sigma_hat / sigma - 1) V^T),
// so U is never formed and a direction whose value is not clamped adds an
// exact zero.  Squaring F costs precision in sigma_min by cond(F)^2, so F
// with lambda_max > 4 lambda_min (cond > 2) takes the SVD route; snow's
// clamped F has cond <= (1 + theta_s) / (1 - theta_c) ~ 1.03 before a step.
// An inverted F (det F < 0, whose reflection the SVD carries on its
// smallest singular value) takes the SVD route as well.
// Cyclic Jacobi sweeps run until the off-diagonal is below 2^-48 of the
// diagonal (at most 5; 2 to 4 for F near a rotation).  Both routes stay
// inline: an out-of-line call, even one never taken, slowed every G2P wave
// (0.69 -> 0.79 ms on the 16.8M-particle three-box step).
__device__ __forceinline__ void snow_clamp_fast(const MatDev& M, float* F, float* Jp, uint32_t* fp64_count) {
  const float lo = 1.0f - M.theta_c, hi = 1.0f + M.theta_s;
  SnowGram g;
  if (snow_inside(F, lo, hi, g)) return;
  float a00 = g.c00, a11 = g.c11, a22 = g.c22, a01 = g.s01, a02 = g.s02, a12 = g.s12;
  float V[9] = {1.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f, 0.0f, 1.0f};
  for (int sweep = 0; sweep < 5; sweep++) {
    const float off = a01 * a01 + a02 * a02 + a12 * a12;
    const float dia = a00 * a00 + a11 * a11 + a22 * a22;
    if (!(off > 3.5527136788005009e-15f * dia)) break;  // 2^-48; also ends on NaN
    sym3_rotate<0, 1>(a00, a11, a01, a02, a12, V);
    sym3_rotate<0, 2>(a00, a22, a02, a01, a12, V);
    sym3_rotate<1, 2>(a11, a22, a12, a01, a02, V);
  }
  const float lmax = fmaxf(a00, fmaxf(a11, a22)), lmin = fminf(a00, fminf(a11, a22));
  if (!(lmin * 4.0f >= lmax) || !(lmax < 1e30f) || !(det3(F) > 0.0f)) {
    snow_clamp_svd(M, F, Jp, fp64_count);
    return;
  }
  const float s[3] = {__builtin_amdgcn_sqrtf(a00), __builtin_amdgcn_sqrtf(a11), __builtin_amdgcn_sqrtf(a22)};
  float sh[3], e[3];
  bool clamped = false;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    sh[d] = fminf(fmaxf(s[d], lo), hi);
    clamped = clamped || sh[d] != s[d];
    e[d] = (sh[d] - s[d]) / s[d];
  }
  if (!clamped) return;
  float jp = *Jp;
#pragma unroll
  for (int d = 0; d < 3; d++) jp = jp * s[d] / sh[d];
  *Jp = jp;
  // G = V diag(e) V^T (symmetric), F <- F + F G
  float G[9];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = a; b < 3; b++) {
      const float v = fmaf(V[3 * a + 0] * e[0], V[3 * b + 0],
                           fmaf(V[3 * a + 1] * e[1], V[3 * b + 1], V[3 * a + 2] * e[2] * V[3 * b + 2]));
      G[3 * a + b] = v;
      G[3 * b + a] = v;
    }
  float out[9];
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++)
      out[3 * r + c] = fmaf(F[3 * r + 0], G[c], fmaf(F[3 * r + 1], G[3 + c], fmaf(F[3 * r + 2], G[6 + c], F[3 * r + c])));
#pragma unroll
  for (int k = 0; k < 9; k++) F[k] = out[k];
}
#endif

}  // namespace
}  // namespace aemh
