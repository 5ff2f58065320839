// q32.h -- Q32.32 fixed-point arithmetic of the all-integer MPM solver, for
// host and device code: FixedPoint, FixedVec3 and FixedMat3 of
// ae_physics/src/fixed_point.rs and the matrix helpers of
// ae_physics/src/mpm.rs (kirchhoff stress, 3x3 SVD).
//
// Values are raw int64 (value = raw / 2^32).  Rules kept bit for bit:
//   +, -   saturate (fixed_point.rs:187-193, 294-316)
//   *      (a * b) as i128 / 2^32, truncating, then `as i64` (wraps)
//          (:63-66, 195-203: the saturating fallback cannot trigger)
//   /      (a * 2^32) as i128 / b, truncating, `as i64`; b = 0 gives 0 or
//          the signed extreme (:318-348)
//   sqrt   floor(sqrt(raw * 2^32)) (:81-88, 856-867)
//   ln, exp, sin, cos, atan2 the series and table of :110-180, 254-289.
// Matrices are 9 raw values, column-major: m[c * 3 + r] is row r of
// column c (x_axis, y_axis, z_axis).
//
// Division is a 96-by-64-bit quotient: one double-precision estimate and
// exact integer correction when the quotient is below 2^50, a plain binary
// long division otherwise.  Dividing by from_i32(k) is exactly the integer
// quotient raw / k, which the callers use for the constant divisors.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define Q32_FN __host__ __device__ inline

namespace q32 {

typedef __int128 i128;
typedef unsigned __int128 u128;

constexpr int64_t ONE = (int64_t)1 << 32;
constexpr int64_t MAXV = INT64_MAX;
constexpr int64_t MINV = INT64_MIN;
constexpr int64_t HALF = (int64_t)1 << 31;      // FP_HALF (mpm.rs:139)
constexpr int64_t ONE_HALF = 6442450944LL;      // FP_ONE_POINT_FIVE (:140)
constexpr int64_t THREE_QUARTERS = 3221225472LL;  // (:141)
constexpr int64_t TWO = 8589934592LL;           // FP_TWO (:142)
constexpr int64_t THREE = 12884901888LL;        // FP_THREE (:143)
constexpr int64_t EPS = 10;                     // FP_EPS (:144)
constexpr int64_t PI = 13493037704LL;           // fixed_point.rs:110
constexpr int64_t TWO_PI = 2 * PI;              // pi * from_i32(2), exact
constexpr int64_t HALF_PI = PI / 2;             // pi / from_i32(2), exact (pi is even)
constexpr int kSinBits = 10;                    // 1024-entry quarter-wave table (:869-883)

Q32_FN int64_t from_int(int32_t v) { return (int64_t)v * ONE; }

Q32_FN int64_t add(int64_t a, int64_t b) {
  int64_t r;
  if (__builtin_add_overflow(a, b, &r)) return b > 0 ? MAXV : MINV;
  return r;
}
Q32_FN int64_t sub(int64_t a, int64_t b) {
  int64_t r;
  if (__builtin_sub_overflow(a, b, &r)) return b < 0 ? MAXV : MINV;
  return r;
}
Q32_FN int64_t neg(int64_t a) { return a == MINV ? MAXV : -a; }  // saturating_neg
Q32_FN int64_t qabs(int64_t a) { return a < 0 ? neg(a) : a; }    // saturating_abs
Q32_FN uint64_t mag(int64_t a) { return a < 0 ? 0 - (uint64_t)a : (uint64_t)a; }

// (a * b) / 2^32 truncated toward zero, low 64 bits.  With a = ah 2^32 + al
// (ah signed, al unsigned 32-bit) and b likewise, the product is
// ah bh 2^64 + (ah bl + al bh) 2^32 + al bl, so its floor over 2^32 modulo
// 2^64 needs only al bl, the two cross products and the low half of ah bh;
// a negative product with a nonzero low word (al bl mod 2^32) rounds up.
Q32_FN int64_t mul(int64_t a, int64_t b) {
  const uint32_t al = (uint32_t)a, bl = (uint32_t)b;
  const uint32_t ahu = (uint32_t)((uint64_t)a >> 32), bhu = (uint32_t)((uint64_t)b >> 32);
  const uint64_t ll = (uint64_t)al * bl;
  // ah bl and al bh as unsigned products; a negative ah (bh) contributes
  // -2^32 bl (-2^32 al), folded into the high word below
  const uint64_t mid = (ll >> 32) + (uint64_t)ahu * bl + (uint64_t)al * bhu;
  const uint32_t hi = ahu * bhu - (a < 0 ? bl : 0u) - (b < 0 ? al : 0u);
  const uint64_t q = mid + ((uint64_t)hi << 32);
  return (int64_t)(q + (uint64_t)(((a ^ b) < 0) & ((uint32_t)ll != 0)));
}
// mul(x, w) for 0 <= w < 2^32 (a B-spline weight or a product of two: the
// weights lie in [0, 0.75]): the same value as mul with fewer partial
// products, since w has no high word
Q32_FN int64_t mul_w(int64_t x, uint32_t w) {
  const uint32_t xl = (uint32_t)x, xhu = (uint32_t)((uint64_t)x >> 32);
  const uint64_t ll = (uint64_t)xl * w;
  const uint64_t mid = (ll >> 32) + (uint64_t)xhu * w;
  const uint64_t q = mid - (x < 0 ? (uint64_t)w << 32 : 0ull);
  return (int64_t)(q + (uint64_t)((x < 0) & ((uint32_t)ll != 0)));
}
// a * b as i128 (never overflows)
Q32_FN i128 wide_mul(int64_t a, int64_t b) { return (i128)a * (i128)b; }

Q32_FN double wide_to_double(u128 n) {
  return (double)(uint64_t)(n >> 64) * 18446744073709551616.0 + (double)(uint64_t)n;
}

// floor(n / d) for n < 2^96, d >= 1; only the low 64 bits are used
Q32_FN uint64_t udiv96(u128 n, uint64_t d) {
  if ((n >> 50) < (u128)d) {  // quotient < 2^50: estimate within one unit
    uint64_t q = (uint64_t)(wide_to_double(n) / (double)d);
    u128 prod = (u128)q * d;
    while (prod > n) {
      q -= 1;
      prod -= d;
    }
    while (n - prod >= (u128)d) {
      q += 1;
      prod += d;
    }
    return q;
  }
  u128 r = 0;
  uint64_t q = 0;  // bits above 63 of the quotient are dropped by the cast
  for (int bit = 95; bit >= 0; --bit) {
    r = (r << 1) | (uint64_t)((n >> bit) & 1);
    q <<= 1;
    if (r >= (u128)d) {
      r -= d;
      q |= 1;
    }
  }
  return q;
}

// floor(n / d) as udiv96, the estimate from a reciprocal of d computed once
// for several numerators (the integer correction makes it exact)
Q32_FN uint64_t udiv96_rcp(u128 n, uint64_t d, double rcp) {
  if ((n >> 50) < (u128)d) {
    uint64_t q = (uint64_t)(wide_to_double(n) * rcp);
    u128 prod = (u128)q * d;
    while (prod > n) {
      q -= 1;
      prod -= d;
    }
    while (n - prod >= (u128)d) {
      q += 1;
      prod += d;
    }
    return q;
  }
  return udiv96(n, d);
}

// (a * 2^32) / b truncated toward zero, low 64 bits
Q32_FN int64_t div(int64_t a, int64_t b) {
  if (b == 0) return a == 0 ? 0 : (a > 0 ? MAXV : MINV);
  const uint64_t q = udiv96((u128)mag(a) << 32, mag(b));
  return (int64_t)(((a < 0) != (b < 0)) ? 0 - q : q);
}
// a / from_int(k) for k > 0: the rational quotient is a / k exactly
Q32_FN int64_t div_int(int64_t a, int64_t k) { return a / k; }

Q32_FN int64_t sqrt(int64_t a) {
  if (a <= 0) return 0;
  const u128 n = (u128)(uint64_t)a << 32;
  uint64_t r = (uint64_t)__builtin_sqrt(wide_to_double(n));
  while ((u128)r * r > n) --r;
  while ((u128)(r + 1) * (r + 1) <= n) ++r;
  return (int64_t)r;
}

Q32_FN int64_t floor(int64_t a) { return a & ~(int64_t)0xffffffff; }
Q32_FN int64_t clamp(int64_t a, int64_t lo, int64_t hi) { return a < lo ? lo : (a > hi ? hi : a); }

// ln: 2 artanh(t), t = (x - 1) / (x + 1), through the t^9 term (:254-271)
Q32_FN int64_t ln(int64_t a) {
  if (a <= 0) return MINV;
  const int64_t t = div(sub(a, ONE), add(a, ONE));
  const int64_t t2 = mul(t, t);
  const int64_t t3 = mul(t2, t), t5 = mul(t3, t2), t7 = mul(t5, t2), t9 = mul(t7, t2);
  int64_t s = add(t, div_int(t3, 3));
  s = add(s, div_int(t5, 5));
  s = add(s, div_int(t7, 7));
  s = add(s, div_int(t9, 9));
  return mul(TWO, s);
}

// exp: Taylor terms 2..11 until a term rounds to zero (:273-289)
Q32_FN int64_t exp(int64_t a) {
  if (a == 0) return ONE;
  int64_t sum = add(ONE, a), term = a;
  for (int k = 2; k < 12; ++k) {
    term = div_int(mul(term, a), k);
    if (qabs(term) < 1) break;
    sum = add(sum, term);
  }
  return sum;
}

// sin from the quarter-wave table with linear interpolation (:110-131)
Q32_FN int64_t sin(int64_t a, const int64_t* tab) {
  int64_t ang = a % TWO_PI;
  if (ang < 0) ang += TWO_PI;  // rem_euclid
  bool negate = false;
  if (ang > PI) {
    ang = TWO_PI - ang;
    negate = true;
  }
  if (ang > HALF_PI) ang = PI - ang;
  constexpr int sh = 32 - kSinBits;
  int64_t idx = ang >> sh;
  if (idx > (1 << kSinBits) - 2) idx = (1 << kSinBits) - 2;
  const int64_t frac = ang & (((int64_t)1 << sh) - 1);
  const int64_t lo = tab[idx], hi = tab[idx + 1];
  const int64_t step = (int64_t)((uint64_t)(hi - lo) * (uint64_t)frac);  // wrapping_mul
  const int64_t v = lo + step / ((int64_t)1 << sh);
  return negate ? -v : v;
}
Q32_FN int64_t cos(int64_t a, const int64_t* tab) { return sin(add(a, HALF_PI), tab); }

Q32_FN int64_t atan_series(int64_t x) {  // :171-180
  const int64_t x2 = mul(x, x);
  const int64_t x3 = mul(x2, x), x5 = mul(x3, x2), x7 = mul(x5, x2), x9 = mul(x7, x2);
  int64_t s = sub(x, div_int(x3, 3));
  s = add(s, div_int(x5, 5));
  s = sub(s, div_int(x7, 7));
  return add(s, div_int(x9, 9));
}
Q32_FN int64_t atan2(int64_t y, int64_t x) {  // :148-169
  if (x == 0 && y == 0) return 0;
  const int64_t ay = qabs(y), ax = qabs(x);
  const int64_t ang = ax >= ay ? atan_series(div(ay, ax)) : sub(HALF_PI, atan_series(div(ax, ay)));
  if (x < 0) return y >= 0 ? sub(PI, ang) : neg(sub(PI, ang));
  return y >= 0 ? ang : neg(ang);
}

// --- FixedVec3 / FixedMat3 -------------------------------------------------
Q32_FN int64_t dot3(int64_t ax, int64_t ay, int64_t az, int64_t bx, int64_t by, int64_t bz) {
  return add(add(mul(ax, bx), mul(ay, by)), mul(az, bz));
}

struct Mat {
  int64_t m[9];  // column-major
  Q32_FN int64_t at(int r, int c) const { return m[c * 3 + r]; }
  Q32_FN int64_t& at(int r, int c) { return m[c * 3 + r]; }
};

Q32_FN Mat mat_zero() {
  Mat o;
  for (int k = 0; k < 9; ++k) o.m[k] = 0;
  return o;
}
Q32_FN Mat mat_diag(int64_t s) {
  Mat o = mat_zero();
  o.m[0] = o.m[4] = o.m[8] = s;
  return o;
}
Q32_FN Mat transpose(const Mat& a) {
  Mat o;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) o.at(r, c) = a.at(c, r);
  return o;
}
Q32_FN Mat mat_add(const Mat& a, const Mat& b) {
  Mat o;
  for (int k = 0; k < 9; ++k) o.m[k] = add(a.m[k], b.m[k]);
  return o;
}
Q32_FN Mat mat_sub(const Mat& a, const Mat& b) {
  Mat o;
  for (int k = 0; k < 9; ++k) o.m[k] = sub(a.m[k], b.m[k]);
  return o;
}
Q32_FN Mat mat_scale(const Mat& a, int64_t s) {
  Mat o;
  for (int k = 0; k < 9; ++k) o.m[k] = mul(a.m[k], s);
  return o;
}
Q32_FN Mat mat_div(const Mat& a, int64_t s) {
  Mat o;
  if (s == 0) {
    for (int k = 0; k < 9; ++k) o.m[k] = div(a.m[k], s);
    return o;
  }
  const uint64_t d = mag(s);
  const double rcp = 1.0 / (double)d;
  for (int k = 0; k < 9; ++k) {
    const uint64_t q = udiv96_rcp((u128)mag(a.m[k]) << 32, d, rcp);
    o.m[k] = (int64_t)(((a.m[k] < 0) != (s < 0)) ? 0 - q : q);
  }
  return o;
}
// (a b)(r, c) = row r of a . column c of b (fixed_point.rs:810-830)
Q32_FN Mat matmul(const Mat& a, const Mat& b) {
  Mat o;
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r)
      o.at(r, c) = dot3(a.at(r, 0), a.at(r, 1), a.at(r, 2), b.at(0, c), b.at(1, c), b.at(2, c));
  return o;
}
// a v, each row summed x, y, z (:832-841)
Q32_FN void matvec(const Mat& a, const int64_t v[3], int64_t o[3]) {
  for (int r = 0; r < 3; ++r) o[r] = add(add(mul(a.at(r, 0), v[0]), mul(a.at(r, 1), v[1])), mul(a.at(r, 2), v[2]));
}

// checked i128 pieces of the determinant; an overflowing piece is 0
// (:711-729).  Products of two i64 always fit, so only the i64 x i128
// products and the sums are checked.
Q32_FN i128 ck_sub(i128 a, i128 b) {
  i128 r;
  return __builtin_sub_overflow(a, b, &r) ? 0 : r;
}
Q32_FN i128 ck_add(i128 a, i128 b) {
  i128 r;
  return __builtin_add_overflow(a, b, &r) ? 0 : r;
}
// a * x, or 0 when the product leaves the i128 range
Q32_FN i128 ck_mul_wide(int64_t a, i128 x) {
  const bool neg_r = (a < 0) != (x < 0);
  const uint64_t ua = mag(a);
  const u128 ux = x < 0 ? (u128)0 - (u128)x : (u128)x;
  const u128 p0 = (u128)ua * (uint64_t)ux;
  const u128 p1 = (u128)ua * (uint64_t)(ux >> 64);
  if ((uint64_t)(p1 >> 64) != 0) return 0;
  const u128 m = p0 + (p1 << 64);
  if (m < p0) return 0;
  const u128 lim = ((u128)1 << 127) - (neg_r ? 0 : 1);
  if (m > lim) return 0;
  return neg_r ? (i128)((u128)0 - m) : (i128)m;
}
Q32_FN int64_t det(const Mat& M) {
  const int64_t a = M.m[0], b = M.m[1], c = M.m[2], d = M.m[3], e = M.m[4], f = M.m[5], g = M.m[6],
                h = M.m[7], i = M.m[8];
  i128 r = ck_mul_wide(a, ck_sub(wide_mul(e, i), wide_mul(f, h)));
  r = ck_sub(r, ck_mul_wide(b, ck_sub(wide_mul(d, i), wide_mul(f, g))));
  r = ck_add(r, ck_mul_wide(c, ck_sub(wide_mul(d, h), wide_mul(e, g))));
  // r / 2^32 / 2^32 truncating is r / 2^64 truncating
  const u128 m = r < 0 ? (u128)0 - (u128)r : (u128)r;
  const uint64_t q = (uint64_t)(m >> 64);
  return (int64_t)(r < 0 ? 0 - q : q);
}
// i128 / 2^32 truncating, as i64
Q32_FN int64_t narrow(i128 x) {
  const u128 m = x < 0 ? (u128)0 - (u128)x : (u128)x;
  const uint64_t q = (uint64_t)(m >> 32);
  return (int64_t)(x < 0 ? 0 - q : q);
}
// adjugate / det, each cofactor narrowed then scaled by 1 / det (:730-762);
// dt = det(M)
Q32_FN Mat inverse_with_det(const Mat& M, int64_t dt) {
  if (dt == 0) return mat_zero();
  const int64_t inv = div(ONE, dt);
  const int64_t a = M.m[0], b = M.m[1], c = M.m[2], d = M.m[3], e = M.m[4], f = M.m[5], g = M.m[6],
                h = M.m[7], i = M.m[8];
  // plain i128 arithmetic in the reference: the products fit, the
  // difference wraps here
  auto w = [](int64_t x, int64_t y, int64_t z, int64_t t) -> i128 {
    return (i128)((u128)wide_mul(x, y) - (u128)wide_mul(z, t));
  };
  Mat o;
  o.m[0] = mul(narrow(w(e, i, f, h)), inv);
  o.m[1] = mul(narrow(w(c, h, b, i)), inv);
  o.m[2] = mul(narrow(w(b, f, c, e)), inv);
  o.m[3] = mul(narrow(w(f, g, d, i)), inv);
  o.m[4] = mul(narrow(w(a, i, c, g)), inv);
  o.m[5] = mul(narrow(w(c, d, a, f)), inv);
  o.m[6] = mul(narrow(w(d, h, e, g)), inv);
  o.m[7] = mul(narrow(w(b, g, a, h)), inv);
  o.m[8] = mul(narrow(w(a, e, b, d)), inv);
  return o;
}

Q32_FN Mat inverse(const Mat& M) { return inverse_with_det(M, det(M)); }

// quadratic B-spline weight (mpm.rs:146-156)
Q32_FN int64_t bspline(int64_t f) {
  const int64_t a = qabs(f);
  const int64_t x = sub(ONE_HALF, a);
  if (a < HALF) return sub(THREE_QUARTERS, mul(f, f));
  if (a < ONE_HALF) return mul(mul(HALF, x), x);
  return 0;
}

// Kirchhoff stress of the neo-Hookean model (mpm.rs:442-452, 608-615):
// (mu (F - F^-T) + lambda ln(Jc) F^-T) F^T / Jc
Q32_FN Mat kirchhoff(const Mat& F, int64_t j, int64_t mu, int64_t lambda, int64_t j_lo,
                     int64_t j_hi) {
  const Mat fit = transpose(inverse_with_det(F, j));  // j = det(F)
  const int64_t jc = clamp(j, j_lo, j_hi);
  const Mat pk = mat_add(mat_scale(mat_sub(F, fit), mu), mat_scale(fit, mul(lambda, ln(jc))));
  return mat_div(matmul(pk, transpose(F)), jc);
}

// trace, deviatoric part and its norm sqrt(sum_diag d^2 + 2 sum_offdiag d^2)
// (mpm.rs:483-494 and the same expression in :518-530, 619-628)
Q32_FN int64_t trace(const Mat& s) { return add(add(s.m[0], s.m[4]), s.m[8]); }
Q32_FN int64_t dev_norm(const Mat& d) {
  const int64_t diag = add(add(mul(d.m[0], d.m[0]), mul(d.m[4], d.m[4])), mul(d.m[8], d.m[8]));
  const int64_t off = add(add(mul(d.m[1], d.m[1]), mul(d.m[2], d.m[2])), mul(d.m[5], d.m[5]));
  return sqrt(add(diag, mul(TWO, off)));
}

// Jacobi sweeps on A^T A (mpm.rs:890-951).  The reference returns the
// rotated matrix, not the accumulated rotations, as the eigenvector matrix;
// that is kept, and the unused rotation product is not formed.
Q32_FN void sym_eigen(int64_t a[3][3], int64_t d[3], const int64_t* tab) {
  d[0] = a[0][0];
  d[1] = a[1][1];
  d[2] = a[2][2];
  for (int it = 0; it < 50; ++it) {
    if (add(add(qabs(a[0][1]), qabs(a[0][2])), qabs(a[1][2])) < EPS) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (qabs(a[p][q]) < EPS) continue;
        const int64_t den = sub(d[p], d[q]);
        if (den == 0) continue;
        const int64_t th = mul(HALF, atan2(div(mul(TWO, a[p][q]), den), ONE));
        const int64_t c = cos(th, tab), s = sin(th, tab);
        const int64_t cc = mul(c, c), ss = mul(s, s), tsc = mul(mul(TWO, s), c);
        const int64_t app = a[p][p], aqq = a[q][q], apq = a[p][q];
        const int r = 3 - p - q;  // the third index
        const int64_t apr = a[p][r], aqr = a[q][r];
        a[p][p] = add(sub(mul(cc, app), mul(tsc, apq)), mul(ss, aqq));
        a[q][q] = add(add(mul(ss, app), mul(tsc, apq)), mul(cc, aqq));
        a[p][q] = a[q][p] = 0;
        a[p][r] = a[r][p] = sub(mul(c, apr), mul(s, aqr));
        a[q][r] = a[r][q] = add(mul(s, apr), mul(c, aqr));
        d[p] = a[p][p];
        d[q] = a[q][q];
      }
  }
}

// fixed_mat3_svd (mpm.rs:794-888): U, the singular values and V
Q32_FN void svd(const Mat& M, Mat& U, int64_t sv[3], Mat& V, const int64_t* tab) {
  int64_t g[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      g[i][j] = add(add(mul(M.at(0, i), M.at(0, j)), mul(M.at(1, i), M.at(1, j))), mul(M.at(2, i), M.at(2, j)));
  int64_t d[3];
  sym_eigen(g, d, tab);
  for (int i = 0; i < 3; ++i) sv[i] = d[i] > 0 ? sqrt(d[i]) : 0;
  // v[i][j] = g[j][i]; V's column c is (v[c][0], v[c][1], v[c][2]), i.e. V(r, c) = g[r][c]
  int64_t u[3][3];
  for (int j = 0; j < 3; ++j) {
    const int64_t si = sv[j] > EPS ? div(ONE, sv[j]) : 0;
    for (int i = 0; i < 3; ++i) {
      int64_t s = 0;
      for (int k = 0; k < 3; ++k) s = add(s, mul(M.at(i, k), g[j][k]));
      u[i][j] = mul(s, si);
    }
  }
  const int64_t du = add(sub(mul(u[0][0], sub(mul(u[1][1], u[2][2]), mul(u[1][2], u[2][1]))),
                             mul(u[0][1], sub(mul(u[1][0], u[2][2]), mul(u[1][2], u[2][0])))),
                         mul(u[0][2], sub(mul(u[1][0], u[2][1]), mul(u[1][1], u[2][0]))));
  if (du < 0) {
    for (int i = 0; i < 3; ++i) u[i][2] = neg(u[i][2]);
    sv[2] = neg(sv[2]);
  }
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      U.at(r, c) = u[r][c];
      V.at(r, c) = g[r][c];
    }
}

}  // namespace q32
